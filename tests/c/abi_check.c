/* Compiled as C23 by tests/test_abi.py: the header is plain C, the library
 * links, and the host-only / failure paths behave without a GPU. */
#include <stdio.h>
#include <string.h>

#include "r3d_hip.h"

int main(void) {
  if (sizeof(r3d_frame_params) != 248 || sizeof(r3d_chunk_header) != 16) return 10;
  r3d_frame_params v[2];
  memset(v, 0, sizeof v);
  v[0].viewport[0] = 64; v[0].viewport[1] = 48;
  v[1].viewport[0] = 31; v[1].viewport[1] = 17; v[1].view_flags = R3D_VIEW_HALF;
  size_t ws = r3d_hip_diff_workspace_bytes(v, 2);
  if (ws != (size_t)(64 * 48 + 16 * 9) * 528u) return 11;
  if (r3d_hip_diff_workspace_bytes(NULL, 2) != 0) return 12;
  if (r3d_hip_decode_chunks(NULL, NULL, NULL, NULL) != R3D_ERR_INVALID) return 13;
  if (r3d_hip_raycast(NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL) != R3D_ERR_INVALID) return 14;
  if (strcmp(r3d_hip_last_error(NULL), "no context") != 0) return 15;
  r3d_hip_ctx *ctx = NULL;
  int rc = r3d_hip_init(&ctx, 0);
  printf("init %d\n", rc);
  if (rc == R3D_OK) { r3d_hip_destroy(ctx); return 0; }
  return (rc == R3D_ERR_DEVICE && ctx == NULL) ? 0 : 16;
}
