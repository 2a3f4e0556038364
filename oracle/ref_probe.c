/* ref_probe.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Compiles the reference codec tools/3ddct/dct3d.c from where it lies
 * (DCT3D_SRC is its path, passed by oracle/Makefile) into oracle/_ref/, and
 * exports entry points that expose its static decode stages, so golden
 * vectors can be taken at each stage of the back half:
 *   - the preamble and range-coded levels of a blob (dct3d.c:542-561),
 *   - the dequantised coefficients (dct3d.c:565-566),
 *   - the inverse transform (dct3d.c:113-118),
 *   - the pre-round values and the u8 output (dct3d.c:568-593, 625-639).
 * Nothing here is part of the product; it is built only where the reference
 * sources are present.
 */
#include DCT3D_SRC

/* Decode the blob's preamble, levels and correction list the way
 * decode_float does before its arithmetic. Returns 1 on success. */
int ref_parse_u8(const uint8_t *blob, size_t len, float *vmin_o, float *vspan_o, float *qstep_o,
                 int32_t *dcq_o, float *lvl, uint32_t *ncorr_o, uint16_t *cpos, float *cdelta) {
  if (len < HDR_BYTES) return 0;
  const uint8_t *p = blob;
  if (get_u8(&p) != MAGIC) return 0;
  if ((int)get_u8(&p) != DT_U8) return 0;
  dct_init(); hfw_init(); scan_init();
  rc_dec d; dec_init(&d, p, len - HDR_BYTES);
  float vmin = dec_fval(&d), vspan = dec_fval(&d), qstep = dec_fval(&d);
  long dcq = dec_seg(&d);
  float dc = (float)dcq / (float)DC_Q;
  if (!(vspan > 0.0f) || !(qstep > 0.0f) || !isfinite(vmin) || !isfinite(dc)) return 0;
  ctx_t cflag; ctx_init(&cflag);
  int has_corr = dec_bit(&d, &cflag);
  dec_coefs(&d, lvl, g_scan);
  uint32_t nc = 0;
  if (has_corr) {
    rc_u32 ncorr = dec_eg(&d) + 1, pos = 0;
    if (ncorr > (rc_u32)N3) ncorr = N3;
    float cq = dec_fval(&d);
    if (!(cq > 0.0f) || !isfinite(cq)) cq = 1.0f;
    for (rc_u32 c = 0; c < ncorr; ++c) {
      pos += dec_eg(&d);
      int neg = dec_bypass(&d);
      rc_u32 mag = dec_eg(&d) + 1;
      if (pos >= (rc_u32)N3) break;
      float delta = (float)mag * cq;
      cpos[nc] = (uint16_t)pos;
      cdelta[nc] = neg ? -delta : delta;
      nc++;
    }
  }
  *vmin_o = vmin; *vspan_o = vspan; *qstep_o = qstep; *dcq_o = (int32_t)dcq; *ncorr_o = nc;
  return 1;
}

/* decode_float itself: the pre-round values (dct3d.c:542-593). */
int ref_decode_float_u8(const uint8_t *blob, size_t len, float *out) {
  return decode_float(blob, len, DT_U8, out);
}

void ref_step_build(float qstep, float *step) { hfw_init(); step_build(qstep, step); }
void ref_basis(float *cm) { dct_init(); memcpy(cm, g_cm, sizeof g_cm); }
float ref_deq_one(float lv, float step) { return deq_one(lv, step); }
void ref_idct3(const float *coef, float *blk) {
  dct_init();
  float a[N3], b[N3];
  dct3_inv(coef, blk, a, b);
}
void ref_fdct3(const float *blk, float *coef) {
  dct_init();
  float a[N3], b[N3];
  dct3_fwd(blk, coef, a, b);
}
float ref_quant_one(float c, float step) { return quant_one(c, step); }
