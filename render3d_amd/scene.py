"""Host-side scene setup, mirroring the reference's small C helpers exactly.

These produce the *inputs* of the raycast call (camera basis, transfer LUT,
skip gate, frame parameters); they run once per frame on the host, in FP32
through the C library's sinf/cosf/tanf so the floats equal the reference's.

* camera basis / orbit: src/core/camera.c:5-75
* transfer LUT build and presets: src/core/transfer.c:5-50
* first visible TF index and skip gate: src/main.c:1238-1242, 6797-6801
* projected-footprint LOD pick and fallback order of a bricks streamer:
  src/core/lod.h:8-26
"""
import ctypes as C
import math

import numpy as np

from ._lib import FrameParams, MODE_FULL

_libm = C.CDLL("libm.so.6")
for _fn in ("sinf", "cosf", "tanf", "sqrtf", "log2f"):
    getattr(_libm, _fn).restype = C.c_float
    getattr(_libm, _fn).argtypes = [C.c_float]

f32 = np.float32


def _sinf(x):
    return f32(_libm.sinf(f32(x)))


def _cosf(x):
    return f32(_libm.cosf(f32(x)))


def _tanf(x):
    return f32(_libm.tanf(f32(x)))


def _v3_len(a):
    return f32(_libm.sqrtf(f32(f32(f32(a[0] * a[0]) + f32(a[1] * a[1])) + f32(a[2] * a[2]))))


def _v3_norm(a):
    n = _v3_len(a)
    if n > 0:
        s = f32(f32(1.0) / n)
        return np.array([a[0] * s, a[1] * s, a[2] * s], f32)
    return np.zeros(3, f32)


def _v3_cross(a, b):
    return np.array([f32(a[1] * b[2]) - f32(a[2] * b[1]), f32(a[2] * b[0]) - f32(a[0] * b[2]),
                     f32(a[0] * b[1]) - f32(a[1] * b[0])], f32)


FOV_Y_DEFAULT = f32(1.04719755)  # 60 degrees, camera.c:9
PITCH_LIMIT = f32(1.55334303)     # camera.c:3


class Camera:
    """r3d_camera (src/core/camera.h:8-17)."""

    def __init__(self, pos=(0.0, 0.0, 0.0)):
        self.pos = np.array(pos, f32)
        self.yaw = f32(0.0)
        self.pitch = f32(0.0)
        self.fov_y = FOV_Y_DEFAULT
        self.target = np.zeros(3, f32)
        self.dist = f32(0.0)

    def forward(self):  # camera.c:12-15
        cp = _cosf(self.pitch)
        return np.array([f32(_sinf(self.yaw) * cp), _sinf(self.pitch), f32(_cosf(self.yaw) * cp)], f32)

    def look(self, dyaw, dpitch):  # camera.c:28-31
        self.yaw = f32(self.yaw + f32(dyaw))
        p = f32(self.pitch + f32(dpitch))
        self.pitch = f32(min(max(p, -PITCH_LIMIT), PITCH_LIMIT))

    def orbit_set(self, target, dist):  # camera.c:33-41
        self.target = np.array(target, f32)
        self.dist = f32(dist)
        self._orbit_sync()

    def orbit_drag(self, dyaw, dpitch):  # camera.c:43-46
        self.look(dyaw, dpitch)
        self._orbit_sync()

    def _orbit_sync(self):
        fwd = self.forward()
        self.pos = np.array([self.target[i] - f32(fwd[i] * self.dist) for i in range(3)], f32)

    def basis(self, aspect):
        """r3d_camera_basis (camera.c:65-75): right/up pre-scaled by tan(fov/2)."""
        fwd = self.forward()
        r = _v3_norm(_v3_cross(fwd, np.array([0, 1, 0], f32)))
        if _v3_len(r) == 0:
            r = np.array([1, 0, 0], f32)
        u = _v3_cross(r, fwd)
        t = _tanf(f32(self.fov_y * f32(0.5)))
        ta = f32(t * f32(aspect))
        return (np.array([r[0] * ta, r[1] * ta, r[2] * ta], f32),
                np.array([u[0] * t, u[1] * t, u[2] * t], f32), fwd)


def tf_build(points):
    """r3d_tf_build (transfer.c:5-23): points = [(value, (r,g,b,a)), ...] sorted."""
    lut = np.zeros((256, 4), np.uint8)
    n = len(points)
    for i in range(256):
        lo, hi = points[0], points[n - 1]
        for k in range(n - 1):
            if points[k][0] <= i <= points[k + 1][0]:
                lo, hi = points[k], points[k + 1]
                break
        span = hi[0] - lo[0]
        t_num = i - lo[0] if span else 0
        t_den = span if span else 1
        for c in range(4):
            a, b = lo[1][c], hi[1][c]
            lut[i, c] = (a * (t_den - t_num) + b * t_num) // t_den
    return lut


def tf_preset(idx):
    """r3d_tf_preset (transfer.c:25-50); returns None past the last preset."""
    if idx == 0:
        return [(0, (0, 0, 0, 0)), (255, (255, 255, 255, 255))]
    if idx == 1:
        return [(0, (0, 0, 0, 0)), (40, (0, 0, 0, 0)), (90, (150, 110, 70, 40)),
                (160, (230, 190, 130, 160)), (255, (255, 250, 240, 255))]
    if idx == 2:
        return [(0, (0, 0, 0, 0)), (120, (0, 0, 0, 0)), (180, (200, 200, 255, 120)),
                (255, (255, 255, 255, 255))]
    return None


TF_PRESET_COUNT = 3


def tf_ramp_40_200():
    """The configured 256-entry piecewise-linear ramp, transparent below 40."""
    return tf_build([(0, (0, 0, 0, 0)), (40, (0, 0, 0, 0)), (200, (255, 255, 255, 255)),
                     (255, (255, 255, 255, 255))])


def tf_min_visible(lut):
    """main.c:1238-1242: first voxel value with nonzero alpha."""
    nz = np.nonzero(lut[:, 3])[0]
    return f32(nz[0]) if nz.size else f32(255.0)


def skip_gate(lut, low_cut=0.0):
    """main.c:6801: max(low_cut, tf_min_visible - 0.5) / 255."""
    return f32(f32(max(f32(low_cut), f32(tf_min_visible(lut) - f32(0.5)))) / f32(255.0))


def tf_unorm(lut):
    """RGBA8 -> FP32 unorm as the library stores it (v * (1/255))."""
    return (lut.astype(f32) * f32(1.0 / 255.0)).astype(f32)


def frame_params(cam, width, height, *, mode=MODE_FULL, step_voxels=1.0, density=1.0, lod_bias=0.0,
                 max_mip=10.0, frame_index=0, threshold=0.0, gate=0.0, view_org=(0, 0), view_flags=0, brick_mode=0):
    """Assemble r3d_frame_params the way main.c:6788-6818 does (identity
    model transform; cube mode, or bricks-LOD with brick_mode = abpa | 0x20000)."""
    right, up, fwd = cam.basis(f32(f32(width) / f32(height)))
    p = FrameParams()
    p.cam_origin[:] = [float(v) for v in cam.pos]
    p.cam_right[:] = [float(v) for v in right]
    p.cam_up[:] = [float(v) for v in up]
    p.cam_forward[:] = [float(v) for v in fwd]
    p.step_voxels = step_voxels
    p.density = density
    p.lod_bias = lod_bias
    p.max_mip = max_mip
    p.viewport[0] = width
    p.viewport[1] = height
    p.mode = mode
    p.frame_index = frame_index
    p.threshold = threshold
    p.skip_gate = gate
    p.vol_r0[:] = [1.0, 0.0, 0.0]
    p.vol_r1[:] = [0.0, 1.0, 0.0]
    p.vol_r2[:] = [0.0, 0.0, 1.0]
    p.view_org = (view_org[0] & 0xFFFF) | ((view_org[1] & 0xFFFF) << 16)
    p.view_flags = view_flags
    p.brick_mode = brick_mode
    return p


def orbit_camera(yaw_deg=45.0, pitch_deg=0.0, dist=2.0, target=(0.5, 0.5, 0.5)):
    cam = Camera()
    cam.yaw = f32(math.radians(yaw_deg))
    cam.pitch = f32(math.radians(pitch_deg))
    cam.orbit_set(target, dist)
    return cam


LOD_NONE = 0xFFFFFFFF


def lod_pick(distance, pixel_cone, base_voxels, nlevels):
    """The level whose voxels match a pixel's footprint at `distance`
    (src/core/lod.h:8-16): floor(log2(max(distance * cone * base_voxels, 1)))
    in FP32, clamped to the pyramid."""
    if not nlevels:
        return 0
    footprint = max(f32(f32(f32(distance) * f32(pixel_cone)) * f32(base_voxels)), f32(1.0))
    level = int(math.floor(_libm.log2f(footprint)))
    return min(max(level, 0), nlevels - 1)


def lod_fallback(desired, ordinal, nlevels):
    """The ordinal-th level a streamer tries for a brick wanted at `desired`:
    desired, the coarser ones in order, then the finer ones going down
    (src/core/lod.h:18-24); LOD_NONE once every level was visited."""
    if not nlevels or desired >= nlevels or ordinal >= nlevels:
        return LOD_NONE
    coarse = nlevels - desired
    if ordinal < coarse:
        return desired + ordinal
    return desired - 1 - (ordinal - coarse)
