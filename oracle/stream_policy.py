"""The bricks streaming policy restated in Python. TEST INFRASTRUCTURE ONLY:
the checker of r3d_hip_bricks_stream_* (tests/ only).

It restates the reference's pump, src/vk/vkbackend.c:
  - candidate filter and order: bricks_candidate 3082-3094, bcand_cmp 3075-3080
    (ties keep their collect order here and in the library: a stable sort);
  - slot choice: bricks_pick_slot 3230-3240;
  - r3d_bricks_stream_begin 5244-5315 (with nothing in flight), the coarsest
    level's stamps (bricks_touch_coarsest 5317-5328);
  - r3d_bricks_stream_box 5330-5382, r3d_bricks_stream_point 5384-5414,
    the cone collect of r3d_bricks_stream 5416-5517 and
    r3d_bricks_stream_submit 5519-5595 (every brick is in the warm tier);
  - r3d_lod_pick, src/core/lod.h:8-17; bricks_axis_bounds 3096-3107;
  - with a compressed warm tier (known=...): a brick's maximum is unknown,
    so not gated, until a fill has measured it (brick_maxk = -1,
    vkbackend.c:3084), and an entry is published only at the noise floor or
    above (5265-5269).
Arithmetic is FP32 in the C's operation order (numpy float32 scalars), so the
decisions equal the library's; log2 is numpy's float32 log2.
"""
import math

import numpy as np

f32 = np.float32
BK = 128
INV = 0xFFFFFFFF
MAX_BATCH = 32
NOISE_FLOOR = 5


def bricks_per_axis(level_dims):
    return [tuple(-(-d // BK) for d in dims) for dims in level_dims]


def brick_maxima(levels):
    """Largest voxel of every 128^3 brick, level by level, x fastest, from
    linear levels [(dims (x, y, z), u8 array)]."""
    out = []
    for (nx, ny, nz), arr in levels:
        a = np.asarray(arr, np.uint8).reshape(nz, ny, nx)
        bx, by, bz = -(-nx // BK), -(-ny // BK), -(-nz // BK)
        for cz in range(bz):
            for cy in range(by):
                for cx in range(bx):
                    out.append(int(a[cz * BK:(cz + 1) * BK, cy * BK:(cy + 1) * BK, cx * BK:(cx + 1) * BK].max()))
    return np.array(out, np.int64)


def gate8(gate):
    g8 = int(f32(f32(gate) * f32(255.0)) + f32(0.5))
    return max(g8, NOISE_FLOOR)


def lod_pick(distance, pixel_cone, base_voxels, nlevels):
    fp = max(f32(f32(f32(distance) * f32(pixel_cone)) * f32(base_voxels)), f32(1.0))
    level = int(math.floor(float(np.log2(f32(fp)))))
    return min(max(level, 0), nlevels - 1)


def axis_bounds(eye, radius, edge, n):
    a = int(np.ceil(f32(f32(f32(eye) - f32(radius)) / f32(edge)) - f32(0.5)))
    b = int(np.floor(f32(f32(f32(eye) + f32(radius)) / f32(edge)) - f32(0.5))) + 1
    return min(max(a, 0), n), min(max(b, 0), n)


class StreamPolicy:
    def __init__(self, level_dims, nslots, bmax, resident=None, known=None, floor=0):
        """bmax: every brick's maximum; known: None (all known) or a flag per
        brick (the residents' are set); floor: the entries' publish floor."""
        self.dims = [tuple(d) for d in level_dims]
        self.nlev = len(self.dims)
        self.nb = bricks_per_axis(self.dims)
        self.off = [0]
        for b in self.nb[:-1]:
            self.off.append(self.off[-1] + b[0] * b[1] * b[2])
        self.nent = self.off[-1] + self.nb[-1][0] * self.nb[-1][1] * self.nb[-1][2]
        self.maxdim = max(self.dims[0])
        self.nslots = nslots
        self.bmax = np.asarray(bmax)
        self.brick_slot = [INV] * self.nent
        self.brick_want = [0] * self.nent
        self.slot_brick = [INV] * nslots
        self.slot_use = [0] * nslots
        self.frame = 0
        self.cands = []
        self.open = False
        self.lod_wanted = [0] * 8
        self.lod_requests = [0] * 8
        self.known = None if known is None else [bool(k) for k in known]
        self.floor = floor
        nxt = 0
        for e in range(self.nent):  # the build's residents, slots in page order
            if resident is None or resident[e]:
                self.brick_slot[e], self.slot_brick[nxt] = nxt, e
                nxt += 1

    def entry(self, l, bx, by, bz):
        return self.off[l] + (bz * self.nb[l][1] + by) * self.nb[l][0] + bx

    def locate(self, e):
        l = max(i for i in range(self.nlev) if self.off[i] <= e)
        k = e - self.off[l]
        return l, k % self.nb[l][0], (k // self.nb[l][0]) % self.nb[l][1], k // (self.nb[l][0] * self.nb[l][1])

    def begin(self):
        self.cands = []
        self.frame += 1
        self.lod_wanted = [0] * 8
        self.open = True
        return True

    def count_wanted(self, level):
        if level < 8:
            self.lod_wanted[level] += 1
            self.lod_requests[level] += 1

    def touch_coarsest(self):
        l = self.nlev - 1
        for i in range(self.nb[l][0] * self.nb[l][1] * self.nb[l][2]):
            s = self.brick_slot[self.off[l] + i]
            if s != INV:
                self.slot_use[s] = self.frame

    def candidate(self, e, d2, priority, g8):
        if (self.known is None or self.known[e]) and self.bmax[e] < g8:
            return
        s = self.brick_slot[e]
        if s != INV:
            self.slot_use[s] = self.frame
            return
        if self.brick_want[e] == self.frame:
            return
        self.brick_want[e] = self.frame
        self.cands.append((f32(d2), e, priority))

    def box(self, lo, hi, pixel_cone, gate):
        if not self.open:
            return
        g8 = gate8(gate)
        maxdim = f32(self.maxdim)
        vpp = max(f32(f32(pixel_cone) * maxdim), f32(1.0))
        desired = min(int(math.floor(float(np.log2(vpp)))), self.nlev - 1)
        self.touch_coarsest()
        while True:
            edge = f32(f32(BK << desired) / maxdim)
            rng = []
            count = 1
            for a in range(3):
                a0 = int(np.floor(f32(f32(lo[a]) / edge)))
                a1 = int(np.floor(f32(f32(hi[a]) / edge))) + 1
                a0 = max(a0, 0)
                a1 = min(a1, self.nb[desired][a])
                a1 = max(a1, a0)
                rng.append((a0, a1))
                count *= a1 - a0
            if count <= 2048 or desired + 1 >= self.nlev:
                break
            desired += 1
        c = [f32(f32(f32(lo[a]) + f32(hi[a])) * f32(0.5)) for a in range(3)]
        for bz in range(*rng[2]):
            for by in range(*rng[1]):
                for bx in range(*rng[0]):
                    d = [f32(f32(f32(q) + f32(0.5)) * edge) - c[a] for a, q in enumerate((bx, by, bz))]
                    d2 = f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))
                    self.count_wanted(desired)
                    if desired + 1 < self.nlev:
                        self.candidate(self.entry(desired + 1, bx >> 1, by >> 1, bz >> 1), d2, 0, g8)
                    self.candidate(self.entry(desired, bx, by, bz), d2, 1, g8)

    def point(self, p, level, gate):
        if not self.open:
            return
        g8 = gate8(gate)
        level = min(level, self.nlev - 1)
        edge = f32(f32(BK << level) / f32(self.maxdim))
        bc = [min(max(int(f32(f32(p[a]) / edge)), 0), self.nb[level][a] - 1) for a in range(3)]
        if level + 1 < self.nlev:
            self.candidate(self.entry(level + 1, bc[0] >> 1, bc[1] >> 1, bc[2] >> 1), 0.0, 0, g8)
        self.count_wanted(level)
        self.candidate(self.entry(level, *bc), 0.0, 1, g8)

    def cone_bricks(self, eye, fwd, half_tan, pixel_cone, slice_z0=0, slice_depth=0, pruned=True):
        """The bricks the cone collect wants, in walk order: (level, bx, by, bz,
        d2). pruned=False walks every brick of every level instead of each
        level's distance-shell box (the same set, by construction)."""
        tanw = f32(f32(f32(half_tan) * f32(1.15)) + f32(1e-3))
        maxdim = f32(self.maxdim)
        lod_factor = max(f32(f32(pixel_cone) * maxdim), f32(1e-6))
        eye = [f32(v) for v in eye]
        fwd = [f32(v) for v in fwd]
        for li in range(self.nlev - 1):
            edge = f32(f32(BK << li) / maxdim)
            brad = f32(f32(0.8660254) * edge)
            outer = f32(f32(f32(2.0 ** (li + 1)) / lod_factor) + brad)
            if pruned:
                (x0, x1), (y0, y1), (z0, z1) = (axis_bounds(eye[a], outer, edge, self.nb[li][a]) for a in range(3))
            else:
                (x0, x1), (y0, y1), (z0, z1) = ((0, self.nb[li][a]) for a in range(3))
            if slice_depth:
                zlast = min(slice_z0 + slice_depth, self.dims[0][2])
                zb = BK << li
                z0, z1 = max(z0, slice_z0 // zb), min(z1, (zlast + zb - 1) // zb)
            for bz in range(z0, z1):
                for by in range(y0, y1):
                    for bx in range(x0, x1):
                        cx = f32(f32(f32(bx) + f32(0.5)) * edge) - eye[0]
                        cy = f32(f32(f32(by) + f32(0.5)) * edge) - eye[1]
                        cz = f32(f32(f32(bz) + f32(0.5)) * edge) - eye[2]
                        d2 = f32(f32(f32(cx * cx) + f32(cy * cy)) + f32(cz * cz))
                        along = f32(f32(f32(cx * fwd[0]) + f32(cy * fwd[1])) + f32(cz * fwd[2]))
                        r0 = f32(edge + brad)
                        vis = d2 < f32(r0 * r0)
                        if not vis and along > 0:
                            perp2 = f32(d2 - f32(along * along))
                            rad = f32(f32(along * tanw) + brad)
                            vis = perp2 < f32(rad * rad)
                        if vis and li == lod_pick(np.sqrt(d2), pixel_cone, maxdim, self.nlev):
                            yield li, bx, by, bz, d2

    def cone(self, eye, fwd, half_tan, pixel_cone, slice_z0, slice_depth, gate):
        g8 = gate8(gate)
        self.touch_coarsest()
        for li, bx, by, bz, d2 in self.cone_bricks(eye, fwd, half_tan, pixel_cone, slice_z0, slice_depth):
            self.count_wanted(li)
            # the parent first
            self.candidate(self.entry(li + 1, bx >> 1, by >> 1, bz >> 1), d2, 0, g8)
            self.candidate(self.entry(li, bx, by, bz), d2, 1, g8)

    def pick_slot(self):
        best, bu = INV, 0xFFFFFFFF
        for s in range(self.nslots):
            if self.slot_brick[s] == INV:
                return s
            if self.slot_use[s] != self.frame and self.slot_use[s] < bu:
                bu, best = self.slot_use[s], s
        return best

    def submit(self, budget):
        """-> (filled [(entry, slot)], evicted [entry])."""
        if not self.open:
            return [], []
        self.open = False
        cands, self.cands = self.cands, []
        if not cands or not budget:
            return [], []
        budget = min(budget, MAX_BATCH)
        cands.sort(key=lambda c: (c[2], float(c[0])))  # stable: ties keep the collect order
        filled, evicted = [], []
        for _, e, _ in cands:
            if len(filled) >= budget:
                break
            if self.brick_slot[e] != INV:
                continue
            s = self.pick_slot()
            if s == INV:
                break
            old = self.slot_brick[s]
            if old != INV:
                self.brick_slot[old] = INV
                evicted.append(old)
            self.slot_brick[s], self.slot_use[s], self.brick_slot[e] = e, self.frame, s
            filled.append((e, s))
        if self.known is not None:
            for e, _ in filled:
                self.known[e] = True
        return filled, evicted

    def stream(self, eye, fwd, half_tan, pixel_cone, gate, budget, slice_z0=0, slice_depth=0):
        self.begin()
        if budget == 0:
            self.open = False
            return [], []
        self.cone(eye, fwd, half_tan, pixel_cone, slice_z0, slice_depth, gate)
        return self.submit(budget)

    def entries(self):
        """Expected page entries (after the 64-word header)."""
        out = np.full(self.nent, INV, np.uint32)
        for e, s in enumerate(self.brick_slot):
            if s != INV and self.bmax[e] >= self.floor:
                out[e] = s | (int(self.bmax[e]) << 24)
        return out
