"""Pins of the oracle's streaming policy (oracle/stream_policy.py), CPU only:
the reference's own lod_pick vectors (tests/test_quick.c:65-67), the cone
collect's distance-shell walk against a walk over every brick, the submit's
ordering, budget, slot and gate rules as invariants over random camera paths
(with a planted wrong slot choice that they catch), convergence to the wanted
set, and small LRU sequences worked by hand."""
import math

import numpy as np
import pytest

from oracle import stream_policy as sp

DIMS5 = [(1024, 768, 512), (512, 384, 256), (256, 192, 128), (128, 96, 64), (64, 48, 32)]


def _bmax(nent, seed, low_frac=0.2):
    rng = np.random.default_rng(seed)
    m = rng.integers(5, 256, nent)
    m[rng.random(nent) < low_frac] = rng.integers(0, 5, 1)[0]
    return m


def _coarsest_only(dims):
    nb = sp.bricks_per_axis(dims)
    n = [x * y * z for x, y, z in nb]
    res = np.zeros(sum(n), np.uint8)
    res[sum(n[:-1]):] = 1
    return res


def _camera(rng, inside=True):
    eye = rng.uniform(0.1, 0.9, 3) if inside else rng.uniform(-0.6, 1.6, 3)
    f = rng.normal(size=3)
    f /= np.linalg.norm(f)
    half_tan = rng.uniform(0.2, 1.0)
    h = int(rng.integers(48, 1081))
    return eye.astype(np.float32), f.astype(np.float32), np.float32(half_tan), np.float32(2 * half_tan / h)


def test_lod_pick_reference_vectors():
    assert sp.lod_pick(0.0, 0.001, 8192.0, 8) == 0
    assert sp.lod_pick(1.0, 0.001, 8192.0, 8) == 3  # 8.2 voxels per pixel
    assert sp.lod_pick(100.0, 0.01, 8192.0, 8) == 7  # clamped to the coarsest
    # the footprint's level boundaries are the powers of two (FP32 log2f
    # rounds a footprint within a few ulps below 2^(k+1) up to k + 1)
    for k in range(7):
        assert sp.lod_pick(2.0 ** k, 1.0, 1.0, 8) == k
        assert sp.lod_pick(2.0 ** (k + 1) * 0.999, 1.0, 1.0, 8) == k


def test_gate8():
    assert sp.gate8(0.0) == 5  # the noise floor
    assert sp.gate8(np.float32(39.5 / 255)) == 40
    assert sp.gate8(1.0) == 255


def _shell_walk_matches(seed, n=60):
    rng = np.random.default_rng(seed)
    pol = sp.StreamPolicy(DIMS5, 222, np.full(222, 200))
    total = 0
    for i in range(n):
        eye, f, ht, pc = _camera(rng, inside=i % 2 == 0)
        sz0, sd = (0, 0) if i % 3 else (int(rng.integers(0, 400)), int(rng.integers(1, 300)))
        a = list(pol.cone_bricks(eye, f, ht, pc, sz0, sd))
        b = list(pol.cone_bricks(eye, f, ht, pc, sz0, sd, pruned=False))
        if a != b:
            return False, total
        total += len(a)
    return True, total


def test_cone_shell_walk_equals_full_walk():
    ok, total = _shell_walk_matches(3)
    assert ok
    assert total > 500  # the cases want bricks at several levels


def test_cone_shell_walk_planted_bound_fails(monkeypatch):
    """The full-walk comparison catches a box bound off by one brick."""
    good = sp.axis_bounds

    def shrunk(eye, radius, edge, n):
        lo, hi = good(eye, radius, edge, n)
        return min(lo + 1, hi), hi

    monkeypatch.setattr(sp, "axis_bounds", shrunk)
    assert not _shell_walk_matches(3)[0]


def _check_frame(pol, before, cands, filled, evicted, budget, g8):
    """Invariants of one submit."""
    assert len(filled) <= min(budget, sp.MAX_BATCH)
    slot_use0, slot_brick0 = before
    free0 = [s for s in range(pol.nslots) if slot_brick0[s] == sp.INV]
    keys = {}
    for d2, e, pr in cands:
        keys.setdefault(e, (pr, float(d2)))
    fk = [keys[e] for e, _ in filled]
    assert fk == sorted(fk), "filled out of (priority, distance) order"
    if filled:  # every strictly better candidate was filled (no skips but residents)
        last = fk[-1]
        got = {e for e, _ in filled}
        for e, k in keys.items():
            if k < last:
                assert e in got or slot_brick0.count(e), e
    for e, s in filled:
        assert pol.bmax[e] >= g8
        assert pol.brick_slot[e] == s and pol.slot_brick[s] == e
    ev_slots = [s for e, s in filled if slot_brick0[s] != sp.INV]
    assert len(ev_slots) == len(evicted)
    # free slots are taken before any eviction, in slot order
    assert [s for _, s in filled[:len(free0)]] == free0[:len(filled)]
    for s in ev_slots:  # only slots not wanted this frame are evicted
        assert slot_use0[s] != pol.frame, s
    # coarsest-level residents are never evicted
    c0 = pol.off[-1]
    assert all(e < c0 for e in evicted)
    # the tables stay inverse maps
    for s, e in enumerate(pol.slot_brick):
        if e != sp.INV:
            assert pol.brick_slot[e] == s
    assert sum(x != sp.INV for x in pol.brick_slot) == sum(x != sp.INV for x in pol.slot_brick)


class _Recording(sp.StreamPolicy):
    def submit(self, budget):
        self.rec = (list(self.slot_use), list(self.slot_brick)), list(self.cands)
        return super().submit(budget)


@pytest.mark.parametrize("seed", [1, 2])
def test_submit_invariants_over_a_camera_path(seed):
    rng = np.random.default_rng(seed)
    bm = _bmax(222, seed)
    pol = _Recording(DIMS5, 12, bm, _coarsest_only(DIMS5))
    fills = evicts = 0
    for i in range(40):
        a = 2 * math.pi * i / 40
        eye = np.array([0.5 + 0.35 * math.cos(a), 0.375, 0.25 + 0.2 * math.sin(a)], np.float32)
        f = np.array([-math.sin(a), 0.1, math.cos(a)], np.float32)
        f /= np.linalg.norm(f)
        budget = [0, 1, 2, 6, 32, 40][i % 6]
        gate = [0.0, 0.3][i % 2]
        filled, evicted = pol.stream(eye, f, np.float32(0.577), np.float32(0.0025), gate, budget)
        if budget:
            before, cands = pol.rec
            _check_frame(pol, before, cands, filled, evicted, budget, sp.gate8(gate))
        else:
            assert not filled
        fills += len(filled)
        evicts += len(evicted)
    assert fills > 60 and evicts > 30


def test_submit_invariants_planted_slot_choice(monkeypatch):
    """The invariants catch a slot choice that evicts the most recently used slot."""

    def mru(self):
        best, bu = sp.INV, -1
        for s in range(self.nslots):
            if self.slot_brick[s] == sp.INV:
                return s
            if self.slot_use[s] > bu:
                bu, best = self.slot_use[s], s
        return best

    monkeypatch.setattr(sp.StreamPolicy, "pick_slot", mru)
    with pytest.raises(AssertionError):
        test_submit_invariants_over_a_camera_path(1)


def test_converges_to_the_wanted_set():
    bm = _bmax(222, 7)
    res = _coarsest_only(DIMS5)
    pol = sp.StreamPolicy(DIMS5, 222, bm, res)
    eye, f = np.array([0.3, 0.4, 0.2], np.float32), np.array([0.6, 0.0, 0.8], np.float32)
    ht, pc, gate = np.float32(0.6), np.float32(0.004), 0.1
    g8 = sp.gate8(gate)
    want = set()
    for li, bx, by, bz, _ in pol.cone_bricks(eye, f, ht, pc, pruned=False):
        for e in (pol.entry(li, bx, by, bz), pol.entry(li + 1, bx >> 1, by >> 1, bz >> 1)):
            if bm[e] >= g8:
                want.add(e)
    missing = [e for e in want if pol.brick_slot[e] == sp.INV]
    assert len(missing) > 20
    frames = 0
    while True:
        filled, evicted = pol.stream(eye, f, ht, pc, gate, 6)
        assert not evicted
        if not filled:
            break
        frames += 1
    assert frames == math.ceil(len(missing) / 6)
    resident = {e for e in range(pol.nent) if pol.brick_slot[e] != sp.INV}
    assert resident == want | {e for e in range(pol.nent) if res[e]}


def test_lru_by_hand_points():
    # 256x128x128, two levels: level 0 holds entries 0, 1, level 1 entry 2 (resident in slot 0)
    dims = [(256, 128, 128), (128, 64, 64)]
    pol = sp.StreamPolicy(dims, 2, np.array([9, 9, 9]), np.array([0, 0, 1]))
    assert pol.brick_slot == [sp.INV, sp.INV, 0]
    pol.begin()
    pol.point((0.1, 0.1, 0.1), 0, 0.0)
    assert pol.submit(8) == ([(0, 1)], [])  # the free slot
    pol.begin()
    pol.point((0.9, 0.1, 0.1), 0, 0.0)
    # slot 0 holds the parent, wanted this frame; slot 1 is the least recent
    assert pol.submit(8) == ([(1, 1)], [0])
    pol.begin()
    pol.point((0.9, 0.1, 0.1), 0, 0.0)
    pol.point((0.1, 0.1, 0.1), 0, 0.0)
    assert pol.submit(8) == ([], [])  # every slot wanted this frame: no thrash
    pol.begin()
    pol.point((0.1, 0.1, 0.1), 0, 0.0)
    assert pol.submit(0) == ([], [])  # no budget
    assert pol.lod_requests[0] == 5 and pol.lod_wanted[0] == 1


def test_box_level_and_parent_first_by_hand():
    dims = [(512, 256, 256), (256, 128, 128), (128, 64, 64)]  # 16 + 2 + 1 entries
    pol = sp.StreamPolicy(dims, 4, np.full(19, 50), np.array([0] * 18 + [1]))
    pol.begin()
    # 3 base voxels per pixel -> level 1; the box covers both level-1 bricks
    pol.box((0.0, 0.0, 0.0), (1.0, 0.5, 0.5), 3.0 / 512, 0.0)
    assert pol.lod_wanted[:3] == [0, 2, 0]
    # the parent (resident) is stamped; both level-1 bricks lie 0.25 from the
    # box centre and fill the free slots 1, 2 in collect order
    assert pol.submit(8) == ([(16, 1), (17, 2)], [])
    pol.begin()
    # 1 voxel per pixel -> level 0: bricks (0, 0, 0) (entry 0, nearest) and
    # (0, 1, 0) (entry 4); their parent 16 is stamped, slot 3 is free and
    # slot 2 (entry 17, last wanted a frame ago) is evicted
    pol.box((0.0, 0.0, 0.0), (0.2, 0.3, 0.2), 1.0 / 512, 0.0)
    assert pol.submit(8) == ([(0, 3), (4, 2)], [17])
    # 8,000 level-0 bricks would exceed 2,048: the box coarsens to level 1
    big = sp.StreamPolicy([(128 * 20,) * 3, (128 * 10,) * 3, (128 * 5,) * 3], 200, np.full(9125, 50),
                          np.array([0] * 9000 + [1] * 125))
    big.begin()
    big.box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, 0.0)
    assert big.lod_wanted[:3] == [0, 1000, 0]
    assert len(big.cands) == 1000  # the parents are resident


def test_unknown_maxima_fill_once_then_gate():
    """A compressed warm tier: a brick's maximum is unknown (not gated) until
    a fill measures it; below the noise floor it stays unpublished, and once
    known it is gated like any other."""
    dims = [(256, 128, 128), (128, 64, 64)]  # entries 0, 1 (level 0), 2 (level 1, resident in slot 0)
    pol = sp.StreamPolicy(dims, 2, np.array([3, 50, 60]), np.array([0, 0, 1]), known=[0, 0, 1], floor=5)
    pol.begin()
    pol.point((0.1, 0.1, 0.1), 0, 0.0)
    assert pol.submit(8) == ([(0, 1)], [])  # unknown maximum: filled
    assert pol.known[0] and pol.entries()[0] == sp.INV and pol.entries()[2] == 0 | 60 << 24
    pol.begin()
    pol.point((0.9, 0.1, 0.1), 0, 0.0)
    assert pol.submit(8) == ([(1, 1)], [0])  # entry 0 is no longer wanted: its slot goes
    pol.begin()
    pol.point((0.1, 0.1, 0.1), 0, 0.0)
    assert pol.submit(8) == ([], [])  # known below the gate: never again
    # with every maximum known up front, entry 0 never takes a slot
    pol = sp.StreamPolicy(dims, 2, np.array([3, 50, 60]), np.array([0, 0, 1]))
    pol.begin()
    pol.point((0.1, 0.1, 0.1), 0, 0.0)
    assert pol.submit(8) == ([], [])
