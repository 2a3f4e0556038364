"""HIP decoder parity: bit-exact against the reference's own outputs (golden
vectors) and against the oracle on generated batches, through the C ABI."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from render3d_amd import workloads

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct3d_golden.npz"))


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_pairs(levels, rng=None):
    """Dense [n,4096] int16 -> (offsets, pairs) with R3D_PAIR packing."""
    offs = [0]
    pairs = []
    for i in range(levels.shape[0]):
        pos = np.nonzero(levels[i])[0].astype(np.uint32)
        if rng is not None:
            rng.shuffle(pos)
        pairs.append(pos | (levels[i][pos].astype(np.int16).view(np.uint16).astype(np.uint32) << 16))
        offs.append(offs[-1] + pos.size)
    p = np.concatenate(pairs).astype(np.uint32) if pairs else np.zeros(0, np.uint32)
    if p.size == 0:
        p = np.zeros(1, np.uint32)
    return np.array(offs, np.uint32), p


def golden_batch():
    n = len(GOLD["blob_names"])
    hdr = orc.hdr_array(GOLD["blob_hdr"][:, 0], GOLD["blob_hdr"][:, 1], GOLD["blob_dcq"], np.arange(n))
    return n, hdr, GOLD["blob_steps"].reshape(-1), GOLD["blob_levels"]


def test_golden_blobs_bit_exact_with_reference_dense(ctx):
    n, hdr, steps, lv = golden_batch()
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(steps), levels=cuda(lv),
                            corr_offsets=cuda(GOLD["blob_corr_off"]), corr_pos=cuda(GOLD["blob_corr_pos"]),
                            corr_delta=cuda(GOLD["blob_corr_delta"]))
    got = out.cpu().numpy()
    bad = [GOLD["blob_names"][i] for i in range(n) if not np.array_equal(got[i], GOLD["blob_u8"][i])]
    assert not bad, f"chunks differing from dct3d_decode_u8: {bad}"


def test_golden_blobs_bit_exact_sparse(ctx):
    n, hdr, steps, lv = golden_batch()
    offs, pairs = to_pairs(lv, np.random.default_rng(1))
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(steps), pairs=cuda(pairs), pair_offsets=cuda(offs),
                            corr_offsets=cuda(GOLD["blob_corr_off"]), corr_pos=cuda(GOLD["blob_corr_pos"]),
                            corr_delta=cuda(GOLD["blob_corr_delta"]))
    assert np.array_equal(out.cpu().numpy(), GOLD["blob_u8"])


def test_linear_table_golden(ctx):
    k = GOLD["lin_levels"].shape[0]
    hdr = orc.hdr_array(GOLD["lin_vmin"], GOLD["lin_vspan"], GOLD["lin_dcq"])
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(GOLD["lin_step"]), levels=cuda(GOLD["lin_levels"]))
    assert np.array_equal(out.cpu().numpy(), GOLD["lin_u8"].reshape(k, 4096))


@pytest.mark.parametrize("n,nsheets,seed", [(64, 3, 11), (128, 6, 1234)])
def test_generated_batch_matches_oracle(ctx, n, nsheets, seed):
    vol = ctx.synth_sheets(n, n, n, nsheets, seed)
    step = cuda(orc.step_linear())
    levels, hdr = ctx.encode_chunks(vol, n, n, n, step)
    out = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels)
    torch.cuda.synchronize()
    ref, status = orc.decode_batch(levels.cpu().numpy(), hdr.cpu().numpy(), orc.step_linear())
    assert status.all()
    got = out.cpu().numpy()
    assert np.array_equal(got, ref), f"{int(np.sum(got != ref))} voxels differ"
    # decoded data resembles the volume it was encoded from (lossy codec)
    vc = vol.cpu().numpy().reshape(n // 16, 16, n // 16, 16, n // 16, 16).transpose(0, 2, 4, 1, 3, 5)
    assert np.mean(np.abs(vc.reshape(-1, 4096).astype(int) - got.astype(int))) < 12


def test_status_and_zero_fill_for_invalid_headers(ctx):
    rng = np.random.default_rng(2)
    lv = (rng.integers(-20, 20, (6, 4096)) * (rng.random((6, 4096)) < 0.05)).astype(np.int16)
    hdr = orc.hdr_array([0, 0, np.nan, 0, 3, 0], [255, 0, 1, 255, 100, -1], [100, 1, 2, 3, 4, 5], [0, 0, 0, 7, 0, 0])
    status = torch.full((6,), -5, dtype=torch.int32, device="cuda")
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(orc.step_linear()), levels=cuda(lv), status=status)
    got = out.cpu().numpy()
    st = status.cpu().numpy()
    assert st.tolist() == [1, 0, 0, 0, 1, 0]
    for i in (1, 2, 3, 5):
        assert not got[i].any()
    for i in (0, 4):
        ok, ref, _ = orc.decode_chunk(lv[i], hdr[i, 0:1].view(np.float32)[0], hdr[i, 1:2].view(np.float32)[0],
                                      hdr[i, 2], orc.step_linear())
        assert ok and np.array_equal(got[i], ref)


def test_extreme_levels(ctx):
    lv = np.zeros((4, 4096), np.int16)
    lv[0, :] = 32767
    lv[1, :] = -32768
    lv[2, ::7] = 32767
    lv[3, 0] = -32768
    hdr = orc.hdr_array([0, 0, 10, -5], [255, 255, 3, 1000], [0, 4000, -100, 12], [0, 0, 0, 0])
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(orc.step_qstep(np.float32(3.0))), levels=cuda(lv))
    ref, _ = orc.decode_batch(lv, hdr, orc.step_qstep(np.float32(3.0)))
    assert np.array_equal(out.cpu().numpy(), ref)


def test_strided_destination_matches_flat(ctx):
    # decode 8 chunks into a 2x2x2-chunk (32^3) volume laid out with padded strides
    vol = ctx.synth_sheets(32, 32, 32, 1, 5)
    step = cuda(orc.step_linear())
    levels, hdr = ctx.encode_chunks(vol, 32, 32, 32, step)
    flat = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels).cpu().numpy()
    sy, sz = 40, 40 * 17
    dst = torch.zeros(8 * 16 * sz + 64, dtype=torch.uint8, device="cuda")
    offs = cuda(np.arange(8, dtype=np.int64) * 16 * sz + 3)
    b = ctx.make_batch(headers=hdr, step_tables=step, out=dst, levels=levels, out_offsets=offs, stride_y=sy,
                       stride_z=sz)
    ctx.decode(b)
    d = dst.cpu().numpy()
    for i in range(8):
        for z in range(16):
            for y in range(16):
                o = 3 + i * 16 * sz + z * sz + y * sy
                assert np.array_equal(d[o:o + 16], flat[i, (z * 16 + y) * 16:(z * 16 + y) * 16 + 16])


def test_invalid_batches_are_rejected_without_launch(ctx):
    import render3d_amd as r3d

    hdr = cuda(orc.hdr_array([0], [255], [0]))
    step = cuda(orc.step_linear())
    lv = torch.zeros(4096 * 2, dtype=torch.int16, device="cuda")
    out = torch.zeros(4096, dtype=torch.uint8, device="cuda")
    b = ctx.make_batch(headers=hdr, step_tables=step, out=out, levels=lv[1:4097])  # misaligned
    with pytest.raises(r3d.R3DError):
        ctx.decode(b)
    b = ctx.make_batch(headers=hdr, step_tables=step, out=out, levels=lv, stride_y=8)
    with pytest.raises(r3d.R3DError):
        ctx.decode(b)
    b = ctx.make_batch(headers=hdr, step_tables=step, out=out, levels=lv)
    with pytest.raises(r3d.R3DError):
        ctx.decode(b, variant=7)
    b.nchunks = 0
    ctx.decode(b)  # empty batch is a no-op


def test_corrupt_sparse_offsets_stay_in_bounds(ctx):
    # offsets past npairs are clamped; the decode still completes
    lv = np.zeros((2, 4096), np.int16)
    lv[:, 0] = 3
    offs, pairs = to_pairs(lv)
    offs = np.array([0, 1, 1 << 30], np.uint32)
    hdr = orc.hdr_array([0, 0], [255, 255], [0, 0])
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(orc.step_linear()), pairs=cuda(pairs),
                            pair_offsets=cuda(offs))
    torch.cuda.synchronize()
    assert out.shape == (2, 4096)


def fp64_reference(levels, hdr, step):
    """FP64 decode of dequantised levels (deq_one in FP32 as dct3d.c:149-154,
    then an FP64 separable iDCT, + dc, denormalise)."""
    cm = orc.basis().astype(np.float64).reshape(16, 16)
    n = levels.shape[0]
    out = np.zeros((n, 4096))
    for i in range(n):
        coef = np.zeros(4096, np.float32)
        orc.load().r3do_dequant(np.ascontiguousarray(levels[i]), np.ascontiguousarray(step, np.float32), coef)
        blk = np.einsum("az,by,cx,abc->zyx", cm, cm, cm, coef.astype(np.float64).reshape(16, 16, 16)).reshape(-1)
        vmin = hdr[i, 0:1].view(np.float32)[0]
        vspan = hdr[i, 1:2].view(np.float32)[0]
        out[i] = (blk + hdr[i, 2] / 16.0) * (vspan / 255.0) + vmin
    return out


def decode_f32(ctx, levels, hdr, step, variant):
    out = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels, variant=variant, f32_out=True)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_exact_f32_output_is_the_pre_round_value(ctx):
    n = 64
    vol = ctx.synth_sheets(n, n, n, 3, 71)
    step = cuda(orc.step_linear())
    levels, hdr = ctx.encode_chunks(vol, n, n, n, step)
    got = decode_f32(ctx, levels, hdr, step, 0)
    lv, hd = levels.cpu().numpy(), hdr.cpu().numpy()
    for i in range(0, lv.shape[0], 7):
        ok, _, pre = orc.decode_chunk(lv[i], hd[i, 0:1].view(np.float32)[0], hd[i, 1:2].view(np.float32)[0],
                                      hd[i, 2], orc.step_linear())
        assert np.array_equal(got[i].view(np.uint32), pre.view(np.uint32)), i


@pytest.mark.parametrize("variant,tol_f32,max_u8,frac_u8", [(1, 2e-3, 1, 2e-3), (3, 2e-3, 1, 2e-3),
                                                             (2, 0.75, 1, 0.6)])
def test_mfma_variants_error_against_fp64(ctx, variant, tol_f32, max_u8, frac_u8):
    n = 64
    vol = ctx.synth_sheets(n, n, n, 3, 1234)
    step = cuda(orc.step_linear())
    levels, hdr = ctx.encode_chunks(vol, n, n, n, step)
    lv, hd = levels.cpu().numpy(), hdr.cpu().numpy()
    got = decode_f32(ctx, levels, hdr, step, variant)
    ref = fp64_reference(lv, hd, orc.step_linear())
    err = np.abs(got - ref)
    assert err.max() < tol_f32
    exact = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels).cpu().numpy().astype(int)
    approx = ctx.decode_chunks(headers=hdr, step_tables=step, levels=levels, variant=variant).cpu().numpy().astype(int)
    d = np.abs(approx - exact)
    print(f"variant {variant}: max |f32 - fp64| = {err.max():.3g}, mean {err.mean():.3g}; "
          f"u8 differing from exact: {(d > 0).mean():.3g}")
    assert d.max() <= max_u8 and (d > 0).mean() <= frac_u8


def test_mfma_sparse_matches_dense(ctx):
    n = 16
    rng = np.random.default_rng(3)
    lv = (rng.integers(-40, 40, (n, 4096)) * (rng.random((n, 4096)) < 0.08)).astype(np.int16)
    hdr = orc.hdr_array(rng.uniform(0, 20, n), rng.uniform(100, 255, n), rng.integers(0, 3000, n))
    offs, pairs = to_pairs(lv, rng)
    for variant in (1, 2, 3):
        a = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(orc.step_linear()), levels=cuda(lv),
                              variant=variant).cpu().numpy()
        b = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(orc.step_linear()), pairs=cuda(pairs),
                              pair_offsets=cuda(offs), variant=variant).cpu().numpy()
        assert np.array_equal(a, b), variant


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_mfma_golden_blobs_with_corrections_within_one(ctx, variant):
    n, hdr, steps, lv = golden_batch()
    out = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(steps), levels=cuda(lv),
                            corr_offsets=cuda(GOLD["blob_corr_off"]), corr_pos=cuda(GOLD["blob_corr_pos"]),
                            corr_delta=cuda(GOLD["blob_corr_delta"]), variant=variant)
    d = np.abs(out.cpu().numpy().astype(int) - GOLD["blob_u8"].astype(int))
    assert d.max() <= 1


def abs_transform_bound(levels, hdr, step):
    """Per voxel (|C| x |C| x |C|) |coef| * vspan / 255: the magnitude every
    rounding error of the separable transform is measured against."""
    cm = np.abs(orc.basis().astype(np.float64).reshape(16, 16))
    out = np.zeros((levels.shape[0], 4096))
    for i in range(levels.shape[0]):
        coef = np.zeros(4096, np.float32)
        orc.load().r3do_dequant(np.ascontiguousarray(levels[i]), np.ascontiguousarray(step, np.float32), coef)
        a = np.einsum("az,by,cx,abc->zyx", cm, cm, cm, np.abs(coef.astype(np.float64)).reshape(16, 16, 16))
        out[i] = a.reshape(-1) * hdr[i, 1:2].view(np.float32)[0] / 255.0
    return out


# Each FP16 rounding (pass-1 input, pass-2 input, pass-3 input, and the basis
# in each pass) contributes at most u16 = 2^-11 times the absolute transform;
# the FP32-level variants get a generous 96 * 2^-24.
BOUND = {1: 96 * 2.0 ** -24, 3: 96 * 2.0 ** -24, 2: 6 * 2.0 ** -11}


@pytest.fixture(scope="module")
def sparse_cases():
    # the batched workload's chunks with the largest absolute transforms, plus a spread
    offs, pairs, hdr = workloads.sparse_laplacian_batch(32768, 20261203)
    idx = [15182, 29724, 11338, 10570, 22645, 25560] + list(np.linspace(0, 32767, 122).astype(np.int64))
    return {"seed17": workloads.sparse_laplacian_batch(192, 17), "batched": workloads.select(offs, pairs, hdr, idx)}


@pytest.mark.parametrize("case", ["seed17", "batched"])
@pytest.mark.parametrize("variant", [1, 3, 2])
def test_mfma_variants_on_small_sparse_levels(ctx, sparse_cases, case, variant):
    # Laplacian levels of scale 12 over a wide range of magnitudes in one chunk
    offs, pairs, hdr = sparse_cases[case]
    step = workloads.step_table_linear()
    got = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(step), pairs=cuda(pairs), pair_offsets=cuda(offs),
                            variant=variant, f32_out=True).cpu().numpy()
    lv = workloads.densify(offs, pairs)
    ref = fp64_reference(lv, hdr, step)
    err = np.abs(got - ref)
    bound = BOUND[variant] * abs_transform_bound(lv, hdr, step) + 1e-5
    print(f"variant {variant} sparse: max |f32 - fp64| = {err.max():.3g}, mean {err.mean():.3g}, "
          f"max err / bound {np.max(err / bound):.3g}")
    assert np.all(err <= bound)


@pytest.mark.parametrize("variant", [1, 3])
def test_split_is_consistent_at_fp16_ties(ctx, variant):
    # level 86 at L1 = 12 (step 25): (86 + 0.2) * 25 rounds to 2155.0 in FP32, an
    # exact tie between the FP16 values 2154 and 2156; single and double rounding
    # of the product pick different sides, and the hi/lo split must still agree
    lv = np.zeros((4, 4096), np.int16)
    lv[0, 2 * 256 + 0 * 16 + 10] = 86
    lv[1, 2 * 256 + 0 * 16 + 10] = -86
    lv[2, 5 * 256 + 3 * 16 + 4] = 86  # L1 = 12 again, another position
    lv[3, ::97] = 86
    hdr = orc.hdr_array([0, 0, 0, 0], [255, 255, 255, 255], [0, 0, 0, 0])
    step = orc.step_linear()
    got = ctx.decode_chunks(headers=cuda(hdr), step_tables=cuda(step), levels=cuda(lv), variant=variant,
                            f32_out=True).cpu().numpy()
    ref = fp64_reference(lv, hdr, step)
    bound = BOUND[variant] * abs_transform_bound(lv, hdr, step) + 1e-5
    assert np.all(np.abs(got - ref) <= bound), np.max(np.abs(got - ref))
