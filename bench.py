"""Benchmark: the 1024^3 / 1920x1080 workload on one MI355X.

One step is the full hot path over one synthetic batch:
  1. decode all 262,144 16^3 chunks (dense int16 levels, step table
     q = 1 + 2(i+j+k)) into flat 4096-byte chunks,
  2. build the bricked volume from them: level 0 with aprons, the 4-level
     mip pyramid and the dilated occupancy grid,
  3. raycast one 1920x1080 view with ray-cone LOD and full gradient shading.
Inputs (the sheet volume and its encoded levels) are generated on the GPU and
resident in HBM before the timed region.

Printed JSON line: value = ray samples per second of the raycast (samples =
loop iterations that fetch a volume sample, counted exactly by the kernel);
fps = 1000 / raycast ms; decode_GBps = algorithmic decode bytes / decode ms;
roofline = the decode kernel against HBM (12,288 B per chunk: 8,192 B of
levels read + 4,096 B of voxels written); cpu_baseline = the oracle's host
raycaster timed on a row sample of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
VALU_PEAK_G = 1228.8  # wave64 VALU instructions/s: the FP32 peak 157.3 TFLOP/s / (2 x 64)
CHUNK_BYTES = 8192 + 4096


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="volume edge (multiple of 16)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--cpu-rows", type=int, default=1, help="CPU baseline renders every Nth row")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-views", type=int, default=8)
    ap.add_argument("--adj-grid", type=int, default=128)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched (graph) workload")
    ap.add_argument("--batched-chunks", type=int, default=32768)
    ap.add_argument("--no-cfg0", action="store_true", help="skip the 256^3 / 512^2 workload")
    ap.add_argument("--no-deblock", action="store_true", help="skip the deblocking workload")
    ap.add_argument("--no-shard", action="store_true", help="skip the zarr shard region-decode workload")
    ap.add_argument("--no-bricks", action="store_true", help="skip the bricks-LOD (page table) workload")
    ap.add_argument("--no-stream", action="store_true", help="skip the bricks streaming-residency workload")
    ap.add_argument("--no-blobs", action="store_true", help="skip the 3ddct blob (entropy stage) workload")
    ap.add_argument("--no-fit", action="store_true", help="skip the differentiable 256^3 fit (200 Adam steps)")
    ap.add_argument("--profile-decode", action="store_true",
                    help="for rocprofv3 --pmc: one 1 GiB device copy (calibration) then one decode launch")
    ap.add_argument("--profile-variant", type=int, default=0,
                    help="decode variant the --profile-decode launch uses (0 exact, 1 MFMA FP32, 2 MFMA FP16, 3 split FP16)")
    ap.add_argument("--profile-blobs", action="store_true",
                    help="build the blob workload, run one 1 GiB copy and one blob parse, then exit (for rocprofv3)")
    ap.add_argument("--profile-raycast", action="store_true",
                    help="for rocprofv3 --pmc: build the volume, one 1 GiB device copy, then one raycast of the frame")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "decode_exact_1024.md"),
                    help="rocprofv3 summary holding the decode kernel's measured HBM bytes per launch")
    return ap.parse_args()


def rows_desc(rs):
    """The rows a CPU-baseline sample renders."""
    return "every row" if rs == 1 else f"every {rs}th row"


def host_cpu():
    """The host's CPU: model (/proc/cpuinfo), online cores (sysconf) and the
    threads the baselines use: the share of host cores this GPU is given
    (OMP_NUM_THREADS, set per GPU on a shared box), else every core."""
    total = os.sysconf("SC_NPROCESSORS_ONLN")
    share = os.environ.get("OMP_NUM_THREADS", "")
    cores = max(1, min(int(share), total)) if share.isdigit() else total
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cores": cores, "cores_total": total, "model": model,
            "cores_note": "OMP_NUM_THREADS (this GPU's share of the host)" if share.isdigit() else "all online cores"}


def median_rate(run, work, runs=5, min_s=1.0):
    """`runs` timed runs of `run` (each repeated until it has taken min_s),
    each giving work / seconds; returns (median, all rates)."""
    import numpy as np
    rates = []
    for _ in range(runs):
        t = time.perf_counter()
        reps = 0
        while True:
            run()
            reps += 1
            if time.perf_counter() - t >= min_s:
                break
        rates.append(reps * work / (time.perf_counter() - t))
    return float(np.median(rates)), [float(r) for r in rates]


def read_traffic(path, key="traffic_bytes_per_launch"):
    """`<key>: N` (HBM bytes or instructions per launch) from a committed
    rocprof summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: one GPU per process; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design, not built)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import render3d_amd as r3d
    from render3d_amd import scene

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = r3d.Context(0)
    n, W, H = args.n, args.width, args.height
    nsheets = max(1, min(48, (n // 2 - 20) // 9 + 1 if n >= 64 else 1))
    seed = 20261203
    stream = torch.cuda.current_stream()

    # ---- inputs, resident in HBM before timing --------------------------------
    t0 = time.time()
    vol_lin = ctx.synth_sheets(n, n, n, nsheets, seed)
    step_tab = torch.arange(46, dtype=torch.float32, device=dev) * 2 + 1  # q = 1 + 2(i+j+k)
    levels, headers = ctx.encode_chunks(vol_lin, n, n, n, step_tab)
    del vol_lin
    nch = levels.shape[0]
    vol = ctx.volume(n, n, n, args.levels)
    flat = torch.empty((nch, 4096), dtype=torch.uint8, device=dev)
    batch = ctx.make_batch(headers=headers, step_tables=step_tab, out=flat, levels=levels)
    lut = scene.tf_ramp_40_200()
    ctx.set_transfer(lut)
    cam = scene.orbit_camera(45.0, 0.0, 2.0)
    view = scene.frame_params(cam, W, H, step_voxels=1.0, gate=scene.skip_gate(lut))
    img = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    setup_s = time.time() - t0

    if args.profile_decode:
        a = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)  # a plain copy of known size: 1 GiB read + 1 GiB written
        torch.cuda.synchronize()
        ctx.decode(batch, variant=args.profile_variant)
        torch.cuda.synchronize()
        print(json.dumps({"profile_decode": True, "variant": args.profile_variant, "chunks": nch,
                          "copy_bytes": 1 << 30}))
        return 0
    if args.profile_blobs:
        del levels, headers, batch, flat
        _, lv, hd, data, offs, lens = make_blobs(ctx, n, nsheets, seed, dev)
        del lv, hd
        a = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        parsed = ctx.parse_blobs(data, offs, lens)
        torch.cuda.synchronize()
        print(json.dumps({"profile_blobs": True, "blobs": int(offs.shape[0]), "blob_bytes": int(lens.long().sum()),
                          "nnz": int(parsed["pair_counts"].long().sum()), "copy_bytes": 1 << 30}))
        return 0
    if args.profile_raycast:
        ctx.decode(batch)
        vol.build_from_chunks(flat)
        a = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        ctx.raycast(vol, [view], W, H, alpha_cutoff=0.99, out=img, counters=counters)
        torch.cuda.synchronize()
        print(json.dumps({"profile_raycast": True, "samples": int(counters[0]), "copy_bytes": 1 << 30}))
        return 0

    def step(timed):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        ctx.decode(batch)
        ev[1].record(stream)
        vol.build_from_chunks(flat)
        ev[2].record(stream)
        ctx.raycast(vol, [view], W, H, alpha_cutoff=0.99, out=img, counters=counters if timed else None)
        ev[3].record(stream)
        return ev

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    counters.zero_()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    events = [step(True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1000.0
    dec_ms = [e[0].elapsed_time(e[1]) for e in events]
    pyr_ms = [e[1].elapsed_time(e[2]) for e in events]
    ray_ms = [e[2].elapsed_time(e[3]) for e in events]
    cnt = counters.cpu().numpy().astype(np.uint64) // np.uint64(args.steps)
    samples, skips, shaded, rays = (int(x) for x in cnt)
    ray_avg = float(np.mean(ray_ms))
    dec_avg = float(np.mean(dec_ms))
    value = samples / (ray_avg / 1000.0)
    dec_bytes = CHUNK_BYTES * nch
    dec_gbps = dec_bytes / (dec_avg / 1000.0) / 1e9

    ray_roof = raycast_roofline(ctx, vol, view, W, H, img, ray_avg, dev)
    variants = decode_variants(ctx, batch, levels, headers, step_tab, flat, stream, dev, max(3, args.steps))
    ctx.decode(batch)  # leave the exact decode in `flat`

    adj = None if args.no_adjoint else adjoint_ratio(ctx, vol, lut, args, W, H, dev)
    bat = None if args.no_batched else batched(ctx, vol, lut, args, dev, stream)
    c0 = None if args.no_cfg0 else cfg0(ctx, args, dev, stream)
    blob = None if args.no_blobs else blob_stage(ctx, args, dev, stream, n, nsheets, seed)
    shd = None if args.no_shard else shard_stage(ctx, args, dev, stream, n, nsheets, seed)
    deb = None
    if shd is not None and not args.no_deblock:
        deb = deblock_stage(ctx, args, dev, stream, shd["_state"]["dst"], BLOB_QSTEP)
    brk = None if args.no_bricks else bricks_stage(ctx, args, dev, stream, n, nsheets, seed, lut, cam, W, H)
    strm = None
    if not args.no_stream:
        strm = stream_stage(ctx, args, dev, stream, n, nsheets, seed, None if shd is None else shd["_state"]["shard"])
    fit = None if args.no_fit else fit_256(ctx, dev)

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(ctx, vol, view, lut, img, args, levels, headers, step_tab, n, W, H)
        if c0 is not None:
            c0["cpu_baseline"] = cpu_cfg0(c0.pop("_state"), args)
        if blob is not None:
            blob["cpu_baseline"], blob["parity_vs_oracle"] = cpu_blobs(blob["_state"], args)
            cpu["decode_reference"] = blob["cpu_baseline"].get("reference")
        if shd is not None:
            shd["cpu_baseline"], shd["parity_vs_reference"] = cpu_shard(shd["_state"], args)
        if deb is not None:
            deb["cpu_baseline"], deb["parity_vs_reference"] = cpu_deblock(deb["_state"], args)
        if brk is not None:
            brk["cpu_baseline"], brk["parity_vs_oracle"] = cpu_bricks(brk["_state"], args)
        if strm is not None and shd is not None and strm.get("shards") is not None:
            c = shd["cpu_baseline"]
            strm["shards"]["cpu_baseline"] = {
                "value": c["value"] / 512.0, "unit": "bricks/s", "cores": c["cores"], "kind": c["kind"],
                "sample": "the reference reader's chunk rate on a 64 x 1024 x 1024 slab of the same shard "
                          "(shard.cpu_baseline) / 512 chunks per 128^3 brick"}
    if c0 is not None:
        c0.pop("_state", None)
    if blob is not None:
        blob.pop("_state", None)
    if shd is not None:
        shd.pop("_state", None)
    if deb is not None:
        deb.pop("_state", None)
    if brk is not None:
        brk.pop("_state", None)

    out = {
        "metric": "ray samples/s + fps @1024³ 1080p; chunk decode GB/s; adjoint/fwd time ratio",
        "value": value,
        "unit": "ray samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sheet volume generated and encoded on the GPU)",
        "config": {"workload": f"{n}^3 u8 sheet volume, {nch} 16^3 chunks dense int16, "
                               f"{args.levels}-level pyramid, {W}x{H} FULL raycast, step 1.0, cutoff 0.99",
                   "volume": n, "chunks": nch, "width": W, "height": H, "levels": args.levels},
        "fps": 1000.0 / ray_avg,
        "raycast_ms": ray_avg,
        "decode_ms": dec_avg,
        "pyramid_ms": float(np.mean(pyr_ms)),
        "decode_GBps": dec_gbps,
        "decode_variants": variants,
        "adjoint_fwd_ratio": None if adj is None else adj["ratio"],
        "adjoint": adj,
        "batched": bat,
        "cfg0": c0,
        "fit": fit,
        "blobs": blob,
        "shard": shd,
        "deblock": deb,
        "bricks": brk,
        "stream": strm,
        "samples_per_frame": samples,
        "skips_per_frame": skips,
        "shaded_per_frame": shaded,
        "rays_per_frame": rays,
        "setup_s": setup_s,
        "roofline": {"kernel": "k_decode_exact", "bound": "hbm", "achieved": dec_gbps,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": dec_gbps / HBM_PEAK_GBPS,
                     "traffic": read_traffic(args.traffic)},
        "raycast_roofline": ray_roof,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
    }
    print(json.dumps(out))
    return 0


def raycast_roofline(ctx, vol, view, W, H, img, ray_ms, dev):
    """The raycast's algorithmic bytes for the timed frame, from one run of
    the measurement instances (same image, checked bit for bit):
      gather bytes     = 8 B per trilinear footprint (16 B when the sample
                         blends two mip levels) x (1 + 6 taps per shaded
                         sample) + 16 B of FP32 output per pixel;
      compulsory bytes = the distinct bricks the footprints read x 4,928 B
                         (16^3 + apron, padded) + the output: the least the
                         frame can move through HBM.
    achieved = compulsory bytes / the timed frame's mean ms, against HBM."""
    import numpy as np
    import torch

    touched = torch.zeros((vol.brick_count + 31) // 32, dtype=torch.int32, device=dev)
    mc = torch.zeros(6, dtype=torch.int64, device=dev)
    img_m = torch.empty_like(img)
    ctx.raycast(vol, [view], W, H, alpha_cutoff=0.99, out=img_m, counters=mc, measure=True, touched=touched)
    torch.cuda.synchronize()
    same = bool(torch.equal(img_m, img))
    samples, skips, shaded, rays, blended, blended_sh = (int(x) for x in mc.cpu().numpy())
    uniq = int(np.unpackbits(touched.cpu().numpy().view(np.uint8)).sum())
    out_bytes = W * H * 16
    gather = 8 * (samples + blended) + 48 * (shaded + blended_sh) + out_bytes
    compulsory = uniq * 4928 + out_bytes
    comp_gbps = compulsory / (ray_ms / 1e3) / 1e9
    prof = os.path.join(ROOT, "profiles", "raycast_1080p.md")
    return {"kernel": "k_raycast_tiers<FULL>", "bound": "hbm", "achieved": comp_gbps, "peak": HBM_PEAK_GBPS,
            "unit": "GB/s", "frac": comp_gbps / HBM_PEAK_GBPS,
            "traffic": read_traffic(prof) if os.path.exists(prof) else None,
            "compulsory_bytes": compulsory, "unique_bricks": uniq, "bricks_in_pyramid": vol.brick_count,
            "gather_bytes": gather, "gather_GBps": gather / (ray_ms / 1e3) / 1e9,
            "samples": samples, "blended": blended, "shaded": shaded, "blended_shaded": blended_sh,
            "measurement_image_equal": same}


def fp64_decode(levels, headers, step_tab):
    """FP64 decode of [k, 4096] int16 levels: dead-zone dequantiser
    (r = (|l| - 1 + 0.8 + 0.4) step), the orthonormal DCT-II basis in FP64,
    separable inverse transform, + dcq/16, denormalise by vspan/255 and vmin."""
    import math

    import torch
    dev = levels.device
    k = torch.arange(16, dtype=torch.float64, device=dev)
    cm = torch.cos(math.pi * (2 * k[None, :] + 1) * k[:, None] / 32)
    cm = cm * torch.where(k == 0, math.sqrt(1 / 16), math.sqrt(2 / 16))[:, None]
    f = (k[:, None, None] + k[None, :, None] + k[None, None, :]).long()
    step = step_tab.double()[f]
    lv = levels.double().view(-1, 16, 16, 16)
    r = (lv.abs() - 1) * step + 0.8 * step + 0.4 * step
    coef = torch.where(lv == 0, torch.zeros_like(lv), torch.sign(lv) * r)
    t = torch.einsum("nabc,cx->nabx", coef, cm)
    t = torch.einsum("nabx,by->nayx", t, cm)
    t = torch.einsum("nayx,az->nzyx", t, cm).reshape(-1, 4096)
    hf = headers.view(torch.float32)
    vmin, vspan = hf[:, 0:1].double(), hf[:, 1:2].double()
    dc = headers[:, 2:3].double() / 16
    return (t + dc) * (vspan / 255) + vmin


def decode_variants(ctx, batch, levels, headers, step_tab, flat, stream, dev, reps):
    """Each decode variant over the whole batch: time, GB/s, HBM fraction,
    voxels whose u8 differs from the exact decode, and max/mean |x - FP64| of
    the value before rounding on a 1024-chunk sample."""
    import numpy as np
    import torch

    import render3d_amd as r3d
    nch = levels.shape[0]
    exact = flat.clone()
    idx = torch.arange(0, nch, max(1, nch // 1024), device=dev)[:1024]
    ls, hs = levels[idx].contiguous(), headers[idx].contiguous()
    ref = fp64_decode(ls, hs, step_tab)
    names = {r3d.R3D_DECODE_EXACT_F32: "exact_f32", r3d.R3D_DECODE_MFMA_F32: "mfma_f32",
             r3d.R3D_DECODE_MFMA_F16: "mfma_f16",
             r3d.R3D_DECODE_MFMA_F16X3: "mfma_f16x3"}
    out = {}
    for v, name in names.items():
        ctx.decode(batch, variant=v)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(stream)
        for i in range(reps):
            ctx.decode(batch, variant=v)
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        ms = float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))
        diff = float((flat != exact).float().mean())
        f32 = ctx.decode_chunks(headers=hs, step_tables=step_tab, levels=ls, variant=v, f32_out=True)
        err = (f32.double() - ref).abs()
        gbps = CHUNK_BYTES * nch / (ms / 1000) / 1e9
        out[name] = {"ms": ms, "GBps": gbps, "hbm_frac": gbps / HBM_PEAK_GBPS, "u8_differs_from_exact": diff,
                     "max_abs_err_vs_fp64": float(err.max()), "mean_abs_err_vs_fp64": float(err.mean())}
    # HBM bytes per launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated)
    for name, doc in (("exact_f32", "decode_exact_1024.md"), ("mfma_f16", "decode_mfma_f16_1024.md")):
        out[name]["traffic"] = read_traffic(os.path.join(ROOT, "profiles", doc))
    return out


def batched(ctx, vol, lut, args, dev, stream):
    """The batched workload over the resident volume: `batched_chunks` sparse
    chunks (Laplacian nonzero counts in 5..400, Laplacian levels of scale 12)
    decoded, and 64 views of mixed resolution (16 each at 256^2, 512^2,
    1280x720, 1920x1080; 53,166,080 rays) rendered into one RGBA8 atlas. The
    render's launch shape comes from the autotuner; decode and render are
    captured into one HIP graph, whose replays are timed."""
    import numpy as np
    import torch

    import render3d_amd as r3d
    from render3d_amd import scene, workloads

    n = args.batched_chunks
    t0 = time.perf_counter()
    offs, pairs, hdr = workloads.sparse_laplacian_batch(n, 20261203)
    gen_s = time.perf_counter() - t0
    nnz = int(offs[-1])
    step = torch.from_numpy(workloads.step_table_linear()).to(dev)
    d_offs, d_pairs, d_hdr = (torch.from_numpy(x).to(dev) for x in (offs, pairs, hdr))
    out = torch.empty((n, 4096), dtype=torch.uint8, device=dev)
    batch = ctx.make_batch(headers=d_hdr, step_tables=step, out=out, pairs=d_pairs, pair_offsets=d_offs)
    views, aw, ah = workloads.mixed_views(gate=scene.skip_gate(lut))
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    plan = ctx.raycast_plan(vol, views, aw, ah, rgba=False, rgba8=True, counters=counters)
    t0 = time.perf_counter()
    tuning = plan.tune()
    tune_s = time.perf_counter() - t0
    chosen = plan.tuning
    reps = max(3, args.steps)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(stream)
        for i in range(reps):
            fn()
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        return float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))

    in_bytes = 4 * nnz + 4 * (n + 1) + 16 * n
    variants = {}
    exact = None
    k = min(1024, n)
    idx = np.linspace(0, n - 1, k).astype(np.int64)
    sub_off = np.zeros(k + 1, np.uint32)
    cnt = np.diff(offs.astype(np.int64))[idx]
    sub_off[1:] = np.cumsum(cnt)
    sub_pairs = np.concatenate([pairs[offs[i]:offs[i + 1]] for i in idx])
    ref = fp64_decode(torch.from_numpy(workloads.densify(sub_off, sub_pairs)).to(dev),
                      torch.from_numpy(hdr[idx]).to(dev), step)
    s_off, s_pairs, s_hdr = (torch.from_numpy(x).to(dev) for x in (sub_off, sub_pairs, np.ascontiguousarray(hdr[idx])))
    for v, name in ((r3d.R3D_DECODE_EXACT_F32, "exact_f32"), (r3d.R3D_DECODE_MFMA_F32, "mfma_f32"),
                    (r3d.R3D_DECODE_MFMA_F16, "mfma_f16"), (r3d.R3D_DECODE_MFMA_F16X3, "mfma_f16x3")):
        ms = timed(lambda: ctx.decode(batch, variant=v))
        if exact is None:
            exact = out.clone()
        f32 = ctx.decode_chunks(headers=s_hdr, step_tables=step, pairs=s_pairs, pair_offsets=s_off, variant=v,
                                f32_out=True)
        err = (f32.double() - ref).abs()
        variants[name] = {"ms": ms, "GBps": (in_bytes + 4096 * n) / (ms / 1e3) / 1e9,
                          "u8_differs_from_exact": float((out != exact).float().mean()),
                          "max_abs_err_vs_fp64": float(err.max()), "mean_abs_err_vs_fp64": float(err.mean())}
    ray_ms = timed(plan.launch)
    # decode (exact) + render captured into one graph
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            ctx.decode(batch)
            plan.launch()
    stream.wait_stream(side)
    torch.cuda.synchronize()
    graph_ms = timed(g.replay)
    counters.zero_()
    g.replay()
    torch.cuda.synchronize()
    samples, skips, shaded, rays = (int(x) for x in counters.cpu().numpy())
    ok = bool(torch.equal(out, exact))
    plan.close()
    return {"chunks": n, "nnz": nnz, "views": len(views), "atlas": [aw, ah], "rays": rays,
            "samples": samples, "shaded": shaded, "skips": skips,
            "graph_ms": graph_ms, "decode_ms": variants["exact_f32"]["ms"], "raycast_ms": ray_ms,
            "samples_per_s": samples / (ray_ms / 1e3), "graph_decode_matches_eager": ok,
            "tuning": {"chosen": chosen, "candidates": tuning, "tune_s": tune_s},
            "decode_variants": variants, "generate_s": gen_s}


BLOB_QSTEP = 8.0  # lodpack's level-0 quality (tools/lodpack/main.c:392)


def make_blobs(ctx, n, nsheets, seed, dev):
    """The n^3 sheet volume quantised with the qstep table and entropy-coded on
    the GPU into one 3ddct blob per chunk, packed back to back as in a shard.
    Returns (step table, levels, headers, bytes, offsets int64, lengths int32)."""
    import torch

    from render3d_amd import workloads

    step = torch.from_numpy(workloads.step_table_qstep(BLOB_QSTEP)).to(dev)
    vol = ctx.synth_sheets(n, n, n, nsheets, seed)
    levels, headers = ctx.encode_chunks(vol, n, n, n, step)
    del vol
    slots, slot_offs, lens = ctx.encode_blobs(levels, headers, BLOB_QSTEP)
    l64 = lens.long()
    offs = torch.cumsum(l64, 0) - l64
    total = int(l64.sum())
    within = torch.arange(total, device=dev) - torch.repeat_interleave(offs, l64)
    data = slots[torch.repeat_interleave(slot_offs, l64) + within].contiguous()
    return step, levels, headers, data, offs, lens


def blob_stage(ctx, args, dev, stream, n, nsheets, seed):
    """The 3ddct blob path at n^3: the sheet volume quantised with the qstep
    table (qstep 8, lodpack's level-0 quality, tools/lodpack/main.c:392) and
    entropy-coded on the GPU into one blob per chunk, packed back to back as in
    a shard. Timed: the parse (blobs -> sparse levels and per-chunk tables),
    the exact decode of the parsed batch, and both captured as one HIP graph.
    The graph's output must equal the decode of the quantised levels."""
    import numpy as np
    import torch

    qstep = BLOB_QSTEP
    step, levels, headers, data, offs, lens = make_blobs(ctx, n, nsheets, seed, dev)
    nch = int(levels.shape[0])
    reps = max(3, args.steps)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(stream)
        for i in range(reps):
            fn()
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        return float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))

    enc_ms = timed(lambda: ctx.encode_blobs(levels, headers, qstep))
    total = int(lens.long().sum())
    direct = ctx.decode_chunks(headers=headers, step_tables=step, levels=levels)
    del levels
    parse_ms = timed(lambda: ctx.parse_blobs(data, offs, lens))
    parsed = ctx.parse_blobs(data, offs, lens)
    out = torch.empty((nch, 4096), dtype=torch.uint8, device=dev)
    batch = ctx.parsed_batch(parsed, out)
    decode_ms = timed(lambda: ctx.decode(batch))
    nnz = int(parsed["pair_counts"].long().sum())
    status_ok = bool((parsed["status"] == 1).all())
    del parsed, batch
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(stream)
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            gp = ctx.parse_blobs(data, offs, lens)
            ctx.decode(ctx.parsed_batch(gp, out))
    stream.wait_stream(side)
    torch.cuda.synchronize()
    graph_ms = timed(g.replay)
    same = bool(torch.equal(out, direct))
    del direct
    # parse kernel: bound by VALU issue (a serial chain of binary decisions per
    # lane); its instruction count per launch is a property of the kernel and
    # this workload, read from the committed counter summary. Algorithmic bytes
    # = blob bytes read + what it writes per chunk (16 B header, 46 x 4 B step
    # table, 2 x 4 B offsets, 2 x 4 B counts, 4 B status) + 4 B per nonzero level
    prof = os.path.join(ROOT, "profiles", "blob_parse_1024.md")
    valu = read_traffic(prof, "valu_insts_per_launch") if n == 1024 else None
    parse_bytes = total + nch * (16 + 184 + 16 + 4) + 4 * nnz
    parse_gbps = parse_bytes / (parse_ms / 1e3) / 1e9
    valu_rate = None if valu is None else valu / (parse_ms / 1e3) / 1e9
    res = {"chunks": nch, "qstep": qstep, "blob_bytes": total, "bytes_per_blob": total / nch, "nnz": nnz,
           "ratio": nch * 4096 / total, "all_parsed": status_ok, "graph_matches_level_decode": same,
           "encode_ms": enc_ms, "parse_ms": parse_ms, "decode_ms": decode_ms, "graph_ms": graph_ms,
           "blobs_per_s": nch / (graph_ms / 1e3), "compressed_GBps": total / (graph_ms / 1e3) / 1e9,
           "voxel_GBps": nch * 4096 / (graph_ms / 1e3) / 1e9,
           "parse_roofline": {"kernel": "k_parse_blobs", "bound": "valu", "achieved": valu_rate,
                              "peak": VALU_PEAK_G, "unit": "G wave64 instr/s",
                              "frac": None if valu_rate is None else valu_rate / VALU_PEAK_G,
                              "algorithmic_GBps": parse_gbps,
                              "traffic": read_traffic(prof) if n == 1024 else None}}
    k = min(4096, nch)
    idx = torch.linspace(0, nch - 1, k, device=dev).long()
    res["_state"] = {"data": data, "offs": offs, "lens": lens, "idx": idx, "out": out[idx].cpu().numpy()}
    return res


def shard_stage(ctx, args, dev, stream, n, nsheets, seed):
    """The zarr shard path at 1024^3 (one shard): the blob workload packed as
    a shard (blobs, index, crc32c) in HBM, and r3d_hip_shard_decode_region of
    the whole shard timed: index crc, index lookups, parse and the exact
    decode straight into the region. Its output must equal the decode of the
    levels."""
    import numpy as np
    import torch

    from render3d_amd import workloads

    if n != 1024:
        return None
    step, levels, headers, data, offs, lens = make_blobs(ctx, n, nsheets, seed, dev)
    shard = workloads.pack_shard(ctx, data, offs, lens)
    del data
    flat = ctx.decode_chunks(headers=headers, step_tables=step, levels=levels)
    del levels
    want = flat.view(64, 64, 64, 16, 16, 16).permute(0, 3, 1, 4, 2, 5).reshape(n, n, n)
    del flat
    dims, region = (n, n, n), (0, 0, 0, n, n, n)
    dst = torch.empty((n, n, n), dtype=torch.uint8, device=dev)
    counts = torch.zeros(4, dtype=torch.int32, device=dev)
    reps = max(3, args.steps)

    def run():
        ctx.shard_decode_region([(shard, 0, 0, 0)], dims, region, out=dst, counts=counts, status=False)

    run()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    ev[0].record(stream)
    for i in range(reps):
        run()
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))
    same = bool(torch.equal(dst, want))
    del want
    c = (counts.cpu().numpy() // (reps + 1)).tolist()
    nbytes = int(shard.numel())
    res = {"shard_bytes": nbytes, "chunks": 64 ** 3, "ms": ms, "chunks_per_s": 64 ** 3 / (ms / 1e3),
           "compressed_GBps": nbytes / (ms / 1e3) / 1e9, "voxel_GBps": n ** 3 / (ms / 1e3) / 1e9,
           "counts_decoded_missing_corrupt_wide": c, "matches_level_decode": same}
    res["_state"] = {"shard": shard, "dst": dst}
    return res


def stream_stage(ctx, args, dev, stream, n, nsheets, seed, shard):
    """Streaming residency (r3d_hip_bricks_stream) at n^3: a 4^3 = 64-slot
    atlas for the args.levels pyramid (585 bricks at 1024^3), only the
    coarsest level resident at build, then 48 frames of the view-cone collect
    along an orbit of radius 0.3 inside the volume (128-pixel-high view, 60
    degree FOV, gate 0, budget 32): each frame fills the nearest missing
    bricks and evicts the least recently wanted ones. Two warm tiers: the
    decoded pyramid in HBM (fill = a 2 MiB copy per brick), and, with the
    shard workload's 1024^3 level-0 shard, the shard store (fill = the brick's
    512 chunks decoded from the shard into the slot; the coarser levels' stores
    are empty). A filling frame's GPU time is taken by HIP events around the
    call on the stream; its host time (policy + launches) by the host clock."""
    import math

    import numpy as np
    import torch

    import render3d_amd as r3d

    lin = ctx.synth_sheets(n, n, n, nsheets, seed)
    nlev = args.levels
    b = r3d.Bricks(ctx, n, n, n, nlev, 4)
    nent = b.tables()[1] - 64
    res_flags = np.zeros(nent, np.uint8)
    res_flags[-((max(n >> (nlev - 1), 1) + 127) // 128) ** 3:] = 1  # the coarsest level
    ht = math.tan(math.radians(30.0))
    pc = float(np.float32(2.0 * ht / 128.0))
    frames = 48

    def run(source):
        if source == "pyramid":
            b.build(lin, res_flags, stream=stream)
        else:
            b.build_from_shards([[(shard, 0, 0, 0)]] + [[] for _ in range(nlev - 1)], res_flags, stream=stream)
        torch.cuda.synchronize()
        gpu_ms, host_ms, fills, evicts = [], [], 0, 0
        for i in range(frames):
            a = 2.0 * math.pi * i / frames
            eye = [0.5 + 0.3 * math.cos(a), 0.5, 0.5 + 0.3 * math.sin(a)]
            fwd = [-math.sin(a), 0.0, math.cos(a)]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            t0 = time.perf_counter()
            b.stream(eye, fwd, ht, pc, 0.0, 32, stream=stream)
            t1 = time.perf_counter()
            e1.record(stream)
            torch.cuda.synchronize()
            st = b.stream_stats()
            if st["inflight"]:
                gpu_ms.append(e0.elapsed_time(e1))
                host_ms.append((t1 - t0) * 1e3)
                fills += st["inflight"]
                evicts += st["evicted_last"]
        g = float(np.sum(gpu_ms))
        return {"frames": frames, "filling_frames": len(gpu_ms), "fills": fills, "evictions": evicts,
                "gpu_ms_per_filling_frame": float(np.mean(gpu_ms)) if gpu_ms else None,
                "host_ms_per_filling_frame": float(np.mean(host_ms)) if host_ms else None,
                "bricks_per_s": fills / (g / 1e3) if g else None, "_gpu_ms": g}

    res = {"slots": 64, "bricks": nent, "view": "128 px high, 60 deg FOV, orbit radius 0.3 inside the volume"}
    pyr = run("pyramid")
    fill_bytes = 2 * 128 ** 3  # a brick read from its level and written to its slot
    pyr["GBps"] = pyr["fills"] * fill_bytes / (pyr.pop("_gpu_ms") / 1e3) / 1e9
    pyr["roofline"] = {"kernel": "k_bk_fill", "bound": "hbm", "achieved": pyr["GBps"], "peak": HBM_PEAK_GBPS,
                       "unit": "GB/s", "frac": pyr["GBps"] / HBM_PEAK_GBPS,
                       "bytes_per_fill": fill_bytes}
    res["pyramid"] = pyr
    del lin
    if shard is not None and n == 1024:
        sh = run("shards")
        g = sh.pop("_gpu_ms")
        nbytes = int(shard.numel())
        per_brick = nbytes / 512.0 + 128 ** 3  # the brick's share of the compressed shard + the slot written
        sh["GBps"] = sh["fills"] * per_brick / (g / 1e3) / 1e9 if g else None
        sh["roofline"] = {"kernel": "shard region decode per brick", "bound": "valu (decode)",
                          "achieved_GBps_of_algorithmic_bytes": sh["GBps"], "peak_hbm": HBM_PEAK_GBPS,
                          "frac_of_hbm": (sh["GBps"] or 0.0) / HBM_PEAK_GBPS,
                          "bytes_per_fill": per_brick,
                          "note": "each fill also re-reads its shard's 4 MiB index for the crc check"}
        res["shards"] = sh
    b.close()
    return res


def cpu_shard(state, args):
    """The reference reader r3d_shard_decode_region (src/shard/shardio.c, built
    from the reference's sources in oracle/_ref) on a 64 x 1024 x 1024 slab
    of the same shard (16,384 chunks), on the host's cores (its own pthread
    pool, host_cpu()'s threads), repeated for at least half of `cpu_seconds`; its
    output against the GPU's slab. The oracle's restatement stands in when
    the reference build is absent."""
    import ctypes as C
    import tempfile

    import numpy as np

    from oracle import oracle as orc

    shard = state["shard"].cpu().numpy()
    gpu = state["dst"][:64].cpu().numpy()
    cores = host_cpu()["cores"]
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libshard_ref.so")
    region = (0, 0, 0, 64, 1024, 1024)
    out = np.zeros(64 * 1024 * 1024, np.uint8)
    if os.path.exists(ref_path):
        ref = C.CDLL(ref_path)

        class Store(C.Structure):
            _fields_ = [("dir", C.c_char * 512), ("nz", C.c_uint64), ("ny", C.c_uint64), ("nx", C.c_uint64)]

        ref.r3d_shard_store_init.argtypes = [C.POINTER(Store), C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint64]
        ref.r3d_shard_decode_region.restype = C.c_int
        ref.r3d_shard_decode_region.argtypes = ([C.POINTER(Store)] + [C.c_uint64] * 3 + [C.c_uint32] * 3 +
                                                [C.c_void_p, C.c_int])
        tmp = tempfile.mkdtemp(prefix="r3d_shard_")
        path = os.path.join(tmp, "0_0_0.shard")
        shard.tofile(path)
        st = Store()
        ref.r3d_shard_store_init(C.byref(st), tmp.encode(), 1024, 1024, 1024)

        def run():
            return ref.r3d_shard_decode_region(C.byref(st), *region, out.ctypes.data, cores)
        kind = "reference"
    else:
        def run():
            rc, d, _, _ = orc.shard_decode_region([(shard, 0, 0, 0)], (1024, 1024, 1024), region, nthreads=cores)
            out[:] = d.reshape(-1)
            return rc
        kind = "port"
    t = time.perf_counter()
    reps = 0
    while True:
        rc = run()
        reps += 1
        if time.perf_counter() - t >= args.cpu_seconds / 2:
            break
    dt = time.perf_counter() - t
    if kind == "reference":
        os.remove(path)
        os.rmdir(tmp)
    cpu = {"value": reps * 16384 / dt, "unit": "chunks/s", "cores": cores, "kind": kind,
           "voxel_GBps": reps * 64 * 1024 * 1024 / dt / 1e9,
           "sample": f"{reps} x r3d_shard_decode_region of a 64x1024x1024 slab (16,384 chunks), {dt:.1f} s"}
    parity = {"slab_equal": bool(np.array_equal(out.reshape(64, 1024, 1024), gpu)), "rc": int(rc)}
    return cpu, parity


def deblock_stage(ctx, args, dev, stream, vol, qstep):
    """Decode-side deblocking of a decoded 1024^3 region (the shard workload's
    output, encoded at qstep): r3d_hip_deblock_u8 in place with step = qstep,
    timed per call on a fresh copy."""
    import numpy as np
    import torch

    reps = max(3, args.steps)
    work = torch.empty_like(vol)
    ms = []
    for i in range(reps + 1):
        work.copy_(vol)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.deblock(work, qstep, 1.0)
        e1.record(stream)
        torch.cuda.synchronize()
        if i:
            ms.append(e0.elapsed_time(e1))
    t = float(np.mean(ms))
    changed = int((work != vol).sum())
    nz, ny, nx = vol.shape
    samples = nz * ny * ((nx - 2) // 16) + nz * ((ny - 2) // 16) * nx + ((nz - 2) // 16) * ny * nx
    # algorithmic bytes: the x seams fall every 16 bytes, so every cache line
    # of the volume is read once; the writes are the changed voxels
    alg = vol.numel() + changed
    res = {"ms": t, "voxels_per_s": vol.numel() / (t / 1e3), "seam_samples": samples, "voxels_changed": changed,
           "roofline": {"kernel": "k_db_tiles", "bound": "hbm", "achieved": alg / (t / 1e3) / 1e9,
                        "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": alg / (t / 1e3) / 1e9 / HBM_PEAK_GBPS,
                        "traffic": read_traffic(os.path.join(ROOT, "profiles", "deblock_tiles_1024.md"))
                        if vol.shape == (1024, 1024, 1024) else None}}
    sub = vol[:256, :256, :256].contiguous()
    g = sub.clone()
    ctx.deblock(g, qstep, 1.0)
    res["_state"] = {"vol": sub.cpu().numpy(), "gpu": g.cpu().numpy(), "qstep": qstep}
    return res


def cpu_deblock(state, args):
    """The reference dct3d_deblock_u8 (tools/3ddct/deblock.c, built from its
    source in oracle/_ref; single-threaded, as the reference is) on a 256^3
    sub-volume, repeated for at least a quarter of `cpu_seconds`; and the GPU
    filter of the same sub-volume against it (the oracle stands in when the
    reference build is absent)."""
    import ctypes as C

    import numpy as np

    from oracle import oracle as orc

    sub, q = state["vol"], state["qstep"]
    path = os.path.join(ROOT, "oracle", "_ref", "libdeblock_ref.so")
    if os.path.exists(path):
        ref = C.CDLL(path)
        ref.dct3d_deblock_u8.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_float]
        dims = (C.c_size_t * 3)(*sub.shape)

        def run():
            v = sub.copy()
            ref.dct3d_deblock_u8(v.ctypes.data, dims, q, 1.0)
            return v
        kind = "reference"
    else:
        def run():
            return orc.deblock(sub, q, 1.0)
        kind = "port"
    t = time.perf_counter()
    reps = 0
    while True:
        want = run()
        reps += 1
        if time.perf_counter() - t >= args.cpu_seconds / 4:
            break
    dt = time.perf_counter() - t
    cpu = {"value": reps * sub.size / dt, "unit": "voxels/s", "cores": 1, "kind": kind,
           "sample": f"{reps} x dct3d_deblock_u8 of a 256^3 sub-volume, {dt:.1f} s"}
    return cpu, {"subvolume_equal": bool(np.array_equal(state["gpu"], want))}


def bricks_stage(ctx, args, dev, stream, n, nsheets, seed, lut, cam, W, H):
    """Bricks-LOD: the n^3 sheet volume as 128^3 bricks of an args.levels
    pyramid (1024^3, 4 levels: 512 + 64 + 8 + 1 = 585 bricks, all resident in
    9^3 = 729 slots), built on the GPU from the linear volume, then the cube
    workload's 1920x1080 FULL frame rendered through the page table
    (raycast.slang R3D_MODE 3 with the LOD flag). Build timed by the host
    clock (it synchronises), frames by HIP events on the stream."""
    import numpy as np
    import torch

    import render3d_amd as r3d
    from render3d_amd import scene

    lin = ctx.synth_sheets(n, n, n, nsheets, seed)
    nb = sum(-(-max(n >> l, 1) // 128) ** 3 for l in range(args.levels))
    abpa = 1
    while abpa ** 3 < nb:
        abpa += 1
    b = r3d.Bricks(ctx, n, n, n, args.levels, abpa)
    reps = max(3, args.steps)
    build_ms = []
    for i in range(reps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        b.build(lin, stream=stream)
        torch.cuda.synchronize()
        if i:
            build_ms.append((time.perf_counter() - t) * 1e3)
    del lin
    view = scene.frame_params(cam, W, H, step_voxels=1.0, gate=scene.skip_gate(lut), brick_mode=b.brick_mode)
    img = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    counters = torch.zeros(4, dtype=torch.int64, device=dev)
    for _ in range(args.warmup):
        b.raycast([view], W, H, alpha_cutoff=0.99, out=img, stream=stream)
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        b.raycast([view], W, H, alpha_cutoff=0.99, out=img, counters=counters, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    cnt = counters.cpu().numpy().astype(np.uint64) // np.uint64(reps)
    samples, skips, shaded, rays = (int(x) for x in cnt)
    t = float(np.mean(ms))
    # build: level 0 read once, each coarser level written and read once, every slot written
    lv = [max(n >> l, 1) ** 3 for l in range(args.levels)]
    build_bytes = lv[0] + 2 * sum(lv[1:]) + nb * 128 ** 3
    bt = float(np.mean(build_ms))
    res = {"bricks": nb, "slots": abpa ** 3, "atlas_bytes": abpa ** 3 * 128 ** 3,
           "build_ms": bt, "build_GBps": build_bytes / (bt / 1e3) / 1e9,
           "raycast_ms": t, "fps": 1000.0 / t, "samples_per_s": samples / (t / 1e3),
           "samples_per_frame": samples, "skips_per_frame": skips, "shaded_per_frame": shaded,
           "rays_per_frame": rays}
    page = b.read_page()
    atlas = b.read_slots(0, nb)
    res["_state"] = {"page": page, "atlas": atlas, "abpa": abpa, "view": view, "lut": lut, "img": img.cpu().numpy(),
                     "W": W, "H": H}
    b.close()
    return res


def cpu_bricks(state, args):
    """The oracle's bricks-LOD raycaster (r3do_raycast_bricks) on the same
    page table and atlas (read back from HBM) and the same frame, every 4th
    row, repeated for at least a quarter of `cpu_seconds`, on host_cpu()'s
    threads; and the GPU frame's rows against it."""
    import numpy as np

    from oracle import oracle as orc
    from render3d_amd import scene

    cores = host_cpu()["cores"]
    nslots = state["abpa"] ** 3
    atlas = np.zeros(nslots * orc.SLOT_VOX, np.uint8)
    atlas[:state["atlas"].size] = state["atlas"].reshape(-1)
    ob = orc.OracleBricks(state["page"], atlas, state["abpa"])
    W, H, rs = state["W"], state["H"], 4
    tf = scene.tf_unorm(state["lut"])
    t = time.perf_counter()
    reps, samples, rays = 0, 0, 0
    while True:
        ref, cnt, _ = orc.raycast_bricks(ob, tf, [state["view"]], W, H, alpha_cutoff=0.99, row0=0, row_step=rs,
                                         nthreads=cores)
        reps += 1
        samples += cnt["samples"]
        rays += cnt["rays"]
        if time.perf_counter() - t >= args.cpu_seconds / 4:
            break
    dt = time.perf_counter() - t
    rows = np.arange(0, H, rs)
    err = float(np.max(np.abs(state["img"][rows] - ref[rows])))
    cpu = {"value": samples / dt, "unit": "ray samples/s", "cores": cores, "kind": "port",
           "sample": f"{reps} x the {W}x{H} bricks-LOD frame ({rows_desc(rs)}), {rays} rays, {samples} samples, "
                     f"{dt:.1f} s"}
    return cpu, {"raycast_max_abs": err, "raycast_rows_checked": int(rows.size)}


def cpu_blobs(state, args):
    """On a 4,096-blob sample of the same blobs: the oracle's blob decode
    (parse, qstep table, decode, on the threads host_cpu() gives) and the
    reference's own dct3d_decode_u8 (built from its source in oracle/_ref,
    -O2 -ffp-contract=off), timed as tools/3ddct/perf.c:55-60 does (passes over
    the blobs, one thread; then the same passes over the host threads), each
    the median of 5 runs of at least 1 s; the oracle's output against the
    GPU's for that sample."""
    import numpy as np

    from oracle import oracle as orc

    idx = state["idx"].cpu().numpy()
    offs = state["offs"].cpu().numpy()[idx]
    lens = state["lens"].cpu().numpy()[idx]
    data = state["data"].cpu().numpy()
    sub = np.concatenate([data[o:o + ln] for o, ln in zip(offs, lens)])
    sub_off = np.concatenate([[0], np.cumsum(lens.astype(np.int64))[:-1]])
    host = host_cpu()
    cores = host["cores"]
    u8, st = orc.decode_blobs(sub, sub_off, lens, nthreads=cores)
    rate, runs = median_rate(lambda: orc.decode_blobs(sub, sub_off, lens, nthreads=cores), idx.size, runs=5,
                             min_s=1.0)
    cpu = {"value": rate, "unit": "blobs/s", "cores": cores, "cores_total": host["cores_total"],
           "model": host["model"], "kind": "port", "voxel_GBps": rate * 4096 / 1e9, "runs": runs,
           "sample": f"{idx.size} blobs ({int(lens.sum())} bytes), median of 5 runs of >= 1 s"}
    fn = orc.reference_decoder()
    if fn is not None:
        def timed(nthreads):
            def run():
                dt, ok = orc.time_decoder(fn, sub, sub_off, lens, 1, nthreads)
                assert ok == idx.size
                run.t += dt
            run.t = 0.0
            return run
        one, one_runs = median_rate(timed(1), idx.size, runs=5, min_s=1.0)
        allc, all_runs = median_rate(timed(cores), idx.size, runs=5, min_s=1.0)
        cpu["reference"] = {"kind": "reference", "function": "dct3d_decode_u8 (tools/3ddct/dct3d.c, with range decoding)",
                            "build": "-O2 -ffp-contract=off, baseline x86-64 (oracle/Makefile target ref)",
                            "single_core": {"value": one, "unit": "blobs/s", "voxel_MBps": one * 4096 / 1e6,
                                            "ns_per_block": 1e9 / one, "runs": one_runs},
                            "all_cores": {"value": allc, "unit": "blobs/s", "cores": cores,
                                          "voxel_GBps": allc * 4096 / 1e9, "runs": all_runs},
                            "sample": f"passes over {idx.size} blobs of the qstep-{BLOB_QSTEP:g} volume, "
                                      f"median of 5 runs of >= 1 s"}
    else:
        cpu["reference"] = None
    parity = {"blobs_checked": int(idx.size), "u8_equal": bool(np.array_equal(u8, state["out"])),
              "status_ok": bool(np.all(st == 1))}
    return cpu, parity


def cfg0(ctx, args, dev, stream):
    """The 256^3 workload: 12-sheet volume with its full pyramid, 512x512,
    step 0.5, cutoff 0.99, FULL and MIP."""
    import numpy as np
    import torch

    import render3d_amd as r3d
    from render3d_amd import scene

    n = 256
    vol = ctx.volume(n, n, n)
    vol.upload(ctx.synth_sheets(n, n, n, 12, 20261203))
    vol.build()
    lut = scene.tf_ramp_40_200()
    ctx.set_transfer(lut)
    cam = scene.orbit_camera(45.0, 0.0, 2.0)
    res, state = {}, {"vol": vol, "lut": lut, "views": {}, "images": {}}
    reps = max(5, args.steps)
    for name, mode in (("full", r3d.MODE_FULL), ("mip", r3d.MODE_MIP)):
        view = scene.frame_params(cam, 512, 512, mode=mode, step_voxels=0.5, gate=scene.skip_gate(lut))
        img = torch.empty((512, 512, 4), dtype=torch.float32, device=dev)
        counters = torch.zeros(4, dtype=torch.int64, device=dev)
        ctx.raycast(vol, [view], 512, 512, alpha_cutoff=0.99, out=img)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(stream)
        for i in range(reps):
            ctx.raycast(vol, [view], 512, 512, alpha_cutoff=0.99, out=img, counters=counters)
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        ms = float(np.mean([ev[i].elapsed_time(ev[i + 1]) for i in range(reps)]))
        samples = int(counters[0]) // reps
        res[name] = {"ms": ms, "samples": samples, "samples_per_s": samples / (ms / 1e3)}
        state["views"][name] = view
        state["images"][name] = img.cpu().numpy()
    ctx.set_transfer(scene.tf_ramp_40_200())
    res["config"] = "256^3 12-sheet volume, full pyramid, 512x512, step 0.5, cutoff 0.99"
    res["_state"] = state
    return res


def fit_256(ctx, dev, steps=200):
    """Differentiable workload: the 256^3 sheet volume, a 64^3 FP32 density
    grid (init 1.0) and an FP32 TF; 16 views at 512^2 in one 2048^2 image;
    targets rendered with a ground-truth TF of Gaussian alpha bumps at 90 and
    170 (sigma 15); L2 loss over RGBA; Adam on TF and grid from the 40->200
    ramp TF. Reports the loss curve and time per forward + adjoint step."""
    import numpy as np
    import torch

    from render3d_amd import scene
    from render3d_amd.autograd import DiffRenderer

    n = 256
    vol = ctx.volume(n, n, n)
    vol.upload(ctx.synth_sheets(n, n, n, 12, 20261203))
    vol.build()
    views = [scene.frame_params(scene.orbit_camera(22.5 * i, 20.0 * (i % 3 - 1), 2.0), 512, 512, step_voxels=0.5,
                                view_org=((i % 4) * 512, (i // 4) * 512)) for i in range(16)]
    R = DiffRenderer(ctx, vol, views, 2048, 2048, alpha_cutoff=0.99)
    x = np.arange(256, dtype=np.float64)
    gt = np.zeros((256, 4))
    gt[:, 0], gt[:, 1], gt[:, 2] = np.clip(x / 255 * 1.1, 0, 1), np.clip(x / 255 * 0.9, 0, 1), np.clip(x / 255 * 0.7, 0, 1)
    gt[:, 3] = np.clip(0.8 * (np.exp(-0.5 * ((x - 90) / 15) ** 2) + np.exp(-0.5 * ((x - 170) / 15) ** 2)), 0, 0.95)
    ones = torch.ones((64, 64, 64), dtype=torch.float32, device=dev)
    target = R.forward_raw(torch.from_numpy(gt.astype(np.float32)).to(dev), ones).clone()
    tf = torch.from_numpy(scene.tf_unorm(scene.tf_ramp_40_200())).to(dev).requires_grad_(True)
    grid = ones.clone().requires_grad_(True)
    opt = torch.optim.Adam([tf, grid], lr=2e-2)
    curve = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(steps):
        opt.zero_grad()
        loss = ((R(tf, grid) - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            tf.clamp_(0.0, 1.0)
            grid.clamp_(min=0.0)
        if it % 20 == 0 or it == steps - 1:
            curve.append(float(loss))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    err = float((tf.detach() - torch.from_numpy(gt.astype(np.float32)).to(dev)).abs().mean())
    return {"views": 16, "size": [512, 512], "grid": 64, "steps": steps, "loss_first": curve[0],
            "loss_last": curve[-1], "loss_curve_every_20": curve, "ms_per_step": dt / steps * 1e3,
            "tf_mean_abs_err_vs_truth": err}


def cpu_cfg0(state, args):
    """The oracle raycaster on the 256^3 frames (same volume read back from
    HBM), on the host's cores; GPU/oracle max-abs difference."""
    import numpy as np

    from oracle import oracle as orc
    from render3d_amd import scene

    vol = state["vol"]
    cores = host_cpu()["cores"]
    lv = [((d), vol.read_level(l)) for l, d in enumerate(vol.level_dims)]
    ov = orc.OracleVolume(None, 256, 256, 256, levels=lv, occ=(vol.odim, vol.read_occupancy()))
    tf = scene.tf_unorm(state["lut"])
    out = {"cores": cores}
    for name, view in state["views"].items():
        t = time.perf_counter()
        ref, cnt, _ = orc.raycast(ov, tf, [view], 512, 512, alpha_cutoff=0.99, nthreads=cores)
        dt = time.perf_counter() - t
        out[name] = {"samples_per_s": cnt["samples"] / dt, "s": dt,
                     "gpu_max_abs": float(np.max(np.abs(state["images"][name] - ref)))}
    return out


def adjoint_ratio(ctx, vol, lut, args, W, H, dev):
    """The 1024^3 differentiable workload: `adj_views` views at WxH, each
    rendered into its own region of one image (two columns), an adj_grid^3
    density grid and the FP32 TF. Timed (mean of 3 after 2 warm-ups):
      - forward-only: the differentiable model's forward without checkpoints
        (r3d_hip_raycast_diff with no workspace), the same image;
      - the forward an adjoint follows (it writes the checkpoints), and the
        adjoint;
      - the product's plain forward (r3d_hip_raycast, FULL, same views and
        FP32 TF, empty-space skipping on, no grid) for context.
    ratio = (checkpointing forward + adjoint) / forward-only."""
    import numpy as np
    import torch

    from render3d_amd import scene
    from render3d_amd.autograd import DiffRenderer

    nv = args.adj_views
    cols = 2
    rows = (nv + cols - 1) // cols
    IW, IH = W * cols, H * rows
    views = [scene.frame_params(scene.orbit_camera(45.0 + 360.0 * i / nv, 10.0 * (i % 3 - 1), 2.0), W, H,
                                step_voxels=1.0, view_org=((i % cols) * W, (i // cols) * H)) for i in range(nv)]
    R = DiffRenderer(ctx, vol, views, IW, IH, alpha_cutoff=0.99)
    tf = torch.from_numpy(scene.tf_unorm(lut)).to(dev)
    g = args.adj_grid
    grid = torch.ones((g, g, g), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    img = torch.empty((IH, IW, 4), dtype=torch.float32, device=dev)
    img_nock = torch.empty_like(img)
    img_plain = torch.empty_like(img)
    dimg = torch.full((IH, IW, 4), 1e-3, dtype=torch.float32, device=dev)
    dtf = torch.empty_like(tf)
    dgrid = torch.zeros_like(grid)
    for _ in range(2):
        R.forward_raw(tf, grid, out=img_nock, checkpoints=False)
        R.forward_raw(tf, grid, out=img)
        R.adjoint_raw(tf, grid, dimg, dtf, dgrid)
        ctx.raycast(vol, views, IW, IH, alpha_cutoff=0.99, tf=tf, out=img_plain)
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    reps = 3
    e = [torch.cuda.Event(enable_timing=True) for _ in range(5 * reps)]
    for r in range(reps):
        e[5 * r].record(stream)
        R.forward_raw(tf, grid, out=img_nock, checkpoints=False)
        e[5 * r + 1].record(stream)
        R.forward_raw(tf, grid, out=img)
        e[5 * r + 2].record(stream)
        R.adjoint_raw(tf, grid, dimg, dtf, dgrid)
        e[5 * r + 3].record(stream)
        ctx.raycast(vol, views, IW, IH, alpha_cutoff=0.99, tf=tf, out=img_plain)
        e[5 * r + 4].record(stream)
    torch.cuda.synchronize()

    def mean(a):
        return float(np.mean([e[5 * r + a].elapsed_time(e[5 * r + a + 1]) for r in range(reps)]))

    fwd_only, fwd, bwd, plain = mean(0), mean(1), mean(2), mean(3)
    same = bool(torch.equal(img, img_nock))
    return {"views": nv, "width": W, "height": H, "image": [IW, IH], "grid": g,
            "forward_only_ms": fwd_only, "forward_ms": fwd, "adjoint_ms": bwd, "plain_forward_ms": plain,
            "ratio": (fwd + bwd) / fwd_only, "ratio_vs_checkpointing_forward": (fwd + bwd) / fwd,
            "ratio_vs_plain_forward": (fwd + bwd) / plain,
            "forward_images_equal": same, "hbm_used_GB": (total - free0) / 1e9,
            "workspace_GB": R.workspace.numel() / 1e9,
            "grad_finite": bool(torch.isfinite(dtf).all() and torch.isfinite(dgrid).all())}


def cpu_baseline(ctx, vol, view, lut, img, args, levels, headers, step_tab, n, W, H):
    """The oracle's host raycaster on the same frame (same volume, pyramid and
    occupancy read back from HBM) and the oracle decoder on a 4,096-chunk
    sample of the same batch, each as the median of 5 runs of at least 2 s
    (decode 1 s) on the threads host_cpu() gives, and again on one core (the
    raycast on every 8th row); the GPU frame's rows against the oracle's."""
    import numpy as np

    from oracle import oracle as orc
    from render3d_amd import scene

    host = host_cpu()
    cores = host["cores"]
    lv = [((d), vol.read_level(l)) for l, d in enumerate(vol.level_dims)]
    occ = (vol.odim, vol.read_occupancy())
    ov = orc.OracleVolume(None, n, n, n, levels=lv, occ=occ)
    rs = args.cpu_rows
    tf = scene.tf_unorm(lut)
    ref, cnt, _ = orc.raycast(ov, tf, [view], W, H, alpha_cutoff=0.99, row0=0, row_step=rs, nthreads=cores)
    g = img.cpu().numpy()
    rows = np.arange(0, H, rs)
    err = float(np.max(np.abs(g[rows] - ref[rows])))
    ray_all, ray_runs = median_rate(
        lambda: orc.raycast(ov, tf, [view], W, H, alpha_cutoff=0.99, row0=0, row_step=rs, nthreads=cores),
        cnt["samples"], runs=5, min_s=args.cpu_seconds / 5)
    _, c1, _ = orc.raycast(ov, tf, [view], W, H, alpha_cutoff=0.99, row0=0, row_step=8, nthreads=1)
    ray_one, ray_one_runs = median_rate(
        lambda: orc.raycast(ov, tf, [view], W, H, alpha_cutoff=0.99, row0=0, row_step=8, nthreads=1),
        c1["samples"], runs=5, min_s=1.0)
    k = min(4096, levels.shape[0])
    idx = np.linspace(0, levels.shape[0] - 1, k).astype(np.int64)
    lvs = levels[idx].cpu().numpy()
    hd = headers[idx].cpu().numpy()
    st = step_tab.cpu().numpy()
    dec_all, dec_runs = median_rate(lambda: orc.decode_batch(lvs, hd, st, nthreads=cores), k * CHUNK_BYTES,
                                    runs=5, min_s=1.0)
    lv1, hd1 = lvs[:512], hd[:512]
    dec_one, _ = median_rate(lambda: orc.decode_batch(lv1, hd1, st, nthreads=1), 512 * CHUNK_BYTES, runs=5,
                             min_s=0.5)
    cpu = {"value": ray_all, "unit": "ray samples/s", "cores": cores, "cores_total": host["cores_total"],
           "model": host["model"], "cores_note": host["cores_note"], "kind": "port",
           "sample": f"the {W}x{H} frame ({rows_desc(rs)}), {cnt['rays']} rays, {cnt['samples']} samples per run; "
                     f"median of 5 runs of >= {args.cpu_seconds / 5:.1f} s",
           "runs": ray_runs,
           "single_core": {"value": ray_one, "unit": "ray samples/s", "runs": ray_one_runs,
                           "sample": f"every 8th row of the frame, {c1['samples']} samples per run, "
                                     f"median of 5 runs of >= 1 s"},
           "decode_GBps": dec_all / 1e9, "decode_runs_GBps": [r / 1e9 for r in dec_runs],
           "decode_single_core_GBps": dec_one / 1e9,
           "decode_sample": f"{k} chunks of the batch (algorithmic bytes {CHUNK_BYTES} per chunk), median of 5 runs "
                            f"of >= 1 s; single core: 512 chunks"}
    parity = {"raycast_max_abs": err, "raycast_rows_checked": int(rows.size)}
    return cpu, parity


if __name__ == "__main__":
    sys.exit(main())
