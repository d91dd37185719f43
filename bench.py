"""Benchmark of the distance-driven projector family on one MI355X.

Workload (the headline metric): fan-beam equal-distance geometry, 512 x 512
image, 736 views over 2*pi x 736 detector bins (dImg 0.0066, dDet 0.0072,
s2r 5.95, d2r 4.906), batch 32 fp32 ellipse phantoms.  One step = one
forward + adjoint iteration of the least-squares data term:
    y = A x,  g = A^T (A x - m)
timed on three paths:
    plain     ctlib_amd.projection + ctlib_amd.projection_t (per-call API)
    autograd  ctlib_amd.autograd.least_squares(x, m).backward()  (batched,
              differentiable; A x and A^T r from one fused library call)
    graph     the fused step captured once in a HIP graph and replayed
`value` is the autograd path's iterations/s.

roofline: the dominant DD kernel's VALU issue rate: its VALU wave-instructions per
launch (rocprofv3 counters, profiles/dd_issue_counters.json) over its mean launch
time from HIP events the library records on the launch stream during the timed
region, against 256 CUs x 2.4 GHz x 1 wave-instruction per CU-cycle.  Its HBM
traffic (profiles/hbm_traffic.json) and the sweep-model equivalent are beside it.
cpu_baseline: the FP64 C++ oracle (oracle/) on the host cores running the same
step slice by slice (bounded at 30 s, scaled to the batch); its gradients also
check the GPU's (oracle_check).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def calib_copy():
    """One 256 MiB device-to-device copy: the reference dispatch for scaling HBM counters."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


# MI355X: 256 CUs at 2.4 GHz; one VALU wave-instruction per cycle per CU (four SIMDs, four
# cycles per wave64 instruction each).
CUS = 256
CLOCK_GHZ = 2.4
VALU_PEAK_GINST = CUS * CLOCK_GHZ


def _profile(name):
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (profiles/hbm_traffic.json), scaled by the copy-kernel calibration; None if absent."""
    data = _profile("hbm_traffic.json")
    try:
        k = data["kernels"][kernel]
        return float(k["fetch_kib"]) * 1024 * data["scale"]["fetch"] + float(k["write_kib"]) * 1024 * data["scale"]["write"]
    except (KeyError, ValueError, TypeError):
        return None


def issue_counters(kernel):
    """Per-launch SQ counters of `kernel` at the metric workload from the committed rocprofv3
    --pmc summary (profiles/dd_issue_counters.json); None if absent.  The instruction counts of
    these branch-free kernels depend only on the geometry and batch, not on the data."""
    data = _profile("dd_issue_counters.json")
    try:
        return data["kernels"][kernel]
    except (KeyError, TypeError):
        return None


def workload_options(views=736):
    from ctlib_amd import phantom
    return phantom.fan_options(views, 736, 512, 512, 0.0066, 0.0072, 5.95, 4.906, 0, dAng=2 * math.pi / views)


def host_cpu():
    """(nproc, CPU model) of the host the baseline runs on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count(), model


def cpu_baseline(x0, m, grad, budget_s=30.0):
    """FP64 oracle on the host cores: the same step, y = A x0 and g = A^T (y - m), slice by slice
    on the workload's own fp32 inputs, until every slice is done or `budget_s` has passed; the rate
    is scaled to the full batch.  The oracle's gradients are also the checker of the GPU's `grad`
    (relative L2 per slice, the largest reported)."""
    from oracle import oracle as O
    o = workload_options().astype(np.float64)
    B = x0.shape[0]
    threads = O.max_threads()
    nproc, model = host_cpu()
    xs = x0[:, 0].double().cpu().numpy()
    ms = m[:, 0].double().cpu().numpy()
    gs = grad[:, 0].double().cpu().numpy()
    t_total, done, worst = 0.0, 0, 0.0
    for s in range(B):
        t0 = time.perf_counter()
        y = O.apply("projection", o, xs[s:s + 1])
        g = O.apply("projection_t", o, y - ms[s:s + 1])
        t_total += time.perf_counter() - t0
        worst = max(worst, float(np.linalg.norm(gs[s] - g[0]) / np.linalg.norm(g[0])))
        done += 1
        if t_total > budget_s:
            break
    per_step = t_total / done * B
    return {"value": 1.0 / per_step, "unit": "iters/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model,
            "sample": f"FP64 oracle y = A x, g = A^T(y - m) on {done} of {B} slices x 736 views "
                      f"({t_total:.2f} s wall on {threads} OpenMP threads), scaled to {B} slices"}, \
        {"slices_checked": done, "max_rel_l2_grad_vs_fp64_oracle": worst}


def sweep_bytes(g, es=4):
    """Sweep-model bytes of one DD operator application on one slice: es * V * (H*W + D)."""
    return es * g["views"] * (g["height"] * g["width"] + g["dets"])


def workload_hetero(args):
    """64 heterogeneous slices: one persistent graph-replayed forward+adjoint launch vs 64 plain calls."""
    import torch

    import ctlib_amd
    from ctlib_amd import batch, phantom, tune

    dev = torch.device("cuda:0")
    opts, imgs = phantom.hetero_slices()
    ot = [torch.from_numpy(o) for o in opts]
    xs = [torch.from_numpy(im)[None, None].to(dev) for im in imgs]
    ms = [ctlib_amd.projection(x * 1.1, o) for x, o in zip(xs, ot)]
    geoms = [ctlib_amd.parse_options(o) for o in ot]
    total_bytes = sum(2 * sweep_bytes(g) for g in geoms)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / steps

    def plain_pairs():
        for x, m, o in zip(xs, ms, ot):
            ctlib_amd.projection_t(ctlib_amd.projection(x, o) - m, o)

    def plain_fused():
        for x, m, o in zip(xs, ms, ot):
            ctlib_amd.residual_step(x, m, o)

    hb = batch.HeteroBatch(ot, device=dev, graph=True)
    sched, tune_ms = tune.tune_batch(hb, reps=3)
    hb.run(xs, ms)
    hb.check()
    t_batch = timed(hb.launch, args.steps, args.warmup)
    hb.check()
    t_pairs = timed(plain_pairs, args.steps, args.warmup)
    t_fused = timed(plain_fused, args.steps, args.warmup)
    return {
        "metric": "heterogeneous batched forward+adjoint, 64 slices: batch iters/s vs 64 plain calls",
        "value": round(1000.0 / t_batch, 3), "unit": "iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_batch, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 slices: size {256,384,512}^2, views U[96,1152], dets {512,736,1024}, "
                               "scan_type {0,1,2} (default_rng(3999)); y_s = A_s x_s, g_s = A_s^T(A_s x_s - m_s)",
                   "items": hb.info()["items"], "schedule": {"blocks_per_cu": sched[0], "order": sched[1]}},
        "plain_pairs_ms": round(t_pairs, 4), "plain_fused_ms": round(t_fused, 4),
        "speedup_vs_plain_pairs": round(t_pairs / t_batch, 3),
        "sweep_pct_hbm": round(total_bytes / (t_batch * 1e-3) / 1e9 / HBM_PEAK_GBS * 100, 2),
        "tuning_ms": {f"{k[0]},{k[1]}": round(v, 4) for k, v in tune_ms.items()},
    }


def workload_ops(args):
    """Every public operator on the reference's fan-ED configuration (512^2, 736 x 736, B=8; fp32 and
    fp64 data) and on fan-EA (512^2, 1024 x 888, B=16): call time, sweep-model rate of its DD kernel
    (library events; element size 4 or 8 bytes), and the FP64 oracle on the host for the same
    operator on a bounded sample of the views."""
    import torch

    import ctlib_amd
    from ctlib_amd import _C, phantom
    from oracle import oracle as O

    dev = torch.device("cuda:0")
    names = ("projection", "projection_t", "backprojection", "backprojection_t", "fbp")
    img_in = ("projection", "backprojection_t")
    out = {"metric": "per-operator calls/s and sweep-model %HBM", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "dtype": "f32 (and f64 where marked)", "data": "synthetic", "configs": {}}
    for cfg, B, dt in ((1, 8, torch.float32), (2, 16, torch.float32), (1, 8, torch.float64)):
        o = phantom.CONFIGS[cfg]["options"]
        ot = torch.from_numpy(o).to(dt)
        g = ctlib_amd.parse_options(ot)
        x = torch.from_numpy(phantom.phantom_batch(B, g["height"], g["width"], config=cfg, scale=0.02)).to(dev, dt)
        y = ctlib_amd.projection(x, ot)
        es = 4.0 if dt == torch.float32 else 8.0
        sweep = es * B * g["views"] * (g["height"] * g["width"] + g["dets"])
        rows = {}
        for name in names:
            f, inp = getattr(ctlib_amd, name), (x if name in img_in else y)
            for _ in range(args.warmup):
                f(inp, ot)
            torch.cuda.synchronize()
            _C.timing_collect()
            _C.timing_enable(True)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                f(inp, ot)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            _C.timing_enable(False)
            st = _C.timing_collect()
            k = "fwd" if name in img_in else "adj"
            k_ms = st[k]["ms"] / max(st[k]["launches"], 1)
            row = {"ms_per_call": round(ms, 4), "calls_per_s": round(1e3 / ms, 2),
                   "dd_kernel_ms": round(k_ms, 4), "dd_kernel_frac_8TBs": round(sweep / (k_ms * 1e-3) / 8e12, 4)}
            if name == "fbp":
                r_ms = st["ramp"]["ms"] / max(st["ramp"]["launches"], 1)
                row["ramp_ms"] = round(r_ms, 4)
                # flops of the dense Toeplitz product 2*B*V*D^2; the kernel skips the zero taps
                # (even offsets), so the MFMA work it executes is half of that
                dense = st["ramp"]["flops"] / max(st["ramp"]["launches"], 1)
                row["ramp_dense_equiv_frac_fp64_mfma"] = round(dense / (r_ms * 1e-3) / 78.6e12, 4)
                row["ramp_executed_frac_fp64_mfma"] = round(0.5 * dense / (r_ms * 1e-3) / 78.6e12, 4)
            if not args.no_cpu and dt == torch.float32:
                # FP64 oracle, same operator, one slice over a view subset, scaled to the call
                vs = max(8, g["views"] // 16)
                o_sub = o.astype(np.float64).copy()
                o_sub[0] = vs
                xin = (x[:1, 0] if name in img_in else y[:1, 0, :vs]).double().cpu().numpy()
                t0 = time.perf_counter()
                O.apply(name, o_sub, xin)
                t_cpu = time.perf_counter() - t0
                row["cpu_oracle_calls_per_s"] = round(1.0 / (t_cpu * (g["views"] / vs) * B), 5)
                row["cpu_sample"] = f"1 slice x {vs} views, {O.max_threads()} threads, scaled"
            rows[name] = row
        tag = f"{'fan-ED' if cfg == 1 else 'fan-EA'} {g['height']}^2 {g['views']}x{g['dets']} B={B}"
        out["configs"][tag + ("" if dt == torch.float32 else " fp64")] = rows
    return out


def workload_small(args):
    """Small batches: every DD operator at batch 1 (and 4 on fan-ED 512^2) on the parallel 256^2,
    fan-ED 512^2 and fan-ED 1024^2 geometries, with the launches split automatically (line chunks
    for line->bin, view chunks for bin->pixel) and unsplit (1 chunk), fp32."""
    import torch

    import ctlib_amd
    from ctlib_amd import _C, phantom

    dev = torch.device("cuda:0")
    names = ("projection", "projection_t", "backprojection", "backprojection_t", "residual_step")
    img_in = ("projection", "backprojection_t")
    out = {"metric": "small-batch ms per call, split (automatic chunks) vs unsplit", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "dtype": "f32", "data": "synthetic", "configs": {}}
    for cfg, B in ((0, 1), (1, 1), (1, 4), (4, 1)):
        o = phantom.CONFIGS[cfg]["options"]
        ot = torch.from_numpy(o)
        g = ctlib_amd.parse_options(ot)
        x = torch.from_numpy(phantom.phantom_batch(B, g["height"], g["width"], config=cfg, scale=0.02)).to(dev)
        y = ctlib_amd.projection(x, ot)
        rows = {}
        for name in names:
            row = {}
            for label, chunks in (("split", 0), ("unsplit", 1)):
                _C.set_plan_chunks(ot, dev.index or 0, chunks, chunks)
                if name == "residual_step":
                    fn = lambda: ctlib_amd.residual_step(x, y, ot)  # noqa: E731
                else:
                    f, inp = getattr(ctlib_amd, name), (x if name in img_in else y)
                    fn = lambda: f(inp, ot)  # noqa: E731
                for _ in range(args.warmup):
                    fn()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    fn()
                torch.cuda.synchronize()
                row[label + "_ms"] = round((time.perf_counter() - t0) * 1e3 / args.steps, 4)
            _C.set_plan_chunks(ot, dev.index or 0, 0, 0)
            row["speedup"] = round(row["unsplit_ms"] / row["split_ms"], 2)
            rows[name] = row
        kind = {0: "parallel", 1: "fan-ED", 4: "fan-ED"}[cfg]
        out["configs"][f"{kind} {g['height']}^2 {g['views']}x{g['dets']} B={B}"] = rows
    return out


def workload_laplacian(args):
    """laplacian (kNN graph weights), fp32: N=16384 points of dimension 49 (7x7 patches), k=16 (the
    register path), and N=16384, d=256, k=128 (the general path: distance rows in chunks, radix
    select).  Compute-bound: 2*N^2*d flops of distances (VALU FMA chains; no N x N matrix is held
    for the first, 128 MiB chunks of rows for the second).  Peak: fp32 FMA without packed math
    (packed FP32 is disabled for all device code), 256 CUs x 64 lanes x 2 flops x 2.4 GHz."""
    import torch

    import ctlib_amd
    dev = torch.device("cuda:0")
    peak = 78.6
    lines = []
    for N, d, k in ((16384, 49, 16), (16384, 256, 128)):
        x = torch.rand(N, d, generator=torch.Generator().manual_seed(7)).to(dev)
        for _ in range(args.warmup):
            ctlib_amd.laplacian(x, k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            W = ctlib_amd.laplacian(x, k)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        flops = 2.0 * N * N * d
        lines.append({"N": N, "d": d, "k": k, "ms_per_call": round(ms, 4), "calls_per_s": round(1e3 / ms, 3),
                      "nnz": int(W._nnz()), "tflops": round(flops / (ms * 1e-3) / 1e12, 2),
                      "frac_fp32_fma_peak": round(flops / (ms * 1e-3) / (peak * 1e12), 4)})
    first = lines[0]
    out = {
        "metric": "laplacian calls/s, N=16384 points, d=49, k=16 -> sparse (N, N)",
        "value": first["calls_per_s"], "unit": "calls/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": first["ms_per_call"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "kNN graph weights of 16384 random points in [0,1]^49, k=16", "nnz": first["nnz"]},
        "roofline": {"bound": "valu", "achieved": first["tflops"], "peak": peak, "unit": "TFLOP/s",
                     "frac": first["frac_fp32_fma_peak"], "traffic": None,
                     "note": "distance flops 2*N^2*d over the whole call; peak = fp32 FMA, not packed"},
        "cases": lines,
    }
    if not args.no_cpu:
        # FP64 numpy restatement: distances and selection for 256 query rows, scaled to N rows
        from oracle import laplacian as L  # noqa: F401  (the restatement this sample follows)
        N = 16384
        xd = torch.rand(N, 49, generator=torch.Generator().manual_seed(7)).double().numpy()
        t0 = time.perf_counter()
        n2 = (xd * xd).sum(1)
        q = xd[:256]
        dd = n2[:256, None] + n2[None, :] - 2.0 * (q @ xd.T)
        idx = np.arange(N)
        for i in range(256):
            np.lexsort((idx, dd[i]))[:16]
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(1.0 / (dt * N / 256), 5), "unit": "calls/s", "cores": 1,
                               "kind": "port", "sample": f"numpy FP64 restatement, 256 of {N} rows ({dt:.2f} s), scaled"}
    return out


def workload_sv(args):
    """backprojection_sv (per-view backprojection) on fan-ED 512^2, 736 views x 736 bins, B=8.

    Output B*V*H*W fp32 = 6.17 GB per call: a store-bound kernel.  roofline.achieved is its
    algorithmic bytes 4*B*V*(H*W + D) (images written + sinogram rows read) over the mean
    launch time of the per-view kernel (HIP events on its stream)."""
    import torch

    import ctlib_amd
    from ctlib_amd import _C, phantom

    dev = torch.device("cuda:0")
    B = 8
    o = workload_options()
    ot = torch.from_numpy(o)
    x = torch.from_numpy(phantom.phantom_batch(B, 512, 512, config=1, scale=0.02)).to(dev)
    y = ctlib_amd.projection(x, ot)
    for _ in range(args.warmup):
        out = ctlib_amd.backprojection_sv(y, ot)
    del out
    torch.cuda.synchronize()
    _C.timing_collect()
    _C.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = ctlib_amd.backprojection_sv(y, ot)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / args.steps
    _C.timing_enable(False)
    st = _C.timing_collect()["sv"]
    k_ms = st["ms"] / max(st["launches"], 1)
    k_bytes = st["algo_bytes"] / max(st["launches"], 1)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    return {
        "metric": "backprojection_sv calls/s, fan-ED 512^2 736x736 B=8 -> [8,736,512,512]",
        "value": round(1000.0 / ms, 3), "unit": "calls/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "fan-ED 512x512, 736 views x 736 dets, B=8: per-view backprojection",
                   "global_batch": B, "output_bytes": B * 736 * 512 * 512 * 4},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic("k_bprj_sv"),
                     "kernel": "k_bprj_sv", "kernel_ms": round(k_ms, 4), "algo_bytes_per_launch": k_bytes},
    }


def workload_unroll(args):
    """50-iteration unrolled gradient descent with learnable step sizes, per-iteration checkpointing."""
    import torch

    import ctlib_amd
    from ctlib_amd import phantom, recon
    from oracle import oracle as O

    dev = torch.device("cuda:0")
    cfg = phantom.CONFIGS[4]
    o = cfg["options"]
    ot = torch.from_numpy(o)
    B, K = cfg["batch"], 50
    x_true = phantom.phantom_batch(B, 1024, 1024, config=4, scale=0.02)
    clean = ctlib_amd.projection(torch.from_numpy(x_true).to(dev), ot).double().cpu().numpy()
    y = np.stack([phantom.poisson_sinogram(clean[b, 0], 5e4, seed=1_000_000 + 4 + b) for b in range(B)])[:, None]
    yt = torch.from_numpy(y).to(dev)
    step = recon.lipschitz_step(ot, (1, 1, 1024, 1024), dev)
    model = recon.UnrolledGD(ot, iters=K, alpha0=step, checkpoint=True).to(dev)
    xt_true = torch.from_numpy(x_true).to(dev)
    # training step: forward through K iterations + backward (recomputes each iteration's projections)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    t0 = time.perf_counter()
    steps = max(1, args.steps // 10)
    for _ in range(steps):
        model.zero_grad()
        loss = ((model(yt) - xt_true) ** 2).mean()
        loss.backward()
    torch.cuda.synchronize()
    t_train = (time.perf_counter() - t0) / steps
    peak = torch.cuda.max_memory_allocated() - base
    with torch.no_grad():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xr = model(yt)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
    out = {
        "metric": "unrolled GD (50 iters, checkpointed) on fan-ED 1024^2, 1440 x 1536, B=4",
        "value": round(K / t_fwd, 3), "unit": "GD iters/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(t_train * 1000, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "x <- x - alpha_k A^T(A x - y), K=50, learnable alpha_k, Poisson I0=5e4",
                   "global_batch": B},
        "train_steps_per_s": round(1.0 / t_train, 4), "peak_activation_mb": round(peak / 1e6, 1),
    }
    if not args.no_cpu:
        # FP64 oracle run of the same 50 iterations on slice 0
        o64 = o.astype(np.float64)
        xo = np.zeros((1, 1024, 1024))
        y0 = y[0:1, 0].astype(np.float64)
        t0 = time.perf_counter()
        for _ in range(K):
            xo = xo - step * O.apply("projection_t", o64, O.apply("projection", o64, xo) - y0)
        t_cpu = time.perf_counter() - t0
        g = xr[0, 0].double().cpu().numpy()
        out["final_rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(g - xo[0]) / np.linalg.norm(xo[0]))
        out["cpu_baseline"] = {"value": round(K / t_cpu / B, 5), "unit": "GD iters/s", "cores": O.max_threads(),
                               "kind": "port", "sample": f"FP64 oracle, 50 iterations on 1 of {B} slices "
                                                         f"({t_cpu:.1f} s), scaled to the batch"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--paths", default="plain,autograd,graph")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for HBM counter scaling)")
    ap.add_argument("--workload", default="metric", choices=("metric", "hetero", "unroll", "sv", "ops", "small",
                                                              "laplacian"),
                    help="metric: the headline fan-ED B=32 step (default); hetero: 64-slice batch; "
                         "unroll: 50-iteration checkpointed reconstruction; sv: backprojection_sv; "
                         "ops: every public operator; small: batch 1 / 4, split vs unsplit launches; "
                         "laplacian: kNN graph weights")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the projector would shard over slices "
                                   "(no exchange) or views (one image all-reduce per adjoint), but no multi-GPU "
                                   "path is built or measured", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.calib_copy:
        calib_copy()
    if args.workload == "hetero":
        print(json.dumps(workload_hetero(args)), flush=True)
        return
    if args.workload == "unroll":
        print(json.dumps(workload_unroll(args)), flush=True)
        return
    if args.workload == "sv":
        print(json.dumps(workload_sv(args)), flush=True)
        return
    if args.workload == "ops":
        print(json.dumps(workload_ops(args)), flush=True)
        return
    if args.workload == "small":
        print(json.dumps(workload_small(args)), flush=True)
        return
    if args.workload == "laplacian":
        print(json.dumps(workload_laplacian(args)), flush=True)
        return

    import torch

    import ctlib_amd
    from ctlib_amd import _C, autograd, phantom

    dev = torch.device("cuda:0")
    B = args.batch
    opts = torch.from_numpy(workload_options())
    x_true = torch.from_numpy(phantom.phantom_batch(B, 512, 512, config=1, scale=0.02)).to(dev)
    m = ctlib_amd.projection(x_true, opts)
    x0 = (0.9 * x_true).contiguous()

    def step_plain():
        y = ctlib_amd.projection(x0, opts)
        return ctlib_amd.projection_t(y - m, opts)

    x_ag = x0.clone().requires_grad_(True)

    def step_autograd():
        x_ag.grad = None
        loss = autograd.least_squares(x_ag, m, opts)
        loss.backward()
        return x_ag.grad

    def step_fused():
        return ctlib_amd.residual_step(x0, m, opts)[1]

    def timed(fn, steps, warmup, collect=False):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        if collect:
            _C.timing_collect()
            _C.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        stats = None
        if collect:
            _C.timing_enable(False)
            stats = _C.timing_collect()
        return (t1 - t0) * 1000.0 / steps, stats

    results = {}
    paths = args.paths.split(",")
    # correctness guard: the three paths compute the same gradient
    g_plain = step_plain()
    g_fused = step_fused()
    g_ag = step_autograd()
    torch.cuda.synchronize()
    ref = torch.linalg.vector_norm(g_plain.double())
    agree = max(float(torch.linalg.vector_norm((g_fused - g_plain).double()) / ref),
                float(torch.linalg.vector_norm((g_ag - g_plain).double()) / ref))

    stats = None
    plain_stats = None
    if "plain" in paths:
        results["plain"], plain_stats = timed(step_plain, args.steps, args.warmup, collect=True)
    results["autograd"], stats = timed(step_autograd, args.steps, args.warmup, collect=True)
    if "graph" in paths:
        step_fused()
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(2):
                step_fused()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            g_graph = step_fused()
        graph.replay()
        torch.cuda.synchronize()
        agree = max(agree, float(torch.linalg.vector_norm((g_graph - g_plain).double()) / ref))
        results["graph"], _ = timed(graph.replay, args.steps, args.warmup)

    ms = results["autograd"]
    # dominant DD kernel from the library's on-stream events
    k = max(("fwd", "adj"), key=lambda c: stats[c]["ms"])
    st = stats[k]
    per_launch_ms = st["ms"] / max(st["launches"], 1)
    per_launch_bytes = st["algo_bytes"] / max(st["launches"], 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    kname = "k_fwd_lt" if k == "fwd" else "k_adj"
    traffic = measured_traffic(kname) if B == 32 else None
    ctr = issue_counters(kname) if B == 32 else None
    # The DD kernels reuse image lines and bin tables from LDS, so what binds them is on chip:
    # the VALU issue rate (instructions counted by rocprofv3, time measured here) and the LDS.
    # The HBM figures are reported beside it.
    roof = {"kernel": "k_fwd_lt (line->bin over prebuilt line tables)" if k == "fwd" else "k_adj (bin->pixel)",
            "kernel_ms": round(per_launch_ms, 4)}
    if ctr:
        valu = ctr["SQ_INSTS_VALU"] / (per_launch_ms * 1e-3) / 1e9
        roof.update({"bound": "valu", "achieved": round(valu, 1), "peak": VALU_PEAK_GINST,
                     "unit": "G VALU wave-instructions/s", "frac": round(valu / VALU_PEAK_GINST, 4),
                     "traffic": traffic,
                     "counters_source": "profiles/dd_issue_counters.json (rocprofv3 --pmc, per launch)",
                     "valu_insts_per_launch": ctr["SQ_INSTS_VALU"],
                     "valu_busy_frac_profiled": ctr.get("valu_busy_frac"),
                     # the LDS pipe, busy for this fraction of the kernel's cycles in the profiled run
                     # (16-byte table reads take 4-8 cycles each, so its instruction rate is no measure)
                     "co_bound": {"unit": "lds", "busy_frac_profiled": ctr.get("lds_busy_frac"),
                                  "bank_conflict_frac_profiled": ctr.get("lds_bank_conflict_frac")}})
    else:
        roof.update({"bound": "valu", "achieved": None, "peak": VALU_PEAK_GINST,
                     "unit": "G VALU wave-instructions/s", "frac": None, "traffic": traffic})
    roof.update({
        "traffic_source": "profiles/hbm_traffic.json (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a 256 MiB copy)",
        "measured_hbm_gbs": round(traffic / (per_launch_ms * 1e-3) / 1e9, 1) if traffic else None,
        "measured_hbm_frac": round(traffic / (per_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        # sweep model: every view streams the whole image once plus its sinogram row
        "sweep_bytes_per_launch": per_launch_bytes,
        "sweep_equiv_frac_8TBs": round(achieved / HBM_PEAK_GBS, 4),
        # compulsory bytes of one operator application: the image and the sinogram once each
        "compulsory_bytes_per_launch": 4 * B * (512 * 512 + 736 * 736)})
    def per_step(st):
        # library kernels per step of a path: launches and milliseconds by class (HIP events)
        return {c: {"launches_per_step": round(st[c]["launches"] / args.steps, 2),
                    "ms_per_step": round(st[c]["ms"] / args.steps, 4)} for c in ("fwd", "adj", "aux")}
    kernels = {"autograd": per_step(stats)}
    if plain_stats:
        kernels["plain"] = per_step(plain_stats)
    step_bytes = stats["fwd"]["algo_bytes"] / max(stats["fwd"]["launches"], 1) + \
        stats["adj"]["algo_bytes"] / max(stats["adj"]["launches"], 1)

    out = {
        "metric": "DD fwd+adjoint iters/s, 512^2 fan-ED B=32, batched-autograd vs plain; %HBM peak",
        "value": round(1000.0 / ms, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"fan-ED 512x512, 736 views x 736 dets, B={B}: y = A x, g = A^T(A x - m)",
                   "global_batch": B, "views": 736, "dets": 736, "image": [512, 512], "scan_type": 0},
        "paths_iters_per_s": {p: round(1000.0 / t, 3) for p, t in results.items()},
        "step_sweep_equiv_pct_8TBs": round(step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS * 100, 2),
        "grad_agreement_rel_l2": agree,
        "kernels": kernels,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"], out["oracle_check"] = cpu_baseline(x0, m, g_ag)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
