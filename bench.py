"""Benchmark: 3D VTI staggered-grid FD on one MI355X.

Workload (the configuration the headline metric is quoted on): 512^3 grid,
dx = 10 m, dt = 0.8 ms, 2000 steps, 8th-order staggered stencil; vp a seeded
smooth random field (Gaussian-filtered noise, sigma 8 cells) over a linear
1800 -> 4200 m/s depth gradient, eps in [0, 0.25] and delta in [0, 0.12]
from smoothed fields; 30-cell absorbing boundary; 15 Hz Ricker at the
surface centre; 256 x 256 receivers (every 2nd cell) at z = 3.

A step is one velocity update plus one stress update of the whole padded
grid (572^3) with the fused source / damping / receiver work.  Inputs are
resident in HBM before the timed region.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM peak; RTM shot time learned vs plain at ≤1% image err"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
BYTES_STRESS = 80              # stress kernel: 12 reads + 8 writes of fp32 per cell
BYTES_VEL = 32                 # velocity kernel: 5 reads + 3 writes
BYTES_STEP = 112


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gaussian_smooth(x, sigma):
    """Separable Gaussian filter (reflect padding) of a (Y, X, Z) cuda tensor,
    as a sum of shifted slices (elementwise kernels only)."""
    import torch
    r = int(4 * sigma + 0.5)
    k = torch.arange(-r, r + 1, dtype=torch.float64)
    k = torch.exp(-0.5 * (k / sigma) ** 2)
    k = (k / k.sum()).tolist()
    for dim in (2, 1, 0):
        n = x.shape[dim]
        lo = x.narrow(dim, 1, r).flip(dim)
        hi = x.narrow(dim, n - 1 - r, r).flip(dim)
        xp = torch.cat([lo, x, hi], dim=dim)
        acc = torch.zeros_like(x)
        for j in range(2 * r + 1):
            acc.add_(xp.narrow(dim, j, n), alpha=k[j])
        del xp
        x = acc
    return x


def make_model(n, seed, device):
    """Synthetic model of the benchmark configuration, on the device, (y, x, z)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)

    def smooth_unit(off):
        g.manual_seed(seed + off)
        a = torch.randn((n, n, n), generator=g, device=device, dtype=torch.float32)
        a = gaussian_smooth(a, 8.0)
        return (a - a.min()) / (a.max() - a.min())

    depth = torch.linspace(1800.0, 4200.0, n, device=device)
    noise = smooth_unit(1) * 2.0 - 1.0
    vp = (depth.view(1, 1, n) + 150.0 * noise).contiguous()
    eps = (0.25 * smooth_unit(2)).contiguous()
    # delta from its own smoothed field, kept <= eps: pseudo-acoustic VTI
    # grows without bound where delta > eps.
    dl = torch.minimum(0.12 * smooth_unit(3), eps).contiguous()
    return vp, eps, dl


def cpu_baseline(vp, eps, dl, n, npml, nt_total, budget_s=12.0):
    """The oracle (C restatement, FP32, OpenMP) on a bounded sample of the
    same workload: the full padded grid for as many steps as fit in budget_s."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    c = n // 2
    p = po.params_gpu3d(n, n, n, 10, 10, 10, npml, 1, 15, 0, c, 0, c, 0, 1, 0, nt_total, 800, 3, 2)
    sim = po.Sim(p, vp, eps, dl, "f32")
    cells = (n + 2 * npml - 8) ** 3
    sim.step(0)          # first touch of the arrays outside the timed loop
    t0 = time.perf_counter()
    steps = 0
    while True:
        sim.step(1 + steps)
        steps += 1
        if time.perf_counter() - t0 > budget_s or steps >= 50:
            break
    dt = time.perf_counter() - t0
    sim.close()
    return {"value": round(cells * steps / dt / 1e9, 4), "unit": "Gcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle/vti_oracle.c FP32 on the full {n}^3 workload ({n + 2 * npml}^3 padded), "
                      f"{steps} steps after the first, {threads} OpenMP threads, {dt:.1f} s"}


def copy_calibration(torch, vti, nbytes=4 << 30, reps=20):
    """Plain copy of nbytes with the library's copy kernel: the achievable
    HBM rate and the known-size reference for FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    src = torch.ones(nbytes // 4, device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    s = torch.cuda.current_stream().cuda_stream
    L = vti.lib()
    L.vti_calib_copy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                     ctypes.c_void_p(s))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        L.vti_calib_copy(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nbytes,
                         ctypes.c_void_p(s))
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    assert torch.equal(dst[:1024], src[:1024])
    print(json.dumps({"copy_bytes": nbytes, "ms": round(ms, 4),
                      "read_plus_write_GBs": round(2 * nbytes / ms / 1e6, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--npml", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--copy-calib", action="store_true",
                    help="time the plain device copy kernel (bandwidth / counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(shots would shard across GPUs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    from geophysics_amd import vti

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.copy_calib:
        copy_calibration(torch, vti)
        return
    n, npml = args.size, args.npml
    nt = max(args.steps, 2000)
    log(f"[bench] building {n}^3 model on {torch.cuda.get_device_name(0)}")
    vp, eps, dl = make_model(n, args.seed, dev)
    torch.cuda.synchronize()

    cfg = vti.cfg_ref3d(n, n, n, 10, 10, 10, npml, 1, 15, nt, 800)
    prop = vti.Propagator(cfg)
    prop.set_model(vp.data_ptr(), eps.data_ptr(), dl.data_ptr(), on_device=True)
    c = n // 2
    prop.set_shot(vti.shot_ref3d(cfg, 0, c, 0, c, 0, 1, 0, rec_depth=3, rec_step=2))
    log(f"[bench] receivers {prop.nrec}, warmup {args.warmup} steps")

    prop.reset()
    prop.step(0, args.warmup)
    prop.sync()
    prop.reset()
    prop.sync()
    torch.cuda.synchronize()

    prop.prof_enable(True)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    chunk = 250
    done = 0
    while done < args.steps:
        k = min(chunk, args.steps - done)
        prop.step(done, k)
        done += k
    prop.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    vel_ms, str_ms, nl = prop.prof_read()
    prop.prof_enable(False)
    rec_max = float(np.abs(prop.record()[:, : args.steps]).max())
    log(f"[bench] done: {t1 - t0:.3f} s, max |record| {rec_max:.4g}")

    cells = (n + 2 * npml - 8) ** 3
    secs = t1 - t0
    gcells = cells * args.steps / secs / 1e9
    str_avg = str_ms / max(nl, 1) / 1e3
    vel_avg = vel_ms / max(nl, 1) / 1e3
    achieved = BYTES_STRESS * cells / str_avg / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "stress3d_traffic.json")
    if os.path.exists(tfile) and n == 512:
        try:
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"3D VTI staggered-grid FD forward, {n}^3 grid, 8th-order, {npml}-cell "
                        "absorbing boundary, dt 0.8 ms, 15 Hz Ricker at surface centre, "
                        f"{prop.nrec} receivers at z=3",
            "grid": [n, n, n], "padded": [n + 2 * npml] * 3, "steps_per_shot": 2000,
            "cells_updated_per_step": cells, "bytes_per_cell_step_algorithmic": BYTES_STEP,
            "pct_hbm_peak_step": round(100.0 * BYTES_STEP * gcells / HBM_PEAK_GBS, 2),
            "vel_kernel_ms": round(vel_avg * 1e3, 4), "stress_kernel_ms": round(str_avg * 1e3, 4),
        },
        "roofline": {
            "bound": "hbm", "kernel": "vti::s3d::stress3d<64, true>",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        },
    }
    prop.close()
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) ...")
        h = [t.cpu().numpy() for t in (vp, eps, dl)]
        del vp, eps, dl
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(h[0], h[1], h[2], n, npml, nt)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
