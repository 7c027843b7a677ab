"""Benchmark: 3D VTI staggered-grid FD on one MI355X.

Workload (the configuration the headline metric is quoted on): 512^3 grid,
dx = 10 m, dt = 0.8 ms, 2000 steps, 8th-order staggered stencil; vp a seeded
smooth random field (Gaussian-filtered noise, sigma 8 cells) over a linear
1800 -> 4200 m/s depth gradient, eps in [0, 0.25] and delta in [0, 0.12]
from smoothed fields; 30-cell absorbing boundary; 15 Hz Ricker at the
surface centre; 256 x 256 receivers (every 2nd cell) at z = 3.

A step is one velocity update plus one stress update of the whole padded
grid (572^3) with the fused source / damping / receiver work.  Inputs are
resident in HBM before the timed region.  Prints one JSON line, which also
carries:

  * cpu_baseline: the oracle (the C restatement, built for the host with
    -O3 -march=native -ffp-contract=off) on a bounded sample of the same
    workload, FP32 and FP64, all host threads and one, with the records of
    the sampled steps compared with the GPU's;
  * config1_2d: the 2D 400 x 400 / 2000-step configuration of the
    reference's CPU path on the GPU and on the host (oracle, and the
    reference CPU path itself where oracle/_ref holds it), with the errors
    between them;
  * learned_rtm: the second half of the metric, 3D RTM per-shot time with
    learned snapshot compression against plain checkpoint recomputation at
    equal memory (384^3, 4000 steps, 3 fitting shots, two shots imaged both
    ways), with the final-image error, and the effective-boundary RTM (band
    reversal) of the same shots.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu] [--no-learned]
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
BYTES_STEP_2D = 84             # 2D step: velocity 16 B read + 8 written, stress 36 read + 24 written


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


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu prints it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))


def native_oracle():
    """The oracle (the checker's C restatement) built for this host with
    -O3 -march=native -ffp-contract=off (bit-identical float results to the
    default build), loaded in place of the default build."""
    import subprocess
    odir = os.path.join(ROOT, "oracle")
    lib = os.path.join(odir, "_native", "liboracle_vti.so")
    subprocess.run(["make", "-s", "-B", "-C", odir, "native"], check=True)
    os.environ["ORC_LIB"] = lib
    from oracle import pyoracle as po
    if po.library_path() != lib:
        raise RuntimeError("the oracle was loaded before the native build")
    return po


CPU_FLAGS = "gcc -O3 -march=native -ffp-contract=off -fopenmp (oracle/Makefile native)"


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def cpu_legs_3d(po, vp, eps, dl, n, npml, nt_total, gpu_rec, budget_s=8.0):
    """The oracle on a bounded sample of the 3D workload, split form (the
    reference arithmetic), FP32 and FP64, all host threads and one thread:
    the full padded grid for as many steps as fit in budget_s (all threads;
    one step on one thread), after one untimed step that first-touches the
    arrays.  The records of the sampled steps are compared with the GPU
    run's."""
    allc = host_threads()
    c = n // 2
    p = po.params_gpu3d(n, n, n, 10, 10, 10, npml, 1, 15, 0, c, 0, c, 0, 1, 0, nt_total, 800, 3, 2)
    cells = (n + 2 * npml - 8) ** 3
    legs, recs = [], {}
    for prec in ("f32", "f64"):
        sim = po.Sim(p, vp, eps, dl, prec)
        po.set_threads(allc)
        rows = [sim.step(0)]
        t0 = time.perf_counter()
        while True:
            rows.append(sim.step(len(rows)))
            if time.perf_counter() - t0 > budget_s or len(rows) > 40:
                break
        dt_all = time.perf_counter() - t0
        k_all = len(rows) - 1
        legs.append({"precision": prec, "threads": allc, "steps": k_all, "seconds": round(dt_all, 2),
                     "gcells": round(cells * k_all / dt_all / 1e9, 4)})
        po.set_threads(1)
        t0 = time.perf_counter()
        rows.append(sim.step(len(rows)))
        dt_one = time.perf_counter() - t0
        legs.append({"precision": prec, "threads": 1, "steps": 1, "seconds": round(dt_one, 2),
                     "gcells": round(cells / dt_one / 1e9, 4)})
        po.set_threads(allc)
        recs[prec] = np.stack(rows, axis=1)
        sim.close()
        log(f"[bench] CPU 3D {prec}: {legs[-2]['gcells']} Gcell/s on {allc} threads, "
            f"{legs[-1]['gcells']} on 1")
    k = min(gpu_rec.shape[1], recs["f32"].shape[1], recs["f64"].shape[1])
    errs = {"steps_compared": k,
            "gpu_vs_cpu_f32": rel(gpu_rec[:, :k], recs["f32"][:, :k]),
            "gpu_vs_cpu_f64": rel(gpu_rec[:, :k], recs["f64"][:, :k]),
            "cpu_f32_vs_cpu_f64": rel(recs["f32"][:, :k], recs["f64"][:, :k])}
    return legs, errs


def layered_1d(n, nlayer, seed, lo, hi):
    rng = np.random.default_rng(seed)
    vals = np.sort(rng.uniform(lo, hi, nlayer))
    return vals[np.minimum(np.arange(n) * nlayer // n, nlayer - 1)]


def bench_config1(torch, vti, po):
    """The 2D configuration the reference's CPU path runs: 400 x 400, dx 10 m,
    dt 1 ms, 2000 steps, 20-cell boundary, 20 Hz Ricker at (x=200, z=2),
    receivers at every x at z=2, 4 flat layers (vp 1500..4000 m/s increasing
    with depth, eps in [0, 0.2], delta in [0, 0.1] kept <= eps), with the
    CPU driver's semantics (CPU_vti2dfd_kernels.c: 7x7 Gaussian source,
    P+P record) on the GPU, the oracle on the host (FP32, FP64; all threads
    and one) and, where it was built, the reference CPU path itself."""
    import subprocess
    import tempfile
    nx = nz = 400
    npd, nt, favg = 20, 2000, 20
    vz = layered_1d(nz, 4, 11, 1500.0, 4000.0)
    ez = layered_1d(nz, 4, 12, 0.0, 0.2)
    dz = np.minimum(layered_1d(nz, 4, 13, 0.0, 0.1), ez)
    vp = np.tile(vz, (nx, 1)).astype(np.float32)
    e = np.tile(ez, (nx, 1)).astype(np.float32)
    d = np.tile(dz, (nx, 1)).astype(np.float32)
    cfg = vti.cfg_ref2d(vti.MODE_CPU_REF, nx, nz, 10, 10, npd, 0, favg, nt, 1000)
    cells = (nx + 2 * npd - 8) * (nz + 2 * npd - 8)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_shot(vti.shot_ref2d(cfg, 1, 200, 0, 2, rec_depth=2))
        prop.run()                       # warm-up
        prop.sync()
        prop.reset()
        prop.sync()
        t0 = time.perf_counter()
        prop.step(0, nt)
        prop.sync()
        t_gpu = time.perf_counter() - t0
        rec = prop.record()
    allc = host_threads()
    p = po.params_cpu2d(nx, nz, 10, 10, npd, 0, favg, 1, 200, 0, 2, nt, 1000, 2)
    legs, recs = [], {}
    for prec in ("f32", "f64"):
        po.set_threads(allc)
        sim = po.Sim(p, vp, e, d, prec)
        t0 = time.perf_counter()
        recs[prec] = sim.run()
        t_all = time.perf_counter() - t0
        sim.close()
        legs.append({"precision": prec, "threads": allc, "steps": nt, "seconds": round(t_all, 3),
                     "gcells": round(cells * nt / t_all / 1e9, 4)})
        po.set_threads(1)
        k = 300
        sim = po.Sim(p, vp, e, d, prec)
        t0 = time.perf_counter()
        for it in range(k):
            sim.step(it, rec=False)
        t_one = time.perf_counter() - t0
        sim.close()
        legs.append({"precision": prec, "threads": 1, "steps": k, "seconds": round(t_one, 3),
                     "gcells": round(cells * k / t_one / 1e9, 4)})
    po.set_threads(allc)
    out = {"workload": "2D VTI FD, 400x400, dx 10 m, dt 1 ms, 2000 steps, 20-cell boundary, 20 Hz Ricker at "
                       "(200, 2), 400 receivers at z=2, 4 flat layers (delta clipped to <= eps), CPU driver "
                       "semantics (CPU_vti2dfd_kernels.c)",
           "cells_per_step": cells,
           "gpu": {"seconds": round(t_gpu, 4), "ms_per_step": round(t_gpu / nt * 1e3, 4),
                   "gcells": round(cells * nt / t_gpu / 1e9, 3)},
           "cpu_oracle": legs, "cpu_flags": CPU_FLAGS,
           "rel_l2_gpu_vs_cpu_f32": rel(rec, recs["f32"]), "gpu_equals_cpu_f32_bitwise": bool(np.array_equal(rec, recs["f32"])),
           "rel_l2_gpu_vs_cpu_f64": rel(rec, recs["f64"]), "rel_l2_cpu_f32_vs_f64": rel(recs["f32"], recs["f64"])}
    ref = os.path.join(ROOT, "oracle", "_ref", "cpu_vti2dfd_ref")
    if os.path.exists(ref):
        # the reference's own CPU path (CPU_vti2dfd_kernels.c, compiled from
        # its source by oracle/Makefile, -O2): a bounded sample of the steps
        k = 300
        with tempfile.TemporaryDirectory() as td:
            for name, a in (("v", vp), ("e", e), ("d", d)):
                a.tofile(os.path.join(td, name))
            cmd = [ref, str(nx), str(nz), "10", "10", str(npd), "0"] + [os.path.join(td, x) for x in "ved"] + \
                [str(favg), "1", "200", "0", "2", os.path.join(td, "shot"), os.path.join(td, "snap"), str(k), "1000"]
            t0 = time.perf_counter()
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
            t_ref = time.perf_counter() - t0
            shot = np.fromfile(os.path.join(td, "shot"), dtype=np.float32).reshape(nx, k)
        # the reference records row npd (z = 1); the GPU run above records
        # z = 2, so compare a GPU run of the reference geometry
        with vti.Propagator(cfg, vp, e, d) as prop:
            prop.set_shot(vti.shot_ref2d(cfg, 1, 200, 0, 2, rec_depth=1))
            prop.reset()
            prop.step(0, k)
            rec1 = prop.record()[:, :k]
        out["reference_cpu_path"] = {
            "steps": k, "seconds": round(t_ref, 3), "gcells": round(cells * k / t_ref / 1e9, 5), "threads": 1,
            "build": "gcc -O2 -ffp-contract=off of 2D-3D-vti-fd-cuda-gtk-GUI/CPU_vti2dfd_kernels.c (oracle/Makefile ref)",
            "gpu_equals_reference_bitwise": bool(np.array_equal(rec1, shot)), "rel_l2_gpu_vs_reference": rel(rec1, shot)}
    log(f"[bench] config 1: GPU {out['gpu']['gcells']} Gcell/s; CPU {legs[0]['gcells']} Gcell/s on {allc} threads")
    return out


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


def layered_fault_model(nx, nz, rng):
    """Flat layers (vp increasing with depth, 1500..3500 m/s; eps in
    [0, 0.2], delta in [0, min(eps, 0.1)]) cut by one dipping normal fault
    with a throw of 5..25 cells.  (nx, nz) arrays, z fastest."""
    nl = int(rng.integers(4, 8))
    tops = np.sort(rng.uniform(0.08, 0.95, nl - 1)) * nz
    vp_l = np.sort(rng.uniform(1500.0, 3500.0, nl))
    ep_l = rng.uniform(0.0, 0.2, nl)
    dl_l = np.minimum(rng.uniform(0.0, 0.1, nl), ep_l)
    z = np.arange(nz, dtype=np.float64)[None, :]
    x = np.arange(nx, dtype=np.float64)[:, None]
    x0 = rng.uniform(0.3, 0.7) * nx
    dip = np.tan(np.radians(rng.uniform(50.0, 75.0)))
    throw = rng.uniform(5.0, 25.0)
    zz = np.where(x > x0 + (z / dip), z - throw, z)          # hanging wall shifted up
    lay = np.searchsorted(tops, np.clip(zz, 0, nz - 1))
    f = lambda v: np.ascontiguousarray(v[lay].astype(np.float32))
    return f(vp_l), f(ep_l), f(dl_l)


def bench_batch_rtm(args, torch, vti):
    """Batched 2D VTI RTM: independent shots of uneven grids in one
    persistent launch per phase (nx in [300, 1200], nz in [200, 800],
    dx 10 m, dt 1 ms, 20-cell boundary, Ricker 10-25 Hz, every surface cell
    a receiver), source field from binomial checkpoints, receiver
    back-propagation, zero-lag imaging with source illumination."""
    nshot = args.shots or 256
    nt = args.nt or 3000
    rng = np.random.default_rng(args.seed)
    cfgs, shots, models = [], [], []
    for k in range(nshot):
        nx, nz = int(rng.integers(300, 1201)), int(rng.integers(200, 801))
        favg = int(rng.integers(10, 26))          # the reference drivers take an integer frequency
        cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 10, 10, 20, 1, favg, nt, 1000)
        cfgs.append(cfg)
        shots.append(vti.shot_ref2d(cfg, 1, nx // 2, 0, 2, rec_depth=2))
        models.append(layered_fault_model(nx, nz, rng))
    log(f"[bench] batch RTM: {nshot} shots, nt {nt}")
    B = vti.Batch(cfgs, shots, models)
    B.set_unsplit(not args.split)
    shape, ms_tune = B.autotune(steps=min(10, nt))
    cfgd = B.get_config()
    log(f"[bench] autotuned tile {cfgd['tile']} grid_mult {cfgd['grid_mult']}: {ms_tune:.3f} ms/step")
    o = vti.rtm_defaults(2)
    o.nsnap = 32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fwd = B.rtm(o, None)
    t1 = time.perf_counter()
    secs = t1 - t0
    cells = B.cells
    updates = cells * (fwd + nt)           # forward (incl. recomputation) + receiver steps
    out = {
        "metric": "batched 2D VTI RTM throughput (Gcell-updates/s over forward, recomputation and receiver steps)",
        "value": round(updates / secs / 1e9, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": nt,
        "warmup": 0, "ms_per_step": round(secs / nt * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nshot} independent 2D shots, nx in [300,1200], nz in [200,800], dx 10 m, "
                               f"nt {nt}, dt 1 ms, Ricker 10-25 Hz, layered models with a dipping fault",
                   "shots": nshot, "cells_per_step": cells, "forward_steps": fwd, "checkpoint_slots": o.nsnap,
                   "stress_update": "split" if args.split else "unsplit interior (both sides; the receiver "
                                                               "injection row split)",
                   "tile": list(cfgd["tile"]), "grid_mult": cfgd["grid_mult"], "seconds": round(secs, 3),
                   "shots_per_s": round(nshot / secs, 3)},
        # one forward step of every shot (the velocity and stress work-list
        # launches), its time from the autotuner's HIP events on the batch's
        # stream; bytes of the reference structure's 2D step (velocity: P, Q,
        # u, w read, u, w written; stress: u, w, vp, eps, delta and the four
        # split fields read, the split fields and P, Q written)
        "roofline": {"bound": "hbm", "kernel": "forward step of all shots (vti_batch velocity + stress launches)",
                     "achieved": round(BYTES_STEP_2D * cells / (ms_tune / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(BYTES_STEP_2D * cells / (ms_tune / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": None, "bytes_per_cell_step": BYTES_STEP_2D, "step_ms": round(ms_tune, 4)},
        "cpu_baseline": None,
    }
    B.close()
    print(json.dumps(out), flush=True)


def bench_fwi_grad(args, torch, vti):
    """2D VTI FWI gradient (FP32 hand-written adjoint, 40 checkpoint slots):
    1000 x 300 grid, dx 12.5 m, 4000 steps, dt 1 ms, shots 15 cells apart;
    true model layered with two Gaussian anomalies (+-300 m/s, radius 20
    cells), starting model the true one smoothed with sigma 15."""
    from scipy.ndimage import gaussian_filter
    nshot = args.shots or 64
    nt = args.nt or 4000
    nx, nz, dx = 1000, 300, 12.5
    rng = np.random.default_rng(args.seed)
    vp, ep, dl = layered_fault_model(nx, nz, rng)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    for sgn, (cx, cz) in ((1.0, (0.35 * nx, 0.5 * nz)), (-1.0, (0.65 * nx, 0.6 * nz))):
        vp = vp + (sgn * 300.0 * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * 20.0 ** 2))).astype(np.float32)
    v0 = gaussian_filter(vp.astype(np.float64), 15.0, mode="nearest")
    cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 12, 12, 20, 0, 10, nt, 1000)
    cfg.dx = cfg.dz = dx                           # the reference helpers take integer spacings
    nstream = max(1, args.streams)
    log(f"[bench] FWI gradient: {nshot} shots, nt {nt}, grid {nx}x{nz}, {nstream} concurrent contexts")
    shots = [20 + 15 * k for k in range(nshot) if 20 + 15 * k < nx - 20]
    done = len(shots)
    # one gradient context (own HIP stream) per worker thread; the library
    # call releases the interpreter lock, so the shots' kernels overlap
    import threading
    lock = threading.Lock()
    acc = {"total": np.zeros((nx, nz)), "misfit": 0.0, "t_obs": 0.0, "t_grad": 0.0, "steps": 0}

    def worker(w):
        with vti.Gradient(cfg, fp64=False) as G:
            for k in range(w, done, nstream):
                shot = vti.shot_ref2d(cfg, 1, shots[k], 0, 2, rec_depth=2)
                G.set_model(vp, ep, dl)
                G.set_shot(shot)
                t0 = time.perf_counter()
                dobs = G.forward()
                t1 = time.perf_counter()
                G.set_model(v0, ep, dl)
                r = G.run(dobs=dobs, nsnap=40)
                t2 = time.perf_counter()
                with lock:
                    acc["t_obs"] += t1 - t0
                    acc["t_grad"] += t2 - t1
                    acc["total"] += r["grad_vp"]
                    acc["misfit"] += r["misfit"]
                    acc["steps"] = r["fwd_steps"]
                log(f"[bench] shot {k}: misfit {r['misfit']:.4e}, {t2 - t1:.2f} s")

    t_wall0 = time.perf_counter()
    threads = [threading.Thread(target=worker, args=(w,)) for w in range(nstream)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    t_wall = time.perf_counter() - t_wall0
    misfit, steps = acc["misfit"], acc["steps"]
    t_obs, t_grad = acc["t_obs"], acc["t_grad"]
    cells = (nx + 40 - 8) * (nz + 40 - 8)
    cpu = None
    if not args.no_cpu:
        # the FP64 oracle on the first shot, a bounded number of steps; the
        # GPU FP32 gradient of the same case for the error statement
        from oracle import pyoracle as po
        nt_c = min(nt, 300)
        cfg_c = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 12, 12, 20, 0, 10, nt_c, 1000)
        cfg_c.dx = cfg_c.dz = dx
        shot_c = vti.shot_ref2d(cfg_c, 1, 20, 0, 2, rec_depth=2)
        p = po.params_gpu2d(nx, nz, 12, 12, 20, 0, 10, 1, 20, 0, 2, nt_c, 1000, 2)
        p.dx = p.dz = dx
        # a random adjoint source: in 300 steps no reflection reaches the
        # surface, so a data residual would be zero
        with vti.Gradient(cfg_c, fp64=False) as G32:
            G32.set_model(v0, ep, dl)
            G32.set_shot(shot_c)
            resid_c = np.random.default_rng(args.seed).standard_normal((G32.nrec, nt_c)) * 1e-3
            r32 = G32.run(resid=resid_c, nsnap=40)
        t0 = time.perf_counter()
        J64, g64, _, _ = po.fwi_grad(po.Sim(p, v0, ep, dl, "f64"), resid=resid_c)
        t_cpu = time.perf_counter() - t0
        err = float(np.linalg.norm(r32["grad_vp"] - g64) / np.linalg.norm(g64))
        cpu = {"value": round(t_cpu * nt / nt_c, 3), "unit": "s/shot", "cores": 1, "kind": "port",
               "sample": f"oracle/vti_oracle.c FP64 gradient (every forward state kept, one thread) of the first "
                         f"shot for {nt_c} steps ({t_cpu:.1f} s), scaled to {nt} steps; GPU FP32 gradient of the "
                         f"same {nt_c}-step case (random adjoint source) against it: rel-L2 {err:.2e}"}
        log(f"[bench] CPU FP64 oracle {t_cpu:.2f} s for {nt_c} steps; GPU FP32 vs FP64 rel-L2 {err:.2e}")
    out = {
        "metric": "2D VTI FWI gradient time per shot (s: wall time / shots), FP32 adjoint with 40 checkpoint slots",
        "value": round(t_wall / done, 4), "unit": "s/shot", "n_gpus": 1, "steps": nt, "warmup": 0,
        "ms_per_step": round(t_wall / done / nt * 1e3, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nx}x{nz} grid, dx {dx} m, {nt} steps, dt 1 ms, {done} shots 15 cells apart, "
                               "layered model with two Gaussian anomalies, start = smoothed (sigma 15)",
                   "shots": done, "forward_steps_per_shot": steps, "misfit": misfit,
                   "observed_data_s_per_shot": round(t_obs / done, 4),
                   "concurrent_contexts": nstream, "wall_s": round(t_wall, 3),
                   "gradient_s_per_shot_one_context": round(t_grad / done, 4),
                   "cells_per_step": cells, "grad_norm": float(np.linalg.norm(acc["total"]))},
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def learned_rtm(args, torch, vti, nshot, nt, n, rows_out=None):
    """3D RTM with learned snapshot compression against plain checkpoint
    recomputation: an n^3 grid, dx 10 m, nt steps, dt 0.8 ms, smoothed
    random vp / eps / delta as the headline workload; snapshots every 4
    steps, compressed with the block basis fitted online on the first 3
    shots' snapshots, applied to the following shots; per-shot wall time and
    image error against the plain path imaging the same steps.  Equal
    memory: the plain path gets as many checkpoint slots as fit in the
    device bytes the learned call held for the stored source field (store
    as allocated, source boxes, work buffers; stats alloc_bytes).  codec
    "sparse": error-bounded sparse coefficients (every snapshot within
    relative L2 error --tol); codec "quant": the same bound with every
    coefficient quantized per block; codec "latent": the autoencoder's 64-channel
    latents with the exact-storage fallback (--exact-frac).  Beside them
    (unless --no-eb) the effective-boundary RTM of the same shot: the source
    field rebuilt backward in time from 4-cell interior bands saved every
    step, timed and compared with the plain image.  Returns the
    result record; each imaged shot is logged (and appended to rows_out)
    as it finishes."""
    npml = args.npml
    sparse = args.codec in ("sparse", "quant")
    log(f"[bench] learned RTM ({args.codec} codec): {n}^3, nt {nt}, {nshot} shots")
    vp, eps, dl = make_model(n, args.seed, torch.device("cuda:0"))
    torch.cuda.synchronize()
    cfg = vti.cfg_ref3d(n, n, n, 10, 10, 10, npml, 1, 15, nt, 800)
    prop = vti.Propagator(cfg)
    prop.set_unsplit(not args.split)      # source-side propagation of both paths
    prop.set_model(vp.data_ptr(), eps.data_ptr(), dl.data_ptr(), on_device=True)
    del vp, eps, dl
    torch.cuda.empty_cache()
    rng = np.random.default_rng(args.seed)
    pos = [(int(rng.integers(n // 4, 3 * n // 4)), int(rng.integers(n // 4, 3 * n // 4))) for _ in range(nshot)]
    A = vti.Autoencoder(crop=64, batch=4, seed=args.seed)
    nsave = (nt + 3) // 4
    max_exact = max(1, int(nsave * args.exact_frac))   # exact-storage fallback budget (latent codec)
    state_bytes = prop.checkpoint_bytes()              # one checkpoint slot of the plain path
    fp32_bytes = nsave * n ** 3 * 4
    rows = []
    t_fit = 0.0
    ntrain = min(3, nshot)
    explained = None
    for k, (cx, cy) in enumerate(pos):
        prop.set_shot(vti.shot_ref3d(cfg, 0, cx, 0, cy, 0, 1, 0, rec_depth=3, rec_step=2))
        if k < ntrain:
            t0 = time.perf_counter()
            prop.ae_fit_shot(A, save_every=4, nblocks=2048, src_box=args.src_box)
            if k == ntrain - 1:
                explained = A.fit_end()
            t_fit += time.perf_counter() - t0
            log(f"[bench] fit shot {k} done ({time.perf_counter() - t0:.1f} s)")
            continue
        lo = vti.learned_defaults()
        lo.src_box = args.src_box
        if sparse:
            lo.codec, lo.tol, lo.floor = (2 if args.codec == "quant" else 1), args.tol, args.floor
            lo.budget = int(fp32_bytes * args.store_frac)
            lo.serial_codec = {"beside": 0, "between": 1, "decode-beside": 2}[args.codec_stream]
        else:
            lo.check_every, lo.tol, lo.ck_every, lo.max_exact = 1, 0.02, 0, max_exact
        ro = prop.rtm_defaults()
        ro.inject_z = npml - 1            # receiver row of Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:334-335
        prop.rtm_reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = prop.rtm_shot_learned(A, ro, lo, None)
        t_learn = time.perf_counter() - t0
        img_l, _ = prop.rtm_image(finalize=False)
        fin_l, _ = prop.rtm_image(finalize=True)
        store, held = st["store_bytes"], st["alloc_bytes"]
        nsnap_equal = max(1, int(held // state_bytes))
        o = prop.rtm_defaults()
        o.image_every = 4
        o.nsnap = nsnap_equal
        o.inject_z = npml - 1
        prop.rtm_reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fsteps = prop.rtm_shot(o)
        prop.sync()
        t_plain = time.perf_counter() - t0
        img_p, _ = prop.rtm_image(finalize=False)
        fin_p, _ = prop.rtm_image(finalize=True)
        err_raw = rel(img_l, img_p)
        err = rel(fin_l, fin_p)
        row = dict(shot=k, t_plain=round(t_plain, 3), t_learn=round(t_learn, 3), speedup=round(t_plain / t_learn, 3),
                   err=err, err_raw=err_raw, fsteps=fsteps, slots=nsnap_equal, store=store, held=held,
                   enc_ms=st["encode_ms"], dec_ms=st["decode_ms"], snap_err=st["mean_checked_err"],
                   exact=st["exact"], recomputed=st["recomputed"])
        if not args.no_eb:   # the third source-state option: reversal from saved boundary bands
            prop.rtm_reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            se = prop.rtm_shot_eb(o)
            prop.sync()
            t_eb = time.perf_counter() - t0
            fin_e, _ = prop.rtm_image(finalize=True)
            row.update(t_eb=round(t_eb, 3), err_eb=rel(fin_e, fin_p), band=se["band_bytes"])
        rows.append(row)
        log("[bench] learned RTM row " + json.dumps(row))
        if rows_out:
            with open(rows_out, "a") as f:
                f.write(json.dumps(row) + "\n")
    prop.close()
    A.close()

    def mean(key):
        v = [r[key] for r in rows]
        return float(np.mean(v)) if v else None

    tp, tl = mean("t_plain"), mean("t_learn")
    return {
        "metric": "3D RTM shot time, learned snapshot compression vs plain checkpoint recomputation at equal memory",
        "speedup": round(tp / tl, 3) if rows else None, "unit": "plain s / learned s",
        "workload": f"3D VTI RTM {n}^3, {nt} steps, dt 0.8 ms, 15 Hz, {ntrain} fitting shots then "
                    f"{len(rows)} shot(s) imaged both ways",
        "codec": args.codec, "block_tol": args.tol if sparse else None, "block_floor": args.floor if sparse else None,
        "codec_stream": args.codec_stream if sparse else None,
        "stress_update": "split" if args.split else "reduced split (both sides; the receiver injection row split)",
        "src_box": args.src_box,
        "plain_s_per_shot": round(tp, 3) if rows else None,
        "learned_s_per_shot": round(tl, 3) if rows else None,
        "final_image_rel_l2_max": max((r["err"] for r in rows), default=None),
        "final_image_rel_l2_mean": mean("err"), "raw_image_rel_l2_mean": mean("err_raw"),
        "snapshot_rel_err_mean": mean("snap_err"),
        "plain_checkpoint_slots": [r["slots"] for r in rows],
        "plain_forward_steps": [r["fsteps"] for r in rows],
        "store_GB_mean": round(mean("store") / 1e9, 3) if rows else None,
        "held_GB_mean": round(mean("held") / 1e9, 3) if rows else None,
        "compression_vs_fp32_snapshots": round(fp32_bytes / mean("store"), 3) if rows else None,
        "fit_s": round(t_fit, 2), "fit_energy_captured_64": explained, "snapshots": nsave,
        "exact_snapshots_mean": mean("exact"), "recomputed_snapshots_mean": mean("recomputed"),
        "effective_boundary": {
            "s_per_shot": round(mean("t_eb"), 3), "speedup_vs_plain": round(tp / mean("t_eb"), 3),
            "final_image_rel_l2_max": max(r["err_eb"] for r in rows), "band_GB": round(mean("band") / 1e9, 3),
            "note": "source field rebuilt backward in time from 4-cell interior bands saved every step "
                    "(vti_rtm_shot_eb): no recomputation, the bands' memory instead of equal memory"}
        if rows and "t_eb" in rows[0] else None,
        "encode_gpu_ms_per_shot": round(mean("enc_ms"), 1) if rows else None,
        "decode_gpu_ms_per_shot": round(mean("dec_ms"), 1) if rows else None,
        "rows": rows,
    }


def bench_rtm3d_learned(args, torch, vti):
    """The learned-vs-plain RTM workload on its own line (--workload
    rtm3d-learned): 384^3, 4000 steps, 19 shots (3 fitting, 16 imaged)."""
    n = args.size if args.size != 512 else 384
    r = learned_rtm(args, torch, vti, args.shots or 19, args.nt or 4000, n, rows_out=args.rows_out)
    out = {
        "metric": r["metric"], "value": r["speedup"], "unit": "speedup (plain s / learned s)", "n_gpus": 1,
        "steps": args.nt or 4000, "warmup": 0,
        "ms_per_step": round(r["learned_s_per_shot"] / (args.nt or 4000) * 1e3, 4) if r["rows"] else None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
        "config": {k: v for k, v in r.items() if k not in ("metric", "speedup", "unit")},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


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
    ap.add_argument("--workload", default="fd3d", choices=["fd3d", "batch-rtm", "fwi-grad", "rtm3d-learned"],
                    help="fd3d: the headline 3D FD line (default); the others time the batched 2D RTM, "
                         "the 2D FWI gradient and the 3D RTM with learned snapshot compression")
    ap.add_argument("--streams", type=int, default=4,
                    help="FWI gradient: shots run concurrently, one context and HIP stream each")
    ap.add_argument("--shots", type=int, default=0, help="shots of the batch / FWI / learned-RTM workloads "
                    "(0: the workload's own count)")
    ap.add_argument("--nt", type=int, default=0, help="time steps of the non-default workloads (0: their own)")
    ap.add_argument("--fused", action="store_true",
                    help="fd3d: one fused kernel per step instead of the velocity and stress kernels")
    ap.add_argument("--split", action="store_true",
                    help="the reference's split stress update everywhere (bit-identical to the oracle) "
                         "instead of the unsplit interior / reduced split")
    ap.add_argument("--exact-frac", type=float, default=0.25,
                    help="learned RTM, latent codec: fraction of snapshots the exact-storage fallback may keep")
    ap.add_argument("--codec", default="quant", choices=["sparse", "quant", "latent"],
                    help="learned RTM: error-bounded block coefficients (sparse: the largest at 15 significant "
                         "bits; quant: all quantized with a per-block step) or autoencoder latents")
    ap.add_argument("--tol", type=float, default=0.0007,
                    help="learned RTM, sparse codec: relative L2 error bound of every stored block")
    ap.add_argument("--src-box", type=int, default=32,
                    help="learned RTM: half-width of the exactly kept box around the source")
    ap.add_argument("--store-frac", type=float, default=0.9,
                    help="learned RTM, sparse codec: most the entry store may take, as a fraction of the fp32 "
                         "snapshots (it is allocated in 4 GB chunks as needed)")
    ap.add_argument("--floor", type=float, default=1e-6,
                    help="learned RTM, sparse codec: blocks below this fraction of the strongest block's "
                         "energy are held to tol of that fraction")
    ap.add_argument("--rows-out", default=None,
                    help="learned RTM: append each imaged shot's result to this file as it finishes")
    ap.add_argument("--no-learned", action="store_true",
                    help="fd3d: leave out the learned-vs-plain 3D RTM record (384^3, 4000 steps, one imaged shot)")
    ap.add_argument("--learned-shots", type=int, default=5,
                    help="fd3d: shots of the learned-vs-plain record (3 fitting, the rest imaged both ways)")
    ap.add_argument("--no-eb", action="store_true",
                    help="learned RTM: leave out the effective-boundary (band reversal) leg")
    ap.add_argument("--codec-stream", default="beside", choices=["beside", "between", "decode-beside"],
                    help="learned RTM: the codec on its own stream beside the steps, on the shot's stream between "
                         "them, or encodes between and decodes beside (the same images)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path "
              "(independent shots would shard across GPUs, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import torch
    from geophysics_amd import vti

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    if args.copy_calib:
        copy_calibration(torch, vti)
        return
    if args.workload == "batch-rtm":
        return bench_batch_rtm(args, torch, vti)
    if args.workload == "fwi-grad":
        return bench_fwi_grad(args, torch, vti)
    if args.workload == "rtm3d-learned":
        return bench_rtm3d_learned(args, torch, vti)
    n, npml = args.size, args.npml
    nt = max(args.steps, 2000)
    log(f"[bench] building {n}^3 model on {torch.cuda.get_device_name(0)}")
    vp, eps, dl = make_model(n, args.seed, dev)
    torch.cuda.synchronize()

    cfg = vti.cfg_ref3d(n, n, n, 10, 10, 10, npml, 1, 15, nt, 800)
    prop = vti.Propagator(cfg)
    prop.set_fused(args.fused)
    prop.set_unsplit(not args.split)
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
    fused = args.fused and args.split
    unsplit = not args.split
    if fused:
        # one kernel per step: the step's algorithmic bytes over its duration
        kernel = "vti::f3d::fused3d<32, true>"
        achieved = BYTES_STEP * cells / str_avg / 1e9
        tfile = os.path.join(ROOT, "profiles", "fused3d_traffic.json")
    elif unsplit:
        kernel = "vti::s3d::stress3d<16, true, true>"
        achieved = BYTES_STRESS * cells / str_avg / 1e9
        tfile = os.path.join(ROOT, "profiles", "stress3d_unsplit_traffic.json")
    else:
        kernel = "vti::s3d::stress3d<16, true>"
        achieved = BYTES_STRESS * cells / str_avg / 1e9
        tfile = os.path.join(ROOT, "profiles", "stress3d_traffic.json")
    traffic = None
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
            "fused_step": fused,
            "stress_update": "reduced split (P, Q only in the interior; split fields of the damped axes only "
                             "in the absorbing layer; records within ~1e-6 rel-L2 of the split form)"
                             if unsplit else "split (bit-identical to the oracle)",
            "vel_kernel_ms": None if fused else round(vel_avg * 1e3, 4),
            "stress_kernel_ms": None if fused else round(str_avg * 1e3, 4),
            "fused_kernel_ms": round(str_avg * 1e3, 4) if fused else None,
        },
        "roofline": {
            "bound": "hbm", "kernel": kernel,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            # measured HBM bytes of the kernel over its duration (achieved counts
            # the reference structure's bytes, which the unsplit forms partly avoid)
            "traffic_rate": round(traffic / str_avg / 1e9, 1) if traffic else None,
        },
    }
    gpu_rec = prop.record()[:, : args.steps].copy()
    prop.close()
    h = [t.cpu().numpy() for t in (vp, eps, dl)]
    del vp, eps, dl
    torch.cuda.empty_cache()
    out["config"]["workload"] += ("; delta clipped to <= eps (pseudo-acoustic VTI grows without bound "
                                  "where delta > eps)")
    out["cpu_baseline"] = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle built for this host) ...")
        po = native_oracle()
        legs, errs = cpu_legs_3d(po, h[0], h[1], h[2], n, npml, nt, gpu_rec)
        main_leg = legs[0]                 # FP32, all host threads
        out["cpu_baseline"] = {
            "value": main_leg["gcells"], "unit": "Gcell-updates/s", "cores": main_leg["threads"], "kind": "port",
            "sample": f"oracle/vti_oracle.c (split form, FP32) on the full {n}^3 workload "
                      f"({n + 2 * npml}^3 padded), {main_leg['steps']} steps after the first, "
                      f"{main_leg['threads']} OpenMP threads, {main_leg['seconds']} s; {cpu_model()}; {CPU_FLAGS}",
            "legs": legs, "cpu_model": cpu_model(), "nproc": len(os.sched_getaffinity(0)),
            "flags": CPU_FLAGS, "records_rel_l2": errs}
        del h
        out["config1_2d"] = bench_config1(torch, vti, po)
    if not args.no_learned:
        try:
            out["learned_rtm"] = learned_rtm(args, torch, vti, max(4, args.learned_shots), 4000, 384,
                                             rows_out=args.rows_out)
        except Exception as exc:          # the headline line is still printed
            log(f"[bench] learned RTM record failed: {exc}")
            out["learned_rtm"] = {"error": str(exc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
