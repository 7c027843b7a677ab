"""Write golden fixtures from the reference CPU path.

Runs the reference CPU_vti2dfd (2D-3D-vti-fd-cuda-gtk-GUI/CPU_vti2dfd_kernels.c
:366-536), compiled from its own source by `make -C oracle ref` into
oracle/_ref/cpu_vti2dfd_ref, on small seeded layered VTI models, and stores
inputs and outputs (shot record, last P snapshot of shot 1) as .npz files in
this directory.  Run in the container that holds the reference:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "cpu_vti2dfd_ref")

# name: (nx, nz, dx, dz, npd, SV, favg, ns, fs, ds, zs, nt, dt_us, seed)
CASES = {
    "cpu2d_sv0": (60, 50, 10, 10, 10, 0, 20, 1, 30, 0, 2, 300, 1000, 1),
    "cpu2d_sv1": (60, 50, 10, 10, 10, 1, 20, 1, 30, 0, 2, 300, 1000, 2),
    "cpu2d_two_shots": (80, 60, 10, 10, 12, 1, 20, 2, 20, 30, 3, 250, 1000, 3),
    "cpu2d_aniso_grid": (128, 96, 8, 6, 20, 1, 25, 1, 64, 0, 5, 600, 500, 4),
}


def layered_model(nx, nz, seed, nlayer=4):
    """Flat layers, x-major / z-fastest float32 (nx, nz).  eps/delta are
    kept strictly positive: the reference pad_vv returns early on a zero
    value (CPU_vti2dfd_kernels.c:244)."""
    rng = np.random.default_rng(seed)
    vpl = np.sort(rng.uniform(1500.0, 4000.0, nlayer))
    el = np.sort(rng.uniform(0.01, 0.2, nlayer))
    dl = np.sort(rng.uniform(0.01, 0.1, nlayer))
    lay = np.minimum(np.arange(nz) * nlayer // nz, nlayer - 1)
    vp = np.tile(vpl[lay], (nx, 1)).astype(np.float32)
    e = np.tile(el[lay], (nx, 1)).astype(np.float32)
    d = np.tile(dl[lay], (nx, 1)).astype(np.float32)
    return vp, e, d


def run_case(name, args):
    nx, nz, dx, dz, npd, SV, favg, ns, fs, ds, zs, nt, dt_us, seed = args
    vp, e, d = layered_model(nx, nz, seed)
    with tempfile.TemporaryDirectory() as td:
        for n, a in (("v", vp), ("e", e), ("d", d)):
            a.tofile(os.path.join(td, n))
        cmd = [REF] + [str(v) for v in (nx, nz, dx, dz, npd, SV)] + \
            [os.path.join(td, n) for n in ("v", "e", "d")] + \
            [str(v) for v in (favg, ns, fs, ds, zs)] + \
            [os.path.join(td, "shot"), os.path.join(td, "snap"), str(nt), str(dt_us)]
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
        shot = np.fromfile(os.path.join(td, "shot"), dtype=np.float32).reshape(ns, nx, nt)
        nnx, nnz = nx + 2 * npd, nz + 2 * npd
        snap = np.fromfile(os.path.join(td, "snap"), dtype=np.float32).reshape(-1, nnx, nnz)
    last = (nt - 1) // 50
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        args=np.array(args[:-1], dtype=np.int64), vp=vp, eps=e, dl=d,
        shot=shot, snap_it=np.int64(last * 50), snap=snap[last])
    print(name, "shot", shot.shape, "absmax %.4g" % np.abs(shot).max())


if __name__ == "__main__":
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    for k, v in CASES.items():
        run_case(k, v)
