"""Write the golden fixture of the OpenMP program Toafd2domp_2.c.

Runs 2D-vti-fd-omp/Toafd2domp_2.c (its main, :40-242: 601 x 301 cells of
5 m, npd = 0, 30 Hz Ricker of gain 1000 at (100, 100), 4001 steps of
0.5 ms), compiled serially from its own source by `make -C oracle ref-omp`
into oracle/_ref/omp2d_ref, on a seeded model, in a scratch directory
holding the three model files it reads (v_601_301.dat, e_601_301.dat,
d_601_301.dat: x-major, z fastest).  Stored in omp2d_601x301.npz:

  * the model as float32 profiles, vp = vz[z] + vx[x] (float32 sum),
    eps = ez[z], delta = dz[z];
  * sha256 of every snapshot frame (P of the interior every 50 steps,
    snap.dat, -0 written as +0) and of the shot record (shot.dat, P+Q at
    row npd: with npd = 0 that row is never updated, so it is all zeros);
  * the frames of steps 100 and 200 cropped around the source, and the
    frame of step 4000 on every 4th cell.

Run in the container that holds the reference (about a minute):

    make -C oracle ref-omp && python tests/golden/make_golden_omp.py
"""
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "omp2d_ref")
NX, NZ, NT = 601, 301, 4001          # Toafd2domp_2.c:62-73


def model(seed=17):
    """Four flat layers plus a lateral vp gradient; eps >= delta (pseudo-
    acoustic VTI is unstable where eps < delta)."""
    rng = np.random.default_rng(seed)
    lay = np.minimum(np.arange(NZ) * 4 // NZ, 3)
    vz = np.sort(rng.uniform(1500.0, 3300.0, 4))[lay].astype(np.float32)
    vx = np.linspace(0.0, 200.0, NX).astype(np.float32)
    el = np.sort(rng.uniform(0.05, 0.2, 4))
    dl = np.minimum(el, rng.uniform(0.0, 0.1, 4))
    return vz, vx, el[lay].astype(np.float32), dl[lay].astype(np.float32)


def expand(vz, vx, ez, dz):
    vp = (vz[None, :] + vx[:, None]).astype(np.float32)
    return vp, np.tile(ez, (NX, 1)), np.tile(dz, (NX, 1))


def frame_sha(a):
    return np.frombuffer(hashlib.sha256((np.asarray(a, np.float32) + np.float32(0)).tobytes()).digest(),
                         dtype=np.uint8)


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref-omp")
    vz, vx, ez, dz = model()
    vp, e, d = expand(vz, vx, ez, dz)
    with tempfile.TemporaryDirectory() as td:
        for name, a in (("v_601_301.dat", vp), ("e_601_301.dat", e), ("d_601_301.dat", d)):
            a.tofile(os.path.join(td, name))
        subprocess.run([REF], cwd=td, check=True, stdout=subprocess.DEVNULL)
        snap = np.fromfile(os.path.join(td, "snap.dat"), dtype=np.float32).reshape(-1, NX, NZ)
        shot = np.fromfile(os.path.join(td, "shot.dat"), dtype=np.float32).reshape(NX, NT)
    nframe = (NT - 1) // 50 + 1
    assert snap.shape[0] == nframe
    shas = np.stack([frame_sha(f) for f in snap])
    np.savez_compressed(
        os.path.join(HERE, "omp2d_601x301.npz"),
        params=np.array([NX, NZ, 5.0, 5.0, 0, 30.0, 1000.0, 100, 100, NT, 0.0005], dtype=np.float64),
        vz=vz, vx=vx, ez=ez, dz=dz, frame_sha=shas, shot_sha=frame_sha(shot),
        shot_absmax=np.float32(np.abs(shot).max()),
        crop100=snap[2, 59:140, 59:140], crop200=snap[4, 19:180, 19:180],
        sub4000=snap[80, ::4, ::4])
    print("frames", snap.shape, "absmax step 200 %.4g, step 4000 %.4g, shot absmax %g"
          % (np.abs(snap[4]).max(), np.abs(snap[80]).max(), np.abs(shot).max()))


if __name__ == "__main__":
    main()
