"""The oracle's Toafd2domp_2.c form against golden vectors written by that
program itself (2D-vti-fd-omp/Toafd2domp_2.c, compiled serially from its
source by `make -C oracle ref-omp`; fixture from
tests/golden/make_golden_omp.py): every one of the 81 snapshot frames of
the 601 x 301 / 4001-step run bit for bit (sha256), and the shot record."""
import hashlib

import numpy as np
import pytest

from conftest import load_golden, rel_l2
from oracle import pyoracle as po


def omp_case():
    g = load_golden("omp2d_601x301")
    nx, nz, dx, dz, npd, favg, pfac, sx, sz, nt, dt = g["params"]
    vp = (g["vz"][None, :] + g["vx"][:, None]).astype(np.float32)
    e = np.tile(g["ez"], (int(nx), 1))
    d = np.tile(g["dz"], (int(nx), 1))
    p = po.params_omp2d(int(nx), int(nz), float(dx), float(dz), int(npd), float(favg), float(pfac),
                        int(sx), int(sz), int(nt), float(np.float32(dt)))
    return g, p, vp, e, d


def frame_sha(a):
    return np.frombuffer(hashlib.sha256((np.asarray(a, np.float32) + np.float32(0)).tobytes()).digest(),
                         dtype=np.uint8)


@pytest.mark.slow
def test_oracle_matches_toafd2domp_every_frame_bitwise():
    g, p, vp, e, d = omp_case()
    sim = po.Sim(p, vp, e, d, "f32")
    nt = p.nt
    shot = np.zeros((p.rec_nx, nt), np.float32)
    for it in range(nt):
        shot[:, it] = sim.step(it)
        if it % 50 == 0:
            P = sim.field("P")
            if it == 100:
                assert np.array_equal(P[59:140, 59:140], g["crop100"]), rel_l2(P[59:140, 59:140], g["crop100"])
            if it == 200:
                assert np.array_equal(P[19:180, 19:180], g["crop200"]), rel_l2(P[19:180, 19:180], g["crop200"])
            if it == 4000:
                assert np.array_equal(P[::4, ::4], g["sub4000"])
            assert np.array_equal(frame_sha(P), g["frame_sha"][it // 50]), f"frame of step {it}"
    assert np.array_equal(frame_sha(shot), g["shot_sha"])
    assert not shot.any()      # row npd = 0 is never updated


def test_oracle_matches_toafd2domp_first_frames():
    """The first 401 steps (frames 0..8) of the same run: a quick form of
    the test above for the default CPU suite."""
    g, p, vp, e, d = omp_case()
    sim = po.Sim(p, vp, e, d, "f32")
    for it in range(401):
        sim.step(it, rec=False)
        if it % 50 == 0:
            assert np.array_equal(frame_sha(sim.field("P")), g["frame_sha"][it // 50]), f"frame of step {it}"


def test_omp_form_differs_from_the_cpu_form():
    """The Toafd2domp_2.c arithmetic is not the CPU driver's: with its own
    time origin, Ricker product order and double pz/qx updates switched to
    the CPU form's, the frames differ (so the golden vectors pin those)."""
    g, p, vp, e, d = omp_case()
    for field, val in (("omp_form", 0), ("wtype", 1), ("t0_kind", 0)):
        q = po.OrcParams.from_buffer_copy(p)
        setattr(q, field, val)
        sim = po.Sim(q, vp, e, d, "f32")
        for it in range(201):
            sim.step(it, rec=False)
        assert not np.array_equal(frame_sha(sim.field("P")), g["frame_sha"][4]), field
