"""The oracle (oracle/vti_oracle.c, float instantiation) against golden
vectors written by the reference CPU path itself
(2D-3D-vti-fd-cuda-gtk-GUI/CPU_vti2dfd_kernels.c, compiled from its source
by oracle/Makefile; fixtures from tests/golden/make_golden.py)."""
import numpy as np
import pytest

from conftest import load_golden, rel_l2
from oracle import pyoracle as po

CASES = ["cpu2d_sv0", "cpu2d_sv1", "cpu2d_two_shots", "cpu2d_aniso_grid"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_cpu_bitwise(name):
    g = load_golden(name)
    nx, nz, dx, dz, npd, SV, favg, ns, fs, ds, zs, nt, dt_us = [int(v) for v in g["args"]]
    for is_ in range(1, ns + 1):
        p = po.params_cpu2d(nx, nz, dx, dz, npd, SV, favg, is_, fs, ds, zs, nt, dt_us)
        sim = po.Sim(p, g["vp"], g["eps"], g["dl"], "f32")
        if is_ == 1:
            # run up to the stored snapshot step and compare the P field
            snap_it = int(g["snap_it"])
            shot = np.zeros((nx, nt), np.float32)
            for it in range(nt):
                shot[:, it] = sim.step(it)
                if it == snap_it:
                    P = sim.field("P")
                    assert np.array_equal(P, g["snap"]), "snapshot P differs"
        else:
            shot = sim.run()
        ref = g["shot"][is_ - 1]
        assert np.array_equal(shot, ref), f"shot {is_}: rel-L2 {rel_l2(shot, ref):.3e}"


def test_fp64_oracle_close_to_reference():
    """FP32 rounding error of the reference path, measured against the FP64
    oracle.  (cpu2d_aniso_grid is not used: without the SV taper its layers
    with eps < delta grow late-time energy, which amplifies any rounding
    difference; pseudo-acoustic VTI is unstable there.)"""
    g = load_golden("cpu2d_sv1")
    nx, nz, dx, dz, npd, SV, favg, ns, fs, ds, zs, nt, dt_us = [int(v) for v in g["args"]]
    p = po.params_cpu2d(nx, nz, dx, dz, npd, SV, favg, 1, fs, ds, zs, nt, dt_us)
    s64 = po.Sim(p, g["vp"], g["eps"], g["dl"], "f64").run()
    err = rel_l2(g["shot"][0], s64)
    assert err < 2e-5, err
