"""Pins of the oracle's adjoint and vp-gradient by mathematics: the
adjoint identity <L a, r> = <a, L^T r> of the linear source-to-data map,
and central finite differences of the (golden-pinned) forward map."""
import numpy as np
import pytest

from conftest import layered_model, smooth_model3d
from oracle import pyoracle as po


def _case2d(nt=160):
    nx, nz, npd = 40, 34, 8
    vp, e, d = layered_model(nx, nz, seed=3, elo=0.1, ehi=0.2, dlo=0.0, dhi=0.1)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 1, 25, 1, 20, 0, 3, nt, 1000, 2)
    return p, vp, e, d


def _case3d(nt=70):
    nx, ny, nz, npml = 16, 14, 15, 5
    vp, e, d = smooth_model3d(nx, ny, nz, seed=2, sigma=2.0)
    p = po.params_gpu3d(nx, ny, nz, 10, 10, 10, npml, 1, 30, 0, 8, 0, 7, 0, 3, 0, nt, 800, 2, 2)
    return p, vp, e, d


@pytest.mark.parametrize("case", [_case2d, _case3d])
def test_oracle_adjoint_dot_product(case):
    p, vp, e, d = case()
    rng = np.random.default_rng(5)
    a = rng.standard_normal(p.nt)
    sim = po.Sim(p, vp, e, d, "f64")
    po.set_source(sim, a)
    La = sim.run()
    r = rng.standard_normal(La.shape)
    sim2 = po.Sim(p, vp, e, d, "f64")
    po.set_source(sim2, a)
    _, _, LTr, _ = po.fwi_grad(sim2, resid=r)
    lhs, rhs = float(np.sum(La * r)), float(np.sum(a * LTr))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("case", [_case2d, _case3d])
def test_oracle_gradient_matches_finite_differences(case):
    p, vp, e, d = case()
    rng = np.random.default_rng(11)
    base = po.Sim(p, vp, e, d, "f64")
    vtrue = vp * (1.0 + 0.03 * rng.standard_normal(vp.shape)).astype(np.float32)
    dobs = po.Sim(p, vtrue, e, d, "f64").run()
    J0, g, _, _ = po.fwi_grad(po.Sim(p, vp, e, d, "f64"), dobs=dobs)

    def J(v):
        # the oracle's model input is float32; set the padded vp in float64
        s = po.Sim(p, vp, e, d, "f64")
        s.set_field("vp", np.pad(v, p.npml, mode="edge"))
        po.copy_tables(s, base)
        return 0.5 * float(np.sum((s.run() - dobs) ** 2))
    h = 0.005   # central differences: error O(h^2), ~1e-9 relative here
    for k in range(3):
        dm = rng.standard_normal(vp.shape)
        fd = (J((vp + h * dm).astype(np.float64)) - J((vp - h * dm).astype(np.float64))) / (2 * h)
        an = float(np.sum(g * dm))
        assert abs(fd - an) <= 1e-7 * abs(an), (k, fd, an)
