"""The oracle's unsplit-interior mode (the product's optional 3D update
where every damping factor is 1) against the reference's split arithmetic
(the oracle's default) and against FP64, on the CPU."""
import numpy as np

from conftest import rel_l2, smooth_model3d
from oracle import pyoracle as po


def test_unsplit_interior_stays_within_1e5_of_the_split_form():
    n, npml, nt = 32, 8, 400
    vp, e, d = smooth_model3d(n, n, n, seed=3)
    d = np.minimum(d, e)
    p = po.params_gpu3d(n, n, n, 10, 10, 10, npml, 1, 20, 0, n // 2, 0, n // 2, 0, 1, 0, nt, 500, 3, 2)
    split = po.Sim(p, vp, e, d, "f32").run()
    f64 = po.Sim(p, vp, e, d, "f64").run()
    p.unsplit = 1
    un = po.Sim(p, vp, e, d, "f32").run()
    assert np.abs(split).max() > 0
    e_us, e_s64, e_u64 = rel_l2(un, split), rel_l2(split, f64), rel_l2(un, f64)
    print(f"unsplit vs split {e_us:.3e}; split vs FP64 {e_s64:.3e}; unsplit vs FP64 {e_u64:.3e}")
    assert 0 < e_us <= 1e-5
    assert e_u64 <= 2 * e_s64          # as close to the FP64 run as the split form
    # the FP64 unsplit run is the split one up to FP64 rounding: same mathematics
    p.unsplit = 1
    un64 = po.Sim(p, vp, e, d, "f64").run()
    assert rel_l2(un64, f64) < 1e-12


def test_unsplit_interior_2d():
    from conftest import layered_model
    nx, nz, npd, nt = 120, 90, 20, 800
    vp, e, d = layered_model(nx, nz, seed=4)
    d = np.minimum(d, e)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 1, 20, 1, 60, 0, 2, nt, 1000)
    split = po.Sim(p, vp, e, d, "f32").run()
    f64 = po.Sim(p, vp, e, d, "f64").run()
    p.unsplit = 1
    un = po.Sim(p, vp, e, d, "f32").run()
    un64 = po.Sim(p, vp, e, d, "f64").run()
    e_us, e_s64, e_u64 = rel_l2(un, split), rel_l2(split, f64), rel_l2(un, f64)
    print(f"2D: unsplit vs split {e_us:.3e}; split vs FP64 {e_s64:.3e}; unsplit vs FP64 {e_u64:.3e}")
    assert 0 < e_us <= 1e-5
    assert e_u64 <= 2 * e_s64
    assert rel_l2(un64, f64) < 1e-12
