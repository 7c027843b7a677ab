"""The oracle's unsplit mode (the product's optional reduced-split update:
P and Q directly in the interior; in 3D, split fields only of the damped
axes in the absorbing layer, full split in its corners) against the
reference's split arithmetic (the oracle's default) and against FP64, on
the CPU."""
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


def _final_fields(p, vp, e, d, prec):
    s = po.Sim(p, vp, e, d, prec)
    s.run()
    out = [s.field(k).copy() for k in ("P", "Q")]
    s.close()
    return out


def test_reduced_split_whole_grid_fp64():
    """Final P and Q over the whole padded grid, absorbing layer included
    (its faces, edges and corners update in three different forms): the
    FP64 reduced-split state equals the FP64 split state to 1e-12, region by
    region."""
    n, npml, nt = 24, 8, 300
    vp, e, d = smooth_model3d(n, n, n, seed=9)
    d = np.minimum(d, e)
    p = po.params_gpu3d(n, n, n, 10, 10, 10, npml, 1, 20, 0, n // 2, 0, n // 2, 0, 1, 0, nt, 500, 3, 2)
    ref = _final_fields(p, vp, e, d, "f64")
    p.unsplit = 1
    red = _final_fields(p, vp, e, d, "f64")
    m = npml + 4
    for a, b in zip(ref, red):
        assert np.abs(a).max() > 0
        for sl in [np.s_[:m, :, :], np.s_[:, :m, :], np.s_[:, :, :m], np.s_[:m, :m, :m],
                   np.s_[m:-m, m:-m, m:-m], np.s_[:, :, :]]:
            assert rel_l2(b[sl], a[sl]) < 1e-12


def test_reduced_split_rtm_fp64():
    """3D RTM whose records are loaded into P, Q at row npml-1 of the
    absorbing layer: in the reduced-split mode (both simulations) the FP64
    image equals the FP64 split image to 1e-12.  The loaded values persist
    in that row's P in the reduced form, but the row's stress result is
    overwritten by the next load before any velocity update reads it and
    lies outside the image."""
    nx, ny, nz, npml, nt = 20, 18, 16, 6, 120
    vp, e, d = smooth_model3d(nx, ny, nz, seed=7)
    d = np.minimum(d, e)
    p = po.params_gpu3d(nx, ny, nz, 10, 10, 10, npml, 1, 25, 0, 10, 0, 9, 0, 1, 0, nt, 800, 1, 2)
    p.rec_kind = 2
    rng = np.random.default_rng(2)
    obs = rng.standard_normal((p.rec_nx * p.rec_ny, nt)) * 0.05
    kw = dict(sv_fwd=1, sv_bwd=0, inject_z=npml - 1, illum_mode=1, obs=obs, prec="f64")
    ref, _, _ = po.rtm(p, vp, e, d, **kw)
    red, _, _ = po.rtm(p, vp, e, d, unsplit_source=True, unsplit_receiver=True, **kw)
    assert np.abs(ref).max() > 0
    assert rel_l2(red, ref) < 1e-12


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


def test_unsplit_rtm_2d_interior_injection_row_fp64():
    """2D RTM whose records are loaded into P, Q at an interior row (the
    first interior row, as the reference's 2D RTM): in the unsplit mode
    (both simulations) that row keeps the split form, so the loaded values
    do not persist in P, and the FP64 image equals the FP64 split image to
    1e-12 (with the row unsplit they would reach the image)."""
    from conftest import layered_model
    nx, nz, npd, nt = 70, 56, 12, 300
    vp, e, d = layered_model(nx, nz, seed=6)
    d = np.minimum(d, e)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 0, 20, 1, 35, 0, 1, nt, 1000)
    assert npd <= p.rec_z < npd + nz           # an interior row
    rng = np.random.default_rng(5)
    obs = rng.standard_normal((p.rec_nx, nt)) * 0.1
    kw = dict(sv_fwd=0, sv_bwd=0, illum_mode=1, obs=obs, prec="f64")
    ref, rill, _ = po.rtm(p, vp, e, d, **kw)
    un, uill, _ = po.rtm(p, vp, e, d, unsplit_source=True, unsplit_receiver=True, **kw)
    assert np.abs(ref).max() > 0
    assert rel_l2(un, ref) < 1e-12
    assert rel_l2(uill, rill) < 1e-12

