"""GPU RTM (checkpointed source-wavefield recomputation) against the
oracle's brute-force RTM (every snapshot stored): bit-exact images."""
import numpy as np
import pytest

from conftest import layered_model, rel_l2, smooth_model3d
from geophysics_amd import vti
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("nsnap", [0, 3, 1])
def test_rtm_2d_matches_oracle_bitwise(nsnap):
    nx, nz, npd, nt = 90, 70, 15, 400
    vp, e, d = layered_model(nx, nz, seed=21, elo=0.1, ehi=0.2, dlo=0.0, dhi=0.1)
    cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 10, 10, npd, 1, 20, nt, 1000)
    cfg.sv_R, cfg.sv_r = 18, 7          # the RTM file's taper (R=18, r=7)
    fs, zs = 45, 1
    rng = np.random.default_rng(1)
    obs = (rng.standard_normal((nx, nt)) * 0.1).astype(np.float32)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_shot(vti.shot_ref2d(cfg, 1, fs, 0, zs))
        o = prop.rtm_defaults()
        o.nsnap = nsnap
        prop.rtm_reset()
        steps = prop.rtm_shot(o, obs)
        img, ill = prop.rtm_image(finalize=False)
        fin, _ = prop.rtm_image(finalize=True)
    _say(f"nsnap={nsnap}: {steps} forward steps for nt={nt}")
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 1, 20, 1, fs, 0, zs, nt, 1000)
    p.sv_R, p.sv_r = 18, 7
    rimg, rill, _ = po.rtm(p, vp, e, d, sv_fwd=1, sv_bwd=0, illum_mode=0, obs=obs)
    assert np.array_equal(ill, rill)
    assert np.array_equal(img, rimg), rel_l2(img, rimg)
    rfin = po.laplace_adj((rimg / rill).astype(np.float32))
    assert np.array_equal(fin, rfin)
    if nsnap == 0:
        assert steps == nt
    else:
        assert steps > nt


def test_rtm_2d_receiver_illumination_bitwise():
    """Illumination mode 1 (source and receiver sides summed, no clamp): the
    receiver side's P^2 + Q^2 is added in the stress update's epilogue; bit
    for bit the oracle's separate pass."""
    nx, nz, npd, nt = 72, 60, 12, 300
    vp, e, d = layered_model(nx, nz, seed=8, elo=0.1, ehi=0.2, dlo=0.0, dhi=0.1)
    cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 10, 10, npd, 0, 20, nt, 1000)
    rng = np.random.default_rng(4)
    obs = (rng.standard_normal((nx, nt)) * 0.1).astype(np.float32)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_shot(vti.shot_ref2d(cfg, 1, 36, 0, 1))
        o = prop.rtm_defaults()
        o.nsnap = 3
        o.illum_mode = 1
        prop.rtm_reset()
        prop.rtm_shot(o, obs)
        img, ill = prop.rtm_image(finalize=False)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 0, 20, 1, 36, 0, 1, nt, 1000)
    rimg, rill, _ = po.rtm(p, vp, e, d, sv_fwd=0, sv_bwd=0, illum_mode=1, obs=obs)
    assert np.abs(rill).max() > 0
    assert np.array_equal(ill, rill), rel_l2(ill, rill)
    assert np.array_equal(img, rimg), rel_l2(img, rimg)


def test_rtm_2d_modelled_record_with_mute():
    """obs=None: the modelled record, muted with the windowed form (tt=20)."""
    nx, nz, npd, nt = 80, 60, 12, 350
    vp, e, d = layered_model(nx, nz, seed=5, elo=0.1, dlo=0.0, dhi=0.1)
    cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 10, 10, npd, 0, 20, nt, 1000)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_shot(vti.shot_ref2d(cfg, 1, 40, 0, 1))
        o = prop.rtm_defaults()
        o.nsnap = 4
        prop.rtm_reset()
        prop.rtm_shot(o, None)
        img, ill = prop.rtm_image(finalize=False)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 0, 20, 1, 40, 0, 1, nt, 1000)
    sim = po.Sim(p, vp, e, d)
    vp1, e1 = float(sim.field("vp")[0, 1]), float(sim.field("eps")[0, 1])
    mute = lambda r: po.mute2d(r, nt, p.dt, p.favg, p.dx, p.dz, 40, 0, 1, 1, vp1, e1, 20)
    rimg, rill, _ = po.rtm(p, vp, e, d, sv_fwd=0, sv_bwd=0, illum_mode=0, mute=mute)
    assert np.array_equal(img, rimg), rel_l2(img, rimg)


def test_rtm_3d_matches_oracle_bitwise():
    nx, ny, nz, npml, nt = 28, 26, 24, 6, 140
    vp, e, d = smooth_model3d(nx, ny, nz, seed=7)
    cfg = vti.cfg_ref3d(nx, ny, nz, 10, 10, 10, npml, 1, 25, nt, 800)
    cfg.rec_kind = vti.REC_PQ            # 3D RTM records P+Q
    rng = np.random.default_rng(2)
    with vti.Propagator(cfg, vp, e, d) as prop:
        sh = vti.shot_ref3d(cfg, 0, 14, 0, 13, 0, 1, 0, rec_depth=1, rec_step=2)
        prop.set_shot(sh)
        obs = (rng.standard_normal((prop.nrec, nt)) * 0.05).astype(np.float32)
        o = prop.rtm_defaults()
        o.nsnap = 5
        o.inject_z = npml - 1               # 3D reference injects at npml-1
        assert prop.checkpoint_bytes() == prop.state_bytes()   # split form: whole-state slots
        prop.rtm_reset()
        prop.rtm_shot(o, obs)
        img, ill = prop.rtm_image(finalize=False)
        fin, _ = prop.rtm_image(finalize=True)
    p = po.params_gpu3d(nx, ny, nz, 10, 10, 10, npml, 1, 25, 0, 14, 0, 13, 0, 1, 0, nt, 800, 1, 2)
    p.rec_kind = 2
    rimg, rill, _ = po.rtm(p, vp, e, d, sv_fwd=1, sv_bwd=0, inject_z=npml - 1, illum_mode=1, obs=obs)
    assert np.array_equal(ill, rill)
    assert np.array_equal(img, rimg), rel_l2(img, rimg)
    # finalize: divide where the illumination is nonzero (3D form,
    # Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:550-551), then the 3D
    # adjoint Laplacian (laplace_3d_filter :555-616)
    q = np.divide(rimg, rill, out=rimg.copy(), where=rill != 0).astype(np.float32)
    rfin = po.laplace_adj(q)
    assert np.abs(rfin).max() > 0
    assert np.array_equal(fin, rfin), rel_l2(fin, rfin)


def test_rtm_3d_unsplit_interior():
    """3D RTM with the unsplit interior (vti_set_unsplit): the records are
    written at row npml-1, inside the damped layer, so both sides use it;
    bit for bit the oracle's RTM with unsplit source and receiver
    simulations, and within 1e-5 rel-L2 of the split RTM image."""
    nx, ny, nz, npml, nt = 28, 26, 24, 6, 140
    vp, e, d = smooth_model3d(nx, ny, nz, seed=7)
    cfg = vti.cfg_ref3d(nx, ny, nz, 10, 10, 10, npml, 1, 25, nt, 800)
    cfg.rec_kind = vti.REC_PQ
    rng = np.random.default_rng(2)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_unsplit(True)
        # trimmed slots: u, v, w, P, Q whole (5 of the 11 volumes) and the
        # split fields of the absorbing layer only (most of this small grid)
        ck, full = prop.checkpoint_bytes(), prop.state_bytes()
        assert 5 * full // 11 < ck < full
        sh = vti.shot_ref3d(cfg, 0, 14, 0, 13, 0, 1, 0, rec_depth=1, rec_step=2)
        prop.set_shot(sh)
        obs = (rng.standard_normal((prop.nrec, nt)) * 0.05).astype(np.float32)
        o = prop.rtm_defaults()
        o.nsnap = 5
        o.inject_z = npml - 1
        prop.rtm_reset()
        prop.rtm_shot(o, obs)
        img, ill = prop.rtm_image(finalize=False)
    p = po.params_gpu3d(nx, ny, nz, 10, 10, 10, npml, 1, 25, 0, 14, 0, 13, 0, 1, 0, nt, 800, 1, 2)
    p.rec_kind = 2
    rimg, rill, _ = po.rtm(p, vp, e, d, sv_fwd=1, sv_bwd=0, inject_z=npml - 1, illum_mode=1, obs=obs,
                           unsplit_source=True, unsplit_receiver=True)
    assert np.array_equal(ill, rill)
    assert np.array_equal(img, rimg), rel_l2(img, rimg)
    simg, _, _ = po.rtm(p, vp, e, d, sv_fwd=1, sv_bwd=0, inject_z=npml - 1, illum_mode=1, obs=obs)
    err = rel_l2(img, simg)
    print(f"3D RTM, unsplit interior: image rel-L2 vs split {err:.3e}", flush=True)
    assert err <= 1e-5


def test_rtm_2d_unsplit_with_interior_injection_row():
    """2D RTM with vti_set_unsplit: the records are loaded into the first
    interior row, which keeps the split form on the receiver side; bit for
    bit the oracle's RTM with unsplit source and receiver simulations, and
    as close to the FP64 image as the split FP32 image (within 1.5x of its
    error: this noise-record image is 2e-4 from FP64 in either form, and
    the FP64 unsplit and split images agree to 1e-13)."""
    nx, nz, npd, nt = 90, 70, 15, 400
    vp, e, d = layered_model(nx, nz, seed=21, elo=0.1, ehi=0.2, dlo=0.0, dhi=0.1)
    cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 10, 10, npd, 0, 20, nt, 1000)
    fs, zs = 45, 1
    rng = np.random.default_rng(1)
    obs = (rng.standard_normal((nx, nt)) * 0.1).astype(np.float32)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_unsplit(True)
        prop.set_shot(vti.shot_ref2d(cfg, 1, fs, 0, zs))
        o = prop.rtm_defaults()
        o.nsnap = 3
        prop.rtm_reset()
        prop.rtm_shot(o, obs)
        img, ill = prop.rtm_image(finalize=False)
    p = po.params_gpu2d(nx, nz, 10, 10, npd, 0, 20, 1, fs, 0, zs, nt, 1000)
    assert npd <= p.rec_z < npd + nz
    rimg, rill, _ = po.rtm(p, vp, e, d, sv_fwd=0, sv_bwd=0, illum_mode=0, obs=obs,
                           unsplit_source=True, unsplit_receiver=True)
    assert np.array_equal(ill, rill)
    assert np.array_equal(img, rimg), rel_l2(img, rimg)
    simg, _, _ = po.rtm(p, vp, e, d, sv_fwd=0, sv_bwd=0, illum_mode=0, obs=obs)
    img64, _, _ = po.rtm(p, vp, e, d, sv_fwd=0, sv_bwd=0, illum_mode=0, obs=obs.astype(np.float64), prec="f64")
    e_us, e_s = rel_l2(img, img64), rel_l2(simg, img64)
    print(f"2D RTM, unsplit: vs split {rel_l2(img, simg):.3e}; vs FP64 {e_us:.3e} (split {e_s:.3e})", flush=True)
    assert e_us <= max(1e-5, 1.5 * e_s)

