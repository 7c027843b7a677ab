"""3D RTM with the source wavefield stored as autoencoder latents, against
the plain checkpointed RTM imaging the same steps (image_every = 4)."""
import numpy as np
import pytest

from conftest import rel_l2, smooth_model3d
from geophysics_amd import vti

pytestmark = pytest.mark.gpu

N, NPML, NT = 32, 6, 200


def _say(*a):
    print(*a, flush=True)


def _setup(seed=7):
    vp, e, d = smooth_model3d(N, N, N, seed=seed)
    cfg = vti.cfg_ref3d(N, N, N, 10, 10, 10, NPML, 1, 25, NT, 800)
    return cfg, vp, e, d


def _shot(cfg, cx, cy):
    return vti.shot_ref3d(cfg, 0, cx, 0, cy, 0, 1, 0, rec_depth=1, rec_step=2)


def _plain(prop, obs, nsnap=6):
    o = prop.rtm_defaults()
    o.image_every = 4
    o.nsnap = nsnap
    prop.rtm_reset()
    prop.rtm_shot(o, obs)
    return prop.rtm_image(finalize=False)


def test_fallback_recomputation_equals_plain_rtm_bitwise():
    """Every checked snapshot above the tolerance is recomputed from the
    fallback checkpoints: with tol < 0 and a check on every snapshot the
    learned path images exactly the plain path's field."""
    cfg, vp, e, d = _setup()
    rng = np.random.default_rng(4)
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        prop.set_shot(_shot(cfg, 16, 16))
        obs = (rng.standard_normal((prop.nrec, NT)) * 0.05).astype(np.float32)
        img_p, ill_p = _plain(prop, obs)
        o = prop.rtm_defaults()
        lo = vti.learned_defaults()
        lo.check_every, lo.tol, lo.ck_every = 1, -1.0, 16
        prop.rtm_reset()
        st = prop.rtm_shot_learned(A, o, lo, obs)
        img_l, ill_l = prop.rtm_image(finalize=False)
    _say("stats:", st)
    assert st["snapshots"] == NT // 4
    assert st["recomputed"] > 0
    assert np.array_equal(ill_l, ill_p)
    assert np.array_equal(img_l, img_p), rel_l2(img_l, img_p)


def test_trained_autoencoder_image_error():
    """Online training on two shots' snapshots, then a third shot imaged
    from latents only (no fallback) against the plain RTM image."""
    cfg, vp, e, d = _setup()
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=16, batch=8, seed=2) as A:
        A.set_lr(2e-3)
        losses = []
        for rep in range(3):
            for cx, cy in ((10, 12), (22, 20)):
                prop.set_shot(_shot(cfg, cx, cy))
                losses.append(prop.ae_train_shot(A, save_every=4, nsteps=4))
        _say("mean training loss per pass:", losses)
        prop.set_shot(_shot(cfg, 16, 16))
        img_p, ill_p = _plain(prop, None)
        o = prop.rtm_defaults()
        lo = vti.learned_defaults()
        lo.check_every, lo.ck_every = 4, 0
        prop.rtm_reset()
        st = prop.rtm_shot_learned(A, o, lo, None)
        img_l, ill_l = prop.rtm_image(finalize=False)
    err = rel_l2(img_l, img_p)
    _say(f"learned vs plain image rel-L2 {err:.3e}; mean snapshot error {st['mean_checked_err']:.3e}; stats {st}")
    assert np.array_equal(ill_l, ill_p)
    assert losses[-1] < losses[0]
    assert err < 0.3
