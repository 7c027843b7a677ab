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


@pytest.mark.parametrize("exact", [0, 1])
def test_fallback_equals_plain_rtm_bitwise(exact):
    """With tol < 0 and a check on every snapshot, every snapshot takes the
    fallback (exact storage, or recomputation from checkpoints): the learned
    path then images exactly the plain path's field."""
    cfg, vp, e, d = _setup()
    rng = np.random.default_rng(4)
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=16, batch=1, seed=1) as A:
        prop.set_shot(_shot(cfg, 16, 16))
        obs = (rng.standard_normal((prop.nrec, NT)) * 0.05).astype(np.float32)
        img_p, ill_p = _plain(prop, obs)
        o = prop.rtm_defaults()
        lo = vti.learned_defaults()
        lo.check_every, lo.tol = 1, -1.0
        if exact:
            lo.max_exact = NT // 4
        else:
            lo.ck_every = 16
        prop.rtm_reset()
        st = prop.rtm_shot_learned(A, o, lo, obs)
        img_l, ill_l = prop.rtm_image(finalize=False)
    _say("stats:", st)
    assert st["snapshots"] == NT // 4
    assert st["over_tol"] > 0
    assert st["exact"] == (st["over_tol"] if exact else 0)
    assert np.array_equal(ill_l, ill_p)
    assert np.array_equal(img_l, img_p), rel_l2(img_l, img_p)


def test_trained_autoencoder_image_error():
    """Online training on two shots' snapshots, then a third shot imaged
    from latents, without and with the exact-storage fallback, against the
    plain RTM image of the same steps."""
    cfg, vp, e, d = _setup()
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=32, batch=2, seed=2) as A:
        A.set_lr(1e-3)
        losses = []
        for rep in range(4):
            for cx, cy in ((10, 12), (22, 20)):
                prop.set_shot(_shot(cfg, cx, cy))
                losses.append(prop.ae_train_shot(A, save_every=4, nsteps=6))
        _say("mean training loss per pass:", losses)
        prop.set_shot(_shot(cfg, 16, 16))
        img_p, ill_p = _plain(prop, None)
        res = {}
        for name, max_exact in (("latent only", 0), ("with exact fallback", NT // 8)):
            lo = vti.learned_defaults()
            lo.check_every, lo.tol, lo.ck_every, lo.max_exact = 1, 0.05, 0, max_exact
            prop.rtm_reset()
            st = prop.rtm_shot_learned(A, prop.rtm_defaults(), lo, None)
            img_l, ill_l = prop.rtm_image(finalize=False)
            assert np.array_equal(ill_l, ill_p)
            res[name] = rel_l2(img_l, img_p)
            _say(f"{name}: image rel-L2 {res[name]:.3e}; mean snapshot error {st['mean_checked_err']:.3e}; {st}")
    assert losses[-1] < losses[0]
    assert res["with exact fallback"] <= res["latent only"]
    assert res["with exact fallback"] < 0.1
