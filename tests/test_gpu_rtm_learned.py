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


def test_exact_fallback_reduces_the_latent_image_error():
    """Property test, not the 1 % bar (the 384^3 benchmark test asserts that
    one): principal-component fit on two shots' snapshots, then a third
    shot imaged from latents, without and with the exact-storage fallback,
    against the plain RTM image of the same steps (raw image, and final
    image divided by the illumination and Laplacian-filtered).  The fit
    samples random blocks, so its captured energy varies from run to run
    (0.9877 to 0.995 measured); the assertion is a floor below that range.
    This 32^3 / 25 Hz case is far harder for 16x compression than the
    15 Hz benchmark (more energy near the grid Nyquist)."""
    cfg, vp, e, d = _setup()
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=32, batch=2, seed=2) as A:
        for cx, cy in ((10, 12), (22, 20)):
            prop.set_shot(_shot(cfg, cx, cy))
            prop.ae_fit_shot(A, save_every=4, nblocks=64)
        explained = A.fit_end()
        _say(f"energy captured by 64 components: {explained:.6f}")
        prop.set_shot(_shot(cfg, 16, 16))
        img_p, ill_p = _plain(prop, None)
        fin_p, _ = prop.rtm_image(finalize=True)
        res = {}
        for name, max_exact in (("latent only", 0), ("with exact fallback", NT // 8)):
            lo = vti.learned_defaults()
            lo.check_every, lo.tol, lo.ck_every, lo.max_exact = 1, 0.02, 0, max_exact
            prop.rtm_reset()
            st = prop.rtm_shot_learned(A, prop.rtm_defaults(), lo, None)
            img_l, ill_l = prop.rtm_image(finalize=False)
            fin_l, _ = prop.rtm_image(finalize=True)
            assert np.array_equal(ill_l, ill_p)
            res[name] = (rel_l2(img_l, img_p), rel_l2(fin_l, fin_p), st)
            _say(f"{name}: raw image rel-L2 {res[name][0]:.3e}, final image rel-L2 {res[name][1]:.3e}; "
                 f"mean snapshot error {st['mean_checked_err']:.3e}; {st}")
    assert explained > 0.98
    assert res["latent only"][2]["mean_checked_err"] < 0.2
    assert 0 < res["with exact fallback"][2]["exact"] <= NT // 8
    assert res["with exact fallback"][0] < res["latent only"][0]
    assert res["with exact fallback"][0] < 5e-3          # raw image (measured 8.1e-4)
    assert res["with exact fallback"][1] < 2e-2          # final image (measured 7.8e-3)

def test_fit_gives_an_orthonormal_projection_network():
    """After the fit the encoder's third layer holds orthonormal rows V and
    the network decodes every 8^3 block to its projection V^T V b (checked
    against the oracle network built from the same V)."""
    import torch
    from oracle import ae_oracle as ae
    cfg, vp, e, d = _setup()
    with vti.Propagator(cfg, vp, e, d) as prop, vti.Autoencoder(crop=32, batch=1, seed=3) as A:
        prop.set_shot(_shot(cfg, 16, 16))
        prop.ae_fit_shot(A, save_every=8, nblocks=64)
        A.fit_end()
        flat = A.params()
        prop.set_shot(_shot(cfg, 12, 20))
        prop.reset()
        prop.step(0, 120)
        P = prop.field("P")                      # padded (nny, nnx, nnz)
        box = np.ascontiguousarray(P[NPML:NPML + N, NPML:NPML + N, NPML:NPML + N])
    sizes = [int(np.prod(ae.kdim(k, ci, co))) for k, ci, co, a in ae.NET]
    offs = np.cumsum([0] + sizes)
    V = flat[offs[2]:offs[3]].reshape(64, 512).astype(np.float64)
    assert np.allclose(V @ V.T, np.eye(64), atol=1e-5)
    with vti.Autoencoder(crop=N, batch=1, seed=4) as B:
        B.set_params(flat)
        s = np.abs(box).max()
        _, rec = B.train_batch((box / s)[None, ..., None], apply=False, want_recon=True)
    ref = ae.forward([flat[offs[i]:offs[i + 1]].reshape(ae.kdim(*ae.NET[i][:3])) for i in range(len(ae.NET))],
                     (box / s)[None, ..., None].astype(np.float32))[-1]
    assert rel_l2(rec, ref) < 1e-2
