"""Batched shots of uneven grids (one persistent launch per phase, graph
blocks, autotuned tiles) against the single-shot path, bit for bit."""
import numpy as np
import pytest

from conftest import layered_model, rel_l2
from geophysics_amd import vti
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SIZES = [(70, 50), (93, 61), (48, 77), (120, 40), (64, 64)]


def _batch(nt=260, sv=1):
    cfgs, shots, models = [], [], []
    for k, (nx, nz) in enumerate(SIZES):
        vp, e, d = layered_model(nx, nz, seed=100 + k, elo=0.1, ehi=0.2, dlo=0.0, dhi=0.1)
        cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, nx, nz, 10, 10, 10 + k, sv, 15 + 2 * k, nt, 1000)
        cfgs.append(cfg)
        shots.append(vti.shot_ref2d(cfg, 1, nx // 2 - k, 0, 2 + k % 2))
        models.append((vp, e, d))
    return cfgs, shots, models


@pytest.mark.parametrize("shape,mult,graph", [(0, 0, 0), (2, 8, 16), (3, 16, 5), (4, 0, 32), (1, 8, 0),
                                               (5, 0, 0), (6, 8, 16), (7, 0, 5), (8, 16, 0), (9, 0, 32)])
def test_batch_forward_records_equal_single_shot(shape, mult, graph):
    cfgs, shots, models = _batch()
    nt = cfgs[0].nt
    with vti.Batch(cfgs, shots, models) as B:
        B.config(shape, mult, graph)
        for rep in range(2):          # replayed graphs with eager work in between
            B.forward(nt, record=True)
            recs = [B.record(s) for s in range(B.n)]
            for s, (cfg, shot, m) in enumerate(zip(cfgs, shots, models)):
                with vti.Propagator(cfg, *m) as prop:
                    prop.set_shot(shot)
                    prop.run()
                    ref = prop.record()
                assert np.array_equal(recs[s], ref), (rep, s, rel_l2(recs[s], ref))


@pytest.mark.parametrize("shape", [0, 5])
def test_batch_rtm_images_equal_single_shot(shape):
    cfgs, shots, models = _batch(nt=220)
    rng = np.random.default_rng(0)
    obs = [(rng.standard_normal((s.rec_nx, c.nt)) * 0.1).astype(np.float32) for c, s in zip(cfgs, shots)]
    with vti.Batch(cfgs, shots, models) as B:
        o = vti.rtm_defaults(2)
        o.nsnap = 3
        B.config(shape, 8, 0)
        steps = B.rtm(o, obs)
        imgs = [B.image(s) for s in range(B.n)]
    for s, (cfg, shot, m) in enumerate(zip(cfgs, shots, models)):
        with vti.Propagator(cfg, *m) as prop:
            prop.set_shot(shot)
            oo = prop.rtm_defaults()
            oo.nsnap = 3
            prop.rtm_reset()
            prop.rtm_shot(oo, obs[s])
            img, ill = prop.rtm_image(finalize=False)
        assert np.array_equal(imgs[s][1], ill), s
        assert np.array_equal(imgs[s][0], img), (s, rel_l2(imgs[s][0], img))
    print("batch RTM forward steps:", steps, flush=True)


def test_batch_autotune_picks_a_shape():
    cfgs, shots, models = _batch(nt=40)
    with vti.Batch(cfgs, shots, models) as B:
        shape, ms = B.autotune(steps=3)
        assert 0 <= shape < 10 and ms > 0
        print("autotuned", B.get_config(), f"{ms:.3f} ms/step", flush=True)


@pytest.mark.parametrize("shape", [2, 7])
def test_batch_rtm_modelled_records_with_mute_equal_single_shot(shape):
    """obs=None: each shot migrates its own modelled record after the
    direct-wave mute (tt=20), as the single-shot RTM does."""
    cfgs, shots, models = _batch(nt=200, sv=0)
    with vti.Batch(cfgs, shots, models) as B:
        o = vti.rtm_defaults(2)
        o.nsnap = 5
        B.config(shape, 0, 0)
        B.rtm(o, None)
        imgs = [B.image(s) for s in range(B.n)]
    for s, (cfg, shot, m) in enumerate(zip(cfgs, shots, models)):
        with vti.Propagator(cfg, *m) as prop:
            prop.set_shot(shot)
            oo = prop.rtm_defaults()
            oo.nsnap = 5
            prop.rtm_reset()
            prop.rtm_shot(oo, None)
            img, _ = prop.rtm_image(finalize=False)
        assert np.array_equal(imgs[s][0], img), (s, rel_l2(imgs[s][0], img))


@pytest.mark.parametrize("shape,graph", [(0, 0), (0, 4), (6, 4), (9, 0)])
def test_batch_rtm_unsplit_equals_single_shot_unsplit(shape, graph):
    """Unsplit interior in the batch (vti_batch_set_unsplit): each shot's
    image and illumination equal the single-shot RTM with vti_set_unsplit
    bit for bit; against the FP64 oracle RTM the unsplit image is as close
    as the split FP32 image (these images correlate noise records, so the
    split image's own FP32 error is ~1e-5)."""
    cfgs, shots, models = _batch(nt=220)
    rng = np.random.default_rng(0)
    obs = [(rng.standard_normal((s.rec_nx, c.nt)) * 0.1).astype(np.float32) for c, s in zip(cfgs, shots)]
    with vti.Batch(cfgs, shots, models) as B:
        o = vti.rtm_defaults(2)
        o.nsnap = 3
        B.config(shape, 8, graph)
        B.set_unsplit(True)
        B.rtm(o, obs)
        imgs = [B.image(s) for s in range(B.n)]
    for s, (cfg, shot, m) in enumerate(zip(cfgs, shots, models)):
        with vti.Propagator(cfg, *m) as prop:
            prop.set_shot(shot)
            oo = prop.rtm_defaults()
            oo.nsnap = 3
            ref = {}
            for us in (True, False):
                prop.set_unsplit(us)
                prop.rtm_reset()
                prop.rtm_shot(oo, obs[s])
                ref[us] = prop.rtm_image(finalize=False)
        assert np.array_equal(imgs[s][1], ref[True][1]), s
        assert np.array_equal(imgs[s][0], ref[True][0]), (s, rel_l2(imgs[s][0], ref[True][0]))
        nx, nz = SIZES[s]                  # the oracle shot of _batch's shot s
        p = po.params_gpu2d(nx, nz, 10, 10, 10 + s, 1, 15 + 2 * s, 1, nx // 2 - s, 0, 2 + s % 2, cfg.nt, 1000)
        i32, _, _ = po.rtm(p, *m, sv_fwd=1, sv_bwd=0, illum_mode=0, obs=obs[s])
        assert np.array_equal(i32, ref[False][0]), s      # the same shot
        i64, _, _ = po.rtm(p, *m, sv_fwd=1, sv_bwd=0, illum_mode=0, obs=obs[s].astype(np.float64), prec="f64")
        e_us, e_sp = rel_l2(imgs[s][0], i64), rel_l2(ref[False][0], i64)
        print(f"shot {s}: unsplit vs FP64 {e_us:.2e}, split vs FP64 {e_sp:.2e}, "
              f"unsplit vs split {rel_l2(imgs[s][0], ref[False][0]):.2e}", flush=True)
        assert e_us <= max(1e-5, 1.5 * e_sp), s
