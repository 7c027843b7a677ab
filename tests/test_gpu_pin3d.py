"""The 3D kernels on every cell of the 512^3 benchmark grid, and the pins of
the point-source form, on the GPU.

The other full-size test starts from rest, so in its 40 steps the wave
covers a few cells around the source.  Here every field of the 572^3
padded grid is seeded with random values first, so that every cell, every
y chunk, every x tile and both faces of the absorbing layer are computed
from nonzero inputs: the split form bit for bit against the oracle (and the
reduced split against the oracle's mode of it), and, with a model and
fields symmetric under x <-> y, the split form's x <-> y symmetry bit for
bit (the y register queue of the kernels against their x stencil)."""
import numpy as np
import pytest

from conftest import rel_l2
from geophysics_amd import vti
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

F3 = ("u", "v", "w", "px", "py", "pz", "qx", "qy", "qz", "P", "Q")
C4 = np.array([1.196289, -0.0797526, 0.009570313, -0.0006975447], np.float32)


def _say(*a):
    print(*a, flush=True)


def _tr(a):
    return a.transpose(1, 0, 2)


def _symmetric_fields(nn, seed):
    """Seeds symmetric under x <-> y: u <-> v, px <-> py, qx <-> qy trade
    places, the rest are symmetric; values ~1e-3 (a few random arrays
    rolled along z, which keeps the symmetry)."""
    rng = np.random.default_rng(seed)
    a = (1e-3 * rng.standard_normal(nn)).astype(np.float32)
    b = a + _tr(a)
    out = {}
    for k, (x, y) in enumerate((("u", "v"), ("px", "py"), ("qx", "qy"))):
        f = np.roll(a, 3 * k + 1, axis=2)
        out[x] = f
        out[y] = np.ascontiguousarray(_tr(f))
    for k, name in enumerate(("w", "pz", "qz", "P", "Q")):
        out[name] = np.roll(b, 5 * k + 2, axis=2) * np.float32(1 + 0.25 * k)
    return out


def _model512(n):
    """Config-2-like smooth model, made symmetric under x <-> y."""
    z = np.arange(n, dtype=np.float32)
    f = (np.sin(0.021 * z + 0.3)[:, None, None] * np.sin(0.021 * z + 0.3)[None, :, None]
         * np.sin(0.013 * z + 1.1)[None, None, :]).astype(np.float32)
    g = (np.cos(0.017 * z)[:, None, None] + np.cos(0.017 * z)[None, :, None]) \
        * np.sin(0.011 * z)[None, None, :]
    depth = np.linspace(1800.0, 4200.0, n, dtype=np.float32)
    vp = (depth[None, None, :] + 150.0 * f).astype(np.float32)
    e = (0.12 + 0.05 * g).astype(np.float32)
    d = np.minimum(e, (0.06 + 0.02 * f).astype(np.float32))
    return vp, e, d


@pytest.mark.parametrize("unsplit", [False, True])
def test_config2_every_cell_seeded_parity_and_symmetry(unsplit):
    n, npml, nt = 512, 30, 4
    vp, e, d = _model512(n)
    assert np.array_equal(vp, _tr(vp)) and np.array_equal(e, _tr(e)) and np.array_equal(d, _tr(d))
    cfg = vti.cfg_ref3d(n, n, n, 10, 10, 10, npml, 1, 15, nt, 800)
    c = n // 2
    _say(f"512^3 seeded, unsplit={unsplit}")
    p = po.params_gpu3d(n, n, n, 10, 10, 10, npml, 1, 15, 0, c, 0, c, 0, 1, 0, nt, 800, 3, 2)
    p.unsplit = 1 if unsplit else 0
    sim = po.Sim(p, vp, e, d, "f32")
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_unsplit(unsplit)
        sh = vti.shot_ref3d(cfg, 0, c, 0, c, 0, 1, 0, rec_depth=3, rec_step=2)
        assert sh.sx == sh.sy and sh.sv_cx == sh.sv_cy
        prop.set_shot(sh)
        prop.reset()
        seeds = _symmetric_fields(prop.shape, 11)
        for k, a in seeds.items():
            prop.set_field(k, a)
            sim.set_field(k, a)
        del seeds
        prop.step(0, nt)
        rec = prop.record()
        ref = np.stack([sim.step(it) for it in range(nt)], axis=1)
        assert np.abs(rec).max() > 0
        assert np.array_equal(rec, ref), rel_l2(rec, ref)
        nr = sh.rec_nx
        r = rec.reshape(nr, nr, nt)
        if not unsplit:
            assert np.array_equal(r, r.transpose(1, 0, 2))
        got = {}
        for k in F3:
            f = prop.field(k)
            assert np.array_equal(f, sim.field(k)), k
            got[k] = f
            if k in ("w", "pz", "qz", "P", "Q"):
                if unsplit:
                    assert rel_l2(f, _tr(f)) <= 1e-6, k
                else:
                    assert np.array_equal(f, _tr(f)), k
        for a, b in (("u", "v"), ("px", "py"), ("qx", "qy")):
            if unsplit:
                assert rel_l2(got[a], _tr(got[b])) <= 1e-6, (a, b)
            else:
                assert np.array_equal(got[a], _tr(got[b])), (a, b)
    _say(f"512^3 seeded, unsplit={unsplit}: every field equals the oracle bit for bit; symmetric")


@pytest.mark.parametrize("ndim", [2, 3])
def test_point_source_one_step_closed_form_on_the_gpu(ndim):
    """One step from rest of the point-source form on the GPU: the source
    value a = amp[0] sits in P and Q before the velocity update, so u = -dt/dx
    c_m a on the m+1 cells before the source and +dt/dx c_m a on the m cells
    after it (v along y, w along z), zero elsewhere (add_source2d
    GPU_vti2dfd_kernels.cu:200-206 before update_vel2d, :423-425)."""
    if ndim == 2:
        vp = np.full((60, 50), 2500.0, np.float32)
        cfg = vti.cfg_ref2d(vti.MODE_GPU_REF, 60, 50, 10, 12, 10, 1, 20, 3, 1000)
        sh = vti.shot_ref2d(cfg, 1, 30, 0, 20)
    else:
        vp = np.full((28, 30, 26), 2500.0, np.float32)
        cfg = vti.cfg_ref3d(30, 28, 26, 10, 11, 12, 8, 1, 25, 3, 800)
        sh = vti.shot_ref3d(cfg, 0, 15, 0, 14, 0, 13, 0)
    e = np.full_like(vp, 0.1)
    d = np.full_like(vp, 0.05)
    with vti.Propagator(cfg, vp, e, d) as prop:
        prop.set_shot(sh)
        prop.reset()
        prop.step(0, 1)
        a = prop.srcamp()[0]
        F = {k: prop.field(k).ravel() for k in (("u", "w") if ndim == 2 else ("u", "v", "w"))}
        nnx, nny, nnz = prop.nnx, prop.nny, prop.nnz
    assert a != 0
    dt = np.float32(cfg.dt)
    center = sh.sz + sh.sx * nnz + (sh.sy * nnz * nnx if ndim == 3 else 0)
    arms = {"u": (cfg.dx, nnz), "w": (cfg.dz, 1)}
    if ndim == 3:
        arms["v"] = (cfg.dy, nnz * nnx)
    for k, (h, stride) in arms.items():
        dtk = np.float32(dt / np.float32(h))
        exp = np.zeros(F[k].size, np.float32)
        for m in range(4):
            t = np.float32(dtk * np.float32(C4[m] * a))
            exp[center - (m + 1) * stride] = -t
            exp[center + m * stride] = t
        assert np.array_equal(F[k], exp), k
