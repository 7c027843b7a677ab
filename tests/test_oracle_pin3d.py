"""Pins of the oracle's 3D arithmetic and of the point-source form that do
not go through a restatement of the same formulas.

The reference has no 3D CPU code and its .cu files cannot be built here, so
the 3D stencil of the oracle (update_vel3d GPU_vti3dfd_kernels.cu:90-129,
update_stress3d :131-201) is tied to arithmetic the reference's own CPU
path fixes, and to symmetries of the equations:

* a field that does not vary along y makes every y-difference exactly 0, so
  the 3D update must reproduce the 2D oracle (pinned bit for bit to the
  compiled CPU_vti2dfd_kernels.c by test_oracle_golden.py) on every y-plane
  that the faces' influence (8 cells per step) has not reached;
* the same with a field constant along x ties the y arithmetic (offsets,
  dt/dy, y tables, the q-system's sqrt(1 + 2 delta)) to the 2D x arithmetic;
* with nx = ny, dx = dy, a model symmetric under x <-> y and a source on the
  diagonal, the fields are symmetric bit for bit (IEEE addition commutes;
  px and py trade places);
* one step from rest of the point-source form (add_source2d
  GPU_vti2dfd_kernels.cu:200-206, before update_vel2d, :423-425) has a
  closed form: u, (v,) w are nonzero only on the 8 cells of each stencil
  arm, +-dt/dx c_m a.

test_planted_mistakes_fail rebuilds the oracle with each mistake of
vti_oracle.c's ORC_PLANT list and checks that these tests then fail.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, layered_model, rel_l2
from oracle import pyoracle as po

C4 = np.array([1.196289, -0.0797526, 0.009570313, -0.0006975447], np.float32)
F2 = ("u", "w", "px", "pz", "qx", "qz", "P", "Q")
F3 = ("u", "v", "w", "px", "py", "pz", "qx", "qy", "qz", "P", "Q")


def _set_axis_tables(dst, axis, src, src_axis):
    """dst's damping tables along `axis` := src's along `src_axis`."""
    for pre in ("coff", "acoff"):
        for k in "12":
            dst.set_table(pre + axis + k, src.table(pre + src_axis + k))


def _reduction_case(prec, invariant_axis, nt=5):
    """2D oracle (CPU form: the golden-pinned tables) against a 3D oracle
    whose fields are constant along `invariant_axis` ('y' or 'x')."""
    rng = np.random.default_rng(7 if invariant_axis == "y" else 8)
    nx2, nz2, npd = 40, 36, 8
    dx2, dz2 = 10, 12
    p2 = po.params_cpu2d(nx2, nz2, dx2, dz2, npd, 0, 20, 1, 20, 0, 10, nt, 500)
    vp2, e2, d2 = layered_model(nx2, nz2, seed=3, elo=0.05, ehi=0.2, dlo=0.0, dhi=0.1)
    s2 = po.Sim(p2, vp2, e2, d2, prec)
    po.set_source(s2, np.zeros(nt))
    nnx2, nnz2 = s2.nnx, s2.nnz
    seed = {k: rng.standard_normal((nnx2, nnz2)).astype(np.float32) for k in F2}
    for k, a in seed.items():
        s2.set_field(k, a)
    margin = 8 * nt + 4
    nlong = 2 * margin + 16 - 2 * npd          # interior cells along the invariant axis
    if invariant_axis == "y":
        nx3, ny3, dx3, dy3 = nx2, nlong, dx2, 11
        vp3 = np.broadcast_to(vp2[None], (ny3, nx2, nz2))
        e3 = np.broadcast_to(e2[None], vp3.shape)
        d3 = np.broadcast_to(d2[None], vp3.shape)
    else:
        nx3, ny3, dx3, dy3 = nlong, nx2, 9, dx2
        vp3 = np.broadcast_to(vp2[:, None], (nx2, nx3, nz2))
        e3 = np.broadcast_to(e2[:, None], vp3.shape)
        d3 = np.broadcast_to(d2[:, None], vp3.shape)
    p3 = po.params_gpu3d(nx3, ny3, nz2, dx3, dy3, dz2, npd, 0, 20, 0, 10, 0, 10, 0, 10, 0,
                         nt, 500, 1, 1)
    s3 = po.Sim(p3, vp3, e3, d3, prec)
    po.set_source(s3, np.zeros(nt))
    nny3, nnx3 = s3.nny, s3.nnx
    # the 2D x axis maps to 3D x (y-invariant) or to 3D y (x-invariant)
    a3 = "x" if invariant_axis == "y" else "y"
    _set_axis_tables(s3, a3, s2, "x")
    _set_axis_tables(s3, "z", s2, "z")
    if invariant_axis == "y":
        names = {"u": "u", "w": "w", "px": "px", "pz": "pz", "qx": "qx", "qz": "qz", "P": "P", "Q": "Q"}
        zero = ("v", "py", "qy")
        for k2, k3 in names.items():
            s3.set_field(k3, np.broadcast_to(seed[k2][None], (nny3, nnx3, nnz2)))
    else:
        names = {"u": "v", "w": "w", "px": "py", "pz": "pz", "qx": "qy", "qz": "qz", "P": "P", "Q": "Q"}
        zero = ("u", "px", "qx")
        for k2, k3 in names.items():
            s3.set_field(k3, np.broadcast_to(seed[k2][:, None], (nny3, nnx3, nnz2)))
    for it in range(nt):
        s2.step(it, rec=False)
        s3.step(it, rec=False)
    sl = slice(margin, (nny3 if invariant_axis == "y" else nnx3) - margin)
    assert sl.stop - sl.start >= 8
    for k2, k3 in names.items():
        f3 = s3.field(k3)
        f3 = f3[sl] if invariant_axis == "y" else f3[:, sl].transpose(1, 0, 2)
        ref = s2.field(k2)
        assert np.abs(ref).max() > 0
        for j in range(f3.shape[0]):
            assert np.array_equal(f3[j], ref), (k3, j)
    for k in zero:
        f3 = s3.field(k)
        f3 = f3[sl] if invariant_axis == "y" else f3[:, sl]
        assert not f3.any(), k


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_y_invariant_3d_equals_golden_pinned_2d(prec):
    _reduction_case(prec, "y")


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_x_invariant_3d_y_axis_equals_2d_x_axis(prec):
    _reduction_case(prec, "x")


def _sym(rng, shape):
    a = rng.standard_normal(shape).astype(np.float32)
    return a + a.transpose(1, 0, 2)


def _tr(a):
    return a.transpose(1, 0, 2)


@pytest.mark.parametrize("unsplit", [0, 1])
def test_xy_transpose_symmetry(unsplit):
    """nx = ny, dx = dy, vp/eps/delta symmetric under x <-> y, source and SV
    taper on the diagonal, initial fields symmetric (u <-> v, px <-> py,
    qx <-> qy), through the absorbing layer: in the split form (the
    reference's arithmetic) every field stays symmetric bit for bit.  The
    reduced-split mode subtracts px before py on the columns where both x
    and y are damped, which does not commute in rounding: there the fields
    are symmetric to 1e-6 rel-L2."""
    rng = np.random.default_rng(5)
    n, nz, npml, nt = 40, 30, 8, 30
    p = po.params_gpu3d(n, n, nz, 10, 10, 12, npml, 1, 25, 0, 20, 0, 20, 0, 3, 0, nt, 600, 2, 3)
    p.unsplit = unsplit
    assert p.sx == p.sy and p.sv_cx == p.sv_cy
    base = np.linspace(1800, 3000, nz, dtype=np.float32)
    vp = (base[None, None, :] + 60.0 * _sym(rng, (n, n, nz))).astype(np.float32)
    e = (0.1 + 0.02 * _sym(rng, (n, n, nz))).astype(np.float32)
    d = np.minimum(e, (0.05 + 0.01 * _sym(rng, (n, n, nz))).astype(np.float32))
    sim = po.Sim(p, vp, e, d, "f32")
    nn = (sim.nny, sim.nnx, sim.nnz)
    for a, b in (("u", "v"), ("px", "py"), ("qx", "qy")):
        f = (1e-3 * rng.standard_normal(nn)).astype(np.float32)
        sim.set_field(a, f)
        sim.set_field(b, _tr(f))
    for k in ("w", "pz", "qz", "P", "Q"):
        sim.set_field(k, 1e-3 * _sym(rng, nn))
    rec = sim.run()
    nr = p.rec_nx
    assert p.rec_ny == nr
    r = rec.reshape(nr, nr, nt)                       # [iry][irx][it]
    assert np.abs(r).max() > 0
    F = {k: sim.field(k) for k in F3}
    pairs = [(r, r.transpose(1, 0, 2), "record")]
    pairs += [(F[k], _tr(F[k]), k) for k in ("w", "pz", "qz", "P", "Q")]
    pairs += [(F[a], _tr(F[b]), a + "/" + b) for a, b in (("u", "v"), ("px", "py"), ("qx", "qy"))]
    for x, y, k in pairs:
        if unsplit:
            assert rel_l2(x, y) <= 1e-6, k
        else:
            assert np.array_equal(x, y), k


def _closed_form_arm(dtx, a, center, stride, ntot):
    """Expected one-step field along one stencil arm: cells center-m-1 carry
    -(dt/dx (c_m a)), cells center+m carry +dt/dx (c_m a)."""
    exp = np.zeros(ntot, np.float32)
    for m in range(4):
        t = np.float32(dtx * np.float32(C4[m] * a))
        exp[center - (m + 1) * stride] = -t
        exp[center + m * stride] = t
    return exp


@pytest.mark.parametrize("ndim", [2, 3])
def test_point_source_one_step_closed_form(ndim):
    """Point-source form, one step from rest: the source value a = amp[0]
    (the Ricker gain, equal to the golden-pinned CPU form's) is in P and Q
    before the velocity update, so u = -dt/dx c_m a on the cells m+1 before
    the source and +dt/dx c_m a on the cells m after it (v along y, w along
    z with dt/dz), every other velocity sample 0; the stress update then
    rebuilds P from the split fields, so a itself does not stay in P."""
    if ndim == 2:
        p = po.params_gpu2d(60, 50, 10, 12, 10, 1, 20, 1, 30, 0, 20, 3, 1000)
        pc = po.params_cpu2d(60, 50, 10, 12, 10, 1, 20, 1, 30, 0, 20, 3, 1000)
        vp, e, d = layered_model(60, 50, seed=1)
    else:
        p = po.params_gpu3d(30, 28, 26, 10, 11, 12, 8, 1, 25, 0, 15, 0, 14, 0, 13, 0, 3, 800, 1, 1)
        pc = None
        rng = np.random.default_rng(2)
        vp = (2500 + 300 * rng.random((28, 30, 26))).astype(np.float32)
        e = np.full_like(vp, 0.1)
        d = np.full_like(vp, 0.05)
    sim = po.Sim(p, vp, e, d, "f32")
    amp = sim.srcamp()
    if pc is not None:
        # same Ricker gains as the CPU form (t = dt first, pfac 10), whose
        # Gaussian source the golden vectors pin
        assert np.array_equal(amp, po.Sim(pc, vp, e, d, "f32").srcamp())
    a = amp[0]
    assert a != 0
    for k in ("coffx1", "coffz1") + (("coffy1",) if ndim == 3 else ()):
        t = sim.table(k)
        idx = {"x": p.sx, "y": p.sy, "z": p.sz}[k[4]]
        assert np.all(t[idx - 4:idx + 5] == 1)       # the arms lie outside the absorbing layer
    P0 = sim.field("P")
    assert not P0.any()
    sim.step(0, rec=False)
    dt = np.float32(p.dt)
    shape = (sim.nny, sim.nnx, sim.nnz) if ndim == 3 else (sim.nnx, sim.nnz)
    ntot = int(np.prod(shape))
    sz, sxs = 1, sim.nnz
    center = p.sz + p.sx * sim.nnz + (p.sy * sim.nnz * sim.nnx if ndim == 3 else 0)
    arms = {"u": (np.float32(dt / np.float32(p.dx)), sxs), "w": (np.float32(dt / np.float32(p.dz)), sz)}
    if ndim == 3:
        arms["v"] = (np.float32(dt / np.float32(p.dy)), sim.nnz * sim.nnx)
    for k, (dtk, stride) in arms.items():
        exp = _closed_form_arm(dtk, a, center, stride, ntot)
        got = sim.field(k).ravel()
        assert np.array_equal(got, exp), k
    P = sim.field("P").ravel()
    assert P[center] != a and P[center] == sim.field("px").ravel()[center] + sim.field("pz").ravel()[center] \
        + (sim.field("py").ravel()[center] if ndim == 3 else 0)


def test_cpu_form_source_one_step_closed_form():
    """CPU form (ptsource CPU_vti2dfd_kernels.c:50-76, added in the stress
    update :173-174): after one step from rest u = w = 0 and P = Q = the
    7x7 footprint pfac w(t) exp(-x^2 - z^2) exactly."""
    p = po.params_cpu2d(60, 50, 10, 12, 10, 1, 20, 1, 30, 0, 20, 3, 1000)
    vp, e, d = layered_model(60, 50, seed=1)
    sim = po.Sim(p, vp, e, d, "f32")
    a = sim.srcamp()[0]
    sim.step(0, rec=False)
    assert not sim.field("u").any() and not sim.field("w").any()
    exp = np.zeros((sim.nnx, sim.nnz), np.float32)
    for dz in range(-3, 4):
        for dx in range(-3, 4):
            exp[p.sx + dx, p.sz + dz] = np.float32(np.float64(a) * np.exp(-dz * dz - dx * dx))
    assert np.array_equal(sim.field("P"), exp)
    assert np.array_equal(sim.field("Q"), exp)


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_fail(plant):
    """Each planted mistake of the oracle (vti_oracle.c ORC_PLANT) makes at
    least one test of this file fail; the unplanted oracle passes them (the
    tests above)."""
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "mutant", "PLANT=%d" % plant], check=True)
    lib = os.path.join(odir, "_mut", "liboracle_vti_plant%d.so" % plant)
    env = dict(os.environ, ORC_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.abspath(__file__), "-k", "not planted"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout
