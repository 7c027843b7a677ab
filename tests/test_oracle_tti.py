"""The oracle's 2D TTI restatement (2D-tti-fd+rtm-cuda/
Toa_gpu_2dtti_fd_1orderfunction.cu update_vel :81-114, update_stress
:116-183), pinned without a golden (the reference has no CPU TTI path and its
.cu cannot be built here):

* tilt 0 reduces to the golden-pinned VTI GPU form (FP64, to association
  rounding) while the absorbing layer is still untouched;
* for linear fields the staggered derivatives are exact up to the
  coefficient sum K = sum_m (2m+1) c_m, so one step gives the rotated
  gradient / divergence of the TTI equations in closed form, per cell, for a
  random tilt field (velocity and stress updates separately);
* every field is damped in every layer by the product of its x and z
  factors (a constant field, no derivatives);
* planted mistakes (a sign of a sin term, cos/sin swapped, the VTI one-axis
  damping) fail these tests."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
C = np.array([1.196289, -0.0797526, 0.009570313, -0.0006975447])
K = float(np.sum(C * (2 * np.arange(4) + 1)))   # sum of the staggered stencil's weights times distance


def _tti(nx=60, nz=50, npml=8, nt=10, prec="f64", seed=0, vp=None, eps=None, dl=None):
    rng = np.random.default_rng(seed)
    p = po.params_tti2d(nx, nz, 5.0, 5.0, npml, 40.0, 10.0, nx // 2, nz // 2, nt, 0.0005)
    if vp is None:
        vp = rng.uniform(1800, 3200, (nx, nz)).astype(np.float32)
        eps = rng.uniform(0.0, 0.25, (nx, nz)).astype(np.float32)
        dl = rng.uniform(0.0, 0.1, (nx, nz)).astype(np.float32)
    s = po.Sim(p, vp, eps, dl, prec)
    return p, s, (vp, eps, dl)


def test_zero_tilt_is_the_vti_gpu_form():
    """Tilt 0: P, Q, u, w over the whole grid equal the VTI GPU form's (whose
    FP32 arithmetic is golden- and oracle-pinned) to FP64 association
    rounding, 20 steps from a centred source while the absorbing layer is
    still at rest (stencil reach 4 cells per step < 90 cells)."""
    nx = nz = 200
    npml, nt = 10, 20
    rng = np.random.default_rng(3)
    vp = rng.uniform(2000, 3000, (nx, nz)).astype(np.float32)
    eps = rng.uniform(0.0, 0.2, (nx, nz)).astype(np.float32)
    dl = rng.uniform(0.0, 0.1, (nx, nz)).astype(np.float32)
    pt = po.params_tti2d(nx, nz, 5.0, 5.0, npml, 40.0, 10.0, 100, 100, nt, 0.0005)
    pv = po.params_gpu2d(nx, nz, 5, 5, npml, 0, 40, 1, 100, 0, 100, nt, 500)
    st, sv = po.Sim(pt, vp, eps, dl, "f64"), po.Sim(pv, vp, eps, dl, "f64")
    st.set_theta(np.zeros((nx, nz), np.float32))
    assert np.array_equal(st.srcamp(), sv.srcamp())
    for it in range(nt):
        st.step(it)
        sv.step(it)
    for name in ("P", "Q", "u", "w"):
        a, b = st.field(name), sv.field(name)
        assert np.abs(b).max() > 0
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), name
    st.close()
    sv.close()


def _interior(s, p):
    """Cells whose damping factors are all 1 and whose stencils stay in the grid."""
    m = np.zeros((s.nnx, s.nnz), bool)
    m[p.npml:p.npml + p.nx, p.npml:p.npml + p.nz] = True
    return m


def test_velocity_update_rotates_the_pressure_gradient():
    """P = a x + b z, Q = c x + d z (cell units), at rest, no source: after
    one step u = -K (cos t dt/dx a - sin t dt/dz b) and
    w = -K (sin t dt/dx c + cos t dt/dz d) in every interior cell, t the
    cell's tilt (the TTI equations' rotated gradient)."""
    p, s, _ = _tti(seed=1)
    rng = np.random.default_rng(11)
    th = rng.uniform(-np.pi / 2, np.pi / 2, (p.nx, p.nz)).astype(np.float32)
    s.set_theta(th)
    s.set_srcamp(np.zeros(p.nt))
    x, z = np.meshgrid(np.arange(s.nnx, dtype=np.float64), np.arange(s.nnz, dtype=np.float64), indexing="ij")
    a, b, c, d = 0.7, -1.3, 2.1, 0.4
    s.set_field("P", a * x + b * z)
    s.set_field("Q", c * x + d * z)
    s.step(0)
    dtx = dtz = np.float64(np.float32(0.0005)) / 5.0
    thp = np.pad(th.astype(np.float64), p.npml, mode="edge")
    cs, sn = np.cos(thp), np.sin(thp)
    m = _interior(s, p)
    ue = -K * (cs * dtx * a - sn * dtz * b)
    we = -K * (sn * dtx * c + cs * dtz * d)
    u, w = s.field("u"), s.field("w")
    assert np.abs(u[m] - ue[m]).max() <= 1e-13 * np.abs(ue[m]).max()
    assert np.abs(w[m] - we[m]).max() <= 1e-13 * np.abs(we[m]).max()
    assert np.abs(sn[m]).max() > 0.5            # the tilt matters
    s.close()


def test_stress_update_rotates_the_velocity_divergence():
    """u = a x + b z, w = c x + d z, P = Q = 0 and split fields 0, no
    source: the velocity update leaves u, w (zero gradient of P, Q), then in
    every interior cell
      P = -vp^2 K [(1+2e)(cos dtx a - sin dtz b) + sqrt(1+2d)(sin dtx c + cos dtz d)]
      Q = -vp^2 K [sqrt(1+2d)(cos dtx a - sin dtz b) + (sin dtx c + cos dtz d)]."""
    p, s, (vp, eps, dl) = _tti(seed=2)
    rng = np.random.default_rng(12)
    th = rng.uniform(-np.pi / 2, np.pi / 2, (p.nx, p.nz)).astype(np.float32)
    s.set_theta(th)
    s.set_srcamp(np.zeros(p.nt))
    x, z = np.meshgrid(np.arange(s.nnx, dtype=np.float64), np.arange(s.nnz, dtype=np.float64), indexing="ij")
    a, b, c, d = -0.9, 1.7, 0.35, -2.2
    s.set_field("u", a * x + b * z)
    s.set_field("w", c * x + d * z)
    s.step(0)
    dtx = dtz = np.float64(np.float32(0.0005)) / 5.0
    pad = lambda f: np.pad(f.astype(np.float64), p.npml, mode="edge")
    thp, v2, e, dd = pad(th), pad(vp) ** 2, pad(eps), pad(dl)
    cs, sn = np.cos(thp), np.sin(thp)
    du = K * (cs * dtx * a - sn * dtz * b)
    dw = K * (sn * dtx * c + cs * dtz * d)
    Pe = -v2 * ((1 + 2 * e) * du + np.sqrt(1 + 2 * dd) * dw)
    Qe = -v2 * (np.sqrt(1 + 2 * dd) * du + dw)
    m = np.zeros_like(_interior(s, p))
    # u, w stay linear where undamped: the stencils of these cells stay there
    m[p.npml + 4:p.npml + p.nx - 4, p.npml + 4:p.npml + p.nz - 4] = True
    P, Q = s.field("P"), s.field("Q")
    assert np.abs(P[m] - Pe[m]).max() <= 1e-12 * np.abs(Pe[m]).max()
    assert np.abs(Q[m] - Qe[m]).max() <= 1e-12 * np.abs(Qe[m]).max()
    s.close()


def test_every_field_damped_by_both_axes():
    """u = w = px = pz = qx = qz = 1, P = Q = 0, vp = 0 (no stress
    increment), no source: after one step every field of every updated cell
    is its value times coff2(x) coff2(z) (velocity) or acoff2(x) acoff2(z)
    (stress) — in the z layer too, where the VTI form damps only w and pz,
    qz, and in the x layer, where it damps only u and px, qx."""
    p, s, _ = _tti(seed=4)
    s.set_theta(np.full((p.nx, p.nz), 0.4, np.float32))
    s.set_srcamp(np.zeros(p.nt))
    one = np.ones((s.nnx, s.nnz))
    for f in ("u", "w", "px", "pz", "qx", "qz"):
        s.set_field(f, one)
    s.set_field("vp", np.zeros((s.nnx, s.nnz)))
    s.step(0)
    cx2, cz2 = s.table("coffx2"), s.table("coffz2")
    ax2, az2 = s.table("acoffx2"), s.table("acoffz2")
    sl = np.s_[4:s.nnx - 4, 4:s.nnz - 4]
    gv = (cx2[:, None] * cz2[None, :])[sl]
    ga = (ax2[:, None] * az2[None, :])[sl]
    assert np.any(cz2 < 1) and np.any(cx2 < 1)
    for f in ("u", "w"):
        assert np.abs(s.field(f)[sl] - gv).max() <= 1e-15, f
    for f in ("px", "pz", "qx", "qz"):
        assert np.abs(s.field(f)[sl] - ga).max() <= 1e-15, f
    s.close()


@pytest.mark.parametrize("plant", [7, 8, 9])
def test_planted_mistakes_fail(plant):
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "mutant", "PLANT=%d" % plant], check=True)
    lib = os.path.join(odir, "_mut", "liboracle_vti_plant%d.so" % plant)
    env = dict(os.environ, ORC_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.abspath(__file__), "-k", "not planted"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout
