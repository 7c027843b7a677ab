"""Poynting-vector angle gathers: the oracle's restatement of Poynting_Adcigs
(Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:718-766, 3D
Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:618-660) pinned by closed forms
and by the physics of a flat reflector, and the product's host-built bin
thresholds against the reference's acosf binning."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2
from oracle import pyoracle as po

_libm = C.CDLL("libm.so.6")
_libm.acosf.restype = C.c_float
_libm.acosf.argtypes = [C.c_float]
_libm.cosf.restype = C.c_float
_libm.cosf.argtypes = [C.c_float]
PI = 3.141592653


def ref_bin(a, dangle):
    """Poynting_Adcigs :754-757 with the host libm's acosf."""
    ang = np.float32(0.5 * float(_libm.acosf(float(a))) * 180.0 / PI)
    return int(float(ang) / (dangle * 1.0))


def test_threshold_bins_equal_the_reference_binning():
    """The kernel bins a by counting thresholds a does not exceed: equal to
    the reference's (int)(0.5 acosf(a) 180 / pi / dangle) for random a and
    for the floats on both sides of every threshold."""
    from geophysics_amd import vti
    rng = np.random.default_rng(1)
    for nangle, dangle in ((60, 1), (30, 3), (91, 1), (7, 10)):
        thr, w = vti.adcig_tables(nangle, dangle)
        assert np.all(np.diff(thr) <= 0)
        a = list(rng.uniform(-1, 1, 3000).astype(np.float32)) + [np.float32(-1), np.float32(1), np.float32(0)]
        for t in thr:
            if t >= -1:
                a += [t, np.nextafter(t, np.float32(2)), np.nextafter(t, np.float32(-2))]
        for x in a:
            x = np.float32(x)
            if not (-1 <= x <= 1):
                continue
            mine = int(np.sum(x <= thr))
            assert mine == ref_bin(x, dangle) or (mine >= nangle and ref_bin(x, dangle) >= nangle), (x, dangle)
        assert np.array_equal(w, np.array([_libm.cosf(np.float32(ia * PI / 180.0)) for ia in range(nangle)],
                                          np.float32))


def _sim2d(nx=24, nz=18, npml=4):
    p = po.params_gpu2d(nx, nz, 10, 10, npml, 0, 20, 1, nx // 2, 0, 3, 10, 1000)
    vp = np.full((nx, nz), 2000.0, np.float32)
    z = np.zeros_like(vp)
    return p, po.Sim(p, vp, z, z, "f32"), po.Sim(p, vp, z, z, "f32")


def test_bins_of_plane_poynting_vectors_closed_form():
    """Source and receiver Poynting vectors of known directions in every
    cell: the gather of each cell holds cos^3(ia) in bin ia = floor(half the
    angle between the vectors, with the reference's 1/(1 - 0.1) cosine
    scale, / dangle), computed here in float64 from the angles alone, and
    nothing elsewhere; cells whose scaled cosine leaves [-1, 1] add nothing."""
    nangle, dangle = 60, 1
    p, s, g = _sim2d()
    nnx, nnz = s.nnx, s.nnz
    rng = np.random.default_rng(3)
    ths = rng.uniform(0, 2 * np.pi, (nnx, nnz))
    dth = rng.uniform(0, np.pi, (nnx, nnz))
    thg = ths + dth
    amp = rng.uniform(0.5, 2.0, (nnx, nnz))
    # S_s = (P u, Q w) = amp (cos ths, sin ths); S_g = (P u, Q w) = (cos thg, sin thg)
    s.set_field("P", amp); s.set_field("Q", amp)
    s.set_field("u", np.cos(ths)); s.set_field("w", np.sin(ths))
    g.set_field("P", np.ones_like(amp)); g.set_field("Q", np.ones_like(amp))
    g.set_field("u", np.cos(thg)); g.set_field("w", np.sin(thg))
    ad = np.zeros(po.adcig_shape(p, nangle), np.float32)
    po.adcig_accumulate(s, g, ad, nangle, dangle)
    ad = ad[0]                                     # (nx, nangle, nz)
    np_ = p.npml
    checked = 0
    for ix in range(p.nx):
        for iz in range(p.nz):
            c = (ix + np_, iz + np_)
            a = np.cos(dth[c]) / 0.9
            col = ad[ix, :, iz]
            if abs(abs(a) - 1) < 1e-5:
                continue                           # too close to the cut for float64 to decide
            if abs(a) > 1:
                assert not col.any()
                continue
            ang = 0.5 * np.degrees(np.arccos(a))
            if abs(ang / dangle - round(ang / dangle)) < 1e-3:
                continue                           # too close to a bin edge for float64 to decide
            ia = int(ang // dangle)
            w = np.float32(_libm.cosf(np.float32(ia * PI / 180.0)))
            exp = np.zeros(nangle, np.float32)
            if ia < nangle:
                exp[ia] = np.float32(amp[c]) * np.float32(1.0) * w * w * w
            assert np.array_equal(col, exp), (ix, iz, ia)
            checked += 1
    assert checked > 0.5 * p.nx * p.nz


def test_gathers_of_a_flat_reflector_are_flat():
    """RTM of one surface shot over a flat reflector: in the cdps around the
    source, the angle bins that hold energy below the shallow zone peak at
    the reflector depth (a flat gather across angles); with the source
    vector negated as in the reference (whose source field runs backward in
    time) they would not.  The stack equals the adjoint Laplacian of the
    float sum over angles, the thinned gathers every d-th gather."""
    nx, nz, npml, nt = 160, 80, 20, 900
    zr = 40
    vp = np.where(np.arange(nz)[None, :] < zr, 2000.0, 2800.0) * np.ones((nx, 1))
    vp = vp.astype(np.float32)
    z = np.zeros_like(vp)
    p = po.params_gpu2d(nx, nz, 10, 10, npml, 0, 25, 1, nx // 2, 0, 1, nt, 1000)
    sim = po.Sim(p, vp, z, z)
    vp1, e1 = float(sim.field("vp")[0, 1]), float(sim.field("eps")[0, 1])
    sim.close()
    mute = lambda r: po.mute2d(r, nt, p.dt, p.favg, p.dx, p.dz, nx // 2, 0, 1, 1, vp1, e1, 20)
    nangle = 60
    img, ill, _, ad = po.rtm(p, vp, z, z, sv_fwd=0, sv_bwd=0, illum_mode=0, mute=mute, nangle=nangle, dangle=1)
    ad = ad[0]
    w = ad[:, :, 10:nz - 5]                 # below the shallow zone of the source and receivers
    emax = np.abs(w).max()
    assert emax > 0
    sx = p.sx - npml
    flat = total = 0
    for ix in range(sx - 40, sx + 41):
        for ia in range(nangle):
            g = w[ix, ia]
            if np.abs(g).max() < 0.1 * emax:
                continue
            total += 1
            flat += abs(10 + int(np.argmax(np.abs(g))) - zr) <= 3
    assert total >= 100 and flat >= 0.95 * total, (flat, total)
    stk = po.adcig_stack(ad)
    ref = np.zeros((nx, nz), np.float32)
    for ia in range(nangle):
        ref = ref + ad[:, ia, :]
    assert np.array_equal(stk, po.laplace_adj(ref))
    thin = po.adcig_thin(ad, 7)
    assert thin.shape == ((nx + 6) // 7, nangle, nz)
    assert np.array_equal(thin, ad[::7])


def test_3d_gathers_reduce_to_2d_for_y_invariant_fields():
    """A 3D accumulation whose fields do not vary along y and have v = 0
    (S_y = 0) equals the 2D accumulation of the same x-z fields, in every
    decimated cdp."""
    nangle, dangle, dcdp = 45, 2, 3
    rng = np.random.default_rng(8)
    p2, s2, g2 = _sim2d(nx=18, nz=14, npml=4)
    nnx, nnz = s2.nnx, s2.nnz
    p3 = po.params_gpu3d(18, 12, 14, 10, 10, 10, 4, 0, 20, 0, 9, 0, 6, 0, 3, 0, 10, 1000, 1, 1)
    vp3 = np.full((12, 18, 14), 2000.0, np.float32)
    z3 = np.zeros_like(vp3)
    s3, g3 = po.Sim(p3, vp3, z3, z3), po.Sim(p3, vp3, z3, z3)
    nny = s3.nny
    for name in ("P", "Q", "u", "w"):
        for s2_, s3_ in ((s2, s3), (g2, g3)):
            f = rng.standard_normal((nnx, nnz)).astype(np.float32)
            s2_.set_field(name, f)
            s3_.set_field(name, np.broadcast_to(f[None], (nny, nnx, nnz)))
    a2 = np.zeros(po.adcig_shape(p2, nangle), np.float32)
    po.adcig_accumulate(s2, g2, a2, nangle, dangle)
    a3 = np.zeros(po.adcig_shape(p3, nangle, dcdp), np.float32)
    po.adcig_accumulate(s3, g3, a3, nangle, dangle, dcdp)
    assert np.abs(a2).max() > 0
    for iya in range(a3.shape[0]):
        assert np.array_equal(a3[iya], a2[0][::dcdp][: a3.shape[1]]), iya
