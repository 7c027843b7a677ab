"""ctypes binding of liboracle_vti.so (the CPU restatement, vti_oracle.c).

TEST INFRASTRUCTURE ONLY: used as the checker by tests/, smoke() and the
cpu_baseline leg of bench.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DEFAULT_LIB = os.path.join(_HERE, "liboracle_vti.so")
# ORC_LIB names another build of the same source (a build with a planted
# mistake, for checking that the tests can fail); it is never built here.
_LIB = os.environ.get("ORC_LIB") or _DEFAULT_LIB


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("ndim", "nx", "ny", "nz", "npml")] + \
        [(n, C.c_float) for n in ("dx", "dy", "dz", "dt")] + \
        [("nt", C.c_int), ("favg", C.c_float), ("pfac", C.c_float)] + \
        [(n, C.c_int) for n in (
            "wtype", "d0_kind", "acoff_kind", "src_kind", "t0_kind",
            "sx", "sy", "sz", "sv", "sv_kind", "sv_R", "sv_r",
            "sv_cx", "sv_cy", "sv_cz",
            "rec_z", "rec_x0", "rec_dx", "rec_nx", "rec_y0", "rec_dy", "rec_ny",
            "rec_kind", "omp_form", "unsplit", "tti")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


FIELDS = {n: i for i, n in enumerate(
    ["u", "v", "w", "px", "py", "pz", "qx", "qy", "qz", "P", "Q", "vp", "eps", "del"])}
TABLES = {n: 20 + i for i, n in enumerate(
    ["coffx1", "coffx2", "coffy1", "coffy2", "coffz1", "coffz2",
     "acoffx1", "acoffx2", "acoffy1", "acoffy2", "acoffz1", "acoffz2"])}
SRCAMP, D0 = 40, 41
FIELDS.update({"cth": 42, "sth": 43})   # TTI: cos / sin of the padded tilt

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB) and _LIB == _DEFAULT_LIB:
            build()
        L = C.CDLL(_LIB)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        pp = C.POINTER(OrcParams)
        for sfx, rp in (("_f32", fp), ("_f64", dp)):
            getattr(L, "orc_create" + sfx).restype = C.c_void_p
            getattr(L, "orc_create" + sfx).argtypes = [pp, fp, fp, fp]
            getattr(L, "orc_step" + sfx).argtypes = [C.c_void_p, C.c_int, rp]
            getattr(L, "orc_run" + sfx).argtypes = [C.c_void_p, rp]
            getattr(L, "orc_get" + sfx).argtypes = [C.c_void_p, C.c_int, rp]
            getattr(L, "orc_set" + sfx).argtypes = [C.c_void_p, C.c_int, rp]
            getattr(L, "orc_destroy" + sfx).argtypes = [C.c_void_p]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_ncell.restype = C.c_long
        L.orc_ncell.argtypes = [C.c_void_p]
        L.orc_params_gpu2d.argtypes = [pp] + [C.c_int] * 14
        L.orc_params_cpu2d.argtypes = [pp] + [C.c_int] * 14
        L.orc_params_gpu3d.argtypes = [pp] + [C.c_int] * 20
        L.orc_params_omp2d.argtypes = [pp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                       C.c_float, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int]
        L.orc_params_tti2d.argtypes = [pp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float,
                                       C.c_float, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int]
        for sfx in ("_f32", "_f64"):
            getattr(L, "orc_set_theta" + sfx).argtypes = [C.c_void_p, fp]
        L.orc_mute2d.argtypes = [C.c_int, C.c_int] + [C.c_float] * 4 + [C.c_int] * 4 + \
            [C.c_float, C.c_float, fp, C.c_int, C.c_int, C.c_int]
        L.orc_mute3d.argtypes = [C.c_int] * 3 + [C.c_float] * 5 + [C.c_int] * 6 + \
            [C.c_float, C.c_float, fp] + [C.c_int] * 6
        for sfx, rp in (("_f32", fp), ("_f64", dp)):
            getattr(L, "orc_rtm" + sfx).argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 5 + [rp] * 4
        L.orc_laplace_adj.argtypes = [C.c_int] * 3 + [fp, fp]
        for sfx, rp in (("_f32", fp), ("_f64", dp)):
            getattr(L, "orc_rtm_ex" + sfx).argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 5 + [rp] * 4 + \
                [C.c_int] * 3 + [rp]
            getattr(L, "orc_adcig" + sfx).argtypes = [C.c_void_p, C.c_void_p] + [C.c_int] * 3 + [rp]
        L.orc_adcig_stack.argtypes = [C.c_int] * 3 + [fp, fp]
        L.orc_adcig_thin.argtypes = [C.c_int] * 4 + [fp, fp]
        for sfx, rp in (("_f32", fp), ("_f64", dp)):
            getattr(L, "orc_fwi_grad" + sfx).argtypes = [C.c_void_p] + [rp] * 6
            getattr(L, "orc_set_source" + sfx).argtypes = [C.c_void_p, rp]
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's loops; returns the number in effect."""
    return int(lib().orc_set_threads(int(n)))


def library_path():
    return _LIB


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def params_gpu2d(nx, nz, dx, dz, npd, SV, favg, is_, fs, ds, zs, nt, dt_us, rec_depth=1):
    p = OrcParams()
    lib().orc_params_gpu2d(C.byref(p), nx, nz, dx, dz, npd, SV, favg, is_, fs, ds, zs,
                           nt, dt_us, rec_depth)
    return p


def params_cpu2d(nx, nz, dx, dz, npd, SV, favg, is_, fs, ds, zs, nt, dt_us, rec_depth=1):
    p = OrcParams()
    lib().orc_params_cpu2d(C.byref(p), nx, nz, dx, dz, npd, SV, favg, is_, fs, ds, zs,
                           nt, dt_us, rec_depth)
    return p


def params_omp2d(nx, nz, dx, dz, npd, favg, pfac, sx, sz, nt, dt, rec_depth=1):
    """Toafd2domp_2.c's main (reference values: 601, 301, 5.0, 5.0, 0, 30,
    1000, 100, 100, 4001, 0.0005)."""
    p = OrcParams()
    lib().orc_params_omp2d(C.byref(p), nx, nz, dx, dz, npd, favg, pfac, sx, sz, nt, dt, rec_depth)
    return p


def params_tti2d(nx, nz, dx, dz, npml, favg, pfac, fs, zs, nt, dt, SV=0, rec_depth=1):
    """main of Toa_gpu_2dtti_fd_1orderfunction.cu (reference values: 600,
    300, 5.0, 5.0, 50, 40, 10, 200, 140, 801, 0.0005, SV off); the tilt is
    set on the Sim with set_theta."""
    p = OrcParams()
    lib().orc_params_tti2d(C.byref(p), nx, nz, dx, dz, npml, favg, pfac, fs, zs, nt, dt, SV, rec_depth)
    return p


def params_gpu3d(nx, ny, nz, dx, dy, dz, npml, SV, favg, is_, fsx, dsx, fsy, dsy,
                 fsz, dsz, nt, dt_us, rec_depth=1, rec_step=1):
    p = OrcParams()
    lib().orc_params_gpu3d(C.byref(p), nx, ny, nz, dx, dy, dz, npml, SV, favg, is_,
                           fsx, dsx, fsy, dsy, fsz, dsz, nt, dt_us, rec_depth, rec_step)
    return p


class Sim:
    """One shot of the oracle engine in float32 ('f32') or float64 ('f64')."""

    def __init__(self, params, vp, eps, dl, prec="f32"):
        self.p = params
        self.prec = prec
        self.dtype = np.float32 if prec == "f32" else np.float64
        L = lib()
        vp = np.ascontiguousarray(vp, dtype=np.float32)
        eps = np.ascontiguousarray(eps, dtype=np.float32)
        dl = np.ascontiguousarray(dl, dtype=np.float32)
        self._keep = (vp, eps, dl)
        self.h = getattr(L, "orc_create_" + prec)(C.byref(params), _fptr(vp), _fptr(eps), _fptr(dl))
        if not self.h:
            raise MemoryError("orc_create failed")
        np_ = params.npml
        self.nnx = params.nx + 2 * np_
        self.nnz = params.nz + 2 * np_
        self.nny = params.ny + 2 * np_ if params.ndim == 3 else 1
        self.nrec = params.rec_nx * (params.rec_ny if params.ndim == 3 else 1)

    def _ptr(self, a):
        return _fptr(a) if self.prec == "f32" else _dptr(a)

    def run(self):
        shot = np.zeros((self.nrec, self.p.nt), dtype=self.dtype)
        getattr(lib(), "orc_run_" + self.prec)(self.h, self._ptr(shot))
        return shot

    def step(self, it, rec=True):
        row = np.zeros(self.nrec, dtype=self.dtype) if rec else None
        getattr(lib(), "orc_step_" + self.prec)(self.h, it, self._ptr(row) if rec else None)
        return row

    def field(self, name):
        out = np.zeros(self.nny * self.nnx * self.nnz, dtype=self.dtype)
        getattr(lib(), "orc_get_" + self.prec)(self.h, FIELDS[name], self._ptr(out))
        shape = (self.nny, self.nnx, self.nnz) if self.p.ndim == 3 else (self.nnx, self.nnz)
        return out.reshape(shape)

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).ravel()
        assert a.size == self.nny * self.nnx * self.nnz
        getattr(lib(), "orc_set_" + self.prec)(self.h, FIELDS[name], self._ptr(a))

    def set_theta(self, theta):
        """TTI tilt of the interior (radians, (nx, nz))."""
        a = np.ascontiguousarray(theta, dtype=np.float32)
        assert a.shape == (self.p.nx, self.p.nz)
        if getattr(lib(), "orc_set_theta_" + self.prec)(self.h, _fptr(a)) != 0:
            raise ValueError("orc_set_theta: not a TTI simulation")

    def table(self, name):
        n = max(self.nnx, self.nny, self.nnz)
        out = np.zeros(n, dtype=self.dtype)
        k = getattr(lib(), "orc_get_" + self.prec)(self.h, TABLES[name], self._ptr(out))
        return out[:k]

    def set_table(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.size == self.table(name).size
        getattr(lib(), "orc_set_" + self.prec)(self.h, TABLES[name], self._ptr(a))

    def set_srcamp(self, amp):
        a = np.ascontiguousarray(amp, dtype=self.dtype)
        assert a.size == self.p.nt
        getattr(lib(), "orc_set_" + self.prec)(self.h, SRCAMP, self._ptr(a))

    def srcamp(self):
        out = np.zeros(self.p.nt, dtype=self.dtype)
        getattr(lib(), "orc_get_" + self.prec)(self.h, SRCAMP, self._ptr(out))
        return out

    def d0(self):
        out = np.zeros(1, dtype=self.dtype)
        getattr(lib(), "orc_get_" + self.prec)(self.h, D0, self._ptr(out))
        return float(out[0])

    def close(self):
        if self.h:
            getattr(lib(), "orc_destroy_" + self.prec)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mute2d(shot, nt, dt, favg, dx, dz, fs, ds, zs, is_, vp1, eps1, tt=100, x0=0, rdx=1):
    """shot: float32 (nrec, nt), muted in place (mute_directwave2d)."""
    assert shot.dtype == np.float32 and shot.flags.c_contiguous
    nrec = shot.shape[0]
    lib().orc_mute2d(nrec, nt, dt, favg, dx, dz, fs, ds, zs, is_, vp1, eps1, _fptr(shot), tt, x0, rdx)
    return shot


def mute3d(shot, nrx, nry, nt, dt, favg, dx, dy, dz, fsx, fsy, dsx, dsy, fsz, is_, vp1, eps1, tt=100,
           window=0, x0=0, rdx=1, y0=0, rdy=1):
    assert shot.dtype == np.float32 and shot.flags.c_contiguous
    lib().orc_mute3d(nrx, nry, nt, dt, favg, dx, dy, dz, fsx, fsy, dsx, dsy, fsz, is_,
                     vp1, eps1, _fptr(shot), tt, window, x0, rdx, y0, rdy)
    return shot


def laplace_adj(img):
    """laplace_filter(adj=1) of an interior image shaped (nx, nz) or (ny, nx, nz)."""
    a = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros_like(a)
    if a.ndim == 2:
        nx, nz, ny = a.shape[0], a.shape[1], 1
    else:
        ny, nx, nz = a.shape
    lib().orc_laplace_adj(nz, nx, ny, _fptr(a), _fptr(out))
    return out


def adcig_shape(params, nangle, dcdp=1):
    """(nya, nxa, nangle, nz) of the gathers: every cdp in 2D, every dcdp-th in 3D."""
    if params.ndim == 3:
        return (params.ny // dcdp, params.nx // dcdp, nangle, params.nz)
    return (1, params.nx, nangle, params.nz)


def adcig_accumulate(src, rcv, adcigs, nangle, dangle, dcdp=1):
    """One Poynting_Adcigs accumulation from the current fields of two Sims
    (adcigs: array of adcig_shape, updated in place)."""
    assert adcigs.dtype == src.dtype and adcigs.flags.c_contiguous
    getattr(lib(), "orc_adcig_" + src.prec)(src.h, rcv.h, nangle, dangle, dcdp, src._ptr(adcigs))
    return adcigs


def adcig_stack(adcigs):
    """stk_adcigs: (nx, nangle, nz) gathers -> Laplacian-filtered stack (nx, nz)."""
    a = np.ascontiguousarray(adcigs, dtype=np.float32)
    nx, na, nz = a.shape
    out = np.zeros((nx, nz), np.float32)
    lib().orc_adcig_stack(nx, nz, na, _fptr(a), _fptr(out))
    return out


def adcig_thin(adcigs, d):
    """adcigs_smiled: the gathers of every d-th cdp."""
    a = np.ascontiguousarray(adcigs, dtype=np.float32)
    nx, na, nz = a.shape
    out = np.zeros(((nx + d - 1) // d, na, nz), np.float32)
    lib().orc_adcig_thin(nx, nz, na, d, _fptr(a), _fptr(out))
    return out


def rtm(params, vp, eps, dl, sv_fwd=1, sv_bwd=0, inject_z=None, illum_mode=0, image_every=1,
        obs=None, mute=None, prec="f32", unsplit_source=False, unsplit_receiver=False,
        nangle=0, dangle=1, dcdp=1, theta=None):
    """One-shot RTM with every forward snapshot kept (no checkpointing).
    theta: the TTI tilt (radians, (nx, nz)) for tti params.

    mute: callable(modelled_record) applied in place before it is used as
    data when obs is None.  unsplit_source / unsplit_receiver: the source /
    receiver simulation in the unsplit interior mode.  Returns (img, illum,
    modelled) shaped like the interior ((nx, nz) or (ny, nx, nz)) and
    (nrec, nt)."""
    ps = OrcParams.from_buffer_copy(params)
    ps.unsplit = 1 if unsplit_source else 0
    pg = OrcParams.from_buffer_copy(params)
    pg.unsplit = 1 if unsplit_receiver else 0
    s = Sim(ps, vp, eps, dl, prec)
    g = Sim(pg, vp, eps, dl, prec)
    if theta is not None:
        s.set_theta(theta)
        g.set_theta(theta)
    dt = s.dtype
    shape = (params.ny, params.nx, params.nz) if params.ndim == 3 else (params.nx, params.nz)
    img = np.zeros(shape, dtype=dt)
    illum = np.zeros(shape, dtype=dt)
    modelled = np.zeros((s.nrec, params.nt), dtype=dt)
    iz = params.rec_z if inject_z is None else inject_z
    if obs is None and mute is not None:
        # the mute applies to the modelled record between the passes: run the
        # forward once to get it, then migrate it as observed data
        tmp = Sim(params, vp, eps, dl, prec)
        if theta is not None:
            tmp.set_theta(theta)
        rec = tmp.run()
        tmp.close()
        mute(rec)
        obs = rec
    o = None if obs is None else np.ascontiguousarray(obs, dtype=dt)
    fn = getattr(lib(), "orc_rtm_ex_" + prec)
    ad = np.zeros(adcig_shape(params, nangle, dcdp), dtype=dt) if nangle > 0 else None
    fn(s.h, g.h, sv_fwd, sv_bwd, iz, illum_mode, image_every,
       s._ptr(o) if o is not None else None, s._ptr(modelled), s._ptr(img), s._ptr(illum),
       nangle, dangle, dcdp, s._ptr(ad) if ad is not None else None)
    s.close()
    g.close()
    if nangle > 0:
        return img, illum, modelled, ad
    return img, illum, modelled


def copy_tables(dst, src):
    """Give Sim dst the damping tables of Sim src (fixed absorbing layer)."""
    for name in TABLES:
        t = np.zeros(dst.nny * dst.nnx * dst.nnz, dtype=dst.dtype)
        k = getattr(lib(), "orc_get_" + src.prec)(src.h, TABLES[name], src._ptr(t))
        if k > 0:
            getattr(lib(), "orc_set_" + dst.prec)(dst.h, TABLES[name], dst._ptr(t))


def fwi_grad(sim, dobs=None, resid=None):
    """Misfit, vp-gradient (interior shape), source gradient (nt), record.
    sim must be fresh or is reset (its fields are overwritten)."""
    p = sim.p
    shape = (p.ny, p.nx, p.nz) if p.ndim == 3 else (p.nx, p.nz)
    dt = sim.dtype
    J = np.zeros(1, dtype=dt)
    g = np.zeros(shape, dtype=dt)
    gs = np.zeros(p.nt, dtype=dt)
    d = np.zeros((sim.nrec, p.nt), dtype=dt)
    o = None if dobs is None else np.ascontiguousarray(dobs, dtype=dt)
    r = None if resid is None else np.ascontiguousarray(resid, dtype=dt)
    getattr(lib(), "orc_fwi_grad_" + sim.prec)(
        sim.h, sim._ptr(o) if o is not None else None, sim._ptr(r) if r is not None else None,
        sim._ptr(J), sim._ptr(g), sim._ptr(gs), sim._ptr(d))
    return float(J[0]), g, gs, d


def set_source(sim, amp):
    a = np.ascontiguousarray(amp, dtype=sim.dtype)
    getattr(lib(), "orc_set_source_" + sim.prec)(sim.h, sim._ptr(a))
