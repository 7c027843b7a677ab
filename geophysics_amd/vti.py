"""Python binding of libvti.so (include/vti.h) — the MI355X VTI FD path.

Mirrors the reference's entry points for this path:

* ``GPU_vti2dfd``, ``GPU_vti3dfd``, ``CPU_vti2dfd`` — the file-in / file-out
  drivers with the reference argument lists
  (2D-3D-vti-fd-cuda-gtk-GUI/GPU_vti2dfd_kernels.cu:301-304,
  GPU_vti3dfd_kernels.cu:371-374, CPU_vti2dfd_kernels.c:366-367).
  The reference returns void; these raise ``VtiError`` on failure.
* ``Propagator`` — the per-time-step structure of those drivers on arrays
  already in memory (velocity update, stress update with absorbing boundary,
  source injection, receiver recording).

Everything computes on the GPU through the HIP library; there is no CPU
path.  Importing works without a GPU; calls that need one raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvti.so")

MODE_GPU_REF = 0
MODE_CPU_REF = 1
MODE_OMP_REF = 2      # 2D-vti-fd-omp/Toafd2domp_2.c
REC_DEFAULT, REC_P, REC_2P, REC_PQ = -1, 0, 1, 2

FIELDS = {n: i for i, n in enumerate(
    ["u", "v", "w", "px", "py", "pz", "qx", "qy", "qz", "P", "Q", "vp", "eps", "del"])}
TABLES = {n: 20 + i for i, n in enumerate(
    ["coffx1", "coffx2", "coffy1", "coffy2", "coffz1", "coffz2",
     "acoffx1", "acoffx2", "acoffy1", "acoffy2", "acoffz1", "acoffz2"])}
T_SRCAMP, T_D0 = 40, 41


class VtiError(RuntimeError):
    pass


class Cfg(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("npml", C.c_int), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("dt", C.c_float), ("nt", C.c_int), ("favg", C.c_float), ("pfac", C.c_float),
                ("wtype", C.c_int), ("mode", C.c_int), ("sv", C.c_int), ("sv_R", C.c_int),
                ("sv_r", C.c_int), ("rec_kind", C.c_int)]


class Shot(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "sx", "sy", "sz", "sv_cx", "sv_cy", "sv_cz", "rec_z",
        "rec_x0", "rec_dx", "rec_nx", "rec_y0", "rec_dy", "rec_ny")]


class RtmOpts(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("sv_forward", "sv_backward", "mute_tt", "inject_z",
                                        "illum_mode", "image_every", "nsnap")]


_lib = None
class LearnedOpts(C.Structure):
    _fields_ = [("save_every", C.c_int), ("check_every", C.c_int), ("tol", C.c_float), ("ck_every", C.c_int),
                ("max_exact", C.c_int), ("src_box", C.c_int), ("codec", C.c_int), ("floor", C.c_float),
                ("budget", C.c_longlong), ("serial_codec", C.c_int)]


class SuGeometry(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("nshot", "ntr", "ns", "dt_us", "fldr0", "scalco")] + \
               [(k, C.c_double) for k in ("sx0", "dsx", "sy0", "dsy", "gx0", "dgx", "dgx_shot", "gy0", "dgy",
                                          "dgy_shot", "sz", "gz")]


FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
VP = C.c_void_p


def _declare(L):
    sig = {
        "vti_strerror": (C.c_char_p, [C.c_int]),
        "vti_last_error": (C.c_char_p, []),
        "vti_cfg_ref2d": (C.c_int, [C.POINTER(Cfg)] + [C.c_int] * 10),
        "vti_cfg_ref3d": (C.c_int, [C.POINTER(Cfg)] + [C.c_int] * 11),
        "vti_shot_ref2d": (C.c_int, [C.POINTER(Shot), C.POINTER(Cfg)] + [C.c_int] * 5),
        "vti_shot_ref3d": (C.c_int, [C.POINTER(Shot), C.POINTER(Cfg)] + [C.c_int] * 9),
        "vti_host_tables": (C.c_int, [C.POINTER(Cfg), FP, FP, FP, FP]),
        "vti_host_source_gains": (C.c_int, [C.POINTER(Cfg), FP]),
        "vti_create": (C.c_int, [C.POINTER(Cfg), C.POINTER(VP)]),
        "vti_destroy": (C.c_int, [VP]),
        "vti_set_stream": (C.c_int, [VP, VP]),
        "vti_set_model": (C.c_int, [VP, VP, VP, VP, C.c_int]),
        "vti_set_shot": (C.c_int, [VP, C.POINTER(Shot)]),
        "vti_reset": (C.c_int, [VP]),
        "vti_update_vel": (C.c_int, [VP, C.c_int]),
        "vti_update_stress": (C.c_int, [VP, C.c_int]),
        "vti_step": (C.c_int, [VP, C.c_int, C.c_int]),
        "vti_run": (C.c_int, [VP]),
        "vti_nrec": (C.c_int, [VP]),
        "vti_mute": (C.c_int, [VP] + [C.c_int] * 5),
        "vti_mute3d": (C.c_int, [VP] + [C.c_int] * 7),
        "vti_get_record": (C.c_int, [VP, FP]),
        "vti_record_device": (VP, [VP]),
        "vti_get_field": (C.c_int, [VP, C.c_int, FP]),
        "vti_set_field": (C.c_int, [VP, C.c_int, FP]),
        "vti_get_table": (C.c_int, [VP, C.c_int, FP]),
        "vti_padded_dims": (C.c_int, [VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "vti_prof_enable": (C.c_int, [VP, C.c_int]),
        "vti_set_fused": (C.c_int, [VP, C.c_int]),
        "vti_set_unsplit": (C.c_int, [VP, C.c_int]),
        "vti_batch_set_unsplit": (C.c_int, [VP, C.c_int]),
        "vti_state_bytes": (C.c_longlong, [VP]),
        "vti_rtm_checkpoint_bytes": (C.c_longlong, [VP]),
        "vti_save_state": (C.c_int, [VP, VP]),
        "vti_load_state": (C.c_int, [VP, VP]),
        "vti_image_accumulate": (C.c_int, [VP, VP, VP, VP, C.c_int]),
        "vti_prof_read": (C.c_int, [VP, DP, DP, C.POINTER(C.c_int)]),
        "vti_calib_copy": (C.c_int, [VP, VP, C.c_size_t, VP]),
        "vti_sync": (C.c_int, [VP]),
        "vti_dropin_status": (C.c_int, []),
        "vti_rtm_defaults": (C.c_int, [C.POINTER(RtmOpts), C.c_int]),
        "vti_rtm_reset": (C.c_int, [VP]),
        "vti_rtm_shot": (C.c_int, [VP, C.POINTER(RtmOpts), FP, C.POINTER(C.c_longlong)]),
        "vti_rtm_shot_eb": (C.c_int, [VP, C.POINTER(RtmOpts), FP, DP]),
        "vti_set_theta": (C.c_int, [VP, FP, C.c_int]),
        "vti_rtm_eb_bytes": (C.c_longlong, [VP]),
        "vti_rtm_image": (C.c_int, [VP, C.c_int, FP, FP]),
        "vti_mute3d_window": (C.c_int, [VP] + [C.c_int] * 7),
        "vti_batch_create": (C.c_int, [C.c_int, C.POINTER(Cfg), C.POINTER(Shot), C.POINTER(FP),
                                       C.POINTER(FP), C.POINTER(FP), C.POINTER(VP)]),
        "vti_batch_destroy": (C.c_int, [VP]),
        "vti_sweep_check": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
        "vti_batch_config": (C.c_int, [VP, C.c_int, C.c_int, C.c_int]),
        "vti_batch_get_config": (C.c_int, [VP] + [C.POINTER(C.c_int)] * 4),
        "vti_batch_autotune": (C.c_int, [VP, C.c_int, DP]),
        "vti_batch_cells": (C.c_longlong, [VP]),
        "vti_batch_forward": (C.c_int, [VP, C.c_int, C.c_int]),
        "vti_batch_rtm": (C.c_int, [VP, C.POINTER(RtmOpts), C.POINTER(FP), C.POINTER(C.c_longlong)]),
        "vti_batch_image": (C.c_int, [VP, C.c_int, C.c_int, FP, FP]),
        "vti_batch_record": (C.c_int, [VP, C.c_int, FP]),
        "vti_ae_create": (C.c_int, [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(VP)]),
        "vti_ae_destroy": (C.c_int, [VP]),
        "vti_ae_nparams": (C.c_longlong, [VP]),
        "vti_ae_set_params": (C.c_int, [VP, FP]),
        "vti_ae_get_params": (C.c_int, [VP, FP]),
        "vti_ae_get_grads": (C.c_int, [VP, FP]),
        "vti_ae_set_lr": (C.c_int, [VP, C.c_float]),
        "vti_ae_train_batch": (C.c_int, [VP, FP, C.c_int, FP, FP]),
        "vti_ae_activation": (C.c_int, [VP, C.c_int, FP]),
        "vti_ae_train_field": (C.c_int, [VP, VP, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int,
                                         C.c_int, C.c_int, FP]),
        "vti_ae_encode": (C.c_int, [VP, VP, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, VP, VP]),
        "vti_ae_decode": (C.c_int, [VP, VP, VP, C.c_int, VP]),
        "vti_ae_latent_elems": (C.c_longlong, [C.c_int]),
        "vti_ae_fit_field": (C.c_int, [VP, VP, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_int]),
        "vti_ae_fit_end": (C.c_int, [VP, DP]),
        "vti_ae_set_basis": (C.c_int, [VP, FP]),
        "vti_ae_get_basis": (C.c_int, [VP, FP]),
        "vti_bc_encode": (C.c_int, [VP, VP, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_float,
                                    C.c_float, VP, C.c_longlong, VP, VP, C.POINTER(C.c_longlong), DP]),
        "vti_bc_encode_noalloc": (C.c_int, [VP, VP, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_float,
                                            C.c_float, VP, C.c_longlong, VP, VP, C.POINTER(C.c_longlong), DP]),
        "vti_bc_decode": (C.c_int, [VP, VP, VP, VP, C.c_int, VP]),
        "vti_bq_encode": (C.c_int, [VP, VP, C.c_longlong, C.c_longlong, C.c_longlong, C.c_int, C.c_float,
                                    C.c_float, VP, C.c_longlong, VP, VP, C.POINTER(C.c_longlong), DP]),
        "vti_bq_decode": (C.c_int, [VP, VP, VP, VP, C.c_int, VP]),
        "vti_bq_max_words": (C.c_longlong, [C.c_int]),
        "vti_codec_decode_image": (C.c_int, [VP, C.c_int, VP, VP, VP, C.c_int, VP, VP, C.c_longlong, C.c_longlong,
                                             C.c_longlong, VP, C.POINTER(C.c_int)]),
        "vti_adcig_config": (C.c_int, [VP, C.c_int, C.c_int, C.c_int]),
        "vti_adcig_elems": (C.c_longlong, [VP]),
        "vti_adcig_tables": (C.c_int, [C.c_int, C.c_int, FP, FP]),
        "vti_adcig_reset": (C.c_int, [VP]),
        "vti_adcig_accumulate": (C.c_int, [VP, VP]),
        "vti_adcig_get": (C.c_int, [VP, C.c_int, FP]),
        "vti_adcig_stack": (C.c_int, [C.c_int, C.c_int, C.c_int, FP, FP]),
        "vti_adcig_thin": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, FP, FP]),
        "vti_sym_eigen": (C.c_int, [C.c_int, DP, DP, DP]),
        "vti_ae_fit_shot": (C.c_int, [VP, VP, C.c_int, C.c_int, C.c_int]),
        "vti_learned_defaults": (C.c_int, [C.POINTER(LearnedOpts)]),
        "vti_rtm_shot_learned": (C.c_int, [VP, VP, C.POINTER(RtmOpts), C.POINTER(LearnedOpts), FP, DP]),
        "vti_ae_train_shot": (C.c_int, [VP, VP, C.c_int, C.c_int, FP]),
        "vti_grad_create": (C.c_int, [C.POINTER(Cfg), C.c_int, C.POINTER(VP)]),
        "vti_grad_destroy": (C.c_int, [VP]),
        "vti_grad_set_model": (C.c_int, [VP, DP, DP, DP, C.c_int]),
        "vti_grad_set_shot": (C.c_int, [VP, C.POINTER(Shot)]),
        "vti_grad_set_source": (C.c_int, [VP, DP]),
        "vti_grad_forward": (C.c_int, [VP, DP]),
        "vti_grad_run": (C.c_int, [VP, DP, DP, C.c_int, DP, DP, DP, DP, C.POINTER(C.c_longlong)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    for name in ("GPU_vti2dfd", "CPU_vti2dfd"):
        fn = getattr(L, name)
        fn.restype = None
        fn.argtypes = [C.c_int] * 6 + [C.c_char_p] * 3 + [C.c_int] * 5 + [C.c_char_p] * 2 + [C.c_int] * 3
    L.vti_omp2dfd.restype = C.c_int
    L.vti_omp2dfd.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, C.c_float,
                              C.c_int, C.c_int, C.c_int, C.c_float] + [C.c_char_p] * 5
    L.vti_tti2dfd.restype = C.c_int
    L.vti_tti2dfd.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, C.c_float,
                              C.c_int, C.c_int, C.c_int, C.c_float] + [C.c_char_p] * 6
    L.vti_bin2su.restype = C.c_int
    L.vti_bin2su.argtypes = [C.c_char_p] * 3 + [C.c_longlong, C.c_int, C.c_int]
    L.vti_su_write.restype = C.c_int
    L.vti_su_write.argtypes = [C.c_char_p, FP, C.POINTER(SuGeometry), C.c_int]
    L.vti_segy_write.restype = C.c_int
    L.vti_segy_write.argtypes = [C.c_char_p, FP, C.POINTER(SuGeometry), C.c_char_p]
    L.GPU_vti3dfd.restype = None
    L.GPU_vti3dfd.argtypes = [C.c_int] * 8 + [C.c_char_p] * 3 + [C.c_int] * 8 + [C.c_char_p] * 2 + \
        [C.c_int] * 3


def lib():
    """Load libvti.so; raise if it has not been built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VtiError(f"{LIB_PATH} is missing: build it with `python -m geophysics_amd.build`")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(code, what=""):
    if code < 0:
        L = lib()
        msg = L.vti_last_error().decode(errors="replace")
        raise VtiError(f"{what}: {L.vti_strerror(code).decode()} ({code}): {msg}")
    return code


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(FP)


def cfg_ref2d(mode, nx, nz, dx, dz, npd, SV, favg, nt, dt_us):
    c = Cfg()
    check(lib().vti_cfg_ref2d(C.byref(c), mode, nx, nz, dx, dz, npd, SV, favg, nt, dt_us), "cfg_ref2d")
    return c


def cfg_ref3d(nx, ny, nz, dx, dy, dz, npml, SV, favg, nt, dt_us):
    c = Cfg()
    check(lib().vti_cfg_ref3d(C.byref(c), nx, ny, nz, dx, dy, dz, npml, SV, favg, nt, dt_us), "cfg_ref3d")
    return c


def sweep_check(nt, nslot):
    """Host-only run of the checkpoint schedule: (forward steps, slots touched)."""
    steps, slots = C.c_longlong(), C.c_int()
    check(lib().vti_sweep_check(nt, nslot, C.byref(steps), C.byref(slots)), "sweep_check")
    return steps.value, slots.value


def learned_defaults():
    o = LearnedOpts()
    check(lib().vti_learned_defaults(C.byref(o)), "learned_defaults")
    return o


def rtm_defaults(ndim=2):
    o = RtmOpts()
    check(lib().vti_rtm_defaults(C.byref(o), ndim), "rtm_defaults")
    return o


def shot_ref2d(cfg, is_, fs, ds, zs, rec_depth=1):
    s = Shot()
    check(lib().vti_shot_ref2d(C.byref(s), C.byref(cfg), is_, fs, ds, zs, rec_depth), "shot_ref2d")
    return s


def shot_ref3d(cfg, is_, fsx, dsx, fsy, dsy, fsz, dsz, rec_depth=1, rec_step=1):
    s = Shot()
    check(lib().vti_shot_ref3d(C.byref(s), C.byref(cfg), is_, fsx, dsx, fsy, dsy, fsz, dsz,
                               rec_depth, rec_step), "shot_ref3d")
    return s


def host_tables(cfg, vp, eps):
    """Damping tables and d0 built by the library's host code (no GPU needed).

    Returns (tables dict, d0)."""
    nn = max(cfg.nx, cfg.ny if cfg.ndim == 3 else 1, cfg.nz) + 2 * cfg.npml
    out = np.zeros((12, nn), dtype=np.float32)
    d0 = np.zeros(1, dtype=np.float32)
    vp, pv = _f32(vp)
    eps, pe = _f32(eps)
    check(lib().vti_host_tables(C.byref(cfg), pv, pe, out.ctypes.data_as(FP), d0.ctypes.data_as(FP)),
          "host_tables")
    names = list(TABLES)
    lens = {"x": cfg.nx + 2 * cfg.npml, "y": (cfg.ny + 2 * cfg.npml) if cfg.ndim == 3 else 1,
            "z": cfg.nz + 2 * cfg.npml}
    return {n: out[i, :lens[n[-2]]].copy() for i, n in enumerate(names)}, float(d0[0])


def host_source_gains(cfg):
    out = np.zeros(cfg.nt, dtype=np.float32)
    check(lib().vti_host_source_gains(C.byref(cfg), out.ctypes.data_as(FP)), "host_source_gains")
    return out


class Propagator:
    """One VTI FD context on the GPU (the per-step structure of GPU_vti2dfd /
    GPU_vti3dfd / CPU_vti2dfd, GPU_vti2dfd_kernels.cu:404-444)."""

    def __init__(self, cfg, vp=None, eps=None, dl=None):
        self.cfg = cfg
        h = VP()
        check(lib().vti_create(C.byref(cfg), C.byref(h)), "vti_create")
        self.h = h
        nnx, nny, nnz = C.c_int(), C.c_int(), C.c_int()
        check(lib().vti_padded_dims(h, C.byref(nnx), C.byref(nny), C.byref(nnz)), "padded_dims")
        self.nnx, self.nny, self.nnz = nnx.value, nny.value, nnz.value
        self.shot = None
        if vp is not None:
            self.set_model(vp, eps, dl)

    @property
    def shape(self):
        return (self.nny, self.nnx, self.nnz) if self.cfg.ndim == 3 else (self.nnx, self.nnz)

    def set_model(self, vp, eps, dl, on_device=False):
        if on_device:   # integer device addresses (e.g. torch tensor.data_ptr())
            check(lib().vti_set_model(self.h, VP(vp), VP(eps), VP(dl), 1), "set_model")
            return
        a, _ = _f32(vp)
        b, _ = _f32(eps)
        d, _ = _f32(dl)
        n = self.cfg.nx * self.cfg.nz * (self.cfg.ny if self.cfg.ndim == 3 else 1)
        if a.size != n or b.size != n or d.size != n:
            raise VtiError(f"model arrays must have {n} cells")
        check(lib().vti_set_model(self.h, a.ctypes.data, b.ctypes.data, d.ctypes.data, 0), "set_model")

    def set_shot(self, shot):
        check(lib().vti_set_shot(self.h, C.byref(shot)), "set_shot")
        self.shot = shot

    def reset(self):
        check(lib().vti_reset(self.h), "reset")

    def update_vel(self, it):
        check(lib().vti_update_vel(self.h, it), "update_vel")

    def update_stress(self, it):
        check(lib().vti_update_stress(self.h, it), "update_stress")

    def step(self, it0, n):
        check(lib().vti_step(self.h, it0, n), "step")

    def run(self):
        check(lib().vti_run(self.h), "run")

    def sync(self):
        check(lib().vti_sync(self.h), "sync")

    @property
    def nrec(self):
        return lib().vti_nrec(self.h)

    def mute(self, fs, ds, zs, is_, tt=100):
        check(lib().vti_mute(self.h, fs, ds, zs, is_, tt), "mute")

    def mute3d(self, fsx, dsx, fsy, dsy, fsz, is_, tt=100):
        check(lib().vti_mute3d(self.h, fsx, dsx, fsy, dsy, fsz, is_, tt), "mute3d")

    def record(self):
        out = np.zeros((self.nrec, self.cfg.nt), dtype=np.float32)
        check(lib().vti_get_record(self.h, out.ctypes.data_as(FP)), "get_record")
        return out

    def field(self, name):
        out = np.zeros(self.shape, dtype=np.float32)
        check(lib().vti_get_field(self.h, FIELDS[name], out.ctypes.data_as(FP)), "get_field")
        return out

    def set_field(self, name, arr):
        a, p = _f32(arr)
        if a.size != int(np.prod(self.shape)):
            raise VtiError("field size mismatch")
        check(lib().vti_set_field(self.h, FIELDS[name], p), "set_field")

    def table(self, name):
        out = np.zeros(max(self.nnx, self.nny, self.nnz, self.cfg.nt, 1), dtype=np.float32)
        n = check(lib().vti_get_table(self.h, TABLES[name], out.ctypes.data_as(FP)), "get_table")
        return out[:n].copy()

    def srcamp(self):
        out = np.zeros(max(self.cfg.nt, 1), dtype=np.float32)
        n = check(lib().vti_get_table(self.h, T_SRCAMP, out.ctypes.data_as(FP)), "get_table")
        return out[:n].copy()

    def d0(self):
        out = np.zeros(1, dtype=np.float32)
        check(lib().vti_get_table(self.h, T_D0, out.ctypes.data_as(FP)), "get_table")
        return float(out[0])

    # ---- reverse-time migration (vti_rtm_*) ----
    def rtm_defaults(self):
        o = RtmOpts()
        check(lib().vti_rtm_defaults(C.byref(o), self.cfg.ndim), "rtm_defaults")
        return o

    def rtm_reset(self):
        check(lib().vti_rtm_reset(self.h), "rtm_reset")

    def rtm_shot(self, opts=None, obs=None):
        """Migrate the current shot; obs: observed record (nrec, nt) or None.
        Returns the number of forward steps taken."""
        o = opts if opts is not None else self.rtm_defaults()
        steps = C.c_longlong(0)
        if obs is not None:
            a, p = _f32(obs)
            if a.shape != (self.nrec, self.cfg.nt):
                raise VtiError(f"obs must be shaped ({self.nrec}, {self.cfg.nt})")
        else:
            p = None
        check(lib().vti_rtm_shot(self.h, C.byref(o), p, C.byref(steps)), "rtm_shot")
        return steps.value

    def set_theta(self, theta, degrees=False):
        """2D TTI tilt of the interior ((nx, nz), radians unless degrees);
        None: back to VTI."""
        if theta is None:
            check(lib().vti_set_theta(self.h, None, 0), "set_theta")
            return
        a, p = _f32(theta)
        if a.shape != (self.cfg.nx, self.cfg.nz):
            raise VtiError(f"theta must be shaped ({self.cfg.nx}, {self.cfg.nz})")
        check(lib().vti_set_theta(self.h, p, 1 if degrees else 0), "set_theta")

    def rtm_shot_eb(self, opts=None, obs=None):
        """Migrate the current shot with the source field rebuilt backward in
        time from saved 4-cell boundary bands (vti_rtm_shot_eb); returns
        {"band_bytes", "source_steps"}."""
        o = opts if opts is not None else self.rtm_defaults()
        p = None
        if obs is not None:
            a, p = _f32(obs)
            if a.shape != (self.nrec, self.cfg.nt):
                raise VtiError(f"obs must be shaped ({self.nrec}, {self.cfg.nt})")
            self._obs_keep = a
        st = np.zeros(2, np.float64)
        check(lib().vti_rtm_shot_eb(self.h, C.byref(o), p, st.ctypes.data_as(DP)), "rtm_shot_eb")
        return {"band_bytes": st[0], "source_steps": int(st[1])}

    def rtm_eb_bytes(self):
        return int(lib().vti_rtm_eb_bytes(self.h))

    def rtm_image(self, finalize=True):
        shape = (self.cfg.ny, self.cfg.nx, self.cfg.nz) if self.cfg.ndim == 3 else (self.cfg.nx, self.cfg.nz)
        img = np.zeros(shape, dtype=np.float32)
        ill = np.zeros(shape, dtype=np.float32)
        check(lib().vti_rtm_image(self.h, 1 if finalize else 0, img.ctypes.data_as(FP),
                                  ill.ctypes.data_as(FP)), "rtm_image")
        return img, ill

    def rtm_shot_learned(self, ae, opts=None, lopts=None, obs=None):
        """RTM of the current shot with the source field stored compressed
        (autoencoder latents, sparse block-transform coefficients with
        lopts.codec = 1, quantized block streams with lopts.codec = 2);
        returns the stats dict."""
        o = opts if opts is not None else self.rtm_defaults()
        lo = lopts if lopts is not None else learned_defaults()
        st = np.zeros(9, np.float64)
        p = None
        if obs is not None:
            a, p = _f32(obs)
            if a.shape != (self.nrec, self.cfg.nt):
                raise VtiError(f"obs must be shaped ({self.nrec}, {self.cfg.nt})")
            self._obs_keep = a
        check(lib().vti_rtm_shot_learned(self.h, ae.h, C.byref(o), C.byref(lo), p, st.ctypes.data_as(DP)),
              "rtm_shot_learned")
        return {"encode_ms": st[0], "decode_ms": st[1], "mean_checked_err": st[2], "over_tol": int(st[3]),
                "snapshots": int(st[4]), "exact": int(st[5]), "store_bytes": st[6], "recomputed": int(st[7]),
                "alloc_bytes": st[8]}

    def ae_fit_shot(self, ae, save_every=4, nblocks=4096, src_box=8):
        check(lib().vti_ae_fit_shot(self.h, ae.h, save_every, nblocks, src_box), "ae_fit_shot")

    def ae_train_shot(self, ae, save_every=4, nsteps=2):
        loss = np.zeros(1, np.float32)
        check(lib().vti_ae_train_shot(self.h, ae.h, save_every, nsteps, loss.ctypes.data_as(FP)), "ae_train_shot")
        return float(loss[0])

    def state_bytes(self):
        return int(lib().vti_state_bytes(self.h))

    def checkpoint_bytes(self):
        """Bytes of one checkpoint slot of rtm_shot under the current settings."""
        return int(lib().vti_rtm_checkpoint_bytes(self.h))

    def save_state(self, dst_ptr):
        """Copy the wavefield state to a device buffer of state_bytes() (waits)."""
        check(lib().vti_save_state(self.h, dst_ptr), "save_state")
        self.sync()

    def load_state(self, src_ptr):
        check(lib().vti_load_state(self.h, src_ptr), "load_state")
        self.sync()

    def image_accumulate(self, src, img_ptr, illum_ptr=None, clamp=False):
        """img += P(src) * P(self), illum += P(src)^2 + Q(src)^2 on device
        arrays of the interior (waits)."""
        check(lib().vti_image_accumulate(src.h, self.h, img_ptr, illum_ptr, 1 if clamp else 0), "image_accumulate")
        self.sync()

    # ---- Poynting-vector angle gathers (vti_adcig_*) ----
    def adcig_config(self, nangle, dangle=1, dcdp=1):
        check(lib().vti_adcig_config(self.h, nangle, dangle, dcdp), "adcig_config")
        self._adcig = (nangle, dangle, dcdp)

    def adcig_shape(self):
        na, _, d = self._adcig
        if self.cfg.ndim == 3:
            return (self.cfg.ny // d, self.cfg.nx // d, na, self.cfg.nz)
        return (self.cfg.nx, na, self.cfg.nz)

    def adcig(self, illum_divide=False):
        out = np.zeros(self.adcig_shape(), np.float32)
        check(lib().vti_adcig_get(self.h, 1 if illum_divide else 0, out.ctypes.data_as(FP)), "adcig_get")
        return out

    def adcig_accumulate(self, src):
        """One gather accumulation: source fields of src, receiver fields of self (waits)."""
        check(lib().vti_adcig_accumulate(src.h, self.h), "adcig_accumulate")
        self.sync()

    def set_unsplit(self, on=True):
        """Point-source form: P, Q updated directly in the interior and, in 3D, only the
        damped axes' split fields kept in the absorbing layer (faster; ~1e-6 from the
        split form)."""
        check(lib().vti_set_unsplit(self.h, 1 if on else 0), "set_unsplit")

    def set_fused(self, on=True):
        """3D: one fused kernel per step, or the separate velocity and stress kernels (default)."""
        check(lib().vti_set_fused(self.h, 1 if on else 0), "set_fused")

    def prof_enable(self, on=True):
        check(lib().vti_prof_enable(self.h, 1 if on else 0), "prof_enable")

    def prof_read(self):
        a, b, n = C.c_double(), C.c_double(), C.c_int()
        check(lib().vti_prof_read(self.h, C.byref(a), C.byref(b), C.byref(n)), "prof_read")
        return a.value, b.value, n.value

    def close(self):
        if getattr(self, "h", None):
            lib().vti_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _b(s):
    return os.fsencode(s)


def GPU_vti2dfd(nx, nz, dx, dz, npd, SV, FNv, FNe, FNd, favg, ns, fs, ds, zs, FNshot, FNsnap,
                nt, dt, run_count=0):
    """GPU_vti2dfd_kernels.cu:301 — 2D GPU driver (files in, shot/snapshot files out)."""
    lib().GPU_vti2dfd(nx, nz, dx, dz, npd, SV, _b(FNv), _b(FNe), _b(FNd), favg, ns, fs, ds, zs,
                      _b(FNshot), _b(FNsnap), nt, dt, run_count)
    check(lib().vti_dropin_status(), "GPU_vti2dfd")


def CPU_vti2dfd(nx, nz, dx, dz, npd, SV, FNv, FNe, FNd, favg, ns, fs, ds, zs, FNshot, FNsnap,
                nt, dt, run_count=0):
    """CPU_vti2dfd_kernels.c:366 — the CPU driver's semantics, computed on the GPU."""
    lib().CPU_vti2dfd(nx, nz, dx, dz, npd, SV, _b(FNv), _b(FNe), _b(FNd), favg, ns, fs, ds, zs,
                      _b(FNshot), _b(FNsnap), nt, dt, run_count)
    check(lib().vti_dropin_status(), "CPU_vti2dfd")


def GPU_vti3dfd(nx, ny, nz, dx, dy, dz, npml, SV, FNv, FNe, FNd, favg, ns, fsx, dsx, fsy, dsy,
                fsz, dsz, FNshot, FNsnap, nt, dt, run_count=0):
    """GPU_vti3dfd_kernels.cu:371 — 3D GPU driver."""
    lib().GPU_vti3dfd(nx, ny, nz, dx, dy, dz, npml, SV, _b(FNv), _b(FNe), _b(FNd), favg, ns,
                      fsx, dsx, fsy, dsy, fsz, dsz, _b(FNshot), _b(FNsnap), nt, dt, run_count)
    check(lib().vti_dropin_status(), "GPU_vti3dfd")


def adcig_tables(nangle, dangle):
    """Bin thresholds and cos^3 weight factors of the angle gathers (host)."""
    t = np.zeros(nangle, np.float32)
    w = np.zeros(nangle, np.float32)
    check(lib().vti_adcig_tables(nangle, dangle, t.ctypes.data_as(FP), w.ctypes.data_as(FP)), "adcig_tables")
    return t, w


def adcig_stack(adcigs):
    """stk_adcigs: (nx, nangle, nz) gathers -> Laplacian-filtered stack (nx, nz)."""
    a, p = _f32(adcigs)
    nx, na, nz = a.shape
    out = np.zeros((nx, nz), np.float32)
    check(lib().vti_adcig_stack(nx, nz, na, p, out.ctypes.data_as(FP)), "adcig_stack")
    return out


def adcig_thin(adcigs, d):
    """adcigs_smiled: the gathers of every d-th cdp."""
    a, p = _f32(adcigs)
    nx, na, nz = a.shape
    out = np.zeros(((nx + d - 1) // d, na, nz), np.float32)
    check(lib().vti_adcig_thin(nx, nz, na, d, p, out.ctypes.data_as(FP)), "adcig_thin")
    return out


def omp2dfd(FNv, FNe, FNd, FNshot, FNsnap, nx=601, nz=301, dx=5.0, dz=5.0, npd=0, favg=30.0,
            pfac=1000.0, sx=100, sz=100, nt=4001, dt=0.0005):
    """main of 2D-vti-fd-omp/Toafd2domp_2.c:40-242 with its hard-coded values
    as defaults (files in, shot / snapshot files out)."""
    check(lib().vti_omp2dfd(nx, nz, dx, dz, npd, favg, pfac, sx, sz, nt, dt, _b(FNv), _b(FNe), _b(FNd),
                            _b(FNshot), _b(FNsnap)), "vti_omp2dfd")


def tti2dfd(FNv, FNe, FNd, FNtheta, FNshot, FNsnap, nx=600, nz=300, dx=5.0, dz=5.0, npml=50, favg=40.0,
            pfac=10.0, fs=200, zs=140, nt=801, dt=0.0005):
    """main of 2D-tti-fd+rtm-cuda/Toa_gpu_2dtti_fd_1orderfunction.cu:313-499
    with its hard-coded values as defaults (files in, shot / snapshot files out)."""
    check(lib().vti_tti2dfd(nx, nz, dx, dz, npml, favg, pfac, fs, zs, nt, dt, _b(FNv), _b(FNe), _b(FNd),
                            _b(FNtheta), _b(FNshot), _b(FNsnap)), "vti_tti2dfd")


def bin2su(su_template, raw, out, ntrace, ns, swap=True):
    """utils-binary2segy/binary2segy.c: template SU headers + raw float
    traces (bytes of every sample reversed when swap) -> out."""
    check(lib().vti_bin2su(_b(su_template), _b(raw), _b(out), ntrace, ns, 1 if swap else 0), "vti_bin2su")


def su_geometry(nshot, ntr, ns, dt_us, **kw):
    """vti_su_geometry with fldr0 = 1, scalco = 1 and zero coordinates unless
    given (metres: sx0, dsx, sy0, dsy, gx0, dgx, dgx_shot, gy0, dgy,
    dgy_shot, sz, gz)."""
    g = SuGeometry(nshot=nshot, ntr=ntr, ns=ns, dt_us=dt_us, fldr0=kw.pop("fldr0", 1), scalco=kw.pop("scalco", 1))
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def _records(rec, g):
    a, p = _f32(rec)
    if a.size != g.nshot * g.ntr * g.ns:
        raise VtiError(f"records must hold nshot x ntr x ns = {g.nshot} x {g.ntr} x {g.ns} samples")
    return a, p


def su_write(path, rec, g, big_endian=False):
    """Records (nshot, ntr, ns) or (ntr, ns), trace-major, to an SU file."""
    a, p = _records(rec, g)
    check(lib().vti_su_write(_b(path), p, C.byref(g), 1 if big_endian else 0), "vti_su_write")


def segy_write(path, rec, g, text=None):
    """Records to a big-endian SEG-Y rev 1 file (IEEE floats)."""
    a, p = _records(rec, g)
    check(lib().vti_segy_write(_b(path), p, C.byref(g), text.encode() if text else None), "vti_segy_write")


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(DP)


class Gradient:
    """Forward, exact discrete adjoint and vp-gradient of the FD path
    (vti_grad_*), FP32 or FP64.  Point-source (GPU) semantics."""

    def __init__(self, cfg, fp64=False):
        self.cfg = cfg
        h = VP()
        check(lib().vti_grad_create(C.byref(cfg), 1 if fp64 else 0, C.byref(h)), "vti_grad_create")
        self.h = h
        self.fp64 = fp64
        self.shot = None

    @property
    def model_shape(self):
        c = self.cfg
        return (c.ny, c.nx, c.nz) if c.ndim == 3 else (c.nx, c.nz)

    @property
    def nrec(self):
        s = self.shot
        return s.rec_nx * (s.rec_ny if self.cfg.ndim == 3 else 1)

    def set_model(self, vp, eps, dl, keep_tables=False):
        a, pa = _f64(vp)
        b, pb = _f64(eps)
        d, pd = _f64(dl)
        check(lib().vti_grad_set_model(self.h, pa, pb, pd, 1 if keep_tables else 0), "grad_set_model")

    def set_shot(self, shot):
        check(lib().vti_grad_set_shot(self.h, C.byref(shot)), "grad_set_shot")
        self.shot = shot

    def set_source(self, amp=None):
        if amp is None:
            check(lib().vti_grad_set_source(self.h, None), "grad_set_source")
        else:
            a, p = _f64(amp)
            check(lib().vti_grad_set_source(self.h, p), "grad_set_source")

    def forward(self):
        d = np.zeros((self.nrec, self.cfg.nt))
        check(lib().vti_grad_forward(self.h, d.ctypes.data_as(DP)), "grad_forward")
        return d

    def run(self, dobs=None, resid=None, nsnap=0):
        """Returns dict(misfit, grad_vp, grad_src, d, fwd_steps)."""
        shape = (self.nrec, self.cfg.nt)
        po = pr = None
        if dobs is not None:
            dobs, po = _f64(dobs)
            assert dobs.shape == shape
        if resid is not None:
            resid, pr = _f64(resid)
            assert resid.shape == shape
        J = C.c_double()
        g = np.zeros(self.model_shape)
        gs = np.zeros(self.cfg.nt)
        d = np.zeros(shape)
        steps = C.c_longlong()
        check(lib().vti_grad_run(self.h, po, pr, nsnap, C.byref(J), g.ctypes.data_as(DP),
                                 gs.ctypes.data_as(DP), d.ctypes.data_as(DP), C.byref(steps)), "grad_run")
        return {"misfit": J.value, "grad_vp": g, "grad_src": gs, "d": d, "fwd_steps": steps.value}

    def close(self):
        if getattr(self, "h", None):
            lib().vti_grad_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent 2D shots of uneven grids advanced together (vti_batch_*)."""

    SHAPES = [(64, 4), (64, 2), (64, 8), (128, 2), (32, 8)]

    def __init__(self, cfgs, shots, models):
        """models: list of (vp, eps, dl) arrays shaped (nx, nz) per shot."""
        n = len(cfgs)
        assert len(shots) == n and len(models) == n
        self.cfgs, self.shots = list(cfgs), list(shots)
        ca = (Cfg * n)(*cfgs)
        sa = (Shot * n)(*shots)
        self._keep = []
        ptrs = []
        for k in range(3):
            arrs = [np.ascontiguousarray(m[k], dtype=np.float32) for m in models]
            self._keep.append(arrs)
            ptrs.append((FP * n)(*[a.ctypes.data_as(FP) for a in arrs]))
        h = VP()
        check(lib().vti_batch_create(n, ca, sa, ptrs[0], ptrs[1], ptrs[2], C.byref(h)), "vti_batch_create")
        self.h = h
        self.n = n

    @property
    def cells(self):
        return lib().vti_batch_cells(self.h)

    def config(self, shape=0, grid_mult=0, graph_steps=0):
        check(lib().vti_batch_config(self.h, shape, grid_mult, graph_steps), "batch_config")

    def get_config(self):
        v = [C.c_int() for _ in range(4)]
        check(lib().vti_batch_get_config(self.h, *[C.byref(x) for x in v]), "batch_get_config")
        return {"tile": (v[0].value, v[1].value), "grid_mult": v[2].value, "graph_steps": v[3].value}

    def set_unsplit(self, on=True):
        """Unsplit interior for every shot on both sides, the receiver injection row split (point-source form)."""
        check(lib().vti_batch_set_unsplit(self.h, 1 if on else 0), "batch_set_unsplit")

    def autotune(self, steps=4):
        ms = C.c_double()
        shape = check(lib().vti_batch_autotune(self.h, steps, C.byref(ms)), "batch_autotune")
        return shape, ms.value

    def forward(self, nsteps, record=True):
        check(lib().vti_batch_forward(self.h, nsteps, 1 if record else 0), "batch_forward")

    def rtm(self, opts, obs=None):
        steps = C.c_longlong()
        if obs is None:
            p = None
        else:
            arrs = [np.ascontiguousarray(o, dtype=np.float32) for o in obs]
            self._obs = arrs
            p = (FP * self.n)(*[a.ctypes.data_as(FP) for a in arrs])
        check(lib().vti_batch_rtm(self.h, C.byref(opts), p, C.byref(steps)), "batch_rtm")
        return steps.value

    def image(self, s, finalize=False):
        c = self.cfgs[s]
        img = np.zeros((c.nx, c.nz), np.float32)
        ill = np.zeros((c.nx, c.nz), np.float32)
        check(lib().vti_batch_image(self.h, s, 1 if finalize else 0, img.ctypes.data_as(FP),
                                    ill.ctypes.data_as(FP)), "batch_image")
        return img, ill

    def record(self, s):
        out = np.zeros((self.shots[s].rec_nx, self.cfgs[s].nt), np.float32)
        check(lib().vti_batch_record(self.h, s, out.ctypes.data_as(FP)), "batch_record")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().vti_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Autoencoder:
    """Snapshot autoencoder (vti_ae_*): BF16-MFMA training and inference."""

    def __init__(self, crop=32, batch=4, seed=0):
        h = VP()
        check(lib().vti_ae_create(crop, batch, seed, C.byref(h)), "vti_ae_create")
        self.h, self.crop, self.batch = h, crop, batch

    @property
    def nparams(self):
        return int(lib().vti_ae_nparams(self.h))

    def set_params(self, flat):
        a, p = _f32(flat)
        assert a.size == self.nparams
        check(lib().vti_ae_set_params(self.h, p), "ae_set_params")

    def params(self):
        out = np.zeros(self.nparams, np.float32)
        check(lib().vti_ae_get_params(self.h, out.ctypes.data_as(FP)), "ae_get_params")
        return out

    def grads(self):
        out = np.zeros(self.nparams, np.float32)
        check(lib().vti_ae_get_grads(self.h, out.ctypes.data_as(FP)), "ae_get_grads")
        return out

    def set_lr(self, lr):
        check(lib().vti_ae_set_lr(self.h, lr), "ae_set_lr")

    def train_batch(self, crops, apply=True, want_recon=False):
        a, p = _f32(crops)
        assert a.size == self.batch * self.crop ** 3
        loss = np.zeros(1, np.float32)
        rec = np.zeros(a.shape, np.float32) if want_recon else None
        check(lib().vti_ae_train_batch(self.h, p, 1 if apply else 0, loss.ctypes.data_as(FP),
                                       rec.ctypes.data_as(FP) if want_recon else None), "ae_train_batch")
        return (float(loss[0]), rec) if want_recon else float(loss[0])

    def activation(self, i, shape):
        out = np.zeros(shape, np.float32)
        check(lib().vti_ae_activation(self.h, i, out.ctypes.data_as(FP)), "ae_activation")
        return out

    def train_field(self, ptr, off, sy, sx, ny, nx, nz, nsteps):
        loss = np.zeros(1, np.float32)
        check(lib().vti_ae_train_field(self.h, ptr, off, sy, sx, ny, nx, nz, nsteps, loss.ctypes.data_as(FP)),
              "ae_train_field")
        return float(loss[0])

    def encode(self, ptr, off, sy, sx, n, latent_ptr, scale_ptr):
        check(lib().vti_ae_encode(self.h, ptr, off, sy, sx, n, latent_ptr, scale_ptr), "ae_encode")

    def decode(self, latent_ptr, scale_ptr, n, out_ptr):
        check(lib().vti_ae_decode(self.h, latent_ptr, scale_ptr, n, out_ptr), "ae_decode")

    def fit_field(self, ptr, off, sy, sx, n, nblocks=4096):
        check(lib().vti_ae_fit_field(self.h, ptr, off, sy, sx, n, nblocks), "ae_fit_field")

    def fit_end(self):
        e = C.c_double()
        check(lib().vti_ae_fit_end(self.h, C.byref(e)), "ae_fit_end")
        return e.value

    def set_basis(self, V):
        """Orthonormal 512 x 512 block basis V[k][j] of the sparse codec."""
        a, p = _f32(V)
        if a.size != 512 * 512:
            raise VtiError("basis must be 512 x 512")
        check(lib().vti_ae_set_basis(self.h, p), "ae_set_basis")

    def basis(self):
        out = np.zeros((512, 512), np.float32)
        check(lib().vti_ae_get_basis(self.h, out.ctypes.data_as(FP)), "ae_get_basis")
        return out

    def bc_encode(self, ptr, off, sy, sx, n, tol, ent, cap, boff, bcnt, floor=1e-8):
        """Sparse codec encode of the n^3 box at device pointer ptr into the
        device buffers ent (cap uint32), boff (uint32), bcnt (uint16) of
        (n/8)^3 blocks; returns (entries needed, error bound)."""
        kept = C.c_longlong()
        err = C.c_double()
        check(lib().vti_bc_encode(self.h, ptr, off, sy, sx, n, tol, floor, ent, cap, boff, bcnt, C.byref(kept),
                                  C.byref(err)), "bc_encode")
        return kept.value, err.value

    def bc_encode_noalloc(self, ptr, off, sy, sx, n, tol, ent, cap, boff, bcnt, floor=1e-8):
        """bc_encode with no device allocation (n <= 128; scratch in a static
        device array)."""
        kept = C.c_longlong()
        err = C.c_double()
        check(lib().vti_bc_encode_noalloc(self.h, ptr, off, sy, sx, n, tol, floor, ent, cap, boff, bcnt,
                                          C.byref(kept), C.byref(err)), "bc_encode_noalloc")
        return kept.value, err.value

    def bc_decode(self, ent, boff, bcnt, n, out):
        check(lib().vti_bc_decode(self.h, ent, boff, bcnt, n, out), "bc_decode")

    def bq_encode(self, ptr, off, sy, sx, n, tol, ent, cap, boff, bcnt, floor=1e-8):
        """Quantized codec encode (same arguments and error bound as
        bc_encode; ent holds 32-bit stream words, at most bq_max_words(n));
        returns (words needed, error bound)."""
        kept = C.c_longlong()
        err = C.c_double()
        check(lib().vti_bq_encode(self.h, ptr, off, sy, sx, n, tol, floor, ent, cap, boff, bcnt, C.byref(kept),
                                  C.byref(err)), "bq_encode")
        return kept.value, err.value

    def bq_decode(self, ent, boff, bcnt, n, out):
        check(lib().vti_bq_decode(self.h, ent, boff, bcnt, n, out), "bq_decode")

    def decode_image(self, codec, ent, boff, bcnt, n, img, rp, off, sy, sx, box=None, geom=None):
        """img += decoded snapshot (the box's values inside geom = (y0, x0,
        z0, by, bx, bz) when box is given) times rp at off + y sy + x sx + z
        (device pointers)."""
        g = None if geom is None else (C.c_int * 6)(*geom)
        check(lib().vti_codec_decode_image(self.h, codec, ent, boff, bcnt, n, img, rp, off, sy, sx, box, g),
              "codec_decode_image")

    @staticmethod
    def bq_max_words(n):
        return int(lib().vti_bq_max_words(n))

    @staticmethod
    def latent_elems(n):
        return int(lib().vti_ae_latent_elems(n))

    def close(self):
        if getattr(self, "h", None):
            lib().vti_ae_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sym_eigen(a):
    """Eigenvalues and eigenvectors (columns) of a symmetric matrix with the
    library's host Jacobi routine (unsorted)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    check(lib().vti_sym_eigen(n, a.ctypes.data_as(DP), w.ctypes.data_as(DP), v.ctypes.data_as(DP)), "sym_eigen")
    return w, v
