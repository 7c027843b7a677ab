// vti_api.hip — C ABI of the VTI FD path (include/vti.h): per-context
// device state, table set-up and the per-time-step call structure of the
// reference drivers (GPU_vti2dfd_kernels.cu:404-444,
// GPU_vti3dfd_kernels.cu:492-539, CPU_vti2dfd_kernels.c:465-510).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "vti_internal.h"

namespace vti {
const char *last_error_cstr();
}

using namespace vti;

namespace vti {
int fail(int code, const std::string &msg)
{
    set_error(msg);
    // a failed hipMalloc leaves its error as the thread's last HIP error;
    // read it here so that the next launch's check does not report it
    if (code == VTI_ERR_ALLOC) (void)hipGetLastError();
    return code;
}
}  // namespace vti

namespace {

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(VTI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int round_up(int v, int m) { return (v + m - 1) / m * m; }

bool is3d(const vti_ctx *c) { return c->cfg.ndim == 3; }

float *table_ptr(vti_ctx *c, int k) { return c->tab + (size_t)k * c->nnmax; }

hipEvent_t next_event(vti_ctx *c)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

}  // namespace

namespace vti {

int state_nfield(const vti_ctx *c) { return c->cfg.ndim == 3 ? 11 : 8; }

int state_alloc(vti_ctx *c, VtiState &s)
{
    const int nf = state_nfield(c);
    float *p = nullptr;
    if (hipMalloc(&p, (size_t)nf * c->nvol * sizeof(float)) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "state allocation");
    static const int order3[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    static const int order2[8] = {0, 2, 3, 5, 6, 8, 9, 10};
    s = VtiState{};
    const int *ord = nf == 11 ? order3 : order2;
    for (int k = 0; k < nf; ++k) s.f[ord[k]] = p + (size_t)k * c->nvol;
    if (launch_zero(p, (size_t)nf * c->nvol, c->stream) != hipSuccess)
        return fail(VTI_ERR_HIP, "state zero");
    return VTI_OK;
}

void state_free(VtiState &s)
{
    // the first present field is the base of the single allocation
    float *base = s.f[0];
    if (base) (void)hipFree(base);
    s = VtiState{};
}

hipError_t state_zero(vti_ctx *c, const VtiState &s)
{
    if (s.f[VTI_F_P] == c->f[VTI_F_P]) c->us_dirty = false;
    for (int k = 0; k < 11; ++k)
        if (s.f[k]) {
            hipError_t e = launch_zero(s.f[k], c->nvol, c->stream);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

hipError_t state_copy(vti_ctx *c, const VtiState &dst, const VtiState &src)
{
    for (int k = 0; k < 11; ++k)
        if (src.f[k]) {
            hipError_t e = hipMemcpyAsync(dst.f[k], src.f[k], c->nvol * sizeof(float),
                                          hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

VtiState main_state(vti_ctx *c)
{
    c->alt_valid = false;   // callers may change the fields outside vti_step
    VtiState s;
    for (int k = 0; k < 11; ++k) s.f[k] = c->f[k];
    return s;
}

VelArgs vel_args_for(vti_ctx *c, const VtiState &s, int it, bool inject)
{
    const int M = c->nnmax;
    VelArgs a{};
    a.g = c->g;
    a.u = s.f[VTI_F_U]; a.v = s.f[VTI_F_V]; a.w = s.f[VTI_F_W];
    a.P = s.f[VTI_F_P]; a.Q = s.f[VTI_F_Q];
    a.cx1 = c->tab + 0 * M; a.cx2 = c->tab + 1 * M;
    a.cy1 = c->tab + 2 * M; a.cy2 = c->tab + 3 * M;
    a.cz1 = c->tab + 4 * M; a.cz2 = c->tab + 5 * M;
    const float dt = c->cfg.dt;
    a.dtx = dt / c->cfg.dx; a.dty = dt / c->cfg.dy; a.dtz = dt / c->cfg.dz;
    a.src = -1;
    a.amp = 0.0f;
    a.cth = c->tti;
    a.sth = c->tti ? c->tti + c->nvol : nullptr;
    if (inject && c->cfg.mode == VTI_MODE_GPU_REF && c->amp[it] != 0.0f) {
        a.src = c->shot.sz + (long long)c->shot.sx * c->g.ldz + (long long)c->shot.sy * c->g.ldxy;
        a.amp = c->amp[it];
    }
    return a;
}

StressArgs stress_args_for(vti_ctx *c, const VtiState &s, int it, bool inject, bool sv, float *rec_row)
{
    const int M = c->nnmax;
    StressArgs a{};
    a.g = c->g;
    a.u = s.f[VTI_F_U]; a.v = s.f[VTI_F_V]; a.w = s.f[VTI_F_W];
    a.px = s.f[VTI_F_PX]; a.py = s.f[VTI_F_PY]; a.pz = s.f[VTI_F_PZ];
    a.qx = s.f[VTI_F_QX]; a.qy = s.f[VTI_F_QY]; a.qz = s.f[VTI_F_QZ];
    a.P = s.f[VTI_F_P]; a.Q = s.f[VTI_F_Q];
    a.vp = c->vp; a.ep = c->ep; a.de = c->de;
    a.ax1 = c->tab + 6 * M; a.ax2 = c->tab + 7 * M;
    a.ay1 = c->tab + 8 * M; a.ay2 = c->tab + 9 * M;
    a.az1 = c->tab + 10 * M; a.az2 = c->tab + 11 * M;
    const float dt = c->cfg.dt;
    a.dtx = dt / c->cfg.dx; a.dty = dt / c->cfg.dy; a.dtz = dt / c->cfg.dz;
    a.cth = c->tti;
    a.sth = c->tti ? c->tti + c->nvol : nullptr;
    a.sv = sv ? 1 : 0;
    a.svR = c->cfg.sv_R;
    a.cx = c->shot.sv_cx; a.cy = c->shot.sv_cy; a.cz = c->shot.sv_cz;
    a.tap = c->dtap;
    a.gsrc = (inject && c->cfg.mode != VTI_MODE_GPU_REF && c->amp[it] != 0.0f) ? 1 : 0;
    a.dpq = c->cfg.mode == VTI_MODE_OMP_REF ? 1 : 0;
    a.gx = c->shot.sx; a.gy = c->shot.sy; a.gz = c->shot.sz;
    a.gamp = c->amp[it];
    a.gw = c->dgw;
    a.rec = rec_row;
    a.rec_kind = c->cfg.rec_kind;
    a.rz = c->shot.rec_z;
    a.rx0 = c->shot.rec_x0; a.rdx = c->shot.rec_dx; a.rnx = c->shot.rec_nx;
    a.ry0 = c->shot.rec_y0; a.rdy = c->shot.rec_dy; a.rny = c->shot.rec_ny;
    a.split_z = -1;
    return a;
}

StressArgs src_stress_args(vti_ctx *c, const VtiState &s, int it, bool sv, float *rec_row)
{
    StressArgs a = stress_args_for(c, s, it, true, sv, rec_row);
    a.unsplit = c->unsplit && c->cfg.mode == VTI_MODE_GPU_REF;
    if (a.unsplit) c->us_dirty = true;
    return a;
}

StressArgs with_illum(vti_ctx *c, StressArgs a, int clamp)
{
    a.illum = c->illum;
    a.illum_clamp = clamp;
    a.inp = c->cfg.npml;
    a.inx = c->cfg.nx;
    a.iny = c->cfg.ndim == 3 ? c->cfg.ny : 1;
    a.inz = c->cfg.nz;
    return a;
}

StressArgs with_snap(vti_ctx *c, StressArgs a, float *snap)
{
    a.snap = snap;
    a.inp = c->cfg.npml;
    a.inx = c->cfg.nx;
    a.iny = c->cfg.ndim == 3 ? c->cfg.ny : 1;
    a.inz = c->cfg.nz;
    return a;
}

StressArgs rcv_stress_args(vti_ctx *c, const VtiState &s, int it, bool sv, int inj_z)
{
    StressArgs a = stress_args_for(c, s, it, false, sv, nullptr);
    a.unsplit = c->unsplit && c->cfg.mode == VTI_MODE_GPU_REF;
    if (a.unsplit) c->us_dirty = true;
    a.split_z = inj_z;
    return a;
}

}  // namespace vti

static VelArgs vel_args(vti_ctx *c, int it) { return vel_args_for(c, main_state(c), it, true); }

static StressArgs stress_args(vti_ctx *c, int it)
{
    float *row = c->nrec > 0 ? c->rec + (size_t)it * c->nrec : nullptr;
    return src_stress_args(c, main_state(c), it, c->cfg.sv != 0, row);
}

extern "C" {

const char *vti_strerror(int code)
{
    switch (code) {
    case VTI_OK: return "ok";
    case VTI_ERR_ARG: return "invalid argument";
    case VTI_ERR_ALLOC: return "allocation failed";
    case VTI_ERR_HIP: return "HIP runtime error";
    case VTI_ERR_IO: return "file I/O error";
    case VTI_ERR_CFL: return "stability or dispersion check failed";
    case VTI_ERR_STATE: return "call out of order";
    case VTI_ERR_NODEV: return "no HIP device";
    default: return "unknown error";
    }
}

const char *vti_last_error(void) { return last_error_cstr(); }

int vti_cfg_ref2d(vti_cfg *cfg, int mode, int nx, int nz, int dx, int dz, int npd, int SV,
                  int favg, int nt, int dt_us)
{
    if (!cfg || (mode != VTI_MODE_GPU_REF && mode != VTI_MODE_CPU_REF && mode != VTI_MODE_OMP_REF))
        return fail(VTI_ERR_ARG, "vti_cfg_ref2d: bad cfg pointer or mode");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->ndim = 2;
    cfg->nx = nx; cfg->ny = 1; cfg->nz = nz; cfg->npml = npd;
    cfg->dx = (float)dx; cfg->dy = (float)dx; cfg->dz = (float)dz;
    cfg->dt = (float)(dt_us * 1.0 / 1000000);   // GPU_vti2dfd_kernels.cu:348
    cfg->nt = nt;
    cfg->favg = (float)favg;
    cfg->pfac = 10.0f;                          // :349, CPU_vti2dfd_kernels.c:382
    cfg->wtype = 1;
    if (mode == VTI_MODE_OMP_REF) {
        cfg->pfac = 1000.0f;                    // Toafd2domp_2.c:70
        cfg->wtype = 4;                         // its Ricker product order, :154
    }
    cfg->mode = mode;
    cfg->sv = SV;
    cfg->rec_kind = VTI_REC_DEFAULT;
    return VTI_OK;
}

int vti_cfg_ref3d(vti_cfg *cfg, int nx, int ny, int nz, int dx, int dy, int dz, int npml,
                  int SV, int favg, int nt, int dt_us)
{
    if (!cfg) return fail(VTI_ERR_ARG, "vti_cfg_ref3d: null cfg");
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->ndim = 3;
    cfg->nx = nx; cfg->ny = ny; cfg->nz = nz; cfg->npml = npml;
    cfg->dx = (float)dx; cfg->dy = (float)dy; cfg->dz = (float)dz;
    cfg->dt = (float)(dt_us * 1.0 / 1000000);   // GPU_vti3dfd_kernels.cu:408
    cfg->nt = nt;
    cfg->favg = (float)favg;
    cfg->pfac = 10.0f;                          // :402
    cfg->wtype = 1;
    cfg->mode = VTI_MODE_GPU_REF;
    cfg->sv = SV;
    cfg->rec_kind = VTI_REC_DEFAULT;
    return VTI_OK;
}

int vti_shot_ref2d(vti_shot *s, const vti_cfg *c, int is, int fs, int ds, int zs, int rec_depth)
{
    if (!s || !c || c->ndim != 2) return fail(VTI_ERR_ARG, "vti_shot_ref2d: bad arguments");
    std::memset(s, 0, sizeof(*s));
    const int npd = c->npml;
    const float fsf = (float)(fs + (is - 1) * ds);             // :185 / :61
    s->sx = (int)(fsf + 0.5) + npd - 1;                        // :202 / :65
    s->sz = (int)((float)zs + 0.5) + npd - 1;                  // :203 / :66
    s->sy = 0;
    if (c->mode == VTI_MODE_GPU_REF) {
        s->sv_cx = fs + (is - 1) * ds + npd;                   // :260 (one cell off the source)
        s->sv_cz = zs + npd;                                   // :261
    } else {
        s->sv_cx = s->sx;                                      // CPU_vti2dfd_kernels.c:137-138
        s->sv_cz = s->sz;
    }
    s->rec_z = rec_depth + npd - 1;                            // :147 / :490 (row npd)
    s->rec_x0 = npd; s->rec_dx = 1; s->rec_nx = c->nx;
    s->rec_y0 = 0; s->rec_dy = 1; s->rec_ny = 1;
    return VTI_OK;
}

int vti_shot_ref3d(vti_shot *s, const vti_cfg *c, int is, int fsx, int dsx, int fsy, int dsy,
                   int fsz, int dsz, int rec_depth, int rec_step)
{
    if (!s || !c || c->ndim != 3 || rec_step < 1) return fail(VTI_ERR_ARG, "vti_shot_ref3d: bad arguments");
    std::memset(s, 0, sizeof(*s));
    const int np = c->npml;
    const float sx = (float)(fsx + is * dsx), sy = (float)(fsy + is * dsy), sz = (float)(fsz + is * dsz);
    s->sx = (int)(sx + np - 1);                                // GPU_vti3dfd_kernels.cu:82-84
    s->sy = (int)(sy + np - 1);
    s->sz = (int)(sz + np - 1);
    s->sv_cx = fsx + is * dsx + np - 1;                        // :155-157
    s->sv_cy = fsy + is * dsy + np - 1;
    s->sv_cz = fsz + is * dsz + np - 1;
    s->rec_z = rec_depth + np - 1;                             // :325 (row npml)
    s->rec_x0 = np; s->rec_dx = rec_step; s->rec_nx = (c->nx + rec_step - 1) / rec_step;
    s->rec_y0 = np; s->rec_dy = rec_step; s->rec_ny = (c->ny + rec_step - 1) / rec_step;
    return VTI_OK;
}

int vti_create(const vti_cfg *cfg, vti_ctx **out)
{
    if (!cfg || !out) return fail(VTI_ERR_ARG, "vti_create: null argument");
    *out = nullptr;
    const vti_cfg &c = *cfg;
    if ((c.ndim != 2 && c.ndim != 3) || c.nx < 1 || c.nz < 1 || (c.ndim == 3 && c.ny < 1) ||
        c.npml < 0 || c.nt < 1 || !(c.dt > 0) || !(c.dx > 0) || !(c.dz > 0) ||
        (c.ndim == 3 && !(c.dy > 0)) || !(c.favg > 0) || c.wtype < 1 || c.wtype > 4 ||
        (c.mode != VTI_MODE_GPU_REF && c.mode != VTI_MODE_CPU_REF && c.mode != VTI_MODE_OMP_REF) ||
        (c.mode == VTI_MODE_OMP_REF && c.ndim != 2) ||
        c.rec_kind < VTI_REC_DEFAULT || c.rec_kind > VTI_REC_PQ)
        return fail(VTI_ERR_ARG, "vti_create: invalid configuration");
    const int nnx = c.nx + 2 * c.npml, nnz = c.nz + 2 * c.npml;
    const int nny = c.ndim == 3 ? c.ny + 2 * c.npml : 1;
    if (nnx < 9 || nnz < 9 || (c.ndim == 3 && nny < 9))
        return fail(VTI_ERR_ARG, "vti_create: padded grid smaller than the 8th-order stencil");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(VTI_ERR_NODEV, "vti_create: no HIP device");

    std::unique_ptr<vti_ctx> x(new (std::nothrow) vti_ctx());
    if (!x) return fail(VTI_ERR_ALLOC, "vti_create: host allocation");
    x->cfg = c;
    if (x->cfg.sv_R <= 0) x->cfg.sv_R = 15;                   // R=15 / Circle_iso=15
    if (x->cfg.sv_r <= 0) x->cfg.sv_r = 4;                    // r=4 (GPU form)
    if (x->cfg.rec_kind == VTI_REC_DEFAULT)
        x->cfg.rec_kind = c.mode == VTI_MODE_CPU_REF ? VTI_REC_2P
                        : (c.mode == VTI_MODE_OMP_REF ? VTI_REC_PQ : VTI_REC_P);
    HIP_TRY(hipGetDevice(&x->device));
    x->g.nnx = nnx; x->g.nny = nny; x->g.nnz = nnz;
    x->g.ldz = round_up(nnz, 32);                             // 128-byte rows
    x->g.ldxy = (long long)x->g.ldz * nnx;
    x->nvol = (size_t)x->g.ldxy * nny;
    x->nnmax = round_up(std::max(nnx, std::max(nny, nnz)), 64);

    const int nfield = c.ndim == 3 ? 11 + 3 : 8 + 3;
    const size_t bytes = (size_t)nfield * x->nvol * sizeof(float);
    if (hipMalloc(&x->arena, bytes) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "vti_create: device allocation of " + std::to_string(bytes) + " bytes");
    float *p = x->arena;
    const int order3[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    const int order2[8] = {0, 2, 3, 5, 6, 8, 9, 10};
    if (c.ndim == 3)
        for (int k : order3) { x->f[k] = p; p += x->nvol; }
    else
        for (int k : order2) { x->f[k] = p; p += x->nvol; }
    x->vp = p; p += x->nvol;
    x->ep = p; p += x->nvol;
    x->de = p; p += x->nvol;
    if (hipMalloc(&x->tab, 12 * (size_t)x->nnmax * sizeof(float)) != hipSuccess ||
        hipMalloc(&x->dgw, 49 * sizeof(double)) != hipSuccess) {
        vti_destroy(x.release());
        return fail(VTI_ERR_ALLOC, "vti_create: table allocation");
    }
    x->htab.assign(12 * (size_t)x->nnmax, 1.0f);
    double gw[49];
    gaussian_weights(gw);
    HIP_TRY(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
    x->own_stream = true;
    HIP_TRY(hipMemcpy(x->dgw, gw, sizeof(gw), hipMemcpyHostToDevice));
    HIP_TRY(launch_zero(x->arena, (size_t)nfield * x->nvol, x->stream));
    HIP_TRY(hipStreamSynchronize(x->stream));
    *out = x.release();
    return VTI_OK;
}

int vti_destroy(vti_ctx *c)
{
    if (!c) return VTI_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    hipFree(c->arena);
    hipFree(c->tab);
    hipFree(c->dgw);
    hipFree(c->dtap);
    hipFree(c->rec);
    hipFree(c->rec_t);
    for (hipEvent_t e : c->ev_acc)
        if (e) hipEventDestroy(e);
    hipFree(c->alt);
    hipFree(c->img);
    hipFree(c->illum);
    hipFree(c->adcig);
    hipFree(c->adcig_tab);
    hipFree(c->tti);
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    delete c;
    return VTI_OK;
}

int vti_set_stream(vti_ctx *c, void *s)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_set_stream: null ctx");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
    if (s) {
        c->stream = static_cast<hipStream_t>(s);
        c->own_stream = false;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    return VTI_OK;
}

int vti_padded_dims(const vti_ctx *c, int *nnx, int *nny, int *nnz)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_padded_dims: null ctx");
    if (nnx) *nnx = c->g.nnx;
    if (nny) *nny = c->g.nny;
    if (nnz) *nnz = c->g.nnz;
    return VTI_OK;
}

// d0 and the 12 damping tables from a host copy of the unpadded model.
// Rows of `h` are M floats apart: coffx1 coffx2 coffy1 coffy2 coffz1 coffz2
// acoffx1 .. acoffz2.
static int host_tables_impl(const vti_cfg &k, const float *hv, int M, float *h, float *d0_out)
{
    const bool three = k.ndim == 3;
    const int ny = three ? k.ny : 1;
    const size_t n = (size_t)k.nx * ny * k.nz;
    const int np = k.npml;
    const int nnx = k.nx + 2 * np, nnz = k.nz + 2 * np, nny = three ? ny + 2 * np : 1;
    auto padded = [&](long long id) {
        // value at reference padded index id (layout iz + ix*nnz + iy*nnz*nnx)
        const long long iz = id % nnz, r = id / nnz, ix = r % nnx, iy = r / nnx;
        const long long jx = ix < np ? 0 : (ix >= np + k.nx ? k.nx - 1 : ix - np);
        const long long jz = iz < np ? 0 : (iz >= np + k.nz ? k.nz - 1 : iz - np);
        long long jy = 0;
        if (three) jy = iy < np ? 0 : (iy >= np + k.ny ? k.ny - 1 : iy - np);
        return hv[jz + jx * k.nz + jy * (long long)k.nz * k.nx];
    };
    for (size_t i = 0; i < n; ++i)
        if (!(hv[i] > 0.0f) || !std::isfinite(hv[i]))
            return fail(VTI_ERR_ARG, "vp must be positive and finite");
    float d0 = 0.0f;
    // Toafd2domp_2.c has no absorbing layer (npd = 0, :67; its updates carry
    // no damping factors, :166-180): d0 = 0 makes every table 1
    if (np > 0 && k.mode != VTI_MODE_OMP_REF) {
        if (k.mode == VTI_MODE_GPU_REF) {
            if (!three) {   // get_d02d GPU_vti2dfd_kernels.cu:105
                const float sdx = k.dx + k.dz;
                d0 = (float)(10.0 * padded((long long)nnx * nnz / 2) * std::log(100000.0) /
                             (2.0 * np * (sdx / 2.0)));
            } else {        // get_d03d GPU_vti3dfd_kernels.cu:205
                const float sdx = k.dx + k.dy + k.dz;
                d0 = (float)(10.0 * padded((long long)nny * nnx * nnz / 2) * std::log(100000.0) /
                             (2.0 * np * (sdx / 3.0)));
            }
        } else {            // get_constant CPU_vti2dfd_kernels.c:185-218
            float vmax = hv[0], vmin = hv[0];
            for (size_t i = 0; i < n; ++i) { vmax = std::max(vmax, hv[i]); vmin = std::min(vmin, hv[i]); }
            d0 = (float)(3.0 * vmax * std::log(100000.0) / (2.0 * np * k.dx));
            const float hmin = k.dx < k.dz ? k.dx : k.dz;
            const float dt_max = (float)(0.5 * hmin / vmax);        // :196
            const float dx_max = (float)(vmin / k.favg * 0.2);      // :197
            if (dx_max < k.dx || dx_max < k.dz || dt_max < k.dt)
                return fail(VTI_ERR_CFL, "dispersion / stability check of get_constant failed "
                                         "(redefine dx, dz or dt)");
        }
    }
    *d0_out = d0;
    const float dtd0 = k.dt * d0;
    const int cpu_form = k.mode == VTI_MODE_CPU_REF;
    for (int i = 0; i < 12 * M; ++i) h[i] = 1.0f;
    axis_tables(k.nx, np, dtd0, cpu_form, h + 0 * M, h + 1 * M, h + 6 * M, h + 7 * M);
    axis_tables(k.nz, np, dtd0, cpu_form, h + 4 * M, h + 5 * M, h + 10 * M, h + 11 * M);
    if (three) axis_tables(k.ny, np, dtd0, cpu_form, h + 2 * M, h + 3 * M, h + 8 * M, h + 9 * M);
    return VTI_OK;
}

int vti_host_tables(const vti_cfg *cfg, const float *vp, const float *eps, float *tables, float *d0)
{
    (void)eps;
    if (!cfg || !vp || !tables || !d0) return fail(VTI_ERR_ARG, "vti_host_tables: null argument");
    const int M = std::max(cfg->nx, std::max(cfg->ndim == 3 ? cfg->ny : 1, cfg->nz)) + 2 * cfg->npml;
    return host_tables_impl(*cfg, vp, M, tables, d0);
}

int vti_host_source_gains(const vti_cfg *cfg, float *amp)
{
    if (!cfg || !amp || cfg->nt < 0) return fail(VTI_ERR_ARG, "vti_host_source_gains: bad argument");
    std::vector<float> a;
    source_gains(*cfg, a);
    std::copy(a.begin(), a.end(), amp);
    return VTI_OK;
}

int vti_set_model(vti_ctx *c, const float *vp, const float *eps, const float *del, int on_device)
{
    if (!c || !vp || !eps || !del) return fail(VTI_ERR_ARG, "vti_set_model: null argument");
    const vti_cfg &k = c->cfg;
    const int ny = is3d(c) ? k.ny : 1;
    const size_t n = (size_t)k.nx * ny * k.nz;
    const int np = k.npml;
    // Host copy of vp: d0 needs one padded value (GPU form) or the
    // maximum (CPU form); the CPU form also runs its stability checks.
    std::vector<float> hv(n), he(n);
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    HIP_TRY(hipMemcpy(hv.data(), vp, n * sizeof(float), kind));
    HIP_TRY(hipMemcpy(he.data(), eps, n * sizeof(float), kind));
    int st = host_tables_impl(k, hv.data(), c->nnmax, c->htab.data(), &c->d0);
    if (st) return st;
    // padded vp[1], eps[1] (ix = 0, iz = 1) used by the direct-wave mute
    const int jz = 1 < np ? 0 : (1 - np < k.nz ? 1 - np : k.nz - 1);
    c->vp1 = hv[jz];
    c->eps1 = he[jz];
    HIP_TRY(hipMemcpyAsync(c->tab, c->htab.data(), c->htab.size() * sizeof(float),
                           hipMemcpyHostToDevice, c->stream));

    // Pad on the device from a staging copy of the unpadded model.
    float *stage = nullptr;
    if (hipMalloc(&stage, n * sizeof(float)) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "vti_set_model: staging buffer");
    const float *srcs[3] = {vp, eps, del};
    float *dsts[3] = {c->vp, c->ep, c->de};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 3 && e == hipSuccess; ++i) {
        e = hipMemcpyAsync(stage, srcs[i], n * sizeof(float),
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = launch_pad_model(c->g, k.ndim, k.nx, ny, k.nz, np, stage, dsts[i], c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    hipFree(stage);
    if (e != hipSuccess) return fail(VTI_ERR_HIP, std::string("vti_set_model: ") + hipGetErrorString(e));
    c->have_model = true;
    return VTI_OK;
}

int vti_set_shot(vti_ctx *c, const vti_shot *s)
{
    if (!c || !s) return fail(VTI_ERR_ARG, "vti_set_shot: null argument");
    const Grid &g = c->g;
    const vti_cfg &k = c->cfg;
    auto inside = [&](int x, int y, int z) {
        return x >= 4 && x < g.nnx - 4 && z >= 4 && z < g.nnz - 4 &&
               (is3d(c) ? (y >= 4 && y < g.nny - 4) : y == 0);
    };
    // The source (and its 7x7 footprint in CPU mode) must lie in the
    // updated region, where the stress update rewrites P and Q every step.
    const int fr = k.mode != VTI_MODE_GPU_REF ? 3 : 0;
    if (!inside(s->sx - fr, s->sy, s->sz - fr) || !inside(s->sx + fr, s->sy, s->sz + fr))
        return fail(VTI_ERR_ARG, "vti_set_shot: source footprint outside the updated region");
    const int ny_rec = is3d(c) ? s->rec_ny : 1;
    if (s->rec_nx < 0 || ny_rec < 0 || s->rec_dx < 1 || (is3d(c) && s->rec_dy < 1))
        return fail(VTI_ERR_ARG, "vti_set_shot: bad receiver grid");
    // Receivers may lie anywhere in the padded grid: one outside the updated
    // region records the P there, which no step changes (0), as the
    // reference does (Toafd2domp_2.c:219 records row npd = 0).
    auto in_grid = [&](int x, int y, int z) {
        return x >= 0 && x < g.nnx && z >= 0 && z < g.nnz && (is3d(c) ? (y >= 0 && y < g.nny) : y == 0);
    };
    if (s->rec_nx > 0 && ny_rec > 0) {
        const int xl = s->rec_x0 + (s->rec_nx - 1) * s->rec_dx;
        const int yl = is3d(c) ? s->rec_y0 + (s->rec_ny - 1) * s->rec_dy : 0;
        const int y0 = is3d(c) ? s->rec_y0 : 0;
        if (!in_grid(s->rec_x0, y0, s->rec_z) || !in_grid(xl, yl, s->rec_z))
            return fail(VTI_ERR_ARG, "vti_set_shot: receivers outside the padded grid");
    }
    const int nrec = s->rec_nx * ny_rec;
    if (nrec != c->nrec || !c->rec) {
        hipFree(c->rec); hipFree(c->rec_t);
        c->rec = c->rec_t = nullptr;
        c->nrec = nrec;
        if (nrec > 0) {
            const size_t nb = (size_t)nrec * k.nt * sizeof(float);
            if (hipMalloc(&c->rec, nb) != hipSuccess || hipMalloc(&c->rec_t, nb) != hipSuccess)
                return fail(VTI_ERR_ALLOC, "vti_set_shot: receiver buffers");
        }
    }
    c->shot = *s;
    if (!is3d(c)) { c->shot.sy = 0; c->shot.sv_cy = 0; c->shot.rec_y0 = 0; c->shot.rec_dy = 1; c->shot.rec_ny = 1; }
    source_gains(k, c->amp);
    std::vector<double> tap;
    taper_table(k.mode == VTI_MODE_CPU_REF, k.sv_R, k.sv_r, tap);
    if (c->tapR != k.sv_R || !c->dtap) {
        hipFree(c->dtap);
        c->dtap = nullptr;
        if (hipMalloc(&c->dtap, tap.size() * sizeof(double)) != hipSuccess)
            return fail(VTI_ERR_ALLOC, "vti_set_shot: taper table");
        c->tapR = k.sv_R;
    }
    HIP_TRY(hipMemcpy(c->dtap, tap.data(), tap.size() * sizeof(double), hipMemcpyHostToDevice));
    c->have_shot = true;
    return VTI_OK;
}

int vti_reset(vti_ctx *c)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_reset: null ctx");
    for (int k = 0; k < 11; ++k)
        if (c->f[k]) HIP_TRY(launch_zero(c->f[k], c->nvol, c->stream));
    c->alt_valid = false;
    c->us_dirty = false;
    if (c->rec) HIP_TRY(launch_zero(c->rec, (size_t)c->nrec * c->cfg.nt, c->stream));
    return VTI_OK;
}

static int check_step(vti_ctx *c, int it)
{
    if (!c) return fail(VTI_ERR_ARG, "null ctx");
    if (!c->have_model || !c->have_shot) return fail(VTI_ERR_STATE, "model and shot must be set before stepping");
    if (!c->unsplit && c->us_dirty)
        return fail(VTI_ERR_STATE, "split-form step after unsplit steps: the interior split fields are stale "
                                   "(call vti_reset first)");
    if (it < 0 || it >= c->cfg.nt) return fail(VTI_ERR_ARG, "time index out of range");
    return VTI_OK;
}

int vti_update_vel(vti_ctx *c, int it)
{
    int st = check_step(c, it);
    if (st) return st;
    HIP_TRY(launch_vel(c->cfg.ndim, vel_args(c, it), c->stream));
    return VTI_OK;
}

int vti_update_stress(vti_ctx *c, int it)
{
    int st = check_step(c, it);
    if (st) return st;
    HIP_TRY(launch_stress(c->cfg.ndim, stress_args(c, it), c->stream));
    return VTI_OK;
}

int vti_step(vti_ctx *c, int it0, int nsteps)
{
    int st = check_step(c, it0);
    if (st) return st;
    if (nsteps < 0 || it0 + nsteps > c->cfg.nt) return fail(VTI_ERR_ARG, "vti_step: range past nt");
    if (c->cfg.ndim == 3 && c->fused && !c->unsplit && nsteps > 0) {
        // fused steps: u, v, w, P, Q alternate between c->f and the second
        // buffers; the cells outside the updated region are read by the
        // stencils but never written, so the second buffers are refreshed
        // from c->f whenever something other than a fused step may have
        // changed the fields
        static const int kAlt[5] = {VTI_F_U, VTI_F_V, VTI_F_W, VTI_F_P, VTI_F_Q};
        if (!c->alt) {
            if (hipMalloc(&c->alt, 5 * c->nvol * sizeof(float)) != hipSuccess)
                return fail(VTI_ERR_ALLOC, "vti_step: second buffers of the fused step");
            for (int k = 0; k < 5; ++k) c->altf[kAlt[k]] = c->alt + (size_t)k * c->nvol;
        }
        if (!c->alt_valid)
            for (int k : kAlt)
                HIP_TRY(hipMemcpyAsync(c->altf[k], c->f[k], c->nvol * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        for (int it = it0; it < it0 + nsteps; ++it) {
            hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
            if (c->prof) {
                e0 = next_event(c); e1 = next_event(c); e2 = next_event(c);
                if (!e0 || !e1 || !e2) return fail(VTI_ERR_HIP, "vti_step: event creation");
                HIP_TRY(hipEventRecord(e0, c->stream));
                HIP_TRY(hipEventRecord(e1, c->stream));   // no separate velocity kernel
            }
            FusedArgs fa{};
            fa.v = vel_args(c, it);
            fa.u0 = c->f[VTI_F_U]; fa.v0 = c->f[VTI_F_V]; fa.w0 = c->f[VTI_F_W];
            fa.v.u = c->altf[VTI_F_U]; fa.v.v = c->altf[VTI_F_V]; fa.v.w = c->altf[VTI_F_W];
            fa.s = stress_args(c, it);
            fa.s.P = c->altf[VTI_F_P]; fa.s.Q = c->altf[VTI_F_Q];
            HIP_TRY(launch_fused(fa, c->stream));
            if (c->prof) HIP_TRY(hipEventRecord(e2, c->stream));
            for (int k : kAlt) std::swap(c->f[k], c->altf[k]);
        }
        c->alt_valid = true;   // both sets hold the same cells outside the updated region
        return VTI_OK;
    }
    for (int it = it0; it < it0 + nsteps; ++it) {
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
        if (c->prof) {
            e0 = next_event(c); e1 = next_event(c); e2 = next_event(c);
            if (!e0 || !e1 || !e2) return fail(VTI_ERR_HIP, "vti_step: event creation");
            HIP_TRY(hipEventRecord(e0, c->stream));
        }
        HIP_TRY(launch_vel(c->cfg.ndim, vel_args(c, it), c->stream));
        if (c->prof) HIP_TRY(hipEventRecord(e1, c->stream));
        HIP_TRY(launch_stress(c->cfg.ndim, stress_args(c, it), c->stream));
        if (c->prof) HIP_TRY(hipEventRecord(e2, c->stream));
    }
    return VTI_OK;
}

long long vti_state_bytes(const vti_ctx *c)
{
    if (!c) return 0;
    long long n = 0;
    for (int k = 0; k < 11; ++k)
        if (c->f[k]) n += (long long)c->nvol * (long long)sizeof(float);
    return n;
}

// The wavefield state (every allocated field of u, v, w, px .. qz, P, Q in
// that order, padded volumes) to / from a device buffer of vti_state_bytes.
int vti_save_state(vti_ctx *c, void *dst)
{
    if (!c || !dst) return fail(VTI_ERR_ARG, "vti_save_state: null argument");
    char *p = static_cast<char *>(dst);
    for (int k = 0; k < 11; ++k)
        if (c->f[k]) {
            HIP_TRY(hipMemcpyAsync(p, c->f[k], c->nvol * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            p += c->nvol * sizeof(float);
        }
    return VTI_OK;
}

int vti_load_state(vti_ctx *c, const void *src)
{
    if (!c || !src) return fail(VTI_ERR_ARG, "vti_load_state: null argument");
    const char *p = static_cast<const char *>(src);
    for (int k = 0; k < 11; ++k)
        if (c->f[k]) {
            HIP_TRY(hipMemcpyAsync(c->f[k], p, c->nvol * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            p += c->nvol * sizeof(float);
        }
    c->alt_valid = false;
    // a state saved with the unsplit update on is valid only with it on
    c->us_dirty = c->unsplit;
    return VTI_OK;
}

int vti_set_unsplit(vti_ctx *c, int on)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_set_unsplit: null ctx");
    if (on && c->cfg.mode != VTI_MODE_GPU_REF)
        return fail(VTI_ERR_ARG, "vti_set_unsplit: point-source (GPU) form only");
    if (on && c->tti) return fail(VTI_ERR_STATE, "vti_set_unsplit: the TTI update is split only");
    c->unsplit = on != 0;
    return VTI_OK;
}

// 2D TTI (2D-tti-fd+rtm-cuda/Toa_gpu_2dtti_fd_1orderfunction.cu): the tilt
// of the interior (x-major, z fastest), in radians or, with degrees, as the
// reference reads it (theta *= pi/180.0 in double, read_file :235); padded
// by edge replication as the model (pad_vv :190-215), cosf / sinf of the
// float angle per cell into two pitched device volumes.  theta NULL: back
// to the VTI update.
int vti_set_theta(vti_ctx *c, const float *theta, int degrees)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_set_theta: null ctx");
    if (!theta) {
        (void)hipFree(c->tti);
        c->tti = nullptr;
        return VTI_OK;
    }
    const vti_cfg &k = c->cfg;
    if (k.ndim != 2 || k.mode != VTI_MODE_GPU_REF)
        return fail(VTI_ERR_ARG, "vti_set_theta: 2D point-source (GPU) form only");
    if (c->unsplit) return fail(VTI_ERR_STATE, "vti_set_theta: the TTI update is split only (vti_set_unsplit(0))");
    const int np = k.npml, nnx = c->g.nnx, nnz = c->g.nnz, ldz = c->g.ldz;
    std::vector<float> h(2 * c->nvol, 0.0f);
    for (int ix = 0; ix < nnx; ++ix)
        for (int iz = 0; iz < nnz; ++iz) {
            const int jx = ix < np ? 0 : (ix >= np + k.nx ? k.nx - 1 : ix - np);
            const int jz = iz < np ? 0 : (iz >= np + k.nz ? k.nz - 1 : iz - np);
            float th = theta[(size_t)jz + (size_t)jx * k.nz];
            if (degrees) th = (float)((double)th * (kPi / 180.0));
            const size_t id = (size_t)iz + (size_t)ix * ldz;
            h[id] = std::cos(th);
            h[c->nvol + id] = std::sin(th);
        }
    if (!c->tti && hipMalloc(&c->tti, 2 * c->nvol * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(VTI_ERR_ALLOC, "vti_set_theta: tilt tables");
    }
    HIP_TRY(hipMemcpy(c->tti, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return VTI_OK;
}

int vti_set_fused(vti_ctx *c, int on)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_set_fused: null ctx");
    c->fused = on != 0;
    return VTI_OK;
}

int vti_run(vti_ctx *c)
{
    int st = vti_reset(c);
    if (st) return st;
    st = vti_step(c, 0, c->cfg.nt);
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VTI_OK;
}

int vti_sync(vti_ctx *c)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_sync: null ctx");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VTI_OK;
}

int vti_nrec(const vti_ctx *c) { return c ? c->nrec : 0; }

float *vti_record_device(vti_ctx *c) { return c ? c->rec : nullptr; }

// Mutes operate on the [it][irec] buffer by transposing to [irec][it],
// muting there with the reference index arithmetic, and transposing back.
static int mute_common(vti_ctx *c, int form, int a0, int a1, int a2, int a3, int a4, int is, int tt)
{
    if (!c || !c->rec) return fail(VTI_ERR_STATE, "vti_mute: no receiver buffer");
    const int nt = c->cfg.nt, nrec = c->nrec, np = c->cfg.npml;
    const vti_shot &s = c->shot;
    HIP_TRY(launch_transpose(c->rec, c->rec_t, nt, nrec, c->stream));
    if (form == 0)
        HIP_TRY(launch_mute2d(c->rec_t, nrec, nt, c->cfg.dt, c->cfg.favg, c->cfg.dx, c->cfg.dz,
                              a0, a1, a2, is, c->vp1, c->eps1, tt, s.rec_x0 - np, s.rec_dx, c->stream));
    else
        HIP_TRY(launch_mute3d(c->rec_t, s.rec_nx, s.rec_ny, nt, c->cfg.dt, c->cfg.favg,
                              c->cfg.dx, c->cfg.dy, c->cfg.dz, a0, a1, a2, a3, a4, is, c->vp1,
                              c->eps1, tt, form == 2, s.rec_x0 - np, s.rec_dx, s.rec_y0 - np,
                              s.rec_dy, c->stream));
    HIP_TRY(launch_transpose(c->rec_t, c->rec, nrec, nt, c->stream));
    return VTI_OK;
}

int vti_mute(vti_ctx *c, int fs, int ds, int zs, int is, int tt)
{
    if (c && c->cfg.ndim != 2) return fail(VTI_ERR_ARG, "vti_mute: 2D only");
    return mute_common(c, 0, fs, ds, zs, 0, 0, is, tt);
}

int vti_mute3d(vti_ctx *c, int fsx, int dsx, int fsy, int dsy, int fsz, int is, int tt)
{
    if (c && c->cfg.ndim != 3) return fail(VTI_ERR_ARG, "vti_mute3d: 3D only");
    return mute_common(c, 1, fsx, dsx, fsy, dsy, fsz, is, tt);
}

int vti_mute3d_window(vti_ctx *c, int fsx, int dsx, int fsy, int dsy, int fsz, int is, int tt)
{
    if (c && c->cfg.ndim != 3) return fail(VTI_ERR_ARG, "vti_mute3d_window: 3D only");
    return mute_common(c, 2, fsx, dsx, fsy, dsy, fsz, is, tt);
}

int vti_get_record(vti_ctx *c, float *dst)
{
    if (!c || !dst) return fail(VTI_ERR_ARG, "vti_get_record: null argument");
    if (!c->rec) return VTI_OK;
    const int nt = c->cfg.nt, nrec = c->nrec;
    HIP_TRY(launch_transpose(c->rec, c->rec_t, nt, nrec, c->stream));
    HIP_TRY(hipMemcpyAsync(dst, c->rec_t, (size_t)nt * nrec * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VTI_OK;
}

static float *field_ptr(vti_ctx *c, int which)
{
    if (which >= VTI_F_U && which <= VTI_F_Q) return c->f[which];
    if (which == VTI_F_VP) return c->vp;
    if (which == VTI_F_EPS) return c->ep;
    if (which == VTI_F_DEL) return c->de;
    return nullptr;
}

int vti_get_field(vti_ctx *c, int which, float *dst)
{
    if (!c || !dst) return fail(VTI_ERR_ARG, "vti_get_field: null argument");
    float *p = field_ptr(c, which);
    if (!p) return fail(VTI_ERR_ARG, "vti_get_field: no such field");
    const Grid &g = c->g;
    HIP_TRY(hipMemcpy2DAsync(dst, g.nnz * sizeof(float), p, g.ldz * sizeof(float),
                             g.nnz * sizeof(float), (size_t)g.nnx * g.nny, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VTI_OK;
}

int vti_set_field(vti_ctx *c, int which, const float *src)
{
    if (!c || !src) return fail(VTI_ERR_ARG, "vti_set_field: null argument");
    float *p = field_ptr(c, which);
    if (!p) return fail(VTI_ERR_ARG, "vti_set_field: no such field");
    c->alt_valid = false;
    const Grid &g = c->g;
    HIP_TRY(hipMemcpy2DAsync(p, g.ldz * sizeof(float), src, g.nnz * sizeof(float),
                             g.nnz * sizeof(float), (size_t)g.nnx * g.nny, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VTI_OK;
}

int vti_get_table(vti_ctx *c, int which, float *dst)
{
    if (!c || !dst) return fail(VTI_ERR_ARG, "vti_get_table: null argument");
    if (which >= VTI_T_COFFX1 && which <= VTI_T_ACOFFZ2) {
        const int k = which - VTI_T_COFFX1;
        const int axis = (k % 6) / 2;
        const int n = axis == 0 ? c->g.nnx : (axis == 1 ? c->g.nny : c->g.nnz);
        std::vector<float> tmp(c->nnmax);
        HIP_TRY(hipMemcpy(tmp.data(), table_ptr(c, k), c->nnmax * sizeof(float), hipMemcpyDeviceToHost));
        std::memcpy(dst, tmp.data(), n * sizeof(float));
        return n;
    }
    if (which == VTI_T_SRCAMP) {
        std::memcpy(dst, c->amp.data(), c->amp.size() * sizeof(float));
        return (int)c->amp.size();
    }
    if (which == VTI_T_D0) { dst[0] = c->d0; return 1; }
    return fail(VTI_ERR_ARG, "vti_get_table: no such table");
}

int vti_prof_enable(vti_ctx *c, int on)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_prof_enable: null ctx");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->prof = on != 0;
    c->ev_used = 0;
    return VTI_OK;
}

int vti_prof_read(vti_ctx *c, double *vel_ms, double *stress_ms, int *nlaunch)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_prof_read: null ctx");
    HIP_TRY(hipStreamSynchronize(c->stream));
    double sv = 0, ss = 0;
    const size_t n = c->ev_used / 3;
    for (size_t i = 0; i < n; ++i) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, c->ev_pool[3 * i], c->ev_pool[3 * i + 1]));
        HIP_TRY(hipEventElapsedTime(&b, c->ev_pool[3 * i + 1], c->ev_pool[3 * i + 2]));
        sv += a; ss += b;
    }
    if (vel_ms) *vel_ms = sv;
    if (stress_ms) *stress_ms = ss;
    if (nlaunch) *nlaunch = (int)n;
    c->ev_used = 0;
    return VTI_OK;
}

int vti_calib_copy(void *dst, const void *src, size_t nbytes, void *stream)
{
    if (!dst || !src) return fail(VTI_ERR_ARG, "vti_calib_copy: null pointer");
    HIP_TRY(launch_copy(dst, src, nbytes, static_cast<hipStream_t>(stream)));
    return VTI_OK;
}

}  // extern "C"
