// vti_rtm.hip — reverse-time migration of one shot at a time on top of the
// FD kernels: forward source wavefield, receiver back-propagation of the
// observed record, zero-lag cross-correlation imaging and illumination
// (RTM() 2D-vti-fd+rtm-cuda/Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1184-1566;
//  3D: 3D-vti-fd+rtm-cuda/Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:686-1041).
//
// Source wavefield.  The reference "reconstructs" the source field in the
// backward loop by re-running the forward operator from the final state
// with a one-cell boundary ring re-injected (:1423-1437), which is not a
// time reversal.  Here the backward loop receives the exact forward field
// P(t) of every step it images, recomputed from full-state checkpoints with
// the binomial (revolve) schedule: with s checkpoint slots a shot of nt
// steps costs at most t*nt forward steps, t the smallest integer with
// C(s+t, s) >= nt.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "vti_internal.h"
#include "vti_rtm.h"

namespace vti {
namespace {

// ------------------------------------------------------------ kernels ----
// shot_record(load) Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:525-528: the observed
// samples of step it overwrite P and Q on the injection row.
__global__ void inject_rec_kernel(Grid g, float *P, float *Q, const float *__restrict__ row, int iz,
                                  int rx0, int rdx, int rnx, int ry0, int rdy, int rny)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rnx * rny) return;
    const int ix = rx0 + (r % rnx) * rdx, iy = ry0 + (r / rnx) * rdy;
    const long long c = iz + (long long)ix * g.ldz + (long long)iy * g.ldxy;
    const float v = row[r];
    P[c] = v;
    Q[c] = v;
}

// cal_migration :679-698 (img += s*g) and cal_illumination :654-677
// (illum += P*P + Q*Q, clamped to 1 where <= 0 when clamp != 0) over the
// interior nx x ny x nz, image layout z fastest.
__global__ void image_kernel(Grid g, int np, int nx, int ny, int nz, const float *__restrict__ sP,
                             const float *__restrict__ gP, float *img)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y * blockDim.y + threadIdx.y, iy = blockIdx.z;
    if (iz >= nz || ix >= nx) return;
    const long long c = (iz + np) + (long long)(ix + np) * g.ldz + (long long)(iy + (g.nny > 1 ? np : 0)) * g.ldxy;
    const long long o = iz + (long long)ix * nz + (long long)iy * nz * nx;
    img[o] += sP[c] * gP[c];
}

__global__ void illum_kernel(Grid g, int np, int nx, int ny, int nz, const float *__restrict__ P,
                             const float *__restrict__ Q, float *illum, int clamp)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y * blockDim.y + threadIdx.y, iy = blockIdx.z;
    if (iz >= nz || ix >= nx) return;
    const long long c = (iz + np) + (long long)(ix + np) * g.ldz + (long long)(iy + (g.nny > 1 ? np : 0)) * g.ldxy;
    const long long o = iz + (long long)ix * nz + (long long)iy * nz * nx;
    const float p = P[c], q = Q[c];
    float v = illum[o] + (p * p + q * q);
    if (clamp && v <= 0.0f) v = 1.0f;
    illum[o] = v;
}

// migration_illum :700-715 (2D: divide) / 3D :539-553 (divide where nonzero)
__global__ void divide_kernel(float *img, const float *__restrict__ illum, size_t n, int guard)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = illum[i];
    if (!guard || d != 0.0f) img[i] /= d;
}

}  // namespace

hipError_t launch_inject_rec(vti_ctx *c, const VtiState &s, const float *row, int iz)
{
    const vti_shot &sh = c->shot;
    const int n = sh.rec_nx * (c->cfg.ndim == 3 ? sh.rec_ny : 1);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(inject_rec_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->g,
                       s.f[VTI_F_P], s.f[VTI_F_Q], row, iz, sh.rec_x0, sh.rec_dx, sh.rec_nx,
                       sh.rec_y0, c->cfg.ndim == 3 ? sh.rec_dy : 1, c->cfg.ndim == 3 ? sh.rec_ny : 1);
    return hipGetLastError();
}

hipError_t launch_image(vti_ctx *c, const float *sP, const float *gP, float *img)
{
    const vti_cfg &k = c->cfg;
    const int ny = k.ndim == 3 ? k.ny : 1;
    dim3 blk(64, 4), grd((k.nz + 63) / 64, (k.nx + 3) / 4, ny);
    hipLaunchKernelGGL(image_kernel, grd, blk, 0, c->stream, c->g, k.npml, k.nx, ny, k.nz, sP, gP, img);
    return hipGetLastError();
}

hipError_t launch_illum(vti_ctx *c, const float *P, const float *Q, float *illum, int clamp)
{
    const vti_cfg &k = c->cfg;
    const int ny = k.ndim == 3 ? k.ny : 1;
    dim3 blk(64, 4), grd((k.nz + 63) / 64, (k.nx + 3) / 4, ny);
    hipLaunchKernelGGL(illum_kernel, grd, blk, 0, c->stream, c->g, k.npml, k.nx, ny, k.nz, P, Q, illum, clamp);
    return hipGetLastError();
}

hipError_t launch_divide(vti_ctx *c, float *img, const float *illum, size_t n, int guard)
{
    hipLaunchKernelGGL(divide_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, img, illum, n, guard);
    return hipGetLastError();
}

// laplace_filter(adj=1) Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:150-202 and the 3D
// form Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:555-616, in the same
// scatter order.
void laplace_adj(int nz, int nx, int ny, const float *in, float *out)
{
    const size_t n = (size_t)nz * nx * ny;
    for (size_t j = 0; j < n; ++j) out[j] = 0.0f;
    const size_t sx = nz, sy = (size_t)nz * nx;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            for (int iz = 0; iz < nz; ++iz) {
                const size_t j = iz + ix * sx + iy * sy;
                if (iz > 0) { out[j - 1] -= in[j]; out[j] += in[j]; }
                if (iz < nz - 1) { out[j + 1] -= in[j]; out[j] += in[j]; }
                if (ix > 0) { out[j - sx] -= in[j]; out[j] += in[j]; }
                if (ix < nx - 1) { out[j + sx] -= in[j]; out[j] += in[j]; }
                if (ny > 1) {
                    if (iy > 0) { out[j - sy] -= in[j]; out[j] += in[j]; }
                    if (iy < ny - 1) { out[j + sy] -= in[j]; out[j] += in[j]; }
                }
            }
}

}  // namespace vti

// ------------------------------------------------------------ C ABI ----
using namespace vti;

namespace {

// ---------------------------------------------------- checkpoint slots ----
// With the point-source form's reduced split in effect (vti_set_unsplit), a
// cell whose stress damping factors are all exactly 1 never reads or writes
// its split fields (px .. qz); the other cells (the absorbing layer) do.  A
// checkpoint of the source state then keeps u, v, w, P, Q as whole volumes
// and the split fields only of those cells: the whole z row of every (y, x)
// outside the undamped [ylo, yhi) x [xlo, xhi), the z ends [0, zlo) and
// [zhi, nnz) of the others, packed row by row.  384^3 with a 30-cell layer:
// 2.5 instead of 3.9 GB per slot.  Otherwise a slot is a full state.
struct CkGeom {
    int nnx, nny, nnz;
    long long ldz, ldxy;
    int ylo, yhi, xlo, xhi, zlo, zhi;
    long long split_n;   // floats of one packed split field
};
struct SplitPtrs {
    float *f[6];
    int n;
};

__device__ __forceinline__ long long ck_row_offset(const CkGeom &q, int y, int x, bool &full)
{
    const bool yin = y >= q.ylo && y < q.yhi, xin = x >= q.xlo && x < q.xhi;
    full = !(yin && xin);
    const long long r = (long long)y * q.nnx + x;
    const long long fp = (long long)min(y, q.ylo) + max(0, y - q.yhi);   // whole-row planes before y
    const long long full_before = fp * q.nnx + ((long long)y - fp) * (q.nnx - (q.xhi - q.xlo)) +
                                  (yin ? (long long)min(x, q.xlo) + max(0, x - q.xhi) : (long long)x);
    return full_before * q.nnz + (r - full_before) * (q.zlo + (q.nnz - q.zhi));
}

// Packs (store) or unpacks the split fields' layer cells; one wavefront per
// (y, x) row.
__global__ __launch_bounds__(256) void ckpt_split_kernel(CkGeom q, SplitPtrs sp, float *packed, int store)
{
    const int lane = threadIdx.x & 63;
    const long long nrows = (long long)q.nny * q.nnx;
    for (long long r = blockIdx.x * 4LL + (threadIdx.x >> 6); r < nrows; r += gridDim.x * 4LL) {
        const int y = (int)(r / q.nnx), x = (int)(r % q.nnx);
        bool full;
        const long long po = ck_row_offset(q, y, x, full);
        const long long ro = (long long)y * q.ldxy + (long long)x * q.ldz;
        const int len = full ? q.nnz : q.zlo + (q.nnz - q.zhi);
        for (int k = 0; k < sp.n; ++k) {
            float *f = sp.f[k] + ro;
            float *pk = packed + (long long)k * q.split_n + po;
            for (int i = lane; i < len; i += 64) {
                const int z = full || i < q.zlo ? i : q.zhi + (i - q.zlo);
                if (store) pk[i] = f[z];
                else f[z] = pk[i];
            }
        }
    }
}

const int kBaseF[5] = {VTI_F_U, VTI_F_V, VTI_F_W, VTI_F_P, VTI_F_Q};
const int kSplitF[6] = {VTI_F_PX, VTI_F_PY, VTI_F_PZ, VTI_F_QX, VTI_F_QY, VTI_F_QZ};

struct CkSlots {
    bool compact = false;
    CkGeom q{};
    int nbase = 0, nsplit = 0;
    long long slot_floats = 0;
    std::vector<float *> mem;      // one allocation per slot
    std::vector<VtiState> view;    // whole-volume fields of each slot (split fields: full slots only)
    ~CkSlots()
    {
        for (float *p : mem) (void)hipFree(p);
    }
};

// The layout vti_rtm_shot uses for the context's current settings.
int ck_layout(vti_ctx *c, CkSlots &k)
{
    k = CkSlots{};
    const int nf = state_nfield(c);
    k.slot_floats = (long long)nf * (long long)c->nvol;
    if (!(c->unsplit && c->cfg.mode == VTI_MODE_GPU_REF) || c->tti) return VTI_OK;
    const int M = c->nnmax;
    std::vector<float> t(12 * (size_t)M);
    if (hipMemcpy(t.data(), c->tab, t.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(VTI_ERR_HIP, "checkpoint layout: tables");
    // the undamped interval of an axis (stress tables exactly 1); no single
    // interval: no compact slots
    auto undamped = [&](int t1, int n, int &lo, int &hi) {
        auto one = [&](int j) { return t[t1 * M + j] == 1.0f && t[(t1 + 1) * M + j] == 1.0f; };
        int i = 0;
        while (i < n && !one(i)) ++i;
        lo = i;
        while (i < n && one(i)) ++i;
        hi = i;
        for (; i < n; ++i)
            if (one(i)) return false;
        return true;
    };
    CkGeom &q = k.q;
    q.nnx = c->g.nnx;
    q.nny = c->cfg.ndim == 3 ? c->g.nny : 1;
    q.nnz = c->g.nnz;
    q.ldz = c->g.ldz;
    q.ldxy = c->g.ldxy;
    if (!undamped(6, q.nnx, q.xlo, q.xhi) || !undamped(10, q.nnz, q.zlo, q.zhi)) return VTI_OK;
    if (c->cfg.ndim == 3) {
        if (!undamped(8, q.nny, q.ylo, q.yhi)) return VTI_OK;
    } else {
        q.ylo = 0, q.yhi = 1;
    }
    const long long mid = (long long)(q.yhi - q.ylo) * (q.xhi - q.xlo);
    const long long rows = (long long)q.nny * q.nnx;
    q.split_n = (rows - mid) * q.nnz + mid * (q.zlo + (q.nnz - q.zhi));
    for (int f : kBaseF) k.nbase += c->f[f] ? 1 : 0;
    for (int f : kSplitF) k.nsplit += c->f[f] ? 1 : 0;
    k.compact = true;
    k.slot_floats = (long long)k.nbase * (long long)c->nvol + (long long)k.nsplit * q.split_n;
    return VTI_OK;
}

int ck_alloc(vti_ctx *c, CkSlots &k, int nslot)
{
    k.mem.assign(nslot, nullptr);
    k.view.assign(nslot, VtiState{});
    for (int i = 0; i < nslot; ++i) {
        if (hipMalloc(&k.mem[i], (size_t)k.slot_floats * sizeof(float)) != hipSuccess) {
            k.mem[i] = nullptr;
            return fail(VTI_ERR_ALLOC, "checkpoint slots");
        }
        if (!k.compact) {   // a full state, laid out as state_alloc lays it out
            static const int order3[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
            static const int order2[8] = {0, 2, 3, 5, 6, 8, 9, 10};
            const int nf = state_nfield(c);
            const int *ord = nf == 11 ? order3 : order2;
            for (int j = 0; j < nf; ++j) k.view[i].f[ord[j]] = k.mem[i] + (size_t)j * c->nvol;
        } else {
            int j = 0;
            for (int f : kBaseF)
                if (c->f[f]) k.view[i].f[f] = k.mem[i] + (size_t)(j++) * c->nvol;
        }
    }
    return VTI_OK;
}

// slot i <- W (store) or W <- slot i
hipError_t ck_copy(vti_ctx *c, const CkSlots &k, int i, const VtiState &W, bool store)
{
    const VtiState &S = k.view[i];
    if (!k.compact) return store ? state_copy(c, S, W) : state_copy(c, W, S);
    for (int f : kBaseF)
        if (S.f[f]) {
            const hipError_t e = hipMemcpyAsync(store ? S.f[f] : W.f[f], store ? W.f[f] : S.f[f],
                                                c->nvol * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) return e;
        }
    SplitPtrs sp{};
    for (int f : kSplitF)
        if (W.f[f]) sp.f[sp.n++] = W.f[f];
    float *packed = k.mem[i] + (size_t)k.nbase * c->nvol;
    const long long rows = (long long)k.q.nny * k.q.nnx;
    hipLaunchKernelGGL(ckpt_split_kernel, dim3((unsigned)std::min<long long>((rows + 3) / 4, 16384)), dim3(256), 0,
                       c->stream, k.q, sp, packed, store ? 1 : 0);
    return hipGetLastError();
}

struct RtmCache {
    VtiState g;                  // receiver wavefield
    CkSlots ck;                  // checkpoints of the source state
};

}  // namespace

extern "C" {

int vti_rtm_defaults(vti_rtm_opts *o, int ndim)
{
    if (!o) return fail(VTI_ERR_ARG, "vti_rtm_defaults: null");
    std::memset(o, 0, sizeof(*o));
    o->sv_forward = 1;                 // Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1387 (true)
    o->sv_backward = 0;                // :1460 (false)
    o->mute_tt = ndim == 3 ? 70 : 20;  // :1409 (20); 3D :mute_directwave(..., 70)
    o->inject_z = -1;                  // the record row (2D :527); 3D reference: npml-1 (:334)
    o->illum_mode = ndim == 3 ? 1 : 0; // 2D: source energy with the <=0 -> 1 clamp; 3D: source+receiver
    o->image_every = 1;
    o->nsnap = 0;
    return VTI_OK;
}

int vti_rtm_reset(vti_ctx *c)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_rtm_reset: null ctx");
    const size_t n = (size_t)c->cfg.nx * (c->cfg.ndim == 3 ? c->cfg.ny : 1) * c->cfg.nz;
    if (!c->img) {
        if (hipMalloc(&c->img, n * sizeof(float)) != hipSuccess ||
            hipMalloc(&c->illum, n * sizeof(float)) != hipSuccess)
            return fail(VTI_ERR_ALLOC, "vti_rtm_reset: image buffers");
        c->nint = n;
    }
    if (launch_zero(c->img, n, c->stream) != hipSuccess || launch_zero(c->illum, n, c->stream) != hipSuccess ||
        (c->adcig && launch_zero(c->adcig, c->adcig_n, c->stream) != hipSuccess))
        return fail(VTI_ERR_HIP, "vti_rtm_reset: zero");
    return VTI_OK;
}

long long vti_rtm_checkpoint_bytes(vti_ctx *c)
{
    if (!c) return 0;
    CkSlots k;
    if (ck_layout(c, k)) return 0;
    return k.slot_floats * (long long)sizeof(float);
}

int vti_rtm_shot(vti_ctx *c, const vti_rtm_opts *o, const float *obs, long long *fwd_steps)
{
    if (!c || !o) return fail(VTI_ERR_ARG, "vti_rtm_shot: null argument");
    if (!c->have_model || !c->have_shot) return fail(VTI_ERR_STATE, "vti_rtm_shot: model and shot first");
    if (c->nrec <= 0) return fail(VTI_ERR_STATE, "vti_rtm_shot: the shot has no receivers");
    if (o->image_every < 1 || o->nsnap < 0) return fail(VTI_ERR_ARG, "vti_rtm_shot: bad options");
    int st;
    if (!c->img && (st = vti_rtm_reset(c))) return st;
    const int nt = c->cfg.nt, nrec = c->nrec;
    const int inj_z = o->inject_z >= 0 ? o->inject_z : c->shot.rec_z;
    if (inj_z < 4 || inj_z >= c->g.nnz - 4) return fail(VTI_ERR_ARG, "vti_rtm_shot: injection row outside the updated region");
    const int nslot = std::max(0, o->nsnap > 0 ? std::min(o->nsnap, nt - 1) : nt - 1);
    RtmCache rc;
    struct Guard {
        RtmCache &r;
        ~Guard() { state_free(r.g); }
    } guard{rc};
    if ((st = state_alloc(c, rc.g))) return st;
    if ((st = ck_layout(c, rc.ck)) || (st = ck_alloc(c, rc.ck, nslot))) return st;
    const VtiState W = main_state(c);
    const int nd = c->cfg.ndim;
    const bool svf = o->sv_forward != 0 && c->cfg.sv != 0;
    const bool svb = o->sv_backward != 0 && c->cfg.sv != 0;
    const int clamp = o->illum_mode == 0;
    auto hip = [](hipError_t e, const char *w) { return e == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, w); };

    SweepOps op;
    op.zero = [&]() { return hip(state_zero(c, W), "zero"); };
    op.store = [&](int i) { return hip(ck_copy(c, rc.ck, i, W, true), "store"); };
    op.restore = [&](int i) { return hip(ck_copy(c, rc.ck, i, W, false), "restore"); };
    op.step = [&](int it) {
        if (launch_vel(nd, vel_args_for(c, W, it, true), c->stream) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
        return hip(launch_stress(nd, src_stress_args(c, W, it, svf, nullptr), c->stream), "stress");
    };
    op.first_step = [&](int it) {
        if (launch_vel(nd, vel_args_for(c, W, it, true), c->stream) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
        float *row = c->rec + (size_t)it * nrec;
        return hip(launch_stress(nd, with_illum(c, src_stress_args(c, W, it, svf, row), clamp), c->stream),
                   "stress");
    };
    op.after_first = [&]() {
        if (obs) {   // observed record [irec][it] replaces the modelled one (:1411-1415)
            if (hipMemcpyAsync(c->rec_t, obs, (size_t)nrec * nt * sizeof(float), hipMemcpyHostToDevice,
                               c->stream) != hipSuccess)
                return fail(VTI_ERR_HIP, "obs upload");
            return hip(launch_transpose(c->rec_t, c->rec, nrec, nt, c->stream), "transpose");
        }
        if (o->mute_tt >= 0) {
            const vti_shot &s = c->shot;
            const int np = c->cfg.npml;
            if (nd == 2) return vti_mute(c, s.sx - np + 1, 0, s.sz - np + 1, 1, o->mute_tt);
            return vti_mute3d_window(c, s.sx - np + 1, 0, s.sy - np + 1, 0, s.sz - np + 1, 0, o->mute_tt);
        }
        return VTI_OK;
    };
    // X_k of the source side: the working state, or a checkpoint slot
    // imaged from where it lies (it is not needed again)
    auto consume_from = [&](const VtiState &S, int k) {
        const int it = k - 1;
        if (launch_inject_rec(c, rc.g, c->rec + (size_t)it * nrec, inj_z) != hipSuccess)
            return fail(VTI_ERR_HIP, "inject");
        if (launch_vel(nd, vel_args_for(c, rc.g, it, false), c->stream) != hipSuccess)
            return fail(VTI_ERR_HIP, "vel");
        // illumination mode 1 adds the receiver side's P^2 + Q^2 in the
        // stress update's epilogue (same float operations as a separate pass)
        StressArgs ra = rcv_stress_args(c, rc.g, it, svb, inj_z);
        if (o->illum_mode == 1) ra = with_illum(c, ra, 0);
        if (launch_stress(nd, ra, c->stream) != hipSuccess) return fail(VTI_ERR_HIP, "stress");
        if (it % o->image_every == 0) {
            const int r = hip(launch_image(c, S.f[VTI_F_P], rc.g.f[VTI_F_P], c->img), "image");
            if (r) return r;
            // Poynting_Adcigs (:1475-1477) on the same source and receiver states
            return hip(launch_adcig(c, S, rc.g), "adcig");
        }
        return VTI_OK;
    };
    op.consume = [&](int k) { return consume_from(W, k); };
    op.consume_slot = [&](int k, int i) { return consume_from(rc.ck.view[i], k); };
    c->rtm_illum_mode = o->illum_mode;
    if ((st = reverse_sweep(nt, nslot, op, fwd_steps))) return st;
    return hip(hipStreamSynchronize(c->stream), "sync");
}

// One imaging step of a caller-driven RTM (cal_migration
// Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:679-698, 3D
// Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:508-521; cal_illumination
// :654-677): img += P_src * P_rcv and illum += P_src^2 + Q_src^2 over the
// interior (device arrays, z fastest, then x, then y), clamp: an
// accumulated illumination <= 0 becomes 1 (the 2D reference's form).  Runs
// on the receiver context's stream after the source context's queued work;
// the source context's later work waits for it.
int vti_image_accumulate(vti_ctx *src, vti_ctx *rcv, float *img, float *illum, int clamp)
{
    if (!src || !rcv || !img) return fail(VTI_ERR_ARG, "vti_image_accumulate: null argument");
    const vti_cfg &a = src->cfg, &b = rcv->cfg;
    if (a.ndim != b.ndim || a.nx != b.nx || a.nz != b.nz || (a.ndim == 3 && a.ny != b.ny) || a.npml != b.npml)
        return fail(VTI_ERR_ARG, "vti_image_accumulate: the contexts' grids differ");
    for (hipEvent_t &e : rcv->ev_acc)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail(VTI_ERR_HIP, "vti_image_accumulate: event");
    if (src != rcv && (hipEventRecord(rcv->ev_acc[0], src->stream) != hipSuccess ||
                       hipStreamWaitEvent(rcv->stream, rcv->ev_acc[0], 0) != hipSuccess))
        return fail(VTI_ERR_HIP, "vti_image_accumulate: ordering after the source context");
    if (launch_image(rcv, src->f[VTI_F_P], rcv->f[VTI_F_P], img) != hipSuccess ||
        (illum && launch_illum(rcv, src->f[VTI_F_P], src->f[VTI_F_Q], illum, clamp ? 1 : 0) != hipSuccess))
        return fail(VTI_ERR_HIP, "vti_image_accumulate: launch");
    if (src != rcv && (hipEventRecord(rcv->ev_acc[1], rcv->stream) != hipSuccess ||
                       hipStreamWaitEvent(src->stream, rcv->ev_acc[1], 0) != hipSuccess))
        return fail(VTI_ERR_HIP, "vti_image_accumulate: ordering of the source context");
    return VTI_OK;
}

int vti_rtm_image(vti_ctx *c, int finalize, float *img, float *illum)
{
    if (!c || !c->img) return fail(VTI_ERR_STATE, "vti_rtm_image: no image");
    const size_t n = c->nint;
    std::vector<float> hi(n), hl(n);
    if (hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(hi.data(), c->img, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hl.data(), c->illum, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(VTI_ERR_HIP, "vti_rtm_image: copy");
    if (illum) std::memcpy(illum, hl.data(), n * sizeof(float));
    if (!img) return VTI_OK;
    if (!finalize) {
        std::memcpy(img, hi.data(), n * sizeof(float));
        return VTI_OK;
    }
    // migration_illum then laplace_filter(adj=1) (:1486-1493)
    const bool guard = c->rtm_illum_mode == 1;
    for (size_t i = 0; i < n; ++i)
        if (!guard || hl[i] != 0.0f) hi[i] /= hl[i];
    laplace_adj(c->cfg.nz, c->cfg.nx, c->cfg.ndim == 3 ? c->cfg.ny : 1, hi.data(), img);
    return VTI_OK;
}

}  // extern "C"
