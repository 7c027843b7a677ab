// vti_grad.hip — exact discrete adjoint of the staggered-grid VTI step and
// the misfit gradient with respect to vp (FWI), in FP32 or FP64.
//
// The reference has no VTI adjoint; its only FWI gradient is the ISO
// 2nd-order one (3D-iso-fd-rtm-fwi-cuda/fwi3d/Toa_gpu_3dFWI.cu:295-311,
// g += s_tt * g_P), computed against a boundary-reconstructed source field.
// Here the gradient is the transpose of the forward step actually run:
//
//   forward step n (region R = cells 4 .. nn-5 on every axis):
//     P[s] += a_n, Q[s] += a_n                         (point source)
//     u  = cx2 u  - cx1 dtx D+x P      (v, w alike)    (D+ f = sum c_m (f[x+m+1]-f[x-m]))
//     px = ax2 px - Cpx D-x u          (py pz qx qy qz) (D- f = sum c_m (f[x+m]-f[x-m-1]))
//     P  = px+py+pz, Q = qx+qy+qz,  d_n = P at the receivers
//   Cpx = acoffx1 vp^2 (1+2e) dtx, Cqx = acoffx1 vp^2 sqrt(1+2d) dtx, ...
//
//   adjoint step n (reverse order), on R:
//     P~[rec] += r_n
//     sx = px~ + P~ (...), P~ = Q~ = 0;  grad_C += s * (D- u_n)   (u_n: forward u after step n)
//     F = Cpx sx + Cqx qx~;  u~ += D+ F;  px~ = ax2 sx
//     H = cx1 dtx u~;  P~ += D- H;  u~ = cx2 u~
//     a~_n = P~[s] + Q~[s]
//
// Because every C is proportional to vp^2, dJ/dvp = -2/vp * sum_n sum_k s_k C_k (D- u_n)_k.
// The absorbing-layer coefficients are held fixed (not differentiated).
// Source-wavefield states are recomputed from checkpoints with the same
// binomial schedule as RTM (reverse_sweep).
//
// For T = float the forward is bit-for-bit the production kernels'
// arithmetic (the per-cell C carries the reference's product order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "vti_internal.h"
#include "vti_rtm.h"

namespace vti {
namespace {

constexpr int TZ = 64, TX = 4;

template <typename T>
__device__ __forceinline__ T cm(int m)
{
    return m == 0 ? T(1.196289f) : (m == 1 ? T(-0.0797526f) : (m == 2 ? T(0.009570313f) : T(-0.0006975447f)));
}
template <>
__device__ __forceinline__ double cm<double>(int m)
{
    return m == 0 ? 1.196289 : (m == 1 ? -0.0797526 : (m == 2 ? 0.009570313 : -0.0006975447));
}

struct GGrid {
    int nnx, nny, nnz, ldz;
    long long ldxy;
    int three;
};

__device__ __forceinline__ bool cell(const GGrid &g, int &ix, int &iy, int &iz, long long &id)
{
    iz = blockIdx.x * TZ + threadIdx.x;
    ix = blockIdx.y * TX + threadIdx.y;
    iy = g.three ? (int)blockIdx.z + 4 : 0;
    if (iz < 4 || iz >= g.nnz - 4 || ix < 4 || ix >= g.nnx - 4) return false;
    id = iz + (long long)ix * g.ldz + (long long)iy * g.ldxy;
    return true;
}

template <typename T>
struct Fields {
    T *u, *v, *w, *px, *py, *pz, *qx, *qy, *qz, *P, *Q;
};

template <typename T>
struct Coef {
    const T *C[6];    // px py pz qx qy qz
    const T *tab;     // 12 tables, M each
    int M;
    T dtx, dty, dtz;
};

// ---- forward --------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void gvel(GGrid g, Fields<T> f, Coef<T> c, long long src, T amp)
{
    int ix, iy, iz;
    long long id;
    if (!cell(g, ix, iy, iz, id)) return;
    const long long sx = g.ldz, sy = g.ldxy;
    auto ldP = [&](long long i) { T v = f.P[i]; if (i == src) v = v + amp; return v; };
    auto ldQ = [&](long long i) { T v = f.Q[i]; if (i == src) v = v + amp; return v; };
    T xx = 0, yy = 0, zz = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (g.three) yy += cm<T>(m) * (ldP(id + (m + 1) * sy) - ldP(id - m * sy));
        xx += cm<T>(m) * (ldP(id + (m + 1) * sx) - ldP(id - m * sx));
        zz += cm<T>(m) * (ldQ(id + m + 1) - ldQ(id - m));
    }
    const T *t = c.tab;
    const int M = c.M;
    f.u[id] = t[1 * M + ix] * f.u[id] - t[0 * M + ix] * c.dtx * xx;
    if (g.three) f.v[id] = t[3 * M + iy] * f.v[id] - t[2 * M + iy] * c.dty * yy;
    f.w[id] = t[5 * M + iz] * f.w[id] - t[4 * M + iz] * c.dtz * zz;
}

template <typename T>
__global__ __launch_bounds__(256) void gstress(GGrid g, Fields<T> f, Coef<T> c, T *rec, int rkind, int rz,
                                               int rx0, int rdx, int rnx, int ry0, int rdy, int rny)
{
    int ix, iy, iz;
    long long id;
    if (!cell(g, ix, iy, iz, id)) return;
    const long long sx = g.ldz, sy = g.ldxy;
    T xx = 0, yy = 0, zz = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (g.three) yy += cm<T>(m) * (f.v[id + m * sy] - f.v[id - (m + 1) * sy]);
        xx += cm<T>(m) * (f.u[id + m * sx] - f.u[id - (m + 1) * sx]);
        zz += cm<T>(m) * (f.w[id + m] - f.w[id - m - 1]);
    }
    const T *t = c.tab;
    const int M = c.M;
    const T ax2 = t[7 * M + ix], az2 = t[11 * M + iz];
    const T px = ax2 * f.px[id] - c.C[0][id] * xx;
    const T pz = az2 * f.pz[id] - c.C[2][id] * zz;
    const T qx = ax2 * f.qx[id] - c.C[3][id] * xx;
    const T qz = az2 * f.qz[id] - c.C[5][id] * zz;
    f.px[id] = px; f.pz[id] = pz; f.qx[id] = qx; f.qz[id] = qz;
    T Pn, Qn;
    if (g.three) {
        const T ay2 = t[9 * M + iy];
        const T py = ay2 * f.py[id] - c.C[1][id] * yy;
        const T qy = ay2 * f.qy[id] - c.C[4][id] * yy;
        f.py[id] = py; f.qy[id] = qy;
        Pn = px + py + pz;
        Qn = qx + qy + qz;
    } else {
        Pn = px + pz;
        Qn = qx + qz;
    }
    f.P[id] = Pn;
    f.Q[id] = Qn;
    if (rec && iz == rz) {
        const int k = ix - rx0, j = g.three ? iy - ry0 : 0;
        if (k >= 0 && k % rdx == 0 && k / rdx < rnx && j >= 0 && j % rdy == 0 && j / rdy < rny) {
            const T r = rkind == 0 ? Pn : (rkind == 1 ? Pn + Pn : Pn + Qn);
            rec[k / rdx + (long long)rnx * (j / rdy)] = r;
        }
    }
}

// Per-cell stress coefficients with the SV taper of the shot, in the
// reference's product order: ((a1*vp)*vp)*(1+2e)*dt/dx.
template <typename T>
__global__ void gcoef(GGrid g, const T *vp, const T *ep, const T *de, const T *tab, int M, T dtx, T dty,
                      T dtz, int sv, int svR, int cx, int cy, int cz, const double *tap, T *C0, T *C1,
                      T *C2, T *C3, T *C4, T *C5)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y, iy = blockIdx.z;
    if (iz >= g.nnz) return;
    const long long id = iz + (long long)ix * g.ldz + (long long)iy * g.ldxy;
    const T vv = vp[id];
    T ee = ep[id], dd = de[id];
    if (sv) {
        const int rx = ix - cx, ry = g.three ? iy - cy : 0, rz = iz - cz;
        const int r2 = rx * rx + ry * ry + rz * rz;
        if (r2 <= svR * svR) {
            ee = (T)(tap[r2] * (double)ee);
            dd = (T)(tap[r2] * (double)dd);
        }
    }
    const T e2 = T(1) + T(2) * ee;
    const T sd = sqrt(T(1) + T(2) * dd);
    const T ax1 = tab[6 * M + ix], ay1 = g.three ? tab[8 * M + iy] : T(0), az1 = tab[10 * M + iz];
    C0[id] = ax1 * vv * vv * e2 * dtx;
    C1[id] = ay1 * vv * vv * e2 * dty;
    C2[id] = az1 * vv * vv * sd * dtz;
    C3[id] = ax1 * vv * vv * sd * dtx;
    C4[id] = ay1 * vv * vv * sd * dty;
    C5[id] = az1 * vv * vv * dtz;
}

// ---- adjoint --------------------------------------------------------------
template <typename T>
__global__ void gadj_rec(GGrid g, Fields<T> a, const T *r, int rkind, int rz, int rx0, int rdx, int rnx,
                         int ry0, int rdy, int rny)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rnx * rny) return;
    const int ix = rx0 + (k % rnx) * rdx, iy = g.three ? ry0 + (k / rnx) * rdy : 0;
    const long long c = rz + (long long)ix * g.ldz + (long long)iy * g.ldxy;
    const T v = r[k];
    if (rkind == 0) a.P[c] += v;
    else if (rkind == 1) a.P[c] += v + v;
    else { a.P[c] += v; a.Q[c] += v; }
}

// stress^T, phase A: distribute P~, Q~ to the splits, accumulate the
// gradient against the forward state fw (after the step), form F.
template <typename T>
__global__ __launch_bounds__(256) void gadj_stress_a(GGrid g, Fields<T> a, Fields<T> fw, Coef<T> c, T *Fx,
                                                     T *Fy, T *Fz, T *grad)
{
    int ix, iy, iz;
    long long id;
    if (!cell(g, ix, iy, iz, id)) return;
    const long long sx = g.ldz, sy = g.ldxy;
    const T Pa = a.P[id], Qa = a.Q[id];
    const T spx = a.px[id] + Pa, spz = a.pz[id] + Pa, sqx = a.qx[id] + Qa, sqz = a.qz[id] + Qa;
    a.P[id] = 0;
    a.Q[id] = 0;
    T xx = 0, yy = 0, zz = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        xx += cm<T>(m) * (fw.u[id + m * sx] - fw.u[id - (m + 1) * sx]);
        zz += cm<T>(m) * (fw.w[id + m] - fw.w[id - m - 1]);
        if (g.three) yy += cm<T>(m) * (fw.v[id + m * sy] - fw.v[id - (m + 1) * sy]);
    }
    const T cpx = c.C[0][id], cpz = c.C[2][id], cqx = c.C[3][id], cqz = c.C[5][id];
    T gsum = (spx * cpx + sqx * cqx) * xx + (spz * cpz + sqz * cqz) * zz;
    Fx[id] = cpx * spx + cqx * sqx;
    Fz[id] = cpz * spz + cqz * sqz;
    const T *t = c.tab;
    const int M = c.M;
    const T ax2 = t[7 * M + ix], az2 = t[11 * M + iz];
    a.px[id] = ax2 * spx; a.pz[id] = az2 * spz;
    a.qx[id] = ax2 * sqx; a.qz[id] = az2 * sqz;
    if (g.three) {
        const T spy = a.py[id] + Pa, sqy = a.qy[id] + Qa;
        const T cpy = c.C[1][id], cqy = c.C[4][id];
        gsum += (spy * cpy + sqy * cqy) * yy;
        Fy[id] = cpy * spy + cqy * sqy;
        const T ay2 = t[9 * M + iy];
        a.py[id] = ay2 * spy; a.qy[id] = ay2 * sqy;
    }
    grad[id] += gsum;
}

// stress^T, phase B: u~ += D+ F (gather form of the transpose of D-).
template <typename T>
__global__ __launch_bounds__(256) void gadj_stress_b(GGrid g, Fields<T> a, const T *Fx, const T *Fy, const T *Fz)
{
    int ix, iy, iz;
    long long id;
    if (!cell(g, ix, iy, iz, id)) return;
    const long long sx = g.ldz, sy = g.ldxy;
    T ux = 0, vy = 0, wz = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        ux += cm<T>(m) * (Fx[id + (m + 1) * sx] - Fx[id - m * sx]);
        wz += cm<T>(m) * (Fz[id + m + 1] - Fz[id - m]);
        if (g.three) vy += cm<T>(m) * (Fy[id + (m + 1) * sy] - Fy[id - m * sy]);
    }
    a.u[id] += ux;
    a.w[id] += wz;
    if (g.three) a.v[id] += vy;
}

// vel^T, phase A: H = coff1*dt/dx * u~, u~ *= coff2.
template <typename T>
__global__ __launch_bounds__(256) void gadj_vel_a(GGrid g, Fields<T> a, Coef<T> c, T *Hx, T *Hy, T *Hz)
{
    int ix, iy, iz;
    long long id;
    if (!cell(g, ix, iy, iz, id)) return;
    const T *t = c.tab;
    const int M = c.M;
    const T ua = a.u[id], wa = a.w[id];
    Hx[id] = t[0 * M + ix] * c.dtx * ua;
    Hz[id] = t[4 * M + iz] * c.dtz * wa;
    a.u[id] = t[1 * M + ix] * ua;
    a.w[id] = t[5 * M + iz] * wa;
    if (g.three) {
        const T va = a.v[id];
        Hy[id] = t[2 * M + iy] * c.dty * va;
        a.v[id] = t[3 * M + iy] * va;
    }
}

// vel^T, phase B: P~ += D- Hx (+ D- Hy), Q~ += D- Hz.
template <typename T>
__global__ __launch_bounds__(256) void gadj_vel_b(GGrid g, Fields<T> a, const T *Hx, const T *Hy, const T *Hz)
{
    int ix, iy, iz;
    long long id;
    if (!cell(g, ix, iy, iz, id)) return;
    const long long sx = g.ldz, sy = g.ldxy;
    T p = 0, q = 0, py = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        p += cm<T>(m) * (Hx[id + m * sx] - Hx[id - (m + 1) * sx]);
        q += cm<T>(m) * (Hz[id + m] - Hz[id - m - 1]);
        if (g.three) py += cm<T>(m) * (Hy[id + m * sy] - Hy[id - (m + 1) * sy]);
    }
    a.P[id] += g.three ? p + py : p;
    a.Q[id] += q;
}

template <typename T>
__global__ void gsrc_grad(const T *P, const T *Q, long long src, T *out)
{
    if (threadIdx.x == 0) *out = P[src] + Q[src];
}

template <typename T>
__global__ void gzero(T *p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = T(0);
}

template <typename T>
__global__ void gpad(GGrid g, int nx, int ny, int nz, int np, const T *src, T *dst)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y, iy = blockIdx.z;
    if (iz >= g.nnz) return;
    const int jx = ix < np ? 0 : (ix >= np + nx ? nx - 1 : ix - np);
    const int jz = iz < np ? 0 : (iz >= np + nz ? nz - 1 : iz - np);
    int jy = 0;
    if (g.three) jy = iy < np ? 0 : (iy >= np + ny ? ny - 1 : iy - np);
    dst[iz + (long long)ix * g.ldz + (long long)iy * g.ldxy] = src[jz + (long long)jx * nz + (long long)jy * nz * nx];
}

}  // namespace

// ---- host tables in precision T ---------------------------------------------
template <typename T>
static void axis_tables_t(int nn_int, int npml, T dtd0, int cpu_form, T *c1, T *c2, T *a1, T *a2)
{
    const int n = nn_int + 2 * npml;
    for (int i = 0; i < n; ++i) {
        if (i >= npml && i < npml + nn_int) {
            c1[i] = c2[i] = a1[i] = a2[i] = T(1);
            continue;
        }
        double xv, xs;
        if (i < npml) {
            xv = (npml - 0.5 - i) / npml;
            xs = ((npml - i) * 1.0) / npml;
        } else {
            xv = (0.5 + i - nn_int - npml) / npml;
            xs = cpu_form ? ((1 + i - nn_int - npml) * 1.0) / npml : ((i - nn_int - npml) * 1.0) / npml;
        }
        const double hv = (double)dtd0 * std::pow(xv, 2.0), hs = (double)dtd0 * std::pow(xs, 2.0);
        c1[i] = (T)(1.0 / (1.0 + hv / 2.0));
        c2[i] = (T)((double)c1[i] * (1.0 - hv / 2.0));
        a1[i] = (T)(1.0 / (1.0 + hs / 2.0));
        a2[i] = (T)((double)(cpu_form ? c1[i] : a1[i]) * (1.0 - hs / 2.0));
    }
}

template <typename T>
static T wavelet_t(T ts, T favg, int wtype)
{
    if (wtype == 1) {
        const T x = favg * ts, xx = x * x;
        return (T)((1 - 2 * kPi * kPi * xx) * std::exp(-(kPi * kPi * xx)));
    }
    if (wtype == 2) {
        const T x = (T)((-4) * favg * favg * kPi * kPi / std::log(0.1));
        const T e = -x * ts * ts;
        return (T)((-2) * kPi * kPi * ts * std::exp((double)e));
    }
    if (wtype == 3) {
        const T x = (T)((-1) * favg * favg * kPi * kPi / std::log(0.1));
        const T e = -x * ts * ts;
        return (T)std::exp((double)e);
    }
    return T(0);
}

}  // namespace vti

using namespace vti;

struct vti_grad {
    virtual ~vti_grad() {}
};

namespace vti {
namespace {

template <typename T>
struct Grad : vti_grad {
    vti_cfg cfg{};
    GGrid g{};
    size_t nvol = 0;
    int M = 0;
    hipStream_t stream = nullptr;
    T *arena = nullptr;
    T *vp = nullptr, *ep = nullptr, *de = nullptr, *C[6] = {}, *tab = nullptr, *grad = nullptr;
    T *tmp[3] = {};
    Fields<T> f{}, a{};
    std::vector<T> htab, amp;
    std::vector<double> hvp;     // interior vp (host) for the final scaling
    bool have_tables = false, have_model = false, have_shot = false, custom_src = false;
    vti_shot shot{};
    int nrec = 0;
    T *rec = nullptr, *res = nullptr, *sgrad = nullptr;
    double *dtap = nullptr;
    std::vector<T *> slots;

    ~Grad() override
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (T *s : slots) (void)hipFree(s);
        (void)hipFree(arena);
        (void)hipFree(tab);
        (void)hipFree(rec);
        (void)hipFree(res);
        (void)hipFree(sgrad);
        (void)hipFree(dtap);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int nstate() const { return g.three ? 11 : 8; }

    static void set_fields(Fields<T> &F, T *base, size_t nvol, bool three)
    {
        T *p[11] = {};
        if (three)
            for (int k = 0; k < 11; ++k) p[k] = base + k * nvol;
        else {
            const int ord[8] = {0, 2, 3, 5, 6, 8, 9, 10};
            for (int k = 0; k < 8; ++k) p[ord[k]] = base + k * nvol;
        }
        F = Fields<T>{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10]};
    }

    int create(const vti_cfg &c)
    {
        cfg = c;
        if (cfg.sv_R <= 0) cfg.sv_R = 15;
        if (cfg.sv_r <= 0) cfg.sv_r = 4;
        if (cfg.rec_kind == VTI_REC_DEFAULT) cfg.rec_kind = c.mode == VTI_MODE_CPU_REF ? VTI_REC_2P : VTI_REC_P;
        if (c.mode != VTI_MODE_GPU_REF) return fail(VTI_ERR_ARG, "vti_grad: the point-source (GPU) form only");
        const bool three = c.ndim == 3;
        g.three = three;
        g.nnx = c.nx + 2 * c.npml;
        g.nnz = c.nz + 2 * c.npml;
        g.nny = three ? c.ny + 2 * c.npml : 1;
        if (g.nnx < 9 || g.nnz < 9 || (three && g.nny < 9)) return fail(VTI_ERR_ARG, "vti_grad: grid too small");
        g.ldz = (g.nnz + 31) / 32 * 32;
        g.ldxy = (long long)g.ldz * g.nnx;
        nvol = (size_t)g.ldxy * g.nny;
        M = std::max(g.nnx, std::max(g.nny, g.nnz));
        const int nf = 3 + 6 + 1 + 3 + 2 * nstate();   // model, C, grad, tmp, fwd, adj
        if (hipMalloc(&arena, (size_t)nf * nvol * sizeof(T)) != hipSuccess) return fail(VTI_ERR_ALLOC, "vti_grad: arena");
        if (hipMalloc(&tab, 12 * (size_t)M * sizeof(T)) != hipSuccess) return fail(VTI_ERR_ALLOC, "vti_grad: tables");
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return fail(VTI_ERR_HIP, "stream");
        T *p = arena;
        vp = p; p += nvol; ep = p; p += nvol; de = p; p += nvol;
        for (int k = 0; k < 6; ++k) { C[k] = p; p += nvol; }
        grad = p; p += nvol;
        for (int k = 0; k < 3; ++k) { tmp[k] = p; p += nvol; }
        set_fields(f, p, nvol, three); p += nstate() * nvol;
        set_fields(a, p, nvol, three);
        zero(arena, (size_t)nf * nvol);
        htab.assign(12 * (size_t)M, T(1));
        return sync();
    }

    void zero(T *p, size_t n) { hipLaunchKernelGGL(gzero<T>, dim3(4096), dim3(256), 0, stream, p, n); }
    int sync() { return hipStreamSynchronize(stream) == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, "vti_grad: sync"); }
    dim3 grid() const { return dim3((g.nnz + TZ - 1) / TZ, (g.nnx + TX - 1) / TX, g.three ? g.nny - 8 : 1); }
    Coef<T> coef() const
    {
        Coef<T> c{};
        for (int k = 0; k < 6; ++k) c.C[k] = C[k];
        c.tab = tab;
        c.M = M;
        c.dtx = (T)cfg.dt / (T)cfg.dx;
        c.dty = (T)cfg.dt / (T)cfg.dy;
        c.dtz = (T)cfg.dt / (T)cfg.dz;
        return c;
    }

    int set_model(const double *hv, const double *he, const double *hd, int keep_tables)
    {
        const int ny = g.three ? cfg.ny : 1;
        const size_t n = (size_t)cfg.nx * ny * cfg.nz;
        std::vector<T> bv(n), be(n), bd(n);
        for (size_t i = 0; i < n; ++i) {
            if (!(hv[i] > 0) || !std::isfinite(hv[i])) return fail(VTI_ERR_ARG, "vti_grad: vp must be positive");
            bv[i] = (T)hv[i]; be[i] = (T)he[i]; bd[i] = (T)hd[i];
        }
        hvp.assign(hv, hv + n);
        if (!have_tables || !keep_tables) {
            // get_d02d / get_d03d (GPU form), evaluated in precision T
            const int np = cfg.npml;
            T d0 = 0;
            if (np > 0) {
                const long long N = g.three ? (long long)g.nny * g.nnx * g.nnz / 2 : (long long)g.nnx * g.nnz / 2;
                const long long iz = N % g.nnz, r = N / g.nnz, ix = r % g.nnx, iy = r / g.nnx;
                auto cl = [&](long long i, int nn) { return i < np ? 0 : (i >= np + nn ? nn - 1 : i - np); };
                const T v = bv[cl(iz, cfg.nz) + cl(ix, cfg.nx) * (long long)cfg.nz +
                               (g.three ? cl(iy, cfg.ny) * (long long)cfg.nz * cfg.nx : 0)];
                if (!g.three) {
                    const T sdx = (T)cfg.dx + (T)cfg.dz;
                    d0 = (T)(10.0 * v * std::log(100000.0) / (2.0 * np * (sdx / 2.0)));
                } else {
                    const T sdx = (T)cfg.dx + (T)cfg.dy + (T)cfg.dz;
                    d0 = (T)(10.0 * v * std::log(100000.0) / (2.0 * np * (sdx / 3.0)));
                }
            }
            const T dtd0 = (T)cfg.dt * d0;
            T *h = htab.data();
            std::fill(htab.begin(), htab.end(), T(1));
            axis_tables_t<T>(cfg.nx, np, dtd0, 0, h + 0 * M, h + 1 * M, h + 6 * M, h + 7 * M);
            axis_tables_t<T>(cfg.nz, np, dtd0, 0, h + 4 * M, h + 5 * M, h + 10 * M, h + 11 * M);
            if (g.three) axis_tables_t<T>(cfg.ny, np, dtd0, 0, h + 2 * M, h + 3 * M, h + 8 * M, h + 9 * M);
            if (hipMemcpy(tab, h, htab.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
                return fail(VTI_ERR_HIP, "tables");
            have_tables = true;
        }
        T *stage = nullptr;
        if (hipMalloc(&stage, n * sizeof(T)) != hipSuccess) return fail(VTI_ERR_ALLOC, "stage");
        const T *srcs[3] = {bv.data(), be.data(), bd.data()};
        T *dsts[3] = {vp, ep, de};
        int st = VTI_OK;
        for (int k = 0; k < 3 && !st; ++k) {
            if (hipMemcpy(stage, srcs[k], n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) st = fail(VTI_ERR_HIP, "upload");
            hipLaunchKernelGGL(gpad<T>, dim3((g.nnz + 63) / 64, g.nnx, g.nny), dim3(64), 0, stream, g, cfg.nx,
                               ny, cfg.nz, cfg.npml, stage, dsts[k]);
            if (!st) st = sync();
        }
        (void)hipFree(stage);
        have_model = !st;
        if (!st && have_shot) st = build_coef();
        return st;
    }

    int build_coef()
    {
        const Coef<T> c = coef();
        hipLaunchKernelGGL(gcoef<T>, dim3((g.nnz + 63) / 64, g.nnx, g.nny), dim3(64), 0, stream, g, vp, ep, de,
                           tab, M, c.dtx, c.dty, c.dtz, cfg.sv, cfg.sv_R, shot.sv_cx, shot.sv_cy, shot.sv_cz,
                           dtap, C[0], C[1], C[2], C[3], C[4], C[5]);
        return sync();
    }

    int set_shot(const vti_shot &s)
    {
        auto inR = [&](int x, int y, int z) {
            return x >= 4 && x < g.nnx - 4 && z >= 4 && z < g.nnz - 4 && (g.three ? (y >= 4 && y < g.nny - 4) : y == 0);
        };
        vti_shot sh = s;
        if (!g.three) { sh.sy = sh.sv_cy = sh.rec_y0 = 0; sh.rec_dy = 1; sh.rec_ny = 1; }
        if (!inR(sh.sx, sh.sy, sh.sz)) return fail(VTI_ERR_ARG, "vti_grad: source outside the updated region");
        if (sh.rec_nx < 1 || sh.rec_ny < 1 || sh.rec_dx < 1 || sh.rec_dy < 1 ||
            !inR(sh.rec_x0, sh.rec_y0, sh.rec_z) ||
            !inR(sh.rec_x0 + (sh.rec_nx - 1) * sh.rec_dx, sh.rec_y0 + (sh.rec_ny - 1) * sh.rec_dy, sh.rec_z))
            return fail(VTI_ERR_ARG, "vti_grad: receivers outside the updated region");
        shot = sh;
        const int nr = sh.rec_nx * sh.rec_ny;
        if (nr != nrec || !rec) {
            (void)hipFree(rec); (void)hipFree(res); rec = res = nullptr;
            nrec = nr;
            const size_t nb = (size_t)nrec * cfg.nt * sizeof(T);
            if (hipMalloc(&rec, nb) != hipSuccess || hipMalloc(&res, nb) != hipSuccess)
                return fail(VTI_ERR_ALLOC, "vti_grad: record buffers");
        }
        if (!sgrad && hipMalloc(&sgrad, (size_t)cfg.nt * sizeof(T)) != hipSuccess)
            return fail(VTI_ERR_ALLOC, "vti_grad: source gradient");
        std::vector<double> tap((size_t)cfg.sv_R * cfg.sv_R + 1, 0.0);
        for (int r2 = 0; r2 <= cfg.sv_R * cfg.sv_R; ++r2)
            if (r2 > cfg.sv_r * cfg.sv_r) {
                const double sr = sizeof(T) == 4 ? (double)std::sqrt((float)r2) : std::sqrt((double)r2);
                tap[r2] = 0.5 * (1 - std::cos(kPi * ((sr - cfg.sv_r) * 4.0 / (cfg.sv_R * 3.0 - 1))));
            }
        (void)hipFree(dtap);
        if (hipMalloc(&dtap, tap.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(dtap, tap.data(), tap.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return fail(VTI_ERR_ALLOC, "vti_grad: taper");
        if (!custom_src) {
            amp.assign(cfg.nt, T(0));
            T t = (T)cfg.dt;
            const T tdelay = (T)(1.0 / cfg.favg);
            for (int it = 0; it < cfg.nt; ++it) {
                if (t <= 2 * tdelay) amp[it] = (T)cfg.pfac * wavelet_t<T>(t - tdelay, (T)cfg.favg, cfg.wtype);
                t += (T)cfg.dt;
            }
        }
        have_shot = true;
        return have_model ? build_coef() : VTI_OK;
    }

    long long src_index() const { return shot.sz + (long long)shot.sx * g.ldz + (long long)shot.sy * g.ldxy; }

    void fwd_step(int it, bool record)
    {
        const Coef<T> c = coef();
        const long long src = amp[it] != T(0) ? src_index() : -1;
        hipLaunchKernelGGL(gvel<T>, grid(), dim3(TZ, TX), 0, stream, g, f, c, src, amp[it]);
        hipLaunchKernelGGL(gstress<T>, grid(), dim3(TZ, TX), 0, stream, g, f, c,
                           record ? rec + (size_t)it * nrec : (T *)nullptr, cfg.rec_kind, shot.rec_z,
                           shot.rec_x0, shot.rec_dx, shot.rec_nx, shot.rec_y0, shot.rec_dy, shot.rec_ny);
    }

    void adj_step(int it, const Fields<T> &fw)
    {
        const Coef<T> c = coef();
        hipLaunchKernelGGL(gadj_rec<T>, dim3((nrec + 255) / 256), dim3(256), 0, stream, g, a, res + (size_t)it * nrec,
                           cfg.rec_kind, shot.rec_z, shot.rec_x0, shot.rec_dx, shot.rec_nx, shot.rec_y0,
                           shot.rec_dy, shot.rec_ny);
        hipLaunchKernelGGL(gadj_stress_a<T>, grid(), dim3(TZ, TX), 0, stream, g, a, fw, c, tmp[0], tmp[1], tmp[2], grad);
        hipLaunchKernelGGL(gadj_stress_b<T>, grid(), dim3(TZ, TX), 0, stream, g, a, (const T *)tmp[0],
                           (const T *)tmp[1], (const T *)tmp[2]);
        hipLaunchKernelGGL(gadj_vel_a<T>, grid(), dim3(TZ, TX), 0, stream, g, a, c, tmp[0], tmp[1], tmp[2]);
        hipLaunchKernelGGL(gadj_vel_b<T>, grid(), dim3(TZ, TX), 0, stream, g, a, (const T *)tmp[0],
                           (const T *)tmp[1], (const T *)tmp[2]);
        hipLaunchKernelGGL(gsrc_grad<T>, dim3(1), dim3(64), 0, stream, (const T *)a.P, (const T *)a.Q, src_index(),
                           sgrad + it);
    }

    T *state_base(const Fields<T> &F) const { return F.u; }

    int forward(double *d_out)
    {
        if (!have_model || !have_shot) return fail(VTI_ERR_STATE, "vti_grad: model and shot first");
        zero(state_base(f), (size_t)nstate() * nvol);
        for (int it = 0; it < cfg.nt; ++it) fwd_step(it, true);
        int st = sync();
        if (st || !d_out) return st;
        return fetch_record(d_out);
    }

    int fetch_record(double *d_out)
    {
        std::vector<T> h((size_t)nrec * cfg.nt);
        if (hipMemcpy(h.data(), rec, h.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(VTI_ERR_HIP, "record copy");
        for (int it = 0; it < cfg.nt; ++it)
            for (int r = 0; r < nrec; ++r) d_out[it + (size_t)cfg.nt * r] = (double)h[(size_t)it * nrec + r];
        return VTI_OK;
    }

    int run(const double *dobs, const double *resid, int nsnap, double *misfit, double *grad_vp,
            double *grad_src, double *d_out, long long *steps)
    {
        if (!have_model || !have_shot) return fail(VTI_ERR_STATE, "vti_grad: model and shot first");
        if (!dobs && !resid) return fail(VTI_ERR_ARG, "vti_grad: observed data or residual needed");
        const int nt = cfg.nt;
        const int nslot = std::max(0, nsnap > 0 ? std::min(nsnap, nt - 1) : nt - 1);
        while ((int)slots.size() < nslot) {
            T *s = nullptr;
            if (hipMalloc(&s, (size_t)nstate() * nvol * sizeof(T)) != hipSuccess)
                return fail(VTI_ERR_ALLOC, "vti_grad: checkpoint slots");
            slots.push_back(s);
        }
        zero(state_base(a), (size_t)nstate() * nvol);
        zero(grad, nvol);
        zero(tmp[0], 3 * nvol);
        auto hip_ok = []() { return hipGetLastError() == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, "launch"); };
        const size_t sbytes = (size_t)nstate() * nvol * sizeof(T);
        SweepOps op;
        op.zero = [&]() { zero(state_base(f), (size_t)nstate() * nvol); return hip_ok(); };
        op.store = [&](int i) {
            return hipMemcpyAsync(slots[i], state_base(f), sbytes, hipMemcpyDeviceToDevice, stream) == hipSuccess
                       ? VTI_OK : fail(VTI_ERR_HIP, "store");
        };
        op.restore = [&](int i) {
            return hipMemcpyAsync(state_base(f), slots[i], sbytes, hipMemcpyDeviceToDevice, stream) == hipSuccess
                       ? VTI_OK : fail(VTI_ERR_HIP, "restore");
        };
        op.step = [&](int it) { fwd_step(it, false); return hip_ok(); };
        op.first_step = [&](int it) { fwd_step(it, true); return hip_ok(); };
        double J = 0;
        op.after_first = [&]() {
            int st = sync();
            if (st) return st;
            std::vector<T> h((size_t)nrec * nt);
            if (hipMemcpy(h.data(), rec, h.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(VTI_ERR_HIP, "record copy");
            for (int it = 0; it < nt; ++it)
                for (int r = 0; r < nrec; ++r) {
                    const size_t k = (size_t)it * nrec + r, kr = it + (size_t)nt * r;
                    const double rr = resid ? resid[kr] : (double)h[k] - dobs[kr];
                    if (d_out) d_out[kr] = (double)h[k];
                    J += 0.5 * rr * rr;
                    h[k] = (T)rr;
                }
            return hipMemcpy(res, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess
                       ? VTI_OK : fail(VTI_ERR_HIP, "residual upload");
        };
        op.consume = [&](int k) { adj_step(k - 1, f); return hip_ok(); };
        int st = reverse_sweep(nt, nslot, op, steps);
        if (st || (st = sync())) return st;
        if (misfit) *misfit = J;
        if (grad_src) {
            std::vector<T> h(nt);
            if (hipMemcpy(h.data(), sgrad, nt * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(VTI_ERR_HIP, "source gradient copy");
            for (int it = 0; it < nt; ++it) grad_src[it] = (double)h[it];
        }
        if (grad_vp) {
            std::vector<T> h(nvol);
            if (hipMemcpy(h.data(), grad, nvol * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(VTI_ERR_HIP, "gradient copy");
            const int np = cfg.npml, ny = g.three ? cfg.ny : 1;
            const size_t ni = (size_t)cfg.nx * ny * cfg.nz;
            std::fill(grad_vp, grad_vp + ni, 0.0);
            for (int iy = 0; iy < g.nny; ++iy)
                for (int ix = 0; ix < g.nnx; ++ix)
                    for (int iz = 0; iz < g.nnz; ++iz) {
                        auto cl = [&](int i, int nn) { return i < np ? 0 : (i >= np + nn ? nn - 1 : i - np); };
                        const size_t j = cl(iz, cfg.nz) + (size_t)cl(ix, cfg.nx) * cfg.nz +
                                         (g.three ? (size_t)cl(iy, cfg.ny) * cfg.nz * cfg.nx : 0);
                        const double gc = (double)h[iz + (size_t)ix * g.ldz + (size_t)iy * g.ldxy];
                        grad_vp[j] += -2.0 * gc / hvp[j];
                    }
        }
        return VTI_OK;
    }
};

}  // namespace
}  // namespace vti

extern "C" {

int vti_grad_create(const vti_cfg *cfg, int fp64, vti_grad **out)
{
    if (!cfg || !out) return fail(VTI_ERR_ARG, "vti_grad_create: null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(VTI_ERR_NODEV, "vti_grad_create: no HIP device");
    if ((cfg->ndim != 2 && cfg->ndim != 3) || cfg->nt < 1 || !(cfg->dt > 0) || cfg->npml < 0 || !(cfg->favg > 0))
        return fail(VTI_ERR_ARG, "vti_grad_create: invalid configuration");
    int st;
    if (fp64) {
        auto *gr = new (std::nothrow) Grad<double>();
        if (!gr) return fail(VTI_ERR_ALLOC, "host");
        if ((st = gr->create(*cfg))) { delete gr; return st; }
        *out = gr;
    } else {
        auto *gr = new (std::nothrow) Grad<float>();
        if (!gr) return fail(VTI_ERR_ALLOC, "host");
        if ((st = gr->create(*cfg))) { delete gr; return st; }
        *out = gr;
    }
    return VTI_OK;
}

int vti_grad_destroy(vti_grad *g)
{
    delete g;
    return VTI_OK;
}

#define GDISPATCH(g, call)                                                  \
    do {                                                                    \
        if (!(g)) return fail(VTI_ERR_ARG, "vti_grad: null handle");        \
        if (auto *gd_ = dynamic_cast<Grad<double> *>(g)) return gd_->call;  \
        if (auto *gf_ = dynamic_cast<Grad<float> *>(g)) return gf_->call;   \
        return fail(VTI_ERR_ARG, "vti_grad: bad handle");                   \
    } while (0)

int vti_grad_set_model(vti_grad *g, const double *vp, const double *eps, const double *del, int keep_tables)
{
    if (!vp || !eps || !del) return fail(VTI_ERR_ARG, "vti_grad_set_model: null");
    GDISPATCH(g, set_model(vp, eps, del, keep_tables));
}

int vti_grad_set_shot(vti_grad *g, const vti_shot *shot)
{
    if (!shot) return fail(VTI_ERR_ARG, "vti_grad_set_shot: null");
    GDISPATCH(g, set_shot(*shot));
}

int vti_grad_set_source(vti_grad *g, const double *amp)
{
    if (!g) return fail(VTI_ERR_ARG, "null");
    if (auto *d = dynamic_cast<Grad<double> *>(g)) {
        d->custom_src = amp != nullptr;
        if (amp) d->amp.assign(amp, amp + d->cfg.nt);
        return d->have_shot && !amp ? d->set_shot(d->shot) : VTI_OK;
    }
    if (auto *f = dynamic_cast<Grad<float> *>(g)) {
        f->custom_src = amp != nullptr;
        if (amp) { f->amp.resize(f->cfg.nt); for (int i = 0; i < f->cfg.nt; ++i) f->amp[i] = (float)amp[i]; }
        return f->have_shot && !amp ? f->set_shot(f->shot) : VTI_OK;
    }
    return fail(VTI_ERR_ARG, "bad handle");
}

int vti_grad_forward(vti_grad *g, double *d)
{
    GDISPATCH(g, forward(d));
}

int vti_grad_run(vti_grad *g, const double *dobs, const double *resid, int nsnap, double *misfit,
                 double *grad_vp, double *grad_src, double *d_out, long long *fwd_steps)
{
    GDISPATCH(g, run(dobs, resid, nsnap, misfit, grad_vp, grad_src, d_out, fwd_steps));
}

}  // extern "C"
