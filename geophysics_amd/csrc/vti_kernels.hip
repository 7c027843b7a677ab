// vti_kernels.hip — gfx950 kernels of the staggered-grid VTI FD path.
//
// Arithmetic follows the reference expressions term by term and in the
// same left-to-right order (this file is compiled with -ffp-contract=off),
// so a step is bit-for-bit the reference's float arithmetic:
//   u = coffx2*u - coffx1*dtx*sum_m c_m (P[x+m+1]-P[x-m])
//   px = acoffx2*px - acoffx1*vp*vp*(1+2e)*dtx*sum_m c_m (u[x+m]-u[x-m-1])
// (GPU_vti2dfd_kernels.cu:229-235, :285-296; GPU_vti3dfd_kernels.cu:114-122,
//  :182-197).
//
// Work layout.  Every volume is z-fastest with a 128-byte-aligned row pitch
// (Grid::ldz).  A workgroup is 64 (z) x 4 (x) threads: one wavefront covers
// 64 consecutive z cells of one x row, so every load and store of a row is a
// single 256-byte coalesced access.  The 3D kernels (vti_stencil3d.h) give
// each thread one (z,x) column and a chunk of the slowest axis y, keeping
// the 8-deep y-stencil of P (velocity kernel) or v (stress kernel) in a
// register queue; x- and z-neighbours are a few KB away and are served by
// L1/L2, and workgroup ids are remapped so that each XCD (workgroups are
// dealt round-robin over the 8 XCDs) receives a contiguous slab of x-rows.
#include <hip/hip_runtime.h>

#include "vti_internal.h"
#include "vti_stencil3d.h"
#include "vti_fused3d.h"

namespace vti {
namespace {

// c[4] = {1.196289,-0.0797526,0.009570313,-0.0006975447}
// (GPU_vti2dfd_kernels.cu:37) — the float values of those decimals.
__device__ __forceinline__ float cst(int m)
{
    return m == 0 ? 1.196289f : (m == 1 ? -0.0797526f : (m == 2 ? 0.009570313f : -0.0006975447f));
}

constexpr int TZ = 64;    // z cells per workgroup (one wavefront)
constexpr int TX = 4;     // x rows per workgroup
// ---------------------------------------------------------------- 2D ----
template <bool INJ>
__global__ __launch_bounds__(256) void vel2d_kernel(VelArgs a)
{
    const Grid g = a.g;
    const int iz = blockIdx.x * TZ + threadIdx.x;
    const int ix = blockIdx.y * TX + threadIdx.y;
    if (iz < 4 || iz >= g.nnz - 4 || ix < 4 || ix >= g.nnx - 4) return;
    const long long id = iz + (long long)ix * g.ldz;
    float xx = 0.0f, zz = 0.0f;
#pragma unroll
    for (int im = 0; im < 4; ++im) {
        const long long i1 = id + (long long)(im + 1) * g.ldz, i0 = id - (long long)im * g.ldz;
        const long long j1 = id + im + 1, j0 = id - im;
        float p1 = a.P[i1], p0 = a.P[i0], q1 = a.Q[j1], q0 = a.Q[j0];
        if (INJ) {   // P[src] += amp, Q[src] += amp before the update (add_source2d :204-205)
            if (i1 == a.src) p1 = p1 + a.amp;
            if (i0 == a.src) p0 = p0 + a.amp;
            if (j1 == a.src) q1 = q1 + a.amp;
            if (j0 == a.src) q0 = q0 + a.amp;
        }
        xx += cst(im) * (p1 - p0);
        zz += cst(im) * (q1 - q0);
    }
    a.u[id] = a.cx2[ix] * a.u[id] - a.cx1[ix] * a.dtx * xx;
    a.w[id] = a.cz2[iz] * a.w[id] - a.cz1[iz] * a.dtz * zz;
}

// Effective eps / delta after the SV taper (GPU_vti2dfd_kernels.cu:260-278):
// the factor depends only on the integer r^2, looked up in a host table.
__device__ __forceinline__ void taper(const StressArgs &a, int r2, float &ee, float &dd)
{
    if (r2 <= a.svR * a.svR) {
        const double f = a.tap[r2];
        ee = (float)(f * (double)ee);
        dd = (float)(f * (double)dd);
    }
}

// US: the unsplit interior (point-source form): where the cell's damping
// factors are all exactly 1, P and Q are updated directly and px .. qz are
// neither read nor written (36 instead of 60 bytes per cell); row a.split_z
// (an RTM receiver side's injection row) keeps the split form.
// DP: the Toafd2domp_2.c form, whose pz and qx terms are double products
// (pow(1+2*del, 0.5) returns double, :176-180), subtracted in double and
// rounded to float once; px and qz stay float as there.
template <bool US, bool DP>
__global__ __launch_bounds__(256) void stress2d_kernel(StressArgs a)
{
    const Grid g = a.g;
    const int iz = blockIdx.x * TZ + threadIdx.x;
    const int ix = blockIdx.y * TX + threadIdx.y;
    if (iz < 4 || iz >= g.nnz - 4 || ix < 4 || ix >= g.nnx - 4) return;
    const long long id = iz + (long long)ix * g.ldz;
    float xx = 0.0f, zz = 0.0f;
#pragma unroll
    for (int im = 0; im < 4; ++im) {
        xx += cst(im) * (a.u[id + (long long)im * g.ldz] - a.u[id - (long long)(im + 1) * g.ldz]);
        zz += cst(im) * (a.w[id + im] - a.w[id - im - 1]);
    }
    const float vv = a.vp[id];
    float ee = a.ep[id], dd = a.de[id];
    if (a.sv) {
        const int rx = ix - a.cx, rz = iz - a.cz;
        taper(a, rx * rx + rz * rz, ee, dd);
    }
    const float e2 = 1.0f + 2.0f * ee;
    const float sd = sqrtf(1.0f + 2.0f * dd);
    const float ax1 = a.ax1[ix], ax2 = a.ax2[ix], az1 = a.az1[iz], az2 = a.az2[iz];
    float Pn, Qn;
    if (US && iz != a.split_z && ax1 == 1.0f && ax2 == 1.0f && az1 == 1.0f && az2 == 1.0f) {
        const float aa = vv * vv;
        Pn = a.P[id] - (aa * e2 * a.dtx * xx + aa * sd * a.dtz * zz);
        Qn = a.Q[id] - (aa * sd * a.dtx * xx + aa * a.dtz * zz);
    } else {
        const float px = ax2 * a.px[id] - ax1 * vv * vv * e2 * a.dtx * xx;
        float pz, qx;
        if (DP) {
            const double sq = sqrt((double)(1.0f + 2.0f * dd));
            pz = (float)((double)(az2 * a.pz[id]) - (double)(az1 * vv * vv) * sq * (double)a.dtz * (double)zz);
            qx = (float)((double)(ax2 * a.qx[id]) - (double)(ax1 * vv * vv) * sq * (double)a.dtx * (double)xx);
        } else {
            pz = az2 * a.pz[id] - az1 * vv * vv * sd * a.dtz * zz;
            qx = ax2 * a.qx[id] - ax1 * vv * vv * sd * a.dtx * xx;
        }
        const float qz = az2 * a.qz[id] - az1 * vv * vv * a.dtz * zz;
        a.px[id] = px; a.pz[id] = pz; a.qx[id] = qx; a.qz[id] = qz;
        Pn = px + pz;
        Qn = qx + qz;
        if (a.gsrc) {   // P = px+pz+s (CPU_vti2dfd_kernels.c:173-174), 7x7 footprint
            const int dx = ix - a.gx, dz = iz - a.gz;
            if (dx >= -3 && dx <= 3 && dz >= -3 && dz <= 3) {
                const float s = (float)((double)a.gamp * a.gw[(dx + 3) * 7 + (dz + 3)]);
                Pn = px + pz + s;
                Qn = qx + qz + s;
            }
        }
    }
    a.P[id] = Pn;
    a.Q[id] = Qn;
    illum_add(a, ix, a.inp, iz, Pn, Qn);
    if (a.rec && iz == a.rz) {
        const int k = ix - a.rx0;
        if (k >= 0 && k % a.rdx == 0 && k / a.rdx < a.rnx) {
            const float r = a.rec_kind == 0 ? Pn : (a.rec_kind == 1 ? Pn + Pn : Pn + Qn);
            a.rec[k / a.rdx] = r;
        }
    }
}

// ------------------------------------------------------------ 2D TTI ----
// Toa_gpu_2dtti_fd_1orderfunction.cu: the x and z derivatives of P and Q
// (velocity) and of u and w (stress) rotated by the tilt theta, every
// field damped by the product of its x and z factors; cos/sin of theta come
// from host-built tables (a.cth, a.sth).  Expressions in the reference's
// left-to-right order:
//   u = cx2*cz2*u - cx1*cz1*(cos*dtx*Px - sin*dtz*Pz)            (:110)
//   w = cx2*cz2*w - cx1*cz1*(sin*dtx*Qx + cos*dtz*Qz)            (:111)
//   px = ax2*az2*px - ax1*az1*vp*vp*(1+2e)*(cos*dtx*ux - sin*dtz*uz)  (:170-171)
//   pz = ax2*az2*pz - ax1*az1*vp*vp*sqrtf(1+2d)*(sin*dtx*wx + cos*dtz*wz)
//   qx = ax2*az2*qx - ax1*az1*vp*vp*sqrtf(1+2d)*(cos*dtx*ux - sin*dtz*uz)
//   qz = ax2*az2*qz - ax1*az1*vp*vp*(sin*dtx*wx + cos*dtz*wz)    (:176-177)
// The point source (add_source before update_vel, :441) is seen by all four
// stencils of the velocity update as P + amp, Q + amp at its cell.
template <bool INJ>
__global__ __launch_bounds__(256) void tti_vel2d_kernel(VelArgs a)
{
    const Grid g = a.g;
    const int iz = blockIdx.x * TZ + threadIdx.x;
    const int ix = blockIdx.y * TX + threadIdx.y;
    if (iz < 4 || iz >= g.nnz - 4 || ix < 4 || ix >= g.nnx - 4) return;
    const long long id = iz + (long long)ix * g.ldz;
    float Px = 0.0f, Pz = 0.0f, Qz = 0.0f, Qx = 0.0f;
#pragma unroll
    for (int im = 0; im < 4; ++im) {
        const long long i1 = id + (long long)(im + 1) * g.ldz, i0 = id - (long long)im * g.ldz;
        const long long j1 = id + im + 1, j0 = id - im;
        float pi1 = a.P[i1], pi0 = a.P[i0], pj1 = a.P[j1], pj0 = a.P[j0];
        float qi1 = a.Q[i1], qi0 = a.Q[i0], qj1 = a.Q[j1], qj0 = a.Q[j0];
        if (INJ) {
            if (i1 == a.src) { pi1 = pi1 + a.amp; qi1 = qi1 + a.amp; }
            if (i0 == a.src) { pi0 = pi0 + a.amp; qi0 = qi0 + a.amp; }
            if (j1 == a.src) { pj1 = pj1 + a.amp; qj1 = qj1 + a.amp; }
            if (j0 == a.src) { pj0 = pj0 + a.amp; qj0 = qj0 + a.amp; }
        }
        Px += cst(im) * (pi1 - pi0);
        Pz += cst(im) * (pj1 - pj0);
        Qz += cst(im) * (qj1 - qj0);
        Qx += cst(im) * (qi1 - qi0);
    }
    const float cs = a.cth[id], sn = a.sth[id];
    const float g2 = a.cx2[ix] * a.cz2[iz], g1 = a.cx1[ix] * a.cz1[iz];
    a.u[id] = g2 * a.u[id] - g1 * (cs * a.dtx * Px - sn * a.dtz * Pz);
    a.w[id] = g2 * a.w[id] - g1 * (sn * a.dtx * Qx + cs * a.dtz * Qz);
}

__global__ __launch_bounds__(256) void tti_stress2d_kernel(StressArgs a)
{
    const Grid g = a.g;
    const int iz = blockIdx.x * TZ + threadIdx.x;
    const int ix = blockIdx.y * TX + threadIdx.y;
    if (iz < 4 || iz >= g.nnz - 4 || ix < 4 || ix >= g.nnx - 4) return;
    const long long id = iz + (long long)ix * g.ldz;
    float ux = 0.0f, uz = 0.0f, wz = 0.0f, wx = 0.0f;
#pragma unroll
    for (int im = 0; im < 4; ++im) {
        ux += cst(im) * (a.u[id + (long long)im * g.ldz] - a.u[id - (long long)(im + 1) * g.ldz]);
        uz += cst(im) * (a.u[id + im] - a.u[id - im - 1]);
        wz += cst(im) * (a.w[id + im] - a.w[id - im - 1]);
        wx += cst(im) * (a.w[id + (long long)im * g.ldz] - a.w[id - (long long)(im + 1) * g.ldz]);
    }
    const float vv = a.vp[id];
    float ee = a.ep[id], dd = a.de[id];
    if (a.sv) {
        const int rx = ix - a.cx, rz = iz - a.cz;
        taper(a, rx * rx + rz * rz, ee, dd);
    }
    const float cs = a.cth[id], sn = a.sth[id];
    const float g2 = a.ax2[ix] * a.az2[iz], g1 = a.ax1[ix] * a.az1[iz];
    const float e2 = 1.0f + 2.0f * ee;
    const float sd = sqrtf(1.0f + 2.0f * dd);
    const float du = cs * a.dtx * ux - sn * a.dtz * uz;   // the rotated x derivative of u
    const float dw = sn * a.dtx * wx + cs * a.dtz * wz;   // the rotated z derivative of w
    const float px = g2 * a.px[id] - g1 * vv * vv * e2 * du;
    const float pz = g2 * a.pz[id] - g1 * vv * vv * sd * dw;
    const float qx = g2 * a.qx[id] - g1 * vv * vv * sd * du;
    const float qz = g2 * a.qz[id] - g1 * vv * vv * dw;
    a.px[id] = px; a.pz[id] = pz; a.qx[id] = qx; a.qz[id] = qz;
    const float Pn = px + pz, Qn = qx + qz;
    a.P[id] = Pn;
    a.Q[id] = Qn;
    illum_add(a, ix, a.inp, iz, Pn, Qn);
    if (a.rec && iz == a.rz) {
        const int k = ix - a.rx0;
        if (k >= 0 && k % a.rdx == 0 && k / a.rdx < a.rnx) {
            const float r = a.rec_kind == 0 ? Pn : (a.rec_kind == 1 ? Pn + Pn : Pn + Qn);
            a.rec[k / a.rdx] = r;
        }
    }
}

// ------------------------------------------------------------ helpers ----
// Edge-replicating pad of an unpadded (ny, nx, nz) model into the padded
// pitched volume: equal to pad_vv2d / pad_vv3d (GPU_vti2dfd_kernels.cu:49-74,
// GPU_vti3dfd_kernels.cu:208-255), which copy x then y then z borders.
__global__ void pad_kernel(Grid g, int ndim, int nx, int ny, int nz, int np,
                           const float *__restrict__ src, float *__restrict__ dst)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y;
    const int iy = blockIdx.z;
    if (iz >= g.ldz) return;
    const long long o = iz + (long long)ix * g.ldz + (long long)iy * g.ldxy;
    if (iz >= g.nnz) { dst[o] = 0.0f; return; }
    const int jx = ix < np ? 0 : (ix >= np + nx ? nx - 1 : ix - np);
    const int jz = iz < np ? 0 : (iz >= np + nz ? nz - 1 : iz - np);
    int jy = 0;
    if (ndim == 3) jy = iy < np ? 0 : (iy >= np + ny ? ny - 1 : iy - np);
    dst[o] = src[jz + (long long)jx * nz + (long long)jy * nz * nx];
}

__global__ void zero_kernel(float4 *p, size_t n4)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void zero_tail_kernel(float *p, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// out[c][r] = in[r][c] through a 32x33 LDS tile.
__global__ void transpose_kernel(const float *__restrict__ in, float *__restrict__ out, int rows, int cols)
{
    __shared__ float tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[k][threadIdx.x] = in[(long long)r * cols + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += blockDim.y) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[(long long)c * rows + r] = tile[threadIdx.x][k];
    }
}

// mute_directwave2d GPU_vti2dfd_kernels.cu:151-174 on a [irec][it] record;
// receiver r sits at model column x0 + r*rdx (the reference: x0 = 0, rdx = 1).
__global__ void mute2d_kernel(float *shot, int nrec, int nt, float dt, float favg, float dx,
                              float dz, int fs, int ds, int zs, int is, float vp1, float eps1, int tt,
                              int x0, int rdx)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)nrec * nt) return;
    const int r = (int)(id / nt), it = (int)(id % nt);
    const int ix = x0 + r * rdx;
    const float mu_x = dx * abs(ix - fs - (is - 1) * ds);
    const float mu_z = dz * zs;
    const float mu_t0 = sqrtf(mu_x * mu_x + mu_z * mu_z) / (vp1 * sqrtf(1 + 2 * eps1));
    const int mu_t = (int)(2.0 / (dt * favg));
    const int mu_nt = (int)(mu_t0 / dt) + mu_t + tt;
    if ((it > (int)(mu_t0 / dt) - tt) && it < mu_nt) shot[id] = 0.0f;
}

// mute_directwave3d GPU_vti3dfd_kernels.cu:342-369 (window = 0: every sample
// before the cut) and the windowed form of the 3D RTM main
// (Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:371-380, window = 1), on a
// [iry][irx][it] record; receiver (irx, iry) at model (x0+irx*rdx, y0+iry*rdy).
__global__ void mute3d_kernel(float *shot, int nrx, int nry, int nt, float dt, float favg,
                              float dx, float dy, float dz, int fsx, int dsx, int fsy,
                              int dsy, int fsz, int is, float vp1, float eps1, int tt, int window,
                              int x0, int rdx, int y0, int rdy)
{
    const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (id >= (long long)nrx * nry * nt) return;
    const int it = (int)(id % nt);
    const long long r = id / nt;
    const int ix = x0 + (int)(r % nrx) * rdx, iy = y0 + (int)(r / nrx) * rdy;
    const float mu_x = dx * abs(ix - fsx - is * dsx);
    const float mu_y = dy * abs(iy - fsy - is * dsy);
    const float mu_z = dz * fsz;
    const float mu_t0 = sqrtf(mu_x * mu_x + mu_y * mu_y + mu_z * mu_z) / (vp1 * sqrtf(1 + 2 * eps1));
    const int mu_t = (int)(2.0 / (dt * favg));
    const int mu_nt = (int)(mu_t0 / dt) + mu_t + tt;
    const bool lo = !window || (it > (int)(mu_t0 / dt) - tt);
    if (lo && it < mu_nt) shot[id] = 0.0f;
}

__global__ void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

// ---------------------------------------------------------- launchers ----
// 3D launches march y in chunks of 16 planes per workgroup (36x more
// workgroups than whole-column marching at 512^3, so the tail of the
// launch is short; 8 halo planes of one field re-read per chunk) with
// non-temporal access to the fields touched once.  Measured at 572^3:
// chunks of 16 beat 64 by 1 % (stress) and 3 % (velocity).
constexpr int kYChunk = 16;

hipError_t launch_vel(int ndim, const VelArgs &a, hipStream_t s)
{
    const bool inj = a.src >= 0;
    if (ndim == 2) {
        dim3 blk(TZ, TX), grd((a.g.nnz + TZ - 1) / TZ, (a.g.nnx + TX - 1) / TX);
        if (a.cth) {
            if (inj) hipLaunchKernelGGL(tti_vel2d_kernel<true>, grd, blk, 0, s, a);
            else hipLaunchKernelGGL(tti_vel2d_kernel<false>, grd, blk, 0, s, a);
        } else if (inj) hipLaunchKernelGGL(vel2d_kernel<true>, grd, blk, 0, s, a);
        else hipLaunchKernelGGL(vel2d_kernel<false>, grd, blk, 0, s, a);
    } else {
        int nzt, nt, ntp;
        const int grd = s3d::grid3d(a.g, kYChunk, nzt, nt, ntp);
        hipLaunchKernelGGL((s3d::vel3d<kYChunk, true>), dim3(grd), dim3(s3d::TZ, s3d::TX), 0, s, a, nzt, nt, ntp);
    }
    return hipGetLastError();
}

// Fused 3D step: y chunks of kFusedChunk planes (the prologue recomputes 7
// planes of v per chunk).
constexpr int kFusedChunk = 32;

hipError_t launch_fused(const FusedArgs &a, hipStream_t s)
{
    int nzt, nt, ntp;
    const int grd = f3d::grid(a.v.g, kFusedChunk, nzt, nt, ntp);
    hipLaunchKernelGGL((f3d::fused3d<kFusedChunk, true>), dim3(grd), dim3(f3d::TZ, f3d::TX), 0, s, a, nzt, nt, ntp);
    return hipGetLastError();
}

hipError_t launch_stress(int ndim, const StressArgs &a, hipStream_t s)
{
    if (ndim == 2) {
        dim3 blk(TZ, TX), grd((a.g.nnz + TZ - 1) / TZ, (a.g.nnx + TX - 1) / TX);
        if (a.cth) hipLaunchKernelGGL(tti_stress2d_kernel, grd, blk, 0, s, a);
        else if (a.dpq) hipLaunchKernelGGL((stress2d_kernel<false, true>), grd, blk, 0, s, a);
        else if (a.unsplit && !a.gsrc) hipLaunchKernelGGL((stress2d_kernel<true, false>), grd, blk, 0, s, a);
        else hipLaunchKernelGGL((stress2d_kernel<false, false>), grd, blk, 0, s, a);
    } else {
        int nzt, nt, ntp;
        const int grd = s3d::grid3d(a.g, kYChunk, nzt, nt, ntp);
        const dim3 blk(s3d::TZ, s3d::TX);
        const bool us = a.unsplit && !a.gsrc;
        if (a.snap && us) hipLaunchKernelGGL((s3d::stress3d<kYChunk, true, true, true>), dim3(grd), blk, 0, s, a, nzt, nt, ntp);
        else if (a.snap) hipLaunchKernelGGL((s3d::stress3d<kYChunk, true, false, true>), dim3(grd), blk, 0, s, a, nzt, nt, ntp);
        else if (us) hipLaunchKernelGGL((s3d::stress3d<kYChunk, true, true>), dim3(grd), blk, 0, s, a, nzt, nt, ntp);
        else hipLaunchKernelGGL((s3d::stress3d<kYChunk, true>), dim3(grd), blk, 0, s, a, nzt, nt, ntp);
    }
    return hipGetLastError();
}

hipError_t launch_pad_model(const Grid &g, int ndim, int nx, int ny, int nz, int npml,
                            const float *src, float *dst, hipStream_t s)
{
    dim3 blk(256), grd((g.ldz + 255) / 256, g.nnx, g.nny);
    hipLaunchKernelGGL(pad_kernel, grd, blk, 0, s, g, ndim, nx, ny, nz, npml, src, dst);
    return hipGetLastError();
}

hipError_t launch_zero(float *p, size_t n, hipStream_t s)
{
    const size_t n4 = n / 4;
    if (n4) {
        const size_t blocks = std::min<size_t>((n4 + 255) / 256, 8192);
        hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<float4 *>(p), n4);
    }
    const size_t tail = n - n4 * 4;
    if (tail) hipLaunchKernelGGL(zero_tail_kernel, dim3(1), dim3(4), 0, s, p + n4 * 4, tail);
    return hipGetLastError();
}

hipError_t launch_transpose(const float *in, float *out, int rows, int cols, hipStream_t s)
{
    dim3 blk(32, 8), grd((cols + 31) / 32, (rows + 31) / 32);
    hipLaunchKernelGGL(transpose_kernel, grd, blk, 0, s, in, out, rows, cols);
    return hipGetLastError();
}

hipError_t launch_mute2d(float *rec_t, int nrec, int nt, float dt, float favg, float dx,
                         float dz, int fs, int ds, int zs, int is, float vp1, float eps1,
                         int tt, int x0, int rdx, hipStream_t s)
{
    const long long n = (long long)nrec * nt;
    hipLaunchKernelGGL(mute2d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       rec_t, nrec, nt, dt, favg, dx, dz, fs, ds, zs, is, vp1, eps1, tt, x0, rdx);
    return hipGetLastError();
}

hipError_t launch_mute3d(float *rec_t, int nrx, int nry, int nt, float dt, float favg,
                         float dx, float dy, float dz, int fsx, int dsx, int fsy, int dsy,
                         int fsz, int is, float vp1, float eps1, int tt, int window,
                         int x0, int rdx, int y0, int rdy, hipStream_t s)
{
    const long long n = (long long)nrx * nry * nt;
    hipLaunchKernelGGL(mute3d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       rec_t, nrx, nry, nt, dt, favg, dx, dy, dz, fsx, dsx, fsy, dsy, fsz, is,
                       vp1, eps1, tt, window, x0, rdx, y0, rdy);
    return hipGetLastError();
}

hipError_t launch_copy(void *dst, const void *src, size_t nbytes, hipStream_t s)
{
    const size_t n4 = nbytes / 16;
    const size_t blocks = std::min<size_t>((n4 + 255) / 256, 16384);
    if (n4)
        hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<const float4 *>(src), reinterpret_cast<float4 *>(dst), n4);
    return hipGetLastError();
}

}  // namespace vti
