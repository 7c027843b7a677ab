// vti_stencil3d.h — the 3D velocity and stress kernels (gfx950).
//
// Included by vti_kernels.hip.  Arithmetic is the reference's expression
// order (update_vel3d GPU_vti3dfd_kernels.cu:114-122, update_stress3d
// :182-197), compiled without FMA contraction.
//
// Layout: each thread owns one (z,x) column of a 64(z) x 4(x) workgroup
// tile and marches y over [y0, y1) (the whole updated range, or a chunk),
// holding the 8-deep y-stencil of P (velocity) or v (stress) in a register
// queue.  Every load of an iteration is issued before any of them is
// consumed (a compiler barrier keeps the scheduler from sinking loads to
// their uses, which otherwise serialises them on HBM latency); x- and
// z-neighbours are served from L1/L2.  Tiles are remapped so each XCD
// works on a contiguous slab of x.
#pragma once

#include <hip/hip_runtime.h>

#include "vti_internal.h"

namespace vti {
// Illumination epilogue of a stress update (cal_illumination
// Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:654-677, 3D
// Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:523-537): the same float
// operations as a separate pass over the stored P and Q.
__device__ __forceinline__ void illum_add(const StressArgs &a, int ix, int iy, int iz, float Pn, float Qn)
{
    if (!a.illum) return;
    const int jx = ix - a.inp, jy = iy - a.inp, jz = iz - a.inp;
    if (jx < 0 || jx >= a.inx || jy < 0 || jy >= a.iny || jz < 0 || jz >= a.inz) return;
    const long long o = jz + (long long)jx * a.inz + (long long)jy * a.inz * a.inx;
    float v = a.illum[o] + (Pn * Pn + Qn * Qn);
    if (a.illum_clamp && v <= 0.0f) v = 1.0f;
    a.illum[o] = v;
}

// The epilogue's copy of P into the dense interior snapshot (the
// illumination's interior), in the kernels instantiated for it only.
template <bool SNAP>
__device__ __forceinline__ void snap_put(const StressArgs &a, int ix, int iy, int iz, float Pn)
{
    if (!SNAP) return;
    const int jx = ix - a.inp, jy = iy - a.inp, jz = iz - a.inp;
    if (jx < 0 || jx >= a.inx || jy < 0 || jy >= a.iny || jz < 0 || jz >= a.inz) return;
    a.snap[jz + (long long)jx * a.inz + (long long)jy * a.inz * a.inx] = Pn;
}

namespace s3d {

__device__ __forceinline__ float cst(int m)
{
    return m == 0 ? 1.196289f : (m == 1 ? -0.0797526f : (m == 2 ? 0.009570313f : -0.0006975447f));
}

constexpr int TZ = 64;
constexpr int TX = 4;
constexpr int NXCD = 8;

__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }

template <bool NT>
__device__ __forceinline__ float ld(const float *p)
{
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT>
__device__ __forceinline__ void st(float *p, float v)
{
    if (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Tile decode shared by both kernels: returns false for idle threads.
struct Col {
    int iz, ix, y0, y1;
};
__device__ __forceinline__ bool decode(const Grid &g, int nzt, int ntile, int ntile_p, int ychunk,
                                       Col &c)
{
    const int nchunk = ychunk > 0 ? (g.nny - 8 + ychunk - 1) / ychunk : 1;
    const int b = blockIdx.x;
    const int per = ntile_p / NXCD;
    const int tl = (b % NXCD) * per + (b / NXCD);   // XCD-contiguous logical id
    if (tl >= ntile * nchunk) return false;
    const int t = tl % ntile, ch = tl / ntile;
    c.iz = (t % nzt) * TZ + threadIdx.x;
    c.ix = (t / nzt) * TX + threadIdx.y;
    c.y0 = 4 + ch * (ychunk > 0 ? ychunk : 0);
    c.y1 = ychunk > 0 ? min(g.nny - 4, c.y0 + ychunk) : g.nny - 4;
    return c.iz >= 4 && c.iz < g.nnz - 4 && c.ix >= 4 && c.ix < g.nnx - 4;
}

// ---------------------------------------------------------------- vel ----
template <int YCHUNK, bool NT>
__global__ __launch_bounds__(256) void vel3d(VelArgs a, int nzt, int ntile, int ntile_p)
{
    const Grid g = a.g;
    Col c;
    if (!decode(g, nzt, ntile, ntile_p, YCHUNK, c)) return;
    const int iz = c.iz, ix = c.ix;
    const long long L = g.ldxy, LZ = g.ldz;
    const long long col = iz + (long long)ix * LZ;
    const float *__restrict__ P = a.P;
    const float *__restrict__ Q = a.Q;
    float *__restrict__ u = a.u;
    float *__restrict__ v = a.v;
    float *__restrict__ w = a.w;
    const float cx1 = a.cx1[ix], cx2 = a.cx2[ix], cz1 = a.cz1[iz], cz2 = a.cz2[iz];
    const float dtx = a.dtx, dty = a.dty, dtz = a.dtz;
    // the point source touches the stencils of a 9x9x9 box only
    const bool near = a.src >= 0 && abs(ix - (int)((a.src / LZ) % g.nnx)) <= 4 &&
                      abs(iz - (int)(a.src % LZ)) <= 4;
    const int sy = a.src >= 0 ? (int)(a.src / L) : -100;
    float q[8];   // q[k] = P at y-3+k
#pragma unroll
    for (int k = 0; k < 8; ++k) q[k] = P[col + (long long)(c.y0 - 3 + k) * L];
    for (int iy = c.y0; iy < c.y1; ++iy) {
        const long long id = col + (long long)iy * L;
        float px[8], qz[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) px[k] = P[id + (long long)(k - 3) * LZ];   // x-3 .. x+4
#pragma unroll
        for (int k = 0; k < 8; ++k) qz[k] = Q[id + (k - 3)];                   // z-3 .. z+4
        const float u0 = ld<NT>(u + id), v0 = ld<NT>(v + id), w0 = ld<NT>(w + id);
        const float qn = iy + 1 < c.y1 ? P[id + 5 * L] : 0.0f;
        compiler_fence();
        float qq[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) qq[k] = q[k];
        if (near && abs(iy - sy) <= 4) {
            // add_source3d (GPU_vti3dfd_kernels.cu:85-86): P[src] += amp, Q[src] += amp
            // before the update, applied to this iteration's copies of the loads.
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (id + (long long)(k - 3) * LZ == a.src) px[k] = px[k] + a.amp;
                if (id + (k - 3) == a.src) qz[k] = qz[k] + a.amp;
                if (col + (long long)(iy - 3 + k) * L == a.src) qq[k] = q[k] + a.amp;
            }
        }
        float xx = 0.0f, yy = 0.0f, zz = 0.0f;
#pragma unroll
        for (int im = 0; im < 4; ++im) {
            yy += cst(im) * (qq[3 + im + 1] - qq[3 - im]);
            xx += cst(im) * (px[3 + im + 1] - px[3 - im]);
            zz += cst(im) * (qz[3 + im + 1] - qz[3 - im]);
        }
        st<NT>(u + id, cx2 * u0 - cx1 * dtx * xx);
        st<NT>(v + id, a.cy2[iy] * v0 - a.cy1[iy] * dty * yy);
        st<NT>(w + id, cz2 * w0 - cz1 * dtz * zz);
#pragma unroll
        for (int k = 0; k < 7; ++k) q[k] = q[k + 1];
        q[7] = qn;
    }
}

// ------------------------------------------------------------- stress ----
// US (reduced split): where every damping factor of the cell is exactly 1
// (outside the absorbing layer) the split fields are redundant: P and Q are
// updated directly, P -= (tx + ty + tz), Q likewise, with the same
// per-direction increments as the split form, and px .. qz are neither read
// nor written there (40 instead of 80 bytes per cell).  A cell of the
// absorbing layer with one or two damped axes keeps only those axes' split
// components and carries the rest in P: P' = (px' + py' + pz') + ((P - px -
// py - pz) - (tx + ty + tz)), each term present only for the kept (px ..) or
// the undamped (tx ..) axes (56 or 72 bytes per cell).  Corners (three damped
// axes) and row a.split_z keep the full split form.  Point-source
// (load-time) form only.
template <int YCHUNK, bool NT, bool US = false, bool SNAP = false>
__global__ __launch_bounds__(256) void stress3d(StressArgs a, int nzt, int ntile, int ntile_p)
{
    const Grid g = a.g;
    Col c;
    if (!decode(g, nzt, ntile, ntile_p, YCHUNK, c)) return;
    const int iz = c.iz, ix = c.ix;
    const long long L = g.ldxy, LZ = g.ldz;
    const long long col = iz + (long long)ix * LZ;
    const float *__restrict__ u = a.u;
    const float *__restrict__ v = a.v;
    const float *__restrict__ w = a.w;
    const bool sv_col = a.sv && abs(ix - a.cx) <= a.svR && abs(iz - a.cz) <= a.svR;
    const bool src_col = a.gsrc && abs(ix - a.gx) <= 3 && abs(iz - a.gz) <= 3;
    int rk = -1;
    if (a.rec && iz == a.rz) {
        const int k = ix - a.rx0;
        if (k >= 0 && k % a.rdx == 0 && k / a.rdx < a.rnx) rk = k / a.rdx;
    }
    const float ax1 = a.ax1[ix], ax2 = a.ax2[ix], az1 = a.az1[iz], az2 = a.az2[iz];
    const float dtx = a.dtx, dty = a.dty, dtz = a.dtz;
    const bool dxc = !(ax1 == 1.0f && ax2 == 1.0f), dzc = !(az1 == 1.0f && az2 == 1.0f);
    const bool col_full = !US || iz == a.split_z;
    const bool col_int = !col_full && !dxc && !dzc;
    float vq[8];   // vq[k] = v at y-4+k
#pragma unroll
    for (int k = 0; k < 8; ++k) vq[k] = v[col + (long long)(c.y0 - 4 + k) * L];
    for (int iy = c.y0; iy < c.y1; ++iy) {
        const long long id = col + (long long)iy * L;
        float ux[8], wz[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) ux[k] = u[id + (long long)(k - 4) * LZ];   // x-4 .. x+3
#pragma unroll
        for (int k = 0; k < 8; ++k) wz[k] = w[id + (k - 4)];                   // z-4 .. z+3
        const float vv = ld<NT>(a.vp + id);
        float ee = ld<NT>(a.ep + id), dd = ld<NT>(a.de + id);
        const float ay1 = a.ay1[iy], ay2 = a.ay2[iy];
        const bool dy = !(ay1 == 1.0f && ay2 == 1.0f);
        const bool un = col_int && !dy;
        // kept split components per axis; full: P, Q rebuilt from them alone
        const bool full = col_full || (dxc && dy && dzc);
        const bool fx = full || dxc, fy = full || dy, fz = full || dzc;
        float opx = 0.0f, opy = 0.0f, opz = 0.0f, oqx = 0.0f, oqy = 0.0f, oqz = 0.0f, oP = 0.0f, oQ = 0.0f;
        if (fx) { opx = ld<NT>(a.px + id); oqx = ld<NT>(a.qx + id); }
        if (fy) { opy = ld<NT>(a.py + id); oqy = ld<NT>(a.qy + id); }
        if (fz) { opz = ld<NT>(a.pz + id); oqz = ld<NT>(a.qz + id); }
        if (!full) { oP = ld<NT>(a.P + id); oQ = ld<NT>(a.Q + id); }
        const float vn = iy + 1 < c.y1 ? v[id + 4 * L] : 0.0f;
        compiler_fence();
        float xx = 0.0f, yy = 0.0f, zz = 0.0f;
#pragma unroll
        for (int im = 0; im < 4; ++im) {
            yy += cst(im) * (vq[4 + im] - vq[3 - im]);
            xx += cst(im) * (ux[4 + im] - ux[3 - im]);
            zz += cst(im) * (wz[4 + im] - wz[3 - im]);
        }
        if (sv_col) {   // SV taper (GPU_vti3dfd_kernels.cu:155-174)
            const int rx = ix - a.cx, ry = iy - a.cy, rz = iz - a.cz;
            const int r2 = rx * rx + ry * ry + rz * rz;
            if (ry <= a.svR && ry >= -a.svR && r2 <= a.svR * a.svR) {
                const double f = a.tap[r2];
                ee = (float)(f * (double)ee);
                dd = (float)(f * (double)dd);
            }
        }
        const float e2 = 1.0f + 2.0f * ee;
        const float sd = sqrtf(1.0f + 2.0f * dd);
        if (un) {
            const float aa = vv * vv;
            const float Pn = oP - (aa * e2 * dtx * xx + aa * e2 * dty * yy + aa * sd * dtz * zz);
            const float Qn = oQ - (aa * sd * dtx * xx + aa * sd * dty * yy + aa * dtz * zz);
            st<NT>(a.P + id, Pn);
            st<NT>(a.Q + id, Qn);
            illum_add(a, ix, iy, iz, Pn, Qn);
            snap_put<SNAP>(a, ix, iy, iz, Pn);
            if (rk >= 0) {
                const int j = iy - a.ry0;
                if (j >= 0 && j % a.rdy == 0 && j / a.rdy < a.rny) {
                    const float r = a.rec_kind == 0 ? Pn : (a.rec_kind == 1 ? Pn + Pn : Pn + Qn);
                    a.rec[rk + (long long)a.rnx * (j / a.rdy)] = r;
                }
            }
        } else if (US && !full) {
            const float aa = vv * vv;
            const float px = fx ? ax2 * opx - ax1 * vv * vv * e2 * dtx * xx : 0.0f;
            const float py = fy ? ay2 * opy - ay1 * vv * vv * e2 * dty * yy : 0.0f;
            const float pz = fz ? az2 * opz - az1 * vv * vv * sd * dtz * zz : 0.0f;
            const float qx = fx ? ax2 * oqx - ax1 * vv * vv * sd * dtx * xx : 0.0f;
            const float qy = fy ? ay2 * oqy - ay1 * vv * vv * sd * dty * yy : 0.0f;
            const float qz = fz ? az2 * oqz - az1 * vv * vv * dtz * zz : 0.0f;
            if (fx) { st<NT>(a.px + id, px); st<NT>(a.qx + id, qx); }
            if (fy) { st<NT>(a.py + id, py); st<NT>(a.qy + id, qy); }
            if (fz) { st<NT>(a.pz + id, pz); st<NT>(a.qz + id, qz); }
            const float rp = ((oP - opx) - opy) - opz;   // unkept components are 0
            const float rq = ((oQ - oqx) - oqy) - oqz;
            const float tp = ((fx ? 0.0f : aa * e2 * dtx * xx) + (fy ? 0.0f : aa * e2 * dty * yy)) +
                             (fz ? 0.0f : aa * sd * dtz * zz);
            const float tq = ((fx ? 0.0f : aa * sd * dtx * xx) + (fy ? 0.0f : aa * sd * dty * yy)) +
                             (fz ? 0.0f : aa * dtz * zz);
            const float Pn = ((px + py) + pz) + (rp - tp);
            const float Qn = ((qx + qy) + qz) + (rq - tq);
            st<NT>(a.P + id, Pn);
            st<NT>(a.Q + id, Qn);
            illum_add(a, ix, iy, iz, Pn, Qn);
            snap_put<SNAP>(a, ix, iy, iz, Pn);
            if (rk >= 0) {
                const int j = iy - a.ry0;
                if (j >= 0 && j % a.rdy == 0 && j / a.rdy < a.rny) {
                    const float r = a.rec_kind == 0 ? Pn : (a.rec_kind == 1 ? Pn + Pn : Pn + Qn);
                    a.rec[rk + (long long)a.rnx * (j / a.rdy)] = r;
                }
            }
        } else {
            const float px = ax2 * opx - ax1 * vv * vv * e2 * dtx * xx;
            const float py = ay2 * opy - ay1 * vv * vv * e2 * dty * yy;
            const float pz = az2 * opz - az1 * vv * vv * sd * dtz * zz;
            const float qx = ax2 * oqx - ax1 * vv * vv * sd * dtx * xx;
            const float qy = ay2 * oqy - ay1 * vv * vv * sd * dty * yy;
            const float qz = az2 * oqz - az1 * vv * vv * dtz * zz;
            st<NT>(a.px + id, px); st<NT>(a.py + id, py); st<NT>(a.pz + id, pz);
            st<NT>(a.qx + id, qx); st<NT>(a.qy + id, qy); st<NT>(a.qz + id, qz);
            float Pn = px + py + pz, Qn = qx + qy + qz;
            if (src_col && iy == a.gy) {
                const float s = (float)((double)a.gamp * a.gw[(ix - a.gx + 3) * 7 + (iz - a.gz + 3)]);
                Pn = px + py + pz + s;
                Qn = qx + qy + qz + s;
            }
            a.P[id] = Pn;
            a.Q[id] = Qn;
            illum_add(a, ix, iy, iz, Pn, Qn);
            snap_put<SNAP>(a, ix, iy, iz, Pn);
            if (rk >= 0) {
                const int j = iy - a.ry0;
                if (j >= 0 && j % a.rdy == 0 && j / a.rdy < a.rny) {
                    const float r = a.rec_kind == 0 ? Pn : (a.rec_kind == 1 ? Pn + Pn : Pn + Qn);
                    a.rec[rk + (long long)a.rnx * (j / a.rdy)] = r;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) vq[k] = vq[k + 1];
        vq[7] = vn;
    }
}

// Grid size (workgroups) of a 3D launch with the given y chunk.
inline int grid3d(const Grid &g, int ychunk, int &nzt, int &ntile, int &ntile_p)
{
    nzt = (g.nnz + TZ - 1) / TZ;
    const int nxt = (g.nnx + TX - 1) / TX;
    ntile = nzt * nxt;
    const int nchunk = ychunk > 0 ? (g.nny - 8 + ychunk - 1) / ychunk : 1;
    ntile_p = (ntile * nchunk + NXCD - 1) / NXCD * NXCD;
    return ntile_p;
}

}  // namespace s3d
}  // namespace vti
