// vti_fused3d.h — one 3D time step (velocity update + stress update) in a
// single y-marching kernel (gfx950).
//
// The two-kernel step reads u, v, w back from HBM right after writing them.
// Here a workgroup (64 z x 4 x columns, a chunk of y planes) computes, plane
// by plane: u1 on its 4 x-rows plus the 7 halo rows the x-stencil of the
// stress update needs, w1 on its z-range plus 7 halo cells, v1 three planes
// ahead of the stress plane (a register queue per column), and then the
// stress update of its own cells from those values (LDS for u1 and w1, the
// queue for v1).  P, Q, u, v, w are double-buffered (read the previous step,
// write the next): a halo value computed by one workgroup is never
// overwritten by another during the step.  Algorithmic HBM traffic: 14
// reads + 11 writes = 100 B per cell instead of 112 B; halo re-reads of P,
// Q, u, w come from L2.
//
// Arithmetic is the two kernels' (vti_stencil3d.h) expression by
// expression; results are bit-identical (checked by the GPU tests).
//
// Measured at 572^3 on MI355X: 7.3 ms per step against 4.47 ms for the two
// kernels.  The 7 halo rows of u1 (and the 18 P rows behind them) on a
// 4-row tile recompute u 2.75 times and re-read P 4.5 times from L2, and
// three barriers per plane serialise the phases; the 12 B per cell saved
// do not pay for that, so the two-kernel step stays the default.
#pragma once

#include <hip/hip_runtime.h>

#include "vti_internal.h"
#include "vti_stencil3d.h"

namespace vti {

namespace f3d {

using s3d::cst;
using s3d::ld;
using s3d::st;

constexpr int TZ = 64, TX = 4, NXCD = 8;
constexpr int HP = 7;                 // P rows on each side needed for u1 on the halo rows
constexpr int UR = TX + 7;            // u1 rows: x0-4 .. x0+TX+2
constexpr int PR = UR + 7;            // P rows: x0-7 .. x0+TX+6+... (u1 row r needs P rows r..r+7)
constexpr int QW = TZ + 16;           // Q cells per row: z0-8 .. z0+TZ+7
constexpr int WW = TZ + 8;            // w1 cells per row: z0-4 .. z0+TZ+3

template <int YCHUNK, bool NT>
__global__ __launch_bounds__(256) void fused3d(FusedArgs a, int nzt, int ntile, int ntile_p)
{
    __shared__ float sP[PR][TZ];
    __shared__ float sU[UR][TZ];
    __shared__ float sQ[TX][QW];
    __shared__ float sW[TX][WW];

    const Grid g = a.v.g;
    const int nchunk = (g.nny - 8 + YCHUNK - 1) / YCHUNK;
    const int per = ntile_p / NXCD;
    const int tl = (blockIdx.x % NXCD) * per + (blockIdx.x / NXCD);
    if (tl >= ntile * nchunk) return;   // workgroup-uniform
    const int t = tl % ntile, ch = tl / ntile;
    const int z0 = (t % nzt) * TZ, x0 = (t / nzt) * TX;
    const int y0 = 4 + ch * YCHUNK, y1 = min(g.nny - 4, y0 + YCHUNK);
    const int tz = threadIdx.x, tx = threadIdx.y, lid = tx * TZ + tz;
    const int iz = z0 + tz, ix = x0 + tx;
    const long long L = g.ldxy, LZ = g.ldz;
    const bool colv = iz >= 4 && iz < g.nnz - 4 && ix >= 4 && ix < g.nnx - 4;   // updated column
    const bool ingrid = iz < g.nnz && ix < g.nnx;
    const long long col = iz + (long long)ix * LZ;
    const long long src = a.v.src;
    const float amp = a.v.amp;
    // point source added to this step's copies of P and Q (add_source3d
    // GPU_vti3dfd_kernels.cu:85-86)
    auto P0 = [&](long long i) { float p = a.v.P[i]; if (i == src) p = p + amp; return p; };
    auto Q0 = [&](long long i) { float q = a.v.Q[i]; if (i == src) q = q + amp; return q; };
    const float dtx = a.v.dtx, dty = a.v.dty, dtz = a.v.dtz;

    // v1 at plane yp of this column (the previous value outside the updated
    // region, which no step changes)
    auto v1_at = [&](int yp, const float *pq8) -> float {
        const long long id = col + (long long)yp * L;
        if (!colv || yp < 4 || yp >= g.nny - 4) return ingrid && yp >= 0 && yp < g.nny ? a.v0[id] : 0.0f;
        float yy = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) yy += cst(m) * (pq8[3 + m + 1] - pq8[3 - m]);
        return a.v.cy2[yp] * a.v0[id] - a.v.cy1[yp] * dty * yy;
    };
    auto P0col = [&](int yp) -> float {
        return ingrid && yp >= 0 && yp < g.nny ? P0(col + (long long)yp * L) : 0.0f;
    };

    // prologue: v1 at planes y0-4 .. y0+2 into the queue; P queue at y0 .. y0+6
    float vq[8], pq[8];
    {
        float w8[8];
        for (int k = 0; k < 7; ++k) {
            const int yp = y0 - 4 + k;
#pragma unroll
            for (int m = 0; m < 8; ++m) w8[m] = P0col(yp - 3 + m);
            vq[k] = v1_at(yp, w8);
            if (colv && yp >= y0 && yp >= 4 && yp < g.nny - 4) st<NT>(a.v.v + col + (long long)yp * L, vq[k]);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) pq[k] = P0col(y0 + k);
    }

    for (int y = y0; y < y1; ++y) {
        const long long pl = (long long)y * L;
        const bool yv = y >= 4 && y < g.nny - 4;
        // ---- previous P and Q of this plane into LDS
        for (int e = lid; e < PR * TZ; e += 256) {
            const int r = e / TZ, z = e % TZ;
            const int x = x0 - HP + r, zz = z0 + z;
            sP[r][z] = x >= 0 && x < g.nnx && zz < g.nnz ? P0(zz + (long long)x * LZ + pl) : 0.0f;
        }
        for (int e = lid; e < TX * QW; e += 256) {
            const int r = e / QW, zi = e % QW;
            const int x = x0 + r, zz = z0 - 8 + zi;
            sQ[r][zi] = x < g.nnx && zz >= 0 && zz < g.nnz ? Q0(zz + (long long)x * LZ + pl) : 0.0f;
        }
        pq[7] = P0col(y + 7);
        __syncthreads();
        // ---- u1 on own + halo rows, w1 on own + halo z cells
        for (int e = lid; e < UR * TZ; e += 256) {
            const int r = e / TZ, z = e % TZ;
            const int x = x0 - 4 + r, zz = z0 + z;
            const long long id = zz + (long long)x * LZ + pl;
            float u1 = 0.0f;
            if (x >= 0 && x < g.nnx && zz < g.nnz) {
                const float u0 = a.u0[id];
                if (yv && x >= 4 && x < g.nnx - 4 && zz >= 4 && zz < g.nnz - 4) {
                    float xx = 0.0f;
#pragma unroll
                    for (int m = 0; m < 4; ++m) xx += cst(m) * (sP[r + 4 + m][z] - sP[r + 3 - m][z]);
                    u1 = a.v.cx2[x] * u0 - a.v.cx1[x] * dtx * xx;
                    if (r >= 4 && r < 4 + TX) st<NT>(a.v.u + id, u1);
                } else {
                    u1 = u0;
                }
            }
            sU[r][z] = u1;
        }
        for (int e = lid; e < TX * WW; e += 256) {
            const int r = e / WW, zi = e % WW;
            const int x = x0 + r, zz = z0 - 4 + zi;
            const long long id = zz + (long long)x * LZ + pl;
            float w1 = 0.0f;
            if (x < g.nnx && zz >= 0 && zz < g.nnz) {
                const float w0 = a.w0[id];
                if (yv && x >= 4 && x < g.nnx - 4 && zz >= 4 && zz < g.nnz - 4) {
                    float zt = 0.0f;
#pragma unroll
                    for (int m = 0; m < 4; ++m) zt += cst(m) * (sQ[r][zi + 5 + m] - sQ[r][zi + 4 - m]);
                    w1 = a.v.cz2[zz] * w0 - a.v.cz1[zz] * dtz * zt;
                    if (zi >= 4 && zi < 4 + TZ) st<NT>(a.v.w + id, w1);
                } else {
                    w1 = w0;
                }
            }
            sW[r][zi] = w1;
        }
        // ---- v1 three planes ahead (column queue)
        vq[7] = v1_at(y + 3, pq);
        if (colv && y + 3 < y1 && y + 3 < g.nny - 4) st<NT>(a.v.v + col + (long long)(y + 3) * L, vq[7]);
        __syncthreads();
        // ---- stress update of the own cell
        if (colv && yv) {
            const StressArgs &s = a.s;
            const long long id = col + pl;
            float xx = 0.0f, yy = 0.0f, zz = 0.0f;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                yy += cst(m) * (vq[4 + m] - vq[3 - m]);
                xx += cst(m) * (sU[tx + 4 + m][tz] - sU[tx + 3 - m][tz]);
                zz += cst(m) * (sW[tx][tz + 4 + m] - sW[tx][tz + 3 - m]);
            }
            const float vv = ld<NT>(s.vp + id);
            float ee = ld<NT>(s.ep + id), dd = ld<NT>(s.de + id);
            const float opx = ld<NT>(s.px + id), opy = ld<NT>(s.py + id), opz = ld<NT>(s.pz + id);
            const float oqx = ld<NT>(s.qx + id), oqy = ld<NT>(s.qy + id), oqz = ld<NT>(s.qz + id);
            if (s.sv && abs(ix - s.cx) <= s.svR && abs(iz - s.cz) <= s.svR) {   // SV taper
                const int rx = ix - s.cx, ry = y - s.cy, rz = iz - s.cz;
                const int r2 = rx * rx + ry * ry + rz * rz;
                if (ry <= s.svR && ry >= -s.svR && r2 <= s.svR * s.svR) {
                    const double f = s.tap[r2];
                    ee = (float)(f * (double)ee);
                    dd = (float)(f * (double)dd);
                }
            }
            const float e2 = 1.0f + 2.0f * ee;
            const float sd = sqrtf(1.0f + 2.0f * dd);
            const float ax1 = s.ax1[ix], ax2 = s.ax2[ix], az1 = s.az1[iz], az2 = s.az2[iz];
            const float ay1 = s.ay1[y], ay2 = s.ay2[y];
            const float px = ax2 * opx - ax1 * vv * vv * e2 * s.dtx * xx;
            const float py = ay2 * opy - ay1 * vv * vv * e2 * s.dty * yy;
            const float pz = az2 * opz - az1 * vv * vv * sd * s.dtz * zz;
            const float qx = ax2 * oqx - ax1 * vv * vv * sd * s.dtx * xx;
            const float qy = ay2 * oqy - ay1 * vv * vv * sd * s.dty * yy;
            const float qz = az2 * oqz - az1 * vv * vv * s.dtz * zz;
            st<NT>(s.px + id, px); st<NT>(s.py + id, py); st<NT>(s.pz + id, pz);
            st<NT>(s.qx + id, qx); st<NT>(s.qy + id, qy); st<NT>(s.qz + id, qz);
            float Pn = px + py + pz, Qn = qx + qy + qz;
            if (s.gsrc && y == s.gy && abs(ix - s.gx) <= 3 && abs(iz - s.gz) <= 3) {
                const float sg = (float)((double)s.gamp * s.gw[(ix - s.gx + 3) * 7 + (iz - s.gz + 3)]);
                Pn = px + py + pz + sg;
                Qn = qx + qy + qz + sg;
            }
            s.P[id] = Pn;
            s.Q[id] = Qn;
            if (s.rec && iz == s.rz) {
                const int k = ix - s.rx0, j = y - s.ry0;
                if (k >= 0 && k % s.rdx == 0 && k / s.rdx < s.rnx && j >= 0 && j % s.rdy == 0 && j / s.rdy < s.rny) {
                    const float rv = s.rec_kind == 0 ? Pn : (s.rec_kind == 1 ? Pn + Pn : Pn + Qn);
                    s.rec[k / s.rdx + (long long)s.rnx * (j / s.rdy)] = rv;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            vq[k] = vq[k + 1];
            pq[k] = pq[k + 1];
        }
        __syncthreads();   // LDS is rewritten by the next plane
    }
}

// Workgroups of a fused launch with the given y chunk (XCD-contiguous ids).
inline int grid(const Grid &g, int ychunk, int &nzt, int &ntile, int &ntile_p)
{
    nzt = (g.nnz + TZ - 1) / TZ;
    const int nxt = (g.nnx + TX - 1) / TX;
    ntile = nzt * nxt;
    const int nchunk = (g.nny - 8 + ychunk - 1) / ychunk;
    ntile_p = (ntile * nchunk + NXCD - 1) / NXCD * NXCD;
    return ntile_p;
}

}  // namespace f3d
}  // namespace vti
