// vti_batch.hip — many independent 2D shots of uneven grid sizes advanced
// together: one persistent work-list launch per phase covers every tile of
// every shot, the time index lives on the device so a block of steps is a
// replayable hipGraph, and the tile shape / grid size are chosen by timing
// (autotuner).  RTM of the whole batch runs in lockstep under one binomial
// checkpoint schedule; each shot gets its own image and illumination.
//
// Per shot the arithmetic is exactly the single-shot path's (GPU form:
// GPU_vti2dfd_kernels.cu:209-299 with the point source of :176-207; RTM
// imaging Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:654-698), so every shot's
// record and image equal vti_step / vti_rtm_shot bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "vti_internal.h"
#include "vti_rtm.h"

namespace vti {
namespace {

// Device descriptor of one shot of the batch.
struct BShot {
    int nnx, nnz, ldz, M;
    int nx, nz, np;
    float dtx, dtz;
    const float *vp, *ep, *de, *tab;
    float *f[8];        // source state: u w px pz qx qz P Q
    float *g[8];        // receiver state
    long long src;
    const float *amp;   // nt source gains
    int sv, svR, cx, cz;
    const double *tap;
    float *rec;         // [it][nrec]
    int nrec, rz, rx0, rdx, inj_z;
    float *img, *illum; // interior nx*nz
    int tile0, ntz, ntx;
    // checkpoint copies of the unsplit interior: first grid row of the shot
    // in the batch's row numbering; rows [xlo, xhi) x columns [zlo, zhi) are
    // the cells whose stress tables are all 1 (no split fields kept there)
    int row0, xlo, xhi, zlo, zhi;
};

__device__ __forceinline__ float cst(int m)
{
    return m == 0 ? 1.196289f : (m == 1 ? -0.0797526f : (m == 2 ? 0.009570313f : -0.0006975447f));
}

// tile -> shot by binary search over tile0 (shots sorted by tile0)
__device__ __forceinline__ int find_shot(const BShot *S, int ns, int t)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void compiler_fence() { asm volatile("" ::: "memory"); }

// First and one-past-last updated row of a thread's strip of XS rows
// (x in [4, nnx - 4)).
template <int TZ, int TXB, int XS>
__device__ __forceinline__ bool strip(const BShot &b, int lt, int &iz, int &x0, int &x1)
{
    iz = (lt % b.ntz) * TZ + threadIdx.x;
    const int xs = (lt / b.ntz) * (TXB * XS) + threadIdx.y * XS;
    x0 = max(xs, 4);
    x1 = min(xs + XS, b.nnx - 4);
    return iz >= 4 && iz < b.nnz - 4 && x0 < x1;
}

// which = 0: source state f, 1: receiver state g.  XS = 1: one cell per
// thread; XS > 1: each thread marches XS consecutive rows (x) of its
// column, holding the 8-deep x-stencil of P in a register queue (one load of
// P per cell instead of eight).  The arithmetic is the same in both.
template <int TZ, int TXB, int XS, int WHICH>
__global__ __launch_bounds__(TZ *TXB) void bvel(const BShot *__restrict__ S, int ns, int ntile,
                                               const int *__restrict__ itp, int inject)
{
    const int it = *itp;
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int s = find_shot(S, ns, t);
        const BShot &b = S[s];
        const int lt = t - b.tile0;
        int iz, x0, x1;
        if (!strip<TZ, TXB, XS>(b, lt, iz, x0, x1)) continue;
        float *const *F = WHICH ? b.g : b.f;
        const float *P = F[6], *Q = F[7];
        const float amp = (inject && WHICH == 0) ? b.amp[it] : 0.0f;
        const long long src = amp != 0.0f ? b.src : -1;
        const float *tb = b.tab;
        const long long L = b.ldz;
        float q[8];   // q[k] = P at row x - 3 + k
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = XS > 1 ? P[iz + (long long)(x0 - 3 + k) * L] : 0.0f;
        for (int ix = x0; ix < x1; ++ix) {
            const long long id = iz + (long long)ix * L;
            float pp[8], qz[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                pp[k] = XS > 1 ? q[k] : P[id + (long long)(k - 3) * L];
                qz[k] = Q[id + k - 3];
            }
            const float u0 = F[0][id], w0 = F[1][id];
            const float qn = XS > 1 && ix + 1 < x1 ? P[id + 5 * L] : 0.0f;
            compiler_fence();
            if (src >= 0) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (id + (long long)(k - 3) * L == src) pp[k] = pp[k] + amp;
                    if (id + k - 3 == src) qz[k] = qz[k] + amp;
                }
            }
            float xx = 0.0f, zz = 0.0f;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                xx += cst(m) * (pp[m + 4] - pp[3 - m]);
                zz += cst(m) * (qz[m + 4] - qz[3 - m]);
            }
            F[0][id] = tb[1 * b.M + ix] * u0 - tb[0 * b.M + ix] * b.dtx * xx;
            F[1][id] = tb[5 * b.M + iz] * w0 - tb[4 * b.M + iz] * b.dtz * zz;
            if (XS > 1) {
#pragma unroll
                for (int k = 0; k < 7; ++k) q[k] = q[k + 1];
                q[7] = qn;
            }
        }
    }
}

// MODE 0: recompute (source); 1: source with record (and, illum != 0, the
// illumination P*P+Q*Q, clamped to 1 where <= 0 when illum == 2);
// 2: receiver (imaging with the source P).  unsplit: P and Q updated
// directly where the cell's damping factors are all 1, the split fields
// only in the absorbing layer and, on the receiver side, on the injection
// row (as stress2d_kernel<true>).
// src_off (MODE 2): floats from each shot's source P to the source P imaged
// (0: the states themselves; a checkpoint slot's offset from the states)
template <int TZ, int TXB, int XS, int MODE>
__global__ __launch_bounds__(TZ *TXB) void bstress(const BShot *__restrict__ S, int ns, int ntile,
                                                  const int *__restrict__ itp, int sv_on, int image_every,
                                                  int illum, int unsplit, long long src_off)
{
    const int it = *itp;
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int s = find_shot(S, ns, t);
        const BShot &b = S[s];
        const int lt = t - b.tile0;
        int iz, x0, x1;
        if (!strip<TZ, TXB, XS>(b, lt, iz, x0, x1)) continue;
        float *const *F = MODE == 2 ? b.g : b.f;
        const long long L = b.ldz;
        float uq[8];   // uq[k] = u at row x - 4 + k
#pragma unroll
        for (int k = 0; k < 8; ++k) uq[k] = XS > 1 ? F[0][iz + (long long)(x0 - 4 + k) * L] : 0.0f;
        for (int ix = x0; ix < x1; ++ix) {
        const long long id = iz + (long long)ix * L;
        float ux[8], wz[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ux[k] = XS > 1 ? uq[k] : F[0][id + (long long)(k - 4) * L];
            wz[k] = F[1][id + k - 4];
        }
        const float un = XS > 1 && ix + 1 < x1 ? F[0][id + 4 * L] : 0.0f;
        compiler_fence();
        float xx = 0.0f, zz = 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            xx += cst(m) * (ux[4 + m] - ux[3 - m]);
            zz += cst(m) * (wz[4 + m] - wz[3 - m]);
        }
        const float vv = b.vp[id];
        float ee = b.ep[id], dd = b.de[id];
        if (sv_on && b.sv) {
            const int rx = ix - b.cx, rz = iz - b.cz, r2 = rx * rx + rz * rz;
            if (r2 <= b.svR * b.svR) {
                const double fct = b.tap[r2];
                ee = (float)(fct * (double)ee);
                dd = (float)(fct * (double)dd);
            }
        }
        const float e2 = 1.0f + 2.0f * ee;
        const float sd = sqrtf(1.0f + 2.0f * dd);
        const float *tb = b.tab;
        const float ax1 = tb[6 * b.M + ix], ax2 = tb[7 * b.M + ix];
        const float az1 = tb[10 * b.M + iz], az2 = tb[11 * b.M + iz];
        float Pn, Qn;
        // receiver side: the injection row keeps the split form, which
        // rebuilds P and Q (the written records do not persist), as in the
        // single-shot RTM
        const bool us = unsplit && !(MODE == 2 && iz == b.inj_z);
        if (us && ax1 == 1.0f && ax2 == 1.0f && az1 == 1.0f && az2 == 1.0f) {
            const float aa = vv * vv;
            Pn = F[6][id] - (aa * e2 * b.dtx * xx + aa * sd * b.dtz * zz);
            Qn = F[7][id] - (aa * sd * b.dtx * xx + aa * b.dtz * zz);
        } else {
            const float px = ax2 * F[2][id] - ax1 * vv * vv * e2 * b.dtx * xx;
            const float pz = az2 * F[3][id] - az1 * vv * vv * sd * b.dtz * zz;
            const float qx = ax2 * F[4][id] - ax1 * vv * vv * sd * b.dtx * xx;
            const float qz = az2 * F[5][id] - az1 * vv * vv * b.dtz * zz;
            F[2][id] = px; F[3][id] = pz; F[4][id] = qx; F[5][id] = qz;
            Pn = px + pz;
            Qn = qx + qz;
        }
        F[6][id] = Pn;
        F[7][id] = Qn;
        if (MODE == 1) {
            if (iz == b.rz) {
                const int k = ix - b.rx0;
                if (k >= 0 && k % b.rdx == 0 && k / b.rdx < b.nrec) b.rec[(long long)it * b.nrec + k / b.rdx] = Pn;
            }
            const int jx = ix - b.np, jz = iz - b.np;
            if (illum && jx >= 0 && jx < b.nx && jz >= 0 && jz < b.nz) {
                const long long o = jz + (long long)jx * b.nz;
                float v = b.illum[o] + (Pn * Pn + Qn * Qn);
                if (illum == 2 && v <= 0.0f) v = 1.0f;
                b.illum[o] = v;
            }
        }
        if (MODE == 2 && it % image_every == 0) {
            const int jx = ix - b.np, jz = iz - b.np;
            if (jx >= 0 && jx < b.nx && jz >= 0 && jz < b.nz)
                b.img[jz + (long long)jx * b.nz] += b.f[6][id + src_off] * Pn;
        }
        if (XS > 1) {
#pragma unroll
            for (int k = 0; k < 7; ++k) uq[k] = uq[k + 1];
            uq[7] = un;
        }
        }
    }
}

// Observed samples of step it overwrite g_P and g_Q on each shot's
// injection row (shot_record(load) Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:525-528).
__global__ void binject(const BShot *__restrict__ S, int ns, const int *__restrict__ roff, int ntot,
                        const int *__restrict__ itp)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ntot) return;
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (roff[mid] <= k) lo = mid;
        else hi = mid - 1;
    }
    const BShot &b = S[lo];
    const int r = k - roff[lo];
    if (r >= b.nrec) return;
    const long long c = b.inj_z + (long long)(b.rx0 + r * b.rdx) * b.ldz;
    const float v = b.rec[(long long)(*itp) * b.nrec + r];
    b.g[6][c] = v;
    b.g[7][c] = v;
}

__global__ void bset_it(int *itp, int v) { if (threadIdx.x == 0) *itp = v; }
__global__ void badd_it(int *itp, int d) { if (threadIdx.x == 0) *itp += d; }

__global__ void bzero(float4 *p, size_t n4)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// grid row -> shot by binary search over row0
__device__ __forceinline__ int find_row(const BShot *S, int ns, int r)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S[mid].row0 <= r) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Checkpoint copy of the source states under the unsplit interior (store:
// states -> slot, else slot -> states; the slot has the states' layout):
// u, w, P and Q whole, the split fields px, pz, qx, qz only outside
// [xlo, xhi) x [zlo, zhi), the cells that never read or write them.  One
// workgroup per grid row of every shot, 16-byte pieces (rows are ldz floats,
// a multiple of 32).
__global__ void bckpt(const BShot *__restrict__ S, int ns, int nrows, const float *__restrict__ base,
                      float *__restrict__ slot, int store)
{
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const BShot &b = S[find_row(S, ns, r)];
        const int ix = r - b.row0;
        const bool mid = ix >= b.xlo && ix < b.xhi;
        const int n4 = b.ldz / 4;
        for (int k = 0; k < 8; ++k) {
            const size_t o = (size_t)(b.f[k] - base) + (size_t)ix * b.ldz;
            const float4 *src = reinterpret_cast<const float4 *>(store ? base + o : slot + o);
            float4 *dst = reinterpret_cast<float4 *>(store ? slot + o : const_cast<float *>(base) + o);
            const bool part = mid && k >= 2 && k <= 5;
            for (int i = threadIdx.x; i < n4; i += blockDim.x) {
                if (part && 4 * i >= b.zlo && 4 * i + 4 <= b.zhi) continue;
                dst[i] = src[i];
            }
        }
    }
}

// Tile shapes the autotuner chooses from: tz x tx threads, each thread
// marching xs rows (the tile covers tz x tx*xs cells).
struct Shape {
    int tz, tx, xs;
};
constexpr Shape kShapes[] = {{64, 4, 1},  {64, 2, 1},  {64, 8, 1},  {128, 2, 1}, {32, 8, 1},
                             {64, 4, 16}, {64, 4, 32}, {64, 2, 32}, {32, 8, 16}, {128, 2, 16}};
constexpr int kNumShapes = (int)(sizeof(kShapes) / sizeof(kShapes[0]));

}  // namespace
}  // namespace vti

using namespace vti;

struct vti_batch {
    int nshot = 0, nt = 0;
    std::vector<vti_cfg> cfgs;
    std::vector<vti_shot> shots;
    std::vector<BShot> hb;          // host copy of the descriptors
    BShot *db = nullptr;            // device descriptors (per tile shape: tile0/ntz/ntx differ)
    int *roff = nullptr;            // receiver offsets per shot (device)
    int nrec_tot = 0;
    int max_nrec = 0;
    float *rec_t = nullptr;         // transpose scratch for the mute
    int *itp = nullptr;             // device time index
    hipStream_t stream = nullptr;
    float *arena = nullptr;         // models, states, tables, images
    size_t arena_floats = 0;
    size_t state_floats = 0;        // floats of all source states (contiguous)
    float *state_base = nullptr, *rstate_base = nullptr;
    float *recbuf = nullptr;
    double *taps = nullptr;
    float *amps = nullptr;
    long long cells = 0;            // updated cells per step, all shots
    int shape = 0, grid_mult = 0;   // autotuned
    int ntile = 0;
    int nrows = 0;                  // grid rows of all shots (checkpoint copies)
    long long src_off = 0;          // receiver steps: where the imaged source P lies (see bstress)
    std::vector<float *> slots;
    std::vector<float> vp1, eps1;   // padded model at (ix 0, iz 1): the direct-wave mute's velocity
    int graph_k = 0;
    hipGraphExec_t gexec[4] = {};   // forward+record, forward, receiver, forward+record+illumination
    int image_every = 1, illum_mode = 0, sv_fwd = 1, sv_bwd = 0;
    long long fwd_steps = 0;
    bool unsplit = false;           // source-side unsplit interior (point-source form)

    ~vti_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &e : gexec)
            if (e) (void)hipGraphExecDestroy(e);
        for (float *s : slots) (void)hipFree(s);
        (void)hipFree(db); (void)hipFree(roff); (void)hipFree(itp); (void)hipFree(arena);
        (void)hipFree(recbuf); (void)hipFree(rec_t); (void)hipFree(taps); (void)hipFree(amps);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int bfail(int code, const std::string &m) { return fail(code, m); }

template <int TZ, int TXB, int XS>
void launch_phase(vti_batch *B, int phase, int sv_on, int grid)
{
    const dim3 blk(TZ, TXB);
    const int us = B->unsplit ? 1 : 0;
    switch (phase) {
    case 0: hipLaunchKernelGGL((bvel<TZ, TXB, XS, 0>), dim3(grid), blk, 0, B->stream, B->db, B->nshot, B->ntile, B->itp, 1); break;
    case 1: hipLaunchKernelGGL((bstress<TZ, TXB, XS, 0>), dim3(grid), blk, 0, B->stream, B->db, B->nshot, B->ntile, B->itp, sv_on, B->image_every, 0, us, 0LL); break;
    case 2: hipLaunchKernelGGL((bstress<TZ, TXB, XS, 1>), dim3(grid), blk, 0, B->stream, B->db, B->nshot, B->ntile, B->itp, sv_on, B->image_every, 0, us, 0LL); break;
    case 3: hipLaunchKernelGGL((bvel<TZ, TXB, XS, 1>), dim3(grid), blk, 0, B->stream, B->db, B->nshot, B->ntile, B->itp, 0); break;
    case 4: hipLaunchKernelGGL((bstress<TZ, TXB, XS, 2>), dim3(grid), blk, 0, B->stream, B->db, B->nshot, B->ntile, B->itp, sv_on, B->image_every, 0, us, B->src_off); break;
    case 5: hipLaunchKernelGGL((bstress<TZ, TXB, XS, 1>), dim3(grid), blk, 0, B->stream, B->db, B->nshot, B->ntile, B->itp, sv_on, B->image_every, 2, us, 0LL); break;
    }
}

void phase(vti_batch *B, int ph, int sv_on)
{
    int grid = B->ntile;
    if (B->grid_mult > 0) grid = std::min(grid, B->grid_mult * 256);
    switch (B->shape) {
    case 0: launch_phase<64, 4, 1>(B, ph, sv_on, grid); break;
    case 1: launch_phase<64, 2, 1>(B, ph, sv_on, grid); break;
    case 2: launch_phase<64, 8, 1>(B, ph, sv_on, grid); break;
    case 3: launch_phase<128, 2, 1>(B, ph, sv_on, grid); break;
    case 4: launch_phase<32, 8, 1>(B, ph, sv_on, grid); break;
    case 5: launch_phase<64, 4, 16>(B, ph, sv_on, grid); break;
    case 6: launch_phase<64, 4, 32>(B, ph, sv_on, grid); break;
    case 7: launch_phase<64, 2, 32>(B, ph, sv_on, grid); break;
    case 8: launch_phase<32, 8, 16>(B, ph, sv_on, grid); break;
    case 9: launch_phase<128, 2, 16>(B, ph, sv_on, grid); break;
    }
}

// tiles of the current shape; rewrites the device descriptors
int set_shape(vti_batch *B, int shape)
{
    B->shape = shape;
    const Shape sh = kShapes[shape];
    int t = 0;
    for (auto &b : B->hb) {
        b.tile0 = t;
        b.ntz = (b.nnz + sh.tz - 1) / sh.tz;
        b.ntx = (b.nnx + sh.tx * sh.xs - 1) / (sh.tx * sh.xs);
        t += b.ntz * b.ntx;
    }
    B->ntile = t;
    for (auto &e : B->gexec)
        if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    return hipMemcpy(B->db, B->hb.data(), B->hb.size() * sizeof(BShot), hipMemcpyHostToDevice) == hipSuccess
               ? VTI_OK : bfail(VTI_ERR_HIP, "batch descriptors");
}

// one step: 0 forward with record, 1 forward (recompute), 2 receiver
// (reverse time), 3 forward with record and illumination (RTM first sweep)
void enqueue_step(vti_batch *B, int kind)
{
    if (kind != 2) {
        phase(B, 0, 0);
        phase(B, kind == 0 ? 2 : (kind == 1 ? 1 : 5), B->sv_fwd);
        hipLaunchKernelGGL(badd_it, dim3(1), dim3(64), 0, B->stream, B->itp, 1);
    } else {
        if (B->nrec_tot > 0)
            hipLaunchKernelGGL(binject, dim3((B->nrec_tot + 255) / 256), dim3(256), 0, B->stream, B->db, B->nshot,
                               B->roff, B->nrec_tot, B->itp);
        phase(B, 3, 0);
        phase(B, 4, B->sv_bwd);
        hipLaunchKernelGGL(badd_it, dim3(1), dim3(64), 0, B->stream, B->itp, -1);
    }
}

// n steps of a kind, replaying the captured block graph where it fits
int run_steps(vti_batch *B, int kind, int n)
{
    const int K = B->graph_k;
    int done = 0;
    if (K > 1) {
        if (!B->gexec[kind]) {
            hipGraph_t gr;
            if (hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return bfail(VTI_ERR_HIP, "graph capture");
            for (int k = 0; k < K; ++k) enqueue_step(B, kind);
            if (hipStreamEndCapture(B->stream, &gr) != hipSuccess) return bfail(VTI_ERR_HIP, "graph end");
            const hipError_t e = hipGraphInstantiate(&B->gexec[kind], gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (e != hipSuccess) return bfail(VTI_ERR_HIP, "graph instantiate");
        }
        for (; done + K <= n; done += K)
            if (hipGraphLaunch(B->gexec[kind], B->stream) != hipSuccess) return bfail(VTI_ERR_HIP, "graph launch");
    }
    for (; done < n; ++done) enqueue_step(B, kind);
    return hipGetLastError() == hipSuccess ? VTI_OK : bfail(VTI_ERR_HIP, "batch launch");
}

void zero_floats(vti_batch *B, float *p, size_t n)
{
    hipLaunchKernelGGL(bzero, dim3(4096), dim3(256), 0, B->stream, reinterpret_cast<float4 *>(p), n / 4);
}

}  // namespace

extern "C" {

int vti_batch_create(int nshot, const vti_cfg *cfgs, const vti_shot *shots, const float *const *vp,
                     const float *const *eps, const float *const *del, vti_batch **out)
{
    if (nshot < 1 || !cfgs || !shots || !vp || !eps || !del || !out) return bfail(VTI_ERR_ARG, "vti_batch_create: bad arguments");
    *out = nullptr;
    std::unique_ptr<vti_batch> B(new (std::nothrow) vti_batch());
    if (!B) return bfail(VTI_ERR_ALLOC, "host");
    B->nshot = nshot;
    B->nt = cfgs[0].nt;
    B->cfgs.assign(cfgs, cfgs + nshot);
    B->shots.assign(shots, shots + nshot);
    // Layout: all source states, then all receiver states (each contiguous,
    // so a checkpoint is one copy), then per shot: vp ep de, image,
    // illumination, 12 tables.  Offsets are 64-float aligned.
    int nrec_tot = 0;
    std::vector<int> roff(nshot);
    std::vector<size_t> vol(nshot), off_state(nshot), off_rest(nshot);
    auto al = [](size_t v) { return (v + 63) / 64 * 64; };
    size_t nstate = 0;
    for (int s = 0; s < nshot; ++s) {
        const vti_cfg &c = cfgs[s];
        if (c.ndim != 2 || c.mode != VTI_MODE_GPU_REF || c.nt != B->nt || c.nx < 1 || c.nz < 1 || c.npml < 0)
            return bfail(VTI_ERR_ARG, "vti_batch_create: 2D GPU-form shots with equal nt only");
        const int nnx = c.nx + 2 * c.npml, nnz = c.nz + 2 * c.npml, ldz = (nnz + 31) / 32 * 32;
        if (nnx < 9 || nnz < 9) return bfail(VTI_ERR_ARG, "vti_batch_create: grid too small");
        const vti_shot &sh = shots[s];
        if (sh.sx < 4 || sh.sx >= nnx - 4 || sh.sz < 4 || sh.sz >= nnz - 4 || sh.rec_nx < 0 || sh.rec_dx < 1 ||
            (sh.rec_nx > 0 && (sh.rec_z < 4 || sh.rec_z >= nnz - 4 || sh.rec_x0 < 4 ||
                               sh.rec_x0 + (sh.rec_nx - 1) * sh.rec_dx >= nnx - 4)))
            return bfail(VTI_ERR_ARG, "vti_batch_create: source or receivers outside the updated region");
        roff[s] = nrec_tot;
        nrec_tot += sh.rec_nx;
        B->max_nrec = std::max(B->max_nrec, sh.rec_nx);
        vol[s] = (size_t)ldz * nnx;
        off_state[s] = nstate;
        nstate += 8 * vol[s];
    }
    size_t pos = 2 * nstate;
    for (int s = 0; s < nshot; ++s) {
        const vti_cfg &c = cfgs[s];
        const int M = std::max(c.nx, c.nz) + 2 * c.npml;
        off_rest[s] = pos;
        pos = al(pos + 3 * vol[s] + 2 * (size_t)c.nx * c.nz + 12 * (size_t)M);
    }
    const size_t nf = pos;
    if (hipMalloc(&B->arena, nf * sizeof(float)) != hipSuccess) return bfail(VTI_ERR_ALLOC, "vti_batch_create: arena");
    B->arena_floats = nf;
    B->state_floats = nstate;
    B->state_base = B->arena;
    B->rstate_base = B->arena + nstate;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) return bfail(VTI_ERR_HIP, "stream");
    zero_floats(B.get(), B->arena, nf);
    // the stream does not synchronise with the blocking copies below
    if (hipStreamSynchronize(B->stream) != hipSuccess) return bfail(VTI_ERR_HIP, "vti_batch_create: zero");
    if (hipMalloc(&B->itp, 64) != hipSuccess || hipMalloc(&B->recbuf, (size_t)std::max(nrec_tot, 1) * B->nt * sizeof(float)) != hipSuccess ||
        hipMalloc(&B->amps, (size_t)nshot * B->nt * sizeof(float)) != hipSuccess ||
        hipMalloc(&B->roff, (size_t)nshot * sizeof(int)) != hipSuccess ||
        hipMalloc(&B->db, (size_t)nshot * sizeof(BShot)) != hipSuccess ||
        hipMalloc(&B->rec_t, (size_t)std::max(B->max_nrec, 1) * B->nt * sizeof(float)) != hipSuccess)
        return bfail(VTI_ERR_ALLOC, "vti_batch_create: buffers");
    B->nrec_tot = nrec_tot;
    if (hipMemcpy(B->roff, roff.data(), nshot * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(VTI_ERR_HIP, "roff");
    std::vector<double> tapall;
    std::vector<size_t> tapoff(nshot);
    for (int s = 0; s < nshot; ++s) {
        vti_cfg &c = B->cfgs[s];
        if (c.sv_R <= 0) c.sv_R = 15;
        if (c.sv_r <= 0) c.sv_r = 4;
        std::vector<double> tap;
        taper_table(0, c.sv_R, c.sv_r, tap);
        tapoff[s] = tapall.size();
        tapall.insert(tapall.end(), tap.begin(), tap.end());
    }
    if (hipMalloc(&B->taps, tapall.size() * sizeof(double)) != hipSuccess ||
        hipMemcpy(B->taps, tapall.data(), tapall.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(VTI_ERR_ALLOC, "taper");
    size_t rec_off = 0;
    std::vector<float> amps((size_t)nshot * B->nt);
    B->hb.resize(nshot);
    B->cells = 0;
    for (int s = 0; s < nshot; ++s) {
        const vti_cfg &c = B->cfgs[s];
        const vti_shot &sh = shots[s];
        BShot &b = B->hb[s];
        b.nnx = c.nx + 2 * c.npml;
        b.nnz = c.nz + 2 * c.npml;
        b.ldz = (b.nnz + 31) / 32 * 32;
        b.M = std::max(c.nx, c.nz) + 2 * c.npml;
        b.nx = c.nx; b.nz = c.nz; b.np = c.npml;
        b.dtx = c.dt / c.dx;
        b.dtz = c.dt / c.dz;
        const size_t v = vol[s];
        for (int k = 0; k < 8; ++k) {
            b.f[k] = B->state_base + off_state[s] + k * v;
            b.g[k] = B->rstate_base + off_state[s] + k * v;
        }
        float *pr = B->arena + off_rest[s];
        float *model = pr;
        b.vp = pr; pr += v;
        b.ep = pr; pr += v;
        b.de = pr; pr += v;
        b.img = pr; pr += (size_t)c.nx * c.nz;
        b.illum = pr; pr += (size_t)c.nx * c.nz;
        float *tab = pr;
        b.tab = pr;
        // d0 (get_d02d GPU_vti2dfd_kernels.cu:105) and tables, as vti_set_model
        std::vector<float> htab(12 * (size_t)b.M, 1.0f);
        float d0 = 0.0f;
        if (c.npml > 0) {
            const long long N = (long long)b.nnx * b.nnz / 2;
            const long long iz = N % b.nnz, ix = N / b.nnz;
            auto cl = [&](long long i, int nn) { return i < c.npml ? 0 : (i >= c.npml + nn ? nn - 1 : i - c.npml); };
            const float vv = vp[s][cl(iz, c.nz) + cl(ix, c.nx) * (long long)c.nz];
            const float sdx = c.dx + c.dz;
            d0 = (float)(10.0 * vv * std::log(100000.0) / (2.0 * c.npml * (sdx / 2.0)));
        }
        const float dtd0 = c.dt * d0;
        const int M = b.M;
        axis_tables(c.nx, c.npml, dtd0, 0, &htab[0 * M], &htab[1 * M], &htab[6 * M], &htab[7 * M]);
        axis_tables(c.nz, c.npml, dtd0, 0, &htab[4 * M], &htab[5 * M], &htab[10 * M], &htab[11 * M]);
        // the undamped range of each axis (the stress tables exactly 1), as
        // the stress kernels test it; not one interval: every cell counts as
        // damped (whole-row checkpoint copies)
        auto undamped = [&](int t1, int t2, int nn, int &lo, int &hi) {
            lo = hi = 0;
            int i = 0;
            auto one = [&](int j) { return htab[t1 * M + j] == 1.0f && htab[t2 * M + j] == 1.0f; };
            while (i < nn && !one(i)) ++i;
            lo = i;
            while (i < nn && one(i)) ++i;
            hi = i;
            for (; i < nn; ++i)
                if (one(i)) { lo = hi = 0; return; }
        };
        undamped(6, 7, b.nnx, b.xlo, b.xhi);
        undamped(10, 11, b.nnz, b.zlo, b.zhi);
        b.row0 = B->nrows;
        B->nrows += b.nnx;
        if (hipMemcpy(tab, htab.data(), htab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return bfail(VTI_ERR_HIP, "tables");
        float *stage = nullptr;
        const size_t n = (size_t)c.nx * c.nz;
        for (size_t i = 0; i < n; ++i)
            if (!(vp[s][i] > 0.0f) || !std::isfinite(vp[s][i])) return bfail(VTI_ERR_ARG, "vti_batch_create: vp must be positive");
        if (hipMalloc(&stage, n * sizeof(float)) != hipSuccess) return bfail(VTI_ERR_ALLOC, "stage");
        Grid g{b.nnx, 1, b.nnz, b.ldz, (long long)b.ldz * b.nnx};
        const int jz1 = 1 < c.npml ? 0 : (1 - c.npml < c.nz ? 1 - c.npml : c.nz - 1);
        B->vp1.push_back(vp[s][jz1]);
        B->eps1.push_back(eps[s][jz1]);
        const float *srcs[3] = {vp[s], eps[s], del[s]};
        float *dsts[3] = {model, model + v, model + 2 * v};
        hipError_t e = hipSuccess;
        for (int k = 0; k < 3 && e == hipSuccess; ++k) {
            e = hipMemcpy(stage, srcs[k], n * sizeof(float), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = launch_pad_model(g, 2, c.nx, 1, c.nz, c.npml, stage, dsts[k], B->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(B->stream);
        }
        (void)hipFree(stage);
        if (e != hipSuccess) return bfail(VTI_ERR_HIP, "model upload");
        std::vector<float> a;
        source_gains(c, a);
        std::copy(a.begin(), a.end(), amps.begin() + (size_t)s * B->nt);
        b.src = sh.sz + (long long)sh.sx * b.ldz;
        b.amp = B->amps + (size_t)s * B->nt;
        b.sv = c.sv; b.svR = c.sv_R; b.cx = sh.sv_cx; b.cz = sh.sv_cz;
        b.tap = B->taps + tapoff[s];
        b.nrec = sh.rec_nx; b.rz = sh.rec_z; b.rx0 = sh.rec_x0; b.rdx = sh.rec_dx;
        b.inj_z = sh.rec_z;
        b.rec = B->recbuf + rec_off;
        rec_off += (size_t)sh.rec_nx * B->nt;
        B->cells += (long long)(b.nnx - 8) * (b.nnz - 8);
    }
    if (hipMemcpy(B->amps, amps.data(), amps.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(VTI_ERR_HIP, "amps");
    int st = set_shape(B.get(), 0);
    if (st) return st;
    if (hipStreamSynchronize(B->stream) != hipSuccess) return bfail(VTI_ERR_HIP, "sync");
    *out = B.release();
    return VTI_OK;
}

int vti_batch_destroy(vti_batch *B)
{
    delete B;
    return VTI_OK;
}

int vti_batch_config(vti_batch *B, int tile_shape, int grid_mult, int graph_steps)
{
    if (!B || tile_shape < 0 || tile_shape >= kNumShapes || grid_mult < 0 ||
        graph_steps < 0)
        return bfail(VTI_ERR_ARG, "vti_batch_config: bad arguments");
    B->grid_mult = grid_mult;
    B->graph_k = graph_steps;
    return set_shape(B, tile_shape);
}

int vti_batch_get_config(vti_batch *B, int *tile_z, int *tile_x, int *grid_mult, int *graph_steps)
{
    if (!B) return bfail(VTI_ERR_ARG, "null");
    if (tile_z) *tile_z = kShapes[B->shape].tz;
    if (tile_x) *tile_x = kShapes[B->shape].tx * kShapes[B->shape].xs;
    if (grid_mult) *grid_mult = B->grid_mult;
    if (graph_steps) *graph_steps = B->graph_k;
    return VTI_OK;
}

long long vti_batch_cells(const vti_batch *B) { return B ? B->cells : 0; }

int vti_batch_set_unsplit(vti_batch *B, int on)
{
    if (!B) return bfail(VTI_ERR_ARG, "vti_batch_set_unsplit: null");
    if (on)
        for (const vti_cfg &c : B->cfgs)
            if (c.mode != VTI_MODE_GPU_REF) return bfail(VTI_ERR_ARG, "vti_batch_set_unsplit: point-source (GPU) form only");
    B->unsplit = on != 0;
    for (auto &e : B->gexec)   // captured launches carry the old setting
        if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    return VTI_OK;
}

// Forward of all shots over nsteps steps from a zero state (the modelled
// records); used by the autotuner and the forward-only benchmark.
int vti_batch_forward(vti_batch *B, int nsteps, int record)
{
    if (!B || nsteps < 0 || nsteps > B->nt) return bfail(VTI_ERR_ARG, "vti_batch_forward: bad arguments");
    zero_floats(B, B->state_base, B->state_floats);
    hipLaunchKernelGGL(bset_it, dim3(1), dim3(64), 0, B->stream, B->itp, 0);
    int st = run_steps(B, record ? 0 : 1, nsteps);
    if (st) return st;
    return hipStreamSynchronize(B->stream) == hipSuccess ? VTI_OK : bfail(VTI_ERR_HIP, "sync");
}

// Times `steps` forward steps of every tile shape x grid multiplier and keeps
// the fastest.  Returns the chosen shape index.
int vti_batch_autotune(vti_batch *B, int steps, double *best_ms)
{
    if (!B || steps < 1 || steps > B->nt || B->nt < 2)
        return bfail(VTI_ERR_ARG, "vti_batch_autotune: need 1 <= steps <= nt and nt >= 2");
    const int nshape = kNumShapes;
    const int mults[] = {0, 8, 16};
    double best = 1e30;
    int bs = 0, bm = 0;
    const int keep_graph = B->graph_k;
    B->graph_k = 0;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int s = 0; s < nshape; ++s)
        for (int m : mults) {
            B->grid_mult = m;
            int st = set_shape(B, s);
            if (st) return st;
            if ((st = vti_batch_forward(B, 2, 0))) return st;
            (void)hipEventRecord(e0, B->stream);
            hipLaunchKernelGGL(bset_it, dim3(1), dim3(64), 0, B->stream, B->itp, 0);
            if ((st = run_steps(B, 1, steps))) return st;
            (void)hipEventRecord(e1, B->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) { best = ms; bs = s; bm = m; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B->grid_mult = bm;
    B->graph_k = keep_graph;
    const int st = set_shape(B, bs);
    if (st) return st;
    zero_floats(B, B->state_base, B->state_floats);
    if (best_ms) *best_ms = best / steps;
    return hipStreamSynchronize(B->stream) == hipSuccess ? bs : bfail(VTI_ERR_HIP, "sync");
}

// RTM of every shot of the batch in lockstep.  obs: per-shot observed
// records [irec][it] (host), or NULL to migrate the modelled records.
int vti_batch_rtm(vti_batch *B, const vti_rtm_opts *o, const float *const *obs, long long *fwd_steps)
{
    if (!B || !o || o->image_every < 1 || o->nsnap < 0 || o->illum_mode != 0)
        return bfail(VTI_ERR_ARG, "vti_batch_rtm: bad arguments (the batch images with source illumination, illum_mode 0)");
    const int nt = B->nt;
    // the taper switches and the imaging interval hold for this call only:
    // vti_batch_forward keeps the batch's own (SV taper on in the forward
    // propagation); illum_mode stays, vti_batch_image reads it
    struct Restore {
        vti_batch *B;
        int sf, sb, ie;
        ~Restore() { B->sv_fwd = sf; B->sv_bwd = sb; B->image_every = ie; }
    } restore{B, B->sv_fwd, B->sv_bwd, B->image_every};
    B->image_every = o->image_every;
    B->illum_mode = o->illum_mode;
    B->sv_fwd = o->sv_forward;
    B->sv_bwd = o->sv_backward;
    const int nslot = std::max(0, o->nsnap > 0 ? std::min(o->nsnap, nt - 1) : nt - 1);
    while ((int)B->slots.size() < nslot) {
        float *p = nullptr;
        if (hipMalloc(&p, B->state_floats * sizeof(float)) != hipSuccess)
            return bfail(VTI_ERR_ALLOC, "vti_batch_rtm: checkpoint slots");
        B->slots.push_back(p);
    }
    // zero receiver states, images, illuminations
    zero_floats(B, B->rstate_base, B->state_floats);
    for (auto &b : B->hb) {
        (void)hipMemsetAsync(b.img, 0, (size_t)b.nx * b.nz * sizeof(float), B->stream);
        (void)hipMemsetAsync(b.illum, 0, (size_t)b.nx * b.nz * sizeof(float), B->stream);
    }
    const size_t sb = B->state_floats * sizeof(float);
    auto ok = [](hipError_t e, const char *w) { return e == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, w); };
    int cur_it = 0;   // host mirror of the device time index for the forward state
    SweepOps op;
    op.zero = [&]() {
        zero_floats(B, B->state_base, B->state_floats);
        hipLaunchKernelGGL(bset_it, dim3(1), dim3(64), 0, B->stream, B->itp, 0);
        cur_it = 0;
        return ok(hipGetLastError(), "zero");
    };
    std::vector<int> slot_it(nslot, 0);
    // checkpoints: whole states, or (unsplit interior) all but the split
    // fields of the cells that never use them
    auto ckpt = [&](int i, int store) {
        if (!B->unsplit)
            return ok(store ? hipMemcpyAsync(B->slots[i], B->state_base, sb, hipMemcpyDeviceToDevice, B->stream)
                            : hipMemcpyAsync(B->state_base, B->slots[i], sb, hipMemcpyDeviceToDevice, B->stream),
                      "checkpoint");
        hipLaunchKernelGGL(bckpt, dim3((unsigned)std::min(B->nrows, 8192)), dim3(256), 0, B->stream, B->db, B->nshot,
                           B->nrows, B->state_base, B->slots[i], store);
        return ok(hipGetLastError(), "checkpoint");
    };
    op.store = [&](int i) { slot_it[i] = cur_it; return ckpt(i, 1); };
    op.restore = [&](int i) {
        cur_it = slot_it[i];
        hipLaunchKernelGGL(bset_it, dim3(1), dim3(64), 0, B->stream, B->itp, cur_it);
        return ckpt(i, 0);
    };
    op.step = [&](int it) {
        if (it != cur_it) return fail(VTI_ERR_STATE, "batch: step index");
        ++cur_it;
        return run_steps(B, 1, 1);
    };
    op.first_step = [&](int it) {
        if (it != cur_it) return fail(VTI_ERR_STATE, "batch: step index");
        ++cur_it;
        return run_steps(B, 3, 1);
    };
    op.after_first = [&]() {
        if (obs)
            for (int s = 0; s < B->nshot; ++s) {
                const BShot &b = B->hb[s];
                if (b.nrec == 0) continue;
                // [irec][it] host -> [it][irec] device
                std::vector<float> t((size_t)b.nrec * nt);
                for (int r = 0; r < b.nrec; ++r)
                    for (int it = 0; it < nt; ++it) t[(size_t)it * b.nrec + r] = obs[s][(size_t)r * nt + it];
                if (hipMemcpyAsync(b.rec, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, B->stream) != hipSuccess)
                    return fail(VTI_ERR_HIP, "obs upload");
                if (hipStreamSynchronize(B->stream) != hipSuccess) return fail(VTI_ERR_HIP, "sync");
            }
        else if (o->mute_tt >= 0)   // the modelled record, muted as vti_rtm_shot does (:1409)
            for (int s = 0; s < B->nshot; ++s) {
                const BShot &b = B->hb[s];
                const vti_cfg &c = B->cfgs[s];
                const vti_shot &sh = B->shots[s];
                if (b.nrec == 0) continue;
                if (launch_transpose(b.rec, B->rec_t, nt, b.nrec, B->stream) != hipSuccess ||
                    launch_mute2d(B->rec_t, b.nrec, nt, c.dt, c.favg, c.dx, c.dz, sh.sx - c.npml + 1, 0,
                                  sh.sz - c.npml + 1, 1, B->vp1[s], B->eps1[s], o->mute_tt, sh.rec_x0 - c.npml,
                                  sh.rec_dx, B->stream) != hipSuccess ||
                    launch_transpose(B->rec_t, b.rec, b.nrec, nt, B->stream) != hipSuccess)
                    return fail(VTI_ERR_HIP, "mute");
            }
        hipLaunchKernelGGL(bset_it, dim3(1), dim3(64), 0, B->stream, B->itp + 1, nt - 1);
        return ok(hipGetLastError(), "after_first");
    };
    // receiver steps use their own device time index (itp[1])
    op.consume = [&](int k) {
        (void)k;
        int *save = B->itp;
        B->itp = save + 1;
        int st = run_steps(B, 2, 1);
        B->itp = save;
        return st;
    };
    // a checkpointed state imaged from its slot (the slot has the states'
    // layout and holds P whole in both checkpoint forms)
    op.consume_slot = [&](int k, int i) {
        B->src_off = ((long long)reinterpret_cast<intptr_t>(B->slots[i]) -
                      (long long)reinterpret_cast<intptr_t>(B->state_base)) / (long long)sizeof(float);
        const int st = op.consume(k);
        B->src_off = 0;
        return st;
    };
    // graphs capture the device-index pointer in effect: keep separate graphs
    for (auto &e : B->gexec)
        if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    const int keep = B->graph_k;
    B->graph_k = 0;   // the reverse sweep advances one step at a time
    int st = reverse_sweep(nt, nslot, op, fwd_steps);
    B->graph_k = keep;
    if (st) return st;
    return hipStreamSynchronize(B->stream) == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, "sync");
}

int vti_batch_image(vti_batch *B, int shot, int finalize, float *img, float *illum)
{
    if (!B || shot < 0 || shot >= B->nshot) return bfail(VTI_ERR_ARG, "vti_batch_image: bad shot");
    const BShot &b = B->hb[shot];
    const size_t n = (size_t)b.nx * b.nz;
    std::vector<float> hi(n), hl(n);
    if (hipStreamSynchronize(B->stream) != hipSuccess ||
        hipMemcpy(hi.data(), b.img, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hl.data(), b.illum, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(VTI_ERR_HIP, "vti_batch_image: copy");
    if (illum) std::memcpy(illum, hl.data(), n * sizeof(float));
    if (!img) return VTI_OK;
    if (!finalize) { std::memcpy(img, hi.data(), n * sizeof(float)); return VTI_OK; }
    const bool guard = B->illum_mode == 1;
    for (size_t i = 0; i < n; ++i)
        if (!guard || hl[i] != 0.0f) hi[i] /= hl[i];
    laplace_adj(b.nz, b.nx, 1, hi.data(), img);
    return VTI_OK;
}

int vti_batch_record(vti_batch *B, int shot, float *dst)
{
    if (!B || shot < 0 || shot >= B->nshot || !dst) return bfail(VTI_ERR_ARG, "vti_batch_record: bad arguments");
    const BShot &b = B->hb[shot];
    std::vector<float> t((size_t)b.nrec * B->nt);
    if (hipStreamSynchronize(B->stream) != hipSuccess ||
        hipMemcpy(t.data(), b.rec, t.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(VTI_ERR_HIP, "vti_batch_record: copy");
    for (int r = 0; r < b.nrec; ++r)
        for (int it = 0; it < B->nt; ++it) dst[(size_t)r * B->nt + it] = t[(size_t)it * b.nrec + r];
    return VTI_OK;
}

}  // extern "C"
