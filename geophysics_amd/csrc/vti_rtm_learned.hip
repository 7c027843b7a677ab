// vti_rtm_learned.hip — 3D RTM with the source wavefield kept compressed
// instead of being recomputed from checkpoints: as autoencoder latents
// (codec 0), or as error-bounded coefficients of the fitted block basis
// (codec 1, vti_bcodec.h: the largest coefficients at 15 significant bits;
// codec 2, vti_bqcodec.h: every coefficient quantized with a per-block
// step; both keep every stored snapshot within relative L2 error tol, the
// source box exact, and recompute snapshots past the byte budget from
// checkpoints).  The rest of this comment is codec 0.
//
// First sweep (the forward modelling of the shot, with record and source
// illumination exactly as vti_rtm_shot): after every step it with
// it % save_every == 0 the interior P box (n^3, n a multiple of 8) is
// encoded into a latent slot (n^3/8 bf16 values + one scale).  Optionally
// every check_every-th stored snapshot is decoded at once and its relative
// L2 error against the true field measured, relative to the largest snapshot
// energy so far.  A box of half-width src_box around the source (the
// grid-scale near field of the point source and the pseudo-acoustic shear
// artefact live there) is kept exactly for every snapshot and masked out of
// what the autoencoder sees.  A snapshot above tol takes the fallback: it is kept exactly (uncompressed interior, up to max_exact of
// them), or, once those slots are used, recomputed in the reverse sweep
// from full-state checkpoints kept every ck_every steps (plain checkpoint
// recomputation).  Every other imaging step decodes its latent.  The receiver side is the same as in
// vti_rtm_shot; imaging steps are it % save_every == 0 (image_every =
// save_every there), so the two paths differ only by the compression error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "vti_ae.h"
#include "vti_internal.h"
#include "vti_nn.h"
#include "vti_rtm.h"

using namespace vti;

namespace {

// img[o] += S[o] * gP[c] over the interior (S dense, z fastest)
__global__ void image_dense_kernel(Grid g, int np, int nx, int ny, int nz, const float *__restrict__ S,
                                   const float *__restrict__ gP, float *img)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ix = blockIdx.y * blockDim.y + threadIdx.y, iy = blockIdx.z;
    if (iz >= nz || ix >= nx) return;
    const long long c = (iz + np) + (long long)(ix + np) * g.ldz + (long long)(iy + np) * g.ldxy;
    const long long o = iz + (long long)ix * nz + (long long)iy * nz * nx;
    img[o] += S[o] * gP[c];
}

// out = the interior n^3 box of P (dense, z fastest)
__global__ void interior_kernel(Grid g, int np, int n, const float *__restrict__ P, float *out)
{
    const long long tot = (long long)n * n * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % n);
        const long long r = i / n;
        const int x = (int)(r % n), y = (int)(r / n);
        out[i] = P[(z + np) + (long long)(x + np) * g.ldz + (long long)(y + np) * g.ldxy];
    }
}

// Box of the interior around the source, [y0, y0+by) x [x0, x0+bx) x [z0, z0+bz).
struct Box {
    int y0, x0, z0, by, bx, bz;
    __host__ __device__ bool inside(int y, int x, int z) const
    {
        return y >= y0 && y < y0 + by && x >= x0 && x < x0 + bx && z >= z0 && z < z0 + bz;
    }
    __host__ __device__ long long size() const { return (long long)by * bx * bz; }
};

// D = the interior n^3 box of P with the source box zeroed (dense, z fastest)
__global__ void masked_interior_kernel(Grid g, int np, int n, const float *__restrict__ P, Box b, float *D)
{
    const long long tot = (long long)n * n * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % n);
        const long long r = i / n;
        const int x = (int)(r % n), y = (int)(r / n);
        D[i] = b.inside(y, x, z) ? 0.0f : P[(z + np) + (long long)(x + np) * g.ldz + (long long)(y + np) * g.ldxy];
    }
}

// out = P over the source box (exact copy kept beside the latent)
__global__ void box_save_kernel(Grid g, int np, const float *__restrict__ P, Box b, float *out)
{
    const long long tot = b.size();
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % b.bz);
        const long long r = i / b.bz;
        const int x = (int)(r % b.bx), y = (int)(r / b.bx);
        out[i] = P[(b.z0 + z + np) + (long long)(b.x0 + x + np) * g.ldz + (long long)(b.y0 + y + np) * g.ldxy];
    }
}

// D (dense n^3) over the source box = the kept exact values
// box of a dense n^3 volume (z fastest) -> out (box layout, z fastest)
__global__ void box_save_dense_kernel(int n, const float *__restrict__ src, Box b, float *out)
{
    const long long tot = b.size();
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % b.bz);
        const long long r = i / b.bz;
        const int x = (int)(r % b.bx), y = (int)(r / b.bx);
        out[i] = src[(b.z0 + z) + (long long)(b.x0 + x) * n + (long long)(b.y0 + y) * n * n];
    }
}

__global__ void box_restore_kernel(int n, Box b, const float *__restrict__ in, float *D)
{
    const long long tot = b.size();
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % b.bz);
        const long long r = i / b.bz;
        const int x = (int)(r % b.bx), y = (int)(r / b.bx);
        D[(b.z0 + z) + (long long)(b.x0 + x) * n + (long long)(b.y0 + y) * n * n] = in[i];
    }
}

// sums of (S - P)^2 and P^2 over the interior (double atomics on two cells)
__global__ void err_kernel(Grid g, int np, int n, const float *__restrict__ S, const float *__restrict__ P,
                           double *acc)
{
    __shared__ double r0[256], r1[256];
    double a = 0.0, b = 0.0;
    const long long tot = (long long)n * n * n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % n);
        const long long r = i / n;
        const int x = (int)(r % n), y = (int)(r / n);
        const float p = P[(z + np) + (long long)(x + np) * g.ldz + (long long)(y + np) * g.ldxy];
        const double d = (double)S[i] - (double)p;
        a += d * d;
        b += (double)p * p;
    }
    r0[threadIdx.x] = a;
    r1[threadIdx.x] = b;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(acc, r0[0]);
        atomicAdd(acc + 1, r1[0]);
    }
}

struct Buf {
    void *p = nullptr;
    ~Buf() { (void)hipFree(p); }
};

// GPU time of bracketed stream sections, summed without stalling the
// stream: a ring of event pairs, each pair read when its slot is reused
// (long complete by then) or by total() after the stream is synchronised.
struct EvRing {
    static constexpr int K = 8;
    hipEvent_t a[K] = {}, b[K] = {};
    bool live[K] = {};
    int k = 0;
    double ms = 0.0;
    bool ok = true;
    EvRing()
    {
        for (int i = 0; i < K; ++i)
            if (hipEventCreate(&a[i]) != hipSuccess || hipEventCreate(&b[i]) != hipSuccess) ok = false;
    }
    ~EvRing()
    {
        for (int i = 0; i < K; ++i) {
            if (a[i]) (void)hipEventDestroy(a[i]);
            if (b[i]) (void)hipEventDestroy(b[i]);
        }
    }
    void take(int i)
    {
        float t = 0.0f;
        if (live[i] && hipEventSynchronize(b[i]) == hipSuccess && hipEventElapsedTime(&t, a[i], b[i]) == hipSuccess)
            ms += t;
        live[i] = false;
    }
    hipError_t begin(hipStream_t s)
    {
        take(k);
        return hipEventRecord(a[k], s);
    }
    hipError_t end(hipStream_t s)
    {
        live[k] = true;
        const hipError_t e = hipEventRecord(b[k], s);
        k = (k + 1) % K;
        return e;
    }
    double total()
    {
        for (int i = 0; i < K; ++i) take(i);
        return ms;
    }
};

// A stream of the codec's own (default priority) with the events that order
// it against the shot's stream, and pinned host slots for the encodes'
// counters.  The destructor waits for the stream before anything it used is
// freed, so it is declared after the buffers it works on.
struct CodecStream {
    hipStream_t s = nullptr;
    hipEvent_t p_ready = nullptr;     // the shot stream has finished the snapshot's stress update
    hipEvent_t p_read = nullptr;      // the encode has read the snapshot
    hipEvent_t done[2] = {};          // encode counters copied (per slot) / decoded snapshot imaged ([0])
    bc::Counters *hc = nullptr;       // [2], pinned
    bool ok = false;
    CodecStream()
    {
        ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&p_ready, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&p_read, hipEventDisableTiming) == hipSuccess;
        for (int i = 0; i < 2 && ok; ++i) ok = hipEventCreateWithFlags(&done[i], hipEventDisableTiming) == hipSuccess;
        ok = ok && hipHostMalloc(reinterpret_cast<void **>(&hc), 2 * sizeof(bc::Counters), hipHostMallocDefault) == hipSuccess;
    }
    ~CodecStream()
    {
        if (s) (void)hipStreamSynchronize(s);
        for (hipEvent_t e : {p_ready, p_read, done[0], done[1]})
            if (e) (void)hipEventDestroy(e);
        if (hc) (void)hipHostFree(hc);
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace

extern "C" {

int vti_learned_defaults(vti_learned_opts *o)
{
    if (!o) return fail(VTI_ERR_ARG, "vti_learned_defaults: null");
    std::memset(o, 0, sizeof(*o));
    o->save_every = 4;
    o->check_every = 8;
    o->tol = 0.05f;
    o->ck_every = 0;
    o->max_exact = 0;
    o->src_box = 8;
    o->codec = 0;
    o->floor = 1e-8f;
    o->budget = 0;
    return VTI_OK;
}

int vti_rtm_shot_learned(vti_ctx *c, vti_ae *A, const vti_rtm_opts *o, const vti_learned_opts *lo, const float *obs,
                         double *stats)
{
    if (!c || !A || !o || !lo) return fail(VTI_ERR_ARG, "vti_rtm_shot_learned: null argument");
    if (!c->have_model || !c->have_shot) return fail(VTI_ERR_STATE, "vti_rtm_shot_learned: model and shot first");
    const vti_cfg &k = c->cfg;
    const int n = k.nx;
    if (k.ndim != 3 || k.ny != n || k.nz != n || n % 8)
        return fail(VTI_ERR_ARG, "vti_rtm_shot_learned: 3D cubic interiors with a side divisible by 8 only");
    if (c->nrec <= 0) return fail(VTI_ERR_STATE, "vti_rtm_shot_learned: the shot has no receivers");
    if (lo->save_every < 1 || lo->check_every < 0 || lo->ck_every < 0 || lo->max_exact < 0 || lo->src_box < 0 ||
        lo->codec < 0 || lo->codec > 2 || lo->budget < 0 || lo->serial_codec < 0 || lo->serial_codec > 2 ||
        (lo->codec >= 1 && (!(lo->tol >= 0.0f) || !(lo->floor >= 0.0f))))
        return fail(VTI_ERR_ARG, "vti_rtm_shot_learned: bad options");
    if (c->adcig_na > 0)
        return fail(VTI_ERR_STATE, "vti_rtm_shot_learned: angle gathers need the full source state "
                                   "(vti_rtm_shot); switch them off with vti_adcig_config(ctx, 0, 0, 0)");
    const bool sparse = lo->codec >= 1;   // codec 1 or 2: block streams in 32-bit words
    const bool quant = lo->codec == 2;
    int st;
    if (!c->img && (st = vti_rtm_reset(c))) return st;
    const int nt = k.nt, nrec = c->nrec, np = k.npml, se = lo->save_every;
    const int inj_z = o->inject_z >= 0 ? o->inject_z : c->shot.rec_z;
    if (inj_z < 4 || inj_z >= c->g.nnz - 4) return fail(VTI_ERR_ARG, "vti_rtm_shot_learned: injection row outside the updated region");
    const long long lat_n = vti_ae_latent_elems(n);
    const int nsave = (nt + se - 1) / se;
    const long long off = np + (long long)np * c->g.ldz + (long long)np * c->g.ldxy;
    hipStream_t s = c->stream;
    Buf lat, scl, dense, dense2, acc, exa, boxes, boffs, bcnts, snapb;
    if ((!sparse && (hipMalloc(&lat.p, (size_t)nsave * lat_n * sizeof(uint16_t)) != hipSuccess ||
                     hipMalloc(&scl.p, (size_t)nsave * 64) != hipSuccess)) ||
        (!sparse && hipMalloc(&dense.p, (size_t)n * n * n * sizeof(float)) != hipSuccess) ||
        hipMalloc(&acc.p, 2 * sizeof(double)) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "vti_rtm_shot_learned: latent store");
    // codec 1 / 2 store: words (entries, or quantized block streams; 4 bytes
    // each) in chunks allocated as the snapshots need them, at most `budget`
    // bytes in all; block tables per snapshot.  A chunk is taken when the
    // current one has room for less than one snapshot's largest possible word
    // count (n^3 entries; (n/8)^3 x bq::kMaxBlockWords stream words), so an
    // encode never runs out of room inside a chunk and the store holds what
    // the words use plus less than one chunk.
    const long long nblk = (long long)(n / 8) * (n / 8) * (n / 8);
    const long long budget = lo->budget > 0 ? lo->budget : (long long)nsave * n * n * n;
    const long long snap_max = quant ? nblk * bq::kMaxBlockWords : (long long)n * n * n;   // words of one snapshot at most
    const long long chunk_ents = std::max(snap_max, (4LL << 30) / 4);   // 4 GB chunks
    if (sparse && (hipMalloc(&boffs.p, (size_t)nsave * nblk * 4) != hipSuccess ||
                   hipMalloc(&bcnts.p, (size_t)nsave * nblk * 2) != hipSuccess))
        return fail(VTI_ERR_ALLOC, "vti_rtm_shot_learned: sparse store");
    struct Chunks {
        std::vector<void *> p;
        ~Chunks() { for (void *q : p) (void)hipFree(q); }
    } chunks;
    long long chunk_room = 0, alloc_ents = 0;   // room left in the current chunk, entries allocated
    uint32_t *cur = nullptr;                    // next free entry of the current chunk
    uint32_t *BOFF = static_cast<uint32_t *>(boffs.p);
    uint16_t *BCNT = static_cast<uint16_t *>(bcnts.p);
    std::vector<uint32_t *> ebase(nsave, nullptr);
    long long used = 0;
    uint16_t *L = static_cast<uint16_t *>(lat.p);
    float *Sc = static_cast<float *>(scl.p);
    float *D = static_cast<float *>(dense.p);
    double *E = static_cast<double *>(acc.p);
    const size_t nn3 = (size_t)n * n * n;
    if (lo->max_exact > 0 && hipMalloc(&exa.p, (size_t)lo->max_exact * nn3 * sizeof(float)) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "vti_rtm_shot_learned: exact fallback store");
    float *X = static_cast<float *>(exa.p);
    // exact box around the source (interior coordinates), kept for every snapshot
    Box bx{0, 0, 0, 0, 0, 0};
    if (lo->src_box > 0) {
        const vti_shot &sh = c->shot;
        auto span = [&](int centre, int &lo0, int &len) {
            lo0 = std::max(0, centre - lo->src_box);
            len = std::min(n, centre + lo->src_box) - lo0;
        };
        span(sh.sy - np, bx.y0, bx.by);
        span(sh.sx - np, bx.x0, bx.bx);
        span(sh.sz - np, bx.z0, bx.bz);
        if (bx.by <= 0 || bx.bx <= 0 || bx.bz <= 0) bx = Box{0, 0, 0, 0, 0, 0};
    }
    const long long box_n = bx.size();
    if ((box_n && hipMalloc(&boxes.p, (size_t)nsave * box_n * sizeof(float)) != hipSuccess) ||
        (!sparse && hipMalloc(&dense2.p, nn3 * sizeof(float)) != hipSuccess))
        return fail(VTI_ERR_ALLOC, "vti_rtm_shot_learned: source boxes");
    float *BX = static_cast<float *>(boxes.p);
    float *D2 = static_cast<float *>(dense2.p);
    double emax = 0.0;   // largest snapshot energy so far: the fallback's reference
    std::vector<int> exact(nsave, -1);
    int nexact = 0;
    // fallback checkpoints
    const int nck = lo->ck_every > 0 ? (nt + lo->ck_every - 1) / lo->ck_every : 0;
    std::vector<VtiState> ck(nck);
    struct CkGuard {
        std::vector<VtiState> &v;
        ~CkGuard() { for (auto &x : v) state_free(x); }
    } ckg{ck};
    for (auto &x : ck)
        if ((st = state_alloc(c, x))) return st;
    VtiState G;
    if ((st = state_alloc(c, G))) return st;
    struct GGuard {
        VtiState &g;
        ~GGuard() { state_free(g); }
    } gg{G};
    const VtiState W = main_state(c);
    const bool svf = o->sv_forward != 0 && k.sv != 0;
    const bool svb = o->sv_backward != 0 && k.sv != 0;
    const int clamp = o->illum_mode == 0;
    auto hip = [](hipError_t e, const char *w) { return e == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, w); };
    std::vector<char> bad(nsave, 0);
    EvRing t_enc, t_dec;   // the codec's GPU time (encode: up to its counters' copy to the host)
    if (!t_enc.ok || !t_dec.ok) return fail(VTI_ERR_HIP, "vti_rtm_shot_learned: events");
    double err_sum = 0.0;
    int nchecked = 0, nbad = 0;
    // Codecs 1, 2: the codec runs on a stream of its own.  The stress update
    // of a snapshot step also writes the interior P into the dense copy SN
    // (the source side forward, the receiver side in reverse); the codec
    // works from the copy while the steps go on: the encode (its counters
    // read on the host one snapshot later, where the next encode's words
    // start), and in reverse the decode fused with the image.  The next
    // snapshot step's stress update, the copy's next writer, waits for that
    // work (4 steps earlier).  serial_codec puts the codec on the shot's
    // stream between the steps, reading P itself (the same results).
    const bool side = sparse && lo->serial_codec == 0;                          // encodes
    const bool side_dec = sparse && (lo->serial_codec == 0 || lo->serial_codec == 2);   // decodes
    CodecStream cs;
    if (sparse && !cs.ok) return fail(VTI_ERR_HIP, "vti_rtm_shot_learned: codec stream");
    if ((side || side_dec) && hipMalloc(&snapb.p, nn3 * sizeof(float)) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "vti_rtm_shot_learned: snapshot copy");
    float *SN = static_cast<float *>(snapb.p);
    hipStream_t es = side ? cs.s : s;        // the encodes' stream
    hipStream_t ds = side_dec ? cs.s : s;    // the decodes' stream
    bc::Counters *hcs = cs.hc;
    int pend = -1;                      // snapshot whose encode has not been settled on the host
    long long pend_cap = 0;
    bool wait_read = false;             // the next snapshot step's stress update waits for an encode's projection
    auto settle = [&]() -> int {        // the host side of the pending encode: place its words
        if (pend < 0) return VTI_OK;
        const int j = pend;
        pend = -1;
        if (hipEventSynchronize(cs.done[j & 1]) != hipSuccess) return fail(VTI_ERR_HIP, "vti_rtm_shot_learned: encode");
        long long need = 0;
        double e = 0.0;
        bc::encode_finish(hcs[j & 1], pend_cap, &need, &e, nullptr);
        if (need > pend_cap) {   // over budget: recompute this snapshot from a checkpoint
            if (!nck) return fail(VTI_ERR_ALLOC, "vti_rtm_shot_learned: sparse store budget exhausted (set ck_every)");
            bad[j] = 1;
            ++nbad;
        } else {
            ebase[j] = cur;
            cur += need;
            chunk_room -= need;
            used += need;
            err_sum += e;
            ++nchecked;
        }
        return VTI_OK;
    };

    // ---- first sweep
    if ((st = hip(state_zero(c, W), "zero"))) return st;
    for (int it = 0; it < nt; ++it) {
        if (nck && it % lo->ck_every == 0 && (st = hip(state_copy(c, ck[it / lo->ck_every], W), "checkpoint")))
            return st;
        if (launch_vel(3, vel_args_for(c, W, it, true), s) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
        float *row = c->rec + (size_t)it * nrec;
        StressArgs sa = with_illum(c, src_stress_args(c, W, it, svf, row), clamp);
        if (side && it % se == 0) {
            if (wait_read) {
                if ((st = hip(hipStreamWaitEvent(s, cs.p_read, 0), "wait"))) return st;
                wait_read = false;
            }
            sa = with_snap(c, sa, SN);
        }
        if (launch_stress(3, sa, s) != hipSuccess) return fail(VTI_ERR_HIP, "stress");
        if (it % se) continue;
        const int j = it / se;
        const float *P = W.f[VTI_F_P];
        if (sparse) {   // error-bounded sparse coefficients; the source box exactly
            if ((st = settle())) return st;
            if (side && ((st = hip(hipEventRecord(cs.p_ready, s), "event")) ||
                         (st = hip(hipStreamWaitEvent(es, cs.p_ready, 0), "wait"))))
                return st;
            if ((st = hip(t_enc.begin(es), "event"))) return st;
            if (box_n && side)
                hipLaunchKernelGGL(box_save_dense_kernel, dim3(64), dim3(256), 0, es, n, SN, bx, BX + (size_t)j * box_n);
            else if (box_n)
                hipLaunchKernelGGL(box_save_kernel, dim3(64), dim3(256), 0, es, c->g, np, P, bx, BX + (size_t)j * box_n);
            const bc::Mask m{bx.y0, bx.x0, bx.z0, bx.by, bx.bx, bx.bz};
            const float *src = side ? SN : P;   // the box to encode: the copy, or P itself
            const long long so = side ? 0 : off, ssy = side ? (long long)n * n : c->g.ldxy, ssx = side ? n : c->g.ldz;
            if (chunk_room < snap_max) {   // a new chunk: 4 GB, or what the budget has left
                const long long take = std::min(chunk_ents, budget / 4 - alloc_ents);
                void *q = nullptr;
                if (take > chunk_room) {
                    if (hipMalloc(&q, (size_t)take * 4) == hipSuccess) {
                        chunks.p.push_back(q);
                        cur = static_cast<uint32_t *>(q);
                        chunk_room = take;
                        alloc_ents += take;
                    } else {
                        (void)hipGetLastError();   // no room on the device: the snapshots past it fall back
                    }
                }
            }
            hipEvent_t rd = side ? cs.p_read : nullptr;
            st = quant ? ae_bq_encode_start(A, src, so, ssy, ssx, n, m, lo->tol, lo->floor, cur, chunk_room,
                                            BOFF + (size_t)j * nblk, BCNT + (size_t)j * nblk, rd, hcs + (j & 1), es)
                       : ae_bc_encode_start(A, src, so, ssy, ssx, n, m, lo->tol, lo->floor, cur, chunk_room,
                                            BOFF + (size_t)j * nblk, BCNT + (size_t)j * nblk, rd, hcs + (j & 1), es);
            if (st) return st;
            if ((st = hip(t_enc.end(es), "event"))) return st;
            if ((st = hip(hipEventRecord(cs.done[j & 1], es), "event"))) return st;
            pend = j;
            pend_cap = chunk_room;
            wait_read = side;
            continue;
        }
        if ((st = hip(t_enc.begin(s), "event"))) return st;
        if (box_n) {   // the source box is kept exactly; the autoencoder sees the rest
            hipLaunchKernelGGL(box_save_kernel, dim3(64), dim3(256), 0, s, c->g, np, P, bx, BX + (size_t)j * box_n);
            hipLaunchKernelGGL(masked_interior_kernel, dim3(2048), dim3(256), 0, s, c->g, np, n, P, bx, D);
            st = ae_encode(A, D, 0, (long long)n * n, n, n, L + (size_t)j * lat_n, Sc + (size_t)j * 16, s);
        } else {
            st = ae_encode(A, P, off, c->g.ldxy, c->g.ldz, n, L + (size_t)j * lat_n, Sc + (size_t)j * 16, s);
        }
        if (st) return st;
        if ((st = hip(t_enc.end(s), "event"))) return st;
        if (lo->check_every > 0 && j % lo->check_every == 0) {
            if ((st = ae_decode(A, L + (size_t)j * lat_n, Sc + (size_t)j * 16, n, D2, s))) return st;
            if (box_n)
                hipLaunchKernelGGL(box_restore_kernel, dim3(64), dim3(256), 0, s, n, bx, BX + (size_t)j * box_n, D2);
            if ((st = hip(nn::launch_fill_f32(reinterpret_cast<float *>(E), 4, 0.0f, s), "fill"))) return st;
            hipLaunchKernelGGL(err_kernel, dim3(1024), dim3(256), 0, s, c->g, np, n, D2, P, E);
            double h[2];
            if (hipMemcpyAsync(h, E, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
                return fail(VTI_ERR_HIP, "error check");
            emax = std::max(emax, h[1]);
            if (h[1] > 0.0) {
                const double e = std::sqrt(h[0] / h[1]);
                err_sum += e;
                ++nchecked;
                // the error against the largest snapshot energy so far: weak
                // snapshots weigh little in the image whatever their relative error
                if (std::sqrt(h[0] / emax) > lo->tol) {
                    ++nbad;
                    if (nexact < lo->max_exact && lo->check_every == 1) {   // keep this snapshot exactly
                        hipLaunchKernelGGL(interior_kernel, dim3(2048), dim3(256), 0, s, c->g, np, n, P, X + (size_t)nexact * nn3);
                        exact[j] = nexact++;
                    } else if (nck) {   // this snapshot and the ones up to the next check are recomputed
                        for (int q = j; q < std::min(nsave, j + lo->check_every); ++q) bad[q] = 1;
                    }
                }
            }
        }
    }
    if ((st = settle())) return st;
    // ---- the record: observed or modelled + mute (as vti_rtm_shot)
    if (obs) {
        if (hipMemcpyAsync(c->rec_t, obs, (size_t)nrec * nt * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess)
            return fail(VTI_ERR_HIP, "obs upload");
        if ((st = hip(launch_transpose(c->rec_t, c->rec, nrec, nt, s), "transpose"))) return st;
    } else if (o->mute_tt >= 0) {
        const vti_shot &sh = c->shot;
        if ((st = vti_mute3d_window(c, sh.sx - np + 1, 0, sh.sy - np + 1, 0, sh.sz - np + 1, 0, o->mute_tt))) return st;
    }
    // ---- reverse sweep
    // codecs 1, 2: snapshot j decoded straight into the image (times the
    // receiver field of its step, the exact box included); overlapped, from
    // the copy of the receiver P the step's stress update wrote, while the
    // receiver side steps on.  Any later write of the image or of the copy
    // is on the shot's stream after the next snapshot step's wait for it.
    bool wait_img = false;
    if ((st = hip(state_zero(c, G), "zero"))) return st;
    for (int it = nt - 1; it >= 0; --it) {
        if (launch_inject_rec(c, G, c->rec + (size_t)it * nrec, inj_z) != hipSuccess) return fail(VTI_ERR_HIP, "inject");
        if (launch_vel(3, vel_args_for(c, G, it, false), s) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
        StressArgs ra = rcv_stress_args(c, G, it, svb, inj_z);
        if (o->illum_mode == 1) ra = with_illum(c, ra, 0);   // receiver illumination in the epilogue
        const int j = it / se;
        const bool coded = sparse && it % se == 0 && exact[j] < 0 && !bad[j];
        if (side_dec && it % se == 0) {
            if (wait_img) {
                if ((st = hip(hipStreamWaitEvent(s, cs.done[0], 0), "wait"))) return st;
                wait_img = false;
            }
            if (coded) ra = with_snap(c, ra, SN);
        }
        if (launch_stress(3, ra, s) != hipSuccess) return fail(VTI_ERR_HIP, "stress");
        if (it % se) continue;
        const dim3 blk(64, 4), grd((k.nz + 63) / 64, (k.nx + 3) / 4, k.ny);
        if (exact[j] >= 0) {
            hipLaunchKernelGGL(image_dense_kernel, grd, blk, 0, s, c->g, np, k.nx, k.ny, k.nz, X + (size_t)exact[j] * nn3,
                               G.f[VTI_F_P], c->img);
        } else if (bad[j]) {   // exact source field from the nearest checkpoint at or before it+1
            const int target = it + 1;
            const int ci = std::min((target - 1) / lo->ck_every, nck - 1);
            if ((st = hip(state_copy(c, W, ck[ci]), "restore"))) return st;
            for (int work = ci * lo->ck_every; work < target; ++work) {
                if (launch_vel(3, vel_args_for(c, W, work, true), s) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
                if (launch_stress(3, src_stress_args(c, W, work, svf, nullptr), s) != hipSuccess)
                    return fail(VTI_ERR_HIP, "stress");
            }
            if ((st = hip(launch_image(c, W.f[VTI_F_P], G.f[VTI_F_P], c->img), "image"))) return st;
        } else if (coded) {
            bc::TileSink sk;
            sk.img = c->img;
            sk.rp = side_dec ? SN : G.f[VTI_F_P];
            sk.off = side_dec ? 0 : off;
            sk.sy = side_dec ? (long long)n * n : c->g.ldxy;
            sk.sx = side_dec ? n : c->g.ldz;
            if (box_n) {
                sk.box = BX + (size_t)j * box_n;
                sk.y0 = bx.y0, sk.x0 = bx.x0, sk.z0 = bx.z0, sk.by = bx.by, sk.bx = bx.bx, sk.bz = bx.bz;
            }
            if (side_dec && ((st = hip(hipEventRecord(cs.p_ready, s), "event")) ||
                             (st = hip(hipStreamWaitEvent(ds, cs.p_ready, 0), "wait"))))
                return st;
            if ((st = hip(t_dec.begin(ds), "event"))) return st;
            st = quant ? ae_bq_decode_sink(A, ebase[j], BOFF + (size_t)j * nblk, BCNT + (size_t)j * nblk, n, sk, ds)
                       : ae_bc_decode_sink(A, ebase[j], BOFF + (size_t)j * nblk, BCNT + (size_t)j * nblk, n, sk, ds);
            if (st) return st;
            if ((st = hip(t_dec.end(ds), "event"))) return st;
            if (side_dec) {
                if ((st = hip(hipEventRecord(cs.done[0], ds), "event"))) return st;
                wait_img = true;
            }
        } else {   // latents: decoded into the dense buffer, then imaged
            if ((st = hip(t_dec.begin(s), "event"))) return st;
            if ((st = ae_decode(A, L + (size_t)j * lat_n, Sc + (size_t)j * 16, n, D, s))) return st;
            if (box_n)
                hipLaunchKernelGGL(box_restore_kernel, dim3(64), dim3(256), 0, s, n, bx, BX + (size_t)j * box_n, D);
            if ((st = hip(t_dec.end(s), "event"))) return st;
            hipLaunchKernelGGL(image_dense_kernel, grd, blk, 0, s, c->g, np, k.nx, k.ny, k.nz, D, G.f[VTI_F_P], c->img);
        }
        if ((st = hip(hipGetLastError(), "image"))) return st;
    }
    c->rtm_illum_mode = o->illum_mode;
    if ((st = hip(hipStreamSynchronize(s), "sync")) || (st = hip(hipStreamSynchronize(cs.s ? cs.s : s), "sync")))
        return st;
    if (stats) {
        stats[0] = t_enc.total();
        stats[1] = t_dec.total();
        stats[2] = nchecked ? err_sum / nchecked : -1.0;
        stats[3] = nbad;
        stats[4] = nsave;
        stats[5] = nexact;
        int nrec_snap = 0;
        for (char b : bad) nrec_snap += b;
        stats[6] = sparse ? (double)used * 4 + (double)nsave * nblk * 6 + (double)nsave * box_n * 4
                          : (double)nsave * lat_n * 2 + (double)nexact * nn3 * 4 + (double)nsave * box_n * 4;
        stats[7] = nrec_snap;
        // device memory held for the stored source field, as allocated: the
        // store (entry chunks and block tables, or latents, scales and exact
        // slots), the source boxes, the work buffers (the codec's
        // coefficients and search results and the snapshot copy, or the
        // latents' two dense n^3 buffers) and the fallback checkpoints
        const double ckb = (double)nck * (double)vti_state_bytes(c);
        stats[8] = sparse ? (double)alloc_ents * 4 + (double)nsave * nblk * 6 + (double)nsave * box_n * 4 +
                                (double)ae_codec_work_bytes(A) + (SN ? (double)nn3 * 4 : 0.0) + ckb
                          : (double)nsave * lat_n * 2 + (double)nsave * 64 + (double)lo->max_exact * nn3 * 4 +
                                (double)nsave * box_n * 4 + 2.0 * nn3 * 4 + ckb;
    }
    return VTI_OK;
}

// Forward modelling of the current shot (no record kept); every
// save_every-th snapshot, the source box of half-width src_box masked out,
// contributes nblocks sampled 8^3 blocks to the autoencoder's
// principal-component fit (finished by vti_ae_fit_end).
int vti_ae_fit_shot(vti_ctx *c, vti_ae *A, int save_every, int nblocks, int src_box)
{
    if (!c || !A || save_every < 1 || nblocks < 1) return fail(VTI_ERR_ARG, "vti_ae_fit_shot: bad arguments");
    if (!c->have_model || !c->have_shot) return fail(VTI_ERR_STATE, "vti_ae_fit_shot: model and shot first");
    const vti_cfg &k = c->cfg;
    if (k.ndim != 3 || k.ny != k.nx || k.nz != k.nx || k.nx % 8)
        return fail(VTI_ERR_ARG, "vti_ae_fit_shot: 3D cubic interiors with a side divisible by 8 only");
    if (src_box < 0) return fail(VTI_ERR_ARG, "vti_ae_fit_shot: src_box >= 0");
    const int np = k.npml, n = k.nx;
    const long long off = np + (long long)np * c->g.ldz + (long long)np * c->g.ldxy;
    // the source box is masked out, as in vti_rtm_shot_learned
    Box bx{0, 0, 0, 0, 0, 0};
    if (src_box > 0) {
        const vti_shot &sh = c->shot;
        auto span = [&](int centre, int &lo0, int &len) {
            lo0 = std::max(0, centre - src_box);
            len = std::min(n, centre + src_box) - lo0;
        };
        span(sh.sy - np, bx.y0, bx.by);
        span(sh.sx - np, bx.x0, bx.bx);
        span(sh.sz - np, bx.z0, bx.bz);
        if (bx.by <= 0 || bx.bx <= 0 || bx.bz <= 0) bx = Box{0, 0, 0, 0, 0, 0};
    }
    Buf dense;
    if (bx.size() && hipMalloc(&dense.p, (size_t)n * n * n * sizeof(float)) != hipSuccess)
        return fail(VTI_ERR_ALLOC, "vti_ae_fit_shot: masked field");
    float *D = static_cast<float *>(dense.p);
    const VtiState W = main_state(c);
    if (state_zero(c, W) != hipSuccess) return fail(VTI_ERR_HIP, "zero");
    for (int it = 0; it < k.nt; ++it) {
        if (launch_vel(3, vel_args_for(c, W, it, true), c->stream) != hipSuccess ||
            launch_stress(3, src_stress_args(c, W, it, k.sv != 0, nullptr), c->stream) != hipSuccess)
            return fail(VTI_ERR_HIP, "step");
        if (it % save_every) continue;
        if (bx.size())
            hipLaunchKernelGGL(masked_interior_kernel, dim3(2048), dim3(256), 0, c->stream, c->g, np, n, W.f[VTI_F_P], bx, D);
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(VTI_ERR_HIP, "sync");   // the fit runs on its own stream
        const int st = bx.size() ? ae_fit_field(A, D, 0, (long long)n * n, n, n, nblocks)
                                 : ae_fit_field(A, W.f[VTI_F_P], off, c->g.ldxy, c->g.ldz, n, nblocks);
        if (st) return st;
    }
    return hipStreamSynchronize(c->stream) == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, "sync");
}

// Forward modelling of the current shot (no record kept) with nsteps Adam
// steps of the autoencoder on every save_every-th snapshot.
int vti_ae_train_shot(vti_ctx *c, vti_ae *A, int save_every, int nsteps, float *mean_loss)
{
    if (!c || !A || save_every < 1 || nsteps < 0) return fail(VTI_ERR_ARG, "vti_ae_train_shot: bad arguments");
    if (!c->have_model || !c->have_shot) return fail(VTI_ERR_STATE, "vti_ae_train_shot: model and shot first");
    const vti_cfg &k = c->cfg;
    if (k.ndim != 3) return fail(VTI_ERR_ARG, "vti_ae_train_shot: 3D only");
    const int np = k.npml;
    const long long off = np + (long long)np * c->g.ldz + (long long)np * c->g.ldxy;
    const VtiState W = main_state(c);
    if (state_zero(c, W) != hipSuccess) return fail(VTI_ERR_HIP, "zero");
    double acc = 0.0;
    int cnt = 0;
    for (int it = 0; it < k.nt; ++it) {
        if (launch_vel(3, vel_args_for(c, W, it, true), c->stream) != hipSuccess ||
            launch_stress(3, src_stress_args(c, W, it, k.sv != 0, nullptr), c->stream) != hipSuccess)
            return fail(VTI_ERR_HIP, "step");
        if (it % save_every) continue;
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(VTI_ERR_HIP, "sync");   // training runs on its own stream
        float l = 0.0f;
        const int st = ae_train_field(A, W.f[VTI_F_P], off, c->g.ldxy, c->g.ldz, k.ny, k.nx, k.nz, nsteps, &l);
        if (st) return st;
        acc += l;
        ++cnt;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(VTI_ERR_HIP, "sync");
    if (mean_loss) *mean_loss = cnt ? (float)(acc / cnt) : 0.0f;
    return VTI_OK;
}

}  // extern "C"
