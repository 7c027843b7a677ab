// vti_bqcodec.hip — quantized sparse block-transform codec (codec 2 of the
// learned 3D RTM; gfx950).
//
// The projection on the fitted orthonormal 512 x 512 block basis is the
// sparse codec's (vti_bcodec.hip encode_kernel, MFMA).  Instead of keeping
// the largest coefficients at 15 significant bits each with a 9-bit index
// (codec 1: 32 bits per kept coefficient), every 8^3 block is quantized
// with its own uniform step: q_j = rint(c_j (1/d)), dead zone included (a
// coefficient below d/2 becomes 0).  One wavefront per block bisects on the
// float bits of d, on a float estimate of the quantization error, down to
// 1/64 of a binade: a step within ~1 % of the largest whose quantization
// error sum_j (c_j - d q_j)^2 fits the block's budget — the budget of
// codec 1, tol^2 x max(block energy, floor x the strongest block's
// energy).  The chosen step's error is then computed exactly (q as stored,
// double sums) and the step lowered until it fits, so the error bound and
// its Parseval argument are codec 1's.  A block is stored
// as words of a bit stream:
//
//   word 0       d (fp32 bits)
//   word 1       widths of coefficient groups 0..5 (5 bits each) | mode << 30
//   word 2       widths of groups 6, 7 (5 bits each) | nonzero count << 10
//   words 3..18  bitmap of the nonzero q_j (bit j%32 of word 3 + j/32)
//   words 19..   the nonzero q_j in coefficient order, each |q| << 1 | sign
//                in its group's width (group g = coefficients 64g..64g+63,
//                width = bits of the group's largest |q| + 1), packed from
//                bit 0 of word 19 upward
//
// (mode 1: a block whose budget even the finest step, 2^-22 of its largest
// coefficient, does not meet keeps its nonzero coefficients as raw fp32
// words from word 19.)  A block within its budget at q = 0 has no words.
// The leading principal components carry most of the energy, so widths per
// group of 64 follow the coefficient sizes closely.  Encode after the
// projection: qsize_kernel (one wavefront per block) finds each block's
// step, header and stream length; qscan_kernel / qtop_kernel place the
// streams in block order by a two-level scan of the lengths (no
// same-address atomics, which serialise across the workgroups);
// qwrite_kernel re-quantizes each block with
// its step, assembles its words in LDS with LDS atomics (values may
// straddle words) and writes them out coalesced.  Decode: every lane finds
// its 8 coefficients' bit offsets with a wave prefix sum, dequantizes d q_j
// (an exact fp32 product, split into bf16 hi + lo) into the LDS tile and
// the tile goes through codec 1's MFMA reconstruction.
#include "vti_bqcodec.h"

#include <algorithm>
#include <cmath>

#include "../../include/vti.h"

namespace vti {
namespace bq {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int PW = 8;                 // blocks (wavefronts) per pack workgroup
constexpr int kHead = 19;             // header + bitmap words

__device__ __forceinline__ bf16x8 as_bf(us8 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ uint16_t to_bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float from_bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// wave-uniform reductions (lane 0's result broadcast)
__device__ __forceinline__ double wave_sum_d(double x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return __shfl(x, 0, 64);
}

__device__ __forceinline__ float wave_max(float x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x = fmaxf(x, __shfl_xor(x, d, 64));
    return __shfl(x, 0, 64);
}

__device__ __forceinline__ float wave_sum_f(float x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return __shfl(x, 0, 64);
}

// the step search's estimate of the quantization error at step d: q from
// the hardware reciprocal, float sums (a few 1e-5 relative; the chosen step
// is checked with qerr below)
__device__ __forceinline__ float qerr_est(const float v[8], float d)
{
    const float inv = __builtin_amdgcn_rcpf(d);
    float s = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float r = v[e] - d * rintf(v[e] * inv);
        s += r * r;
    }
    return wave_sum_f(s);
}

// quantization error of the lane's 8 coefficients at step d, wave-summed:
// q = rint(c * (1/d)) (one division per block) as the stream stores it, the
// error summed in double
__device__ __forceinline__ double qerr(const float v[8], float d)
{
    const float inv = 1.0f / d;
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float r = v[e] - d * rintf(v[e] * inv);
        s += (double)(r * r);
    }
    return wave_sum_d(s);
}

// Per-block results of the step search (struct of arrays in Work::aux) and
// the per-group sums of the placement scan (groups of kScanGroup blocks).
constexpr int kScanGroup = 1024;
struct Aux {
    float *d;           // step (0: raw or no words)
    uint32_t *h1, *h2;  // header words 1 and 2
    uint32_t *words;    // stream length (0: within its budget at q = 0)
    double *err, *eb;   // the block's squared error and energy
    uint32_t *gw, *gp;  // group word totals, group bases
    double *ge, *gd;    // group energy and dropped-energy sums
};

Aux aux_view(void *p, long long nblk)
{
    const long long ng = (nblk + kScanGroup - 1) / kScanGroup;
    char *c = static_cast<char *>(p);
    Aux a;
    a.d = reinterpret_cast<float *>(c);
    a.h1 = reinterpret_cast<uint32_t *>(c + 4 * nblk);
    a.h2 = reinterpret_cast<uint32_t *>(c + 8 * nblk);
    a.words = reinterpret_cast<uint32_t *>(c + 12 * nblk);
    a.err = reinterpret_cast<double *>(c + 16 * nblk);
    a.eb = reinterpret_cast<double *>(c + 24 * nblk);
    a.gw = reinterpret_cast<uint32_t *>(c + 32 * nblk);
    a.gp = reinterpret_cast<uint32_t *>(c + 32 * nblk + 4 * ng);
    a.ge = reinterpret_cast<double *>(c + 32 * nblk + 8 * ng);
    a.gd = reinterpret_cast<double *>(c + 32 * nblk + 16 * ng);
    return a;
}

long long aux_bytes(long long nblk)
{
    const long long ng = (nblk + kScanGroup - 1) / kScanGroup;
    return 32 * nblk + 24 * ng;
}

// The quantization of the lane's 8 coefficients at step d (mode 0) or as raw
// values (mode 1): q, the nonzero mask, the group's width (mode 0).
__device__ __forceinline__ void quantize(const float v[8], float d, int mode, int q[8], unsigned &mask, int &width)
{
    mask = 0u;
    width = 0;
    if (mode) {
#pragma unroll
        for (int e = 0; e < 8; ++e) mask |= (v[e] != 0.0f ? 1u : 0u) << e;
        return;
    }
    const float inv = 1.0f / d;
    int amq = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        q[e] = (int)rintf(v[e] * inv);
        mask |= (q[e] != 0 ? 1u : 0u) << e;
        amq = max(amq, abs(q[e]));
    }
    // the group of 8 lanes' largest |q|
#pragma unroll
    for (int s = 1; s < 8; s <<= 1) amq = max(amq, __shfl_xor(amq, s, 64));
    width = amq ? 33 - __clz(amq) : 0;   // magnitude bits + sign
}

// bits of the lane's values and their offset in the stream (wave prefix)
__device__ __forceinline__ unsigned lane_bits(unsigned mask, int width, int mode, int l, unsigned &total)
{
    const int nz = __popc(mask);
    const unsigned mine = mode ? 32u * nz : (unsigned)(nz * width);
    unsigned incl = mine;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const unsigned t = __shfl_up(incl, s, 64);
        if (l >= s) incl += t;
    }
    total = __shfl(incl, 63, 64);
    return incl - mine;
}

// Step search, one wavefront per block: the step (or raw mode), header
// words and stream length of every block, its squared error and energy.
__global__ __launch_bounds__(64 * PW) void qsize_kernel(const float *__restrict__ coef, long long nblk, double tol2,
                                                        float flo, Aux ax, const bc::Counters *cnt)
{
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long long b = (long long)blockIdx.x * PW + w;
    if (b >= nblk) return;   // wavefront-uniform
    float v[8];
    const float *c = coef + b * 512 + 8 * l;
    double e2 = 0.0;
    float amx = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        v[e] = c[e];
        e2 += (double)(v[e] * v[e]);
        amx = fmaxf(amx, fabsf(v[e]));
    }
    const double eblk = wave_sum_d(e2);
    const float mx = wave_max(amx);
    const double emax = (double)__uint_as_float(cnt->emax_bits);
    const double budget = tol2 * fmax(eblk, (double)flo * emax);
    double err = eblk;   // every coefficient quantizes to 0: no words
    float d = 0.0f;
    int mode = 0;
    unsigned words = 0u, h1 = 0u, h2 = 0u;
    if (eblk > budget) {
        const float dmin = fmaxf(mx * 0x1p-22f, 1.17549435e-38f);
        // at dmin every coefficient's error is within dmin (half a step and
        // the rounding of d q), so the exact check only runs for budgets
        // below 512 dmin^2 (tol below ~1e-5)
        if (512.0 * (double)dmin * (double)dmin > budget && qerr(v, dmin) > budget) {
            mode = 1;   // raw fp32 values
            err = 0.0;
        } else {
            // bisection on the float bits of d down to 1/64 of a binade
            // (steps within ~1 % of the largest admissible one) on the
            // estimate, then the exact error of the step chosen; a step the
            // estimate admitted wrongly is lowered by 1/64 binade (dmin
            // itself is admissible)
            unsigned lo = __float_as_uint(dmin), hi = __float_as_uint(mx * 4.0f);
            const float fb = (float)budget;
            while (hi - lo > (1u << 17)) {
                const unsigned mid = lo + (hi - lo) / 2u;
                if (qerr_est(v, __uint_as_float(mid)) <= fb) lo = mid;
                else hi = mid;
            }
            d = __uint_as_float(lo);
            err = qerr(v, d);
            while (err > budget && lo > __float_as_uint(dmin)) {
                lo = lo - (1u << 17) > __float_as_uint(dmin) ? lo - (1u << 17) : __float_as_uint(dmin);
                d = __uint_as_float(lo);
                err = qerr(v, d);
            }
        }
        int q[8];
        unsigned mask;
        int width;
        quantize(v, d, mode, q, mask, width);
        unsigned vbits;
        (void)lane_bits(mask, width, mode, l, vbits);
        int nzt = __popc(mask);
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) nzt += __shfl_xor(nzt, s, 64);
        words = (unsigned)kHead + (vbits + 31u) / 32u;
        for (int g = 0; g < 8; ++g) {   // every lane reads the widths of the 8 groups
            const unsigned wg = (unsigned)__shfl(width, 8 * g, 64);
            if (g < 6) h1 |= wg << (5 * g);
            else h2 |= wg << (5 * (g - 6));
        }
        h1 |= (unsigned)mode << 30;
        h2 |= (unsigned)nzt << 10;
    }
    if (l == 0) {
        ax.d[b] = d;
        ax.h1[b] = h1;
        ax.h2[b] = h2;
        ax.words[b] = words;
        ax.err[b] = err;
        ax.eb[b] = eblk;
    }
}

// Placement, no atomics: each group of kScanGroup blocks scans its stream
// lengths (boff = offset within the group), totals its words, energy and
// error; one workgroup then scans the group totals.
__global__ __launch_bounds__(1024) void qscan_kernel(Aux ax, long long nblk, uint32_t *boff)
{
    __shared__ unsigned ws[16];
    __shared__ double se[16], sd[16];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const long long b = (long long)blockIdx.x * kScanGroup + t;
    const unsigned x = b < nblk ? ax.words[b] : 0u;
    double e = b < nblk ? ax.eb[b] : 0.0, r = b < nblk ? ax.err[b] : 0.0;
    unsigned incl = x;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const unsigned u = __shfl_up(incl, s, 64);
        if (l >= s) incl += u;
    }
    e = wave_sum_d(e);
    r = wave_sum_d(r);
    if (l == 63) ws[w] = incl;
    if (l == 0) {
        se[w] = e;
        sd[w] = r;
    }
    __syncthreads();
    unsigned before = 0u;
    for (int k = 0; k < w; ++k) before += ws[k];
    if (b < nblk) boff[b] = before + incl - x;
    if (t == 0) {
        unsigned tot = 0u;
        double es = 0.0, ds = 0.0;
        for (int k = 0; k < 16; ++k) {
            tot += ws[k];
            es += se[k];
            ds += sd[k];
        }
        ax.gw[blockIdx.x] = tot;
        ax.ge[blockIdx.x] = es;
        ax.gd[blockIdx.x] = ds;
    }
}

__global__ __launch_bounds__(1024) void qtop_kernel(Aux ax, long long ng, unsigned long long cap, bc::Counters *cnt)
{
    __shared__ unsigned long long wsum[16];
    __shared__ double we[16], wd[16];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const long long per = (ng + 1023) / 1024, g0 = t * per, g1 = g0 + per < ng ? g0 + per : ng;
    unsigned long long sum = 0ull;
    double es = 0.0, ds = 0.0;
    for (long long g = g0; g < g1; ++g) {
        sum += ax.gw[g];
        es += ax.ge[g];
        ds += ax.gd[g];
    }
    // exclusive scan of the thread totals in thread order (exact integers),
    // energy and dropped energy summed in a fixed tree (deterministic)
    unsigned long long incl = sum;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const unsigned long long u = __shfl_up(incl, s, 64);
        if (l >= s) incl += u;
    }
    es = wave_sum_d(es);
    ds = wave_sum_d(ds);
    if (l == 63) wsum[w] = incl;
    if (l == 0) {
        we[w] = es;
        wd[w] = ds;
    }
    __syncthreads();
    unsigned long long before = 0ull;
    for (int k = 0; k < w; ++k) before += wsum[k];
    unsigned long long run = before + incl - sum;
    for (long long g = g0; g < g1; ++g) {
        ax.gp[g] = (uint32_t)run;
        run += ax.gw[g];
    }
    if (t == 0) {
        unsigned long long tot = 0ull;
        double et = 0.0, dt = 0.0;
        for (int k = 0; k < 16; ++k) {
            tot += wsum[k];
            et += we[k];
            dt += wd[k];
        }
        cnt->cursor = (unsigned)tot;   // < 2^32: (n/8)^3 x kMaxBlockWords is checked on the host
        cnt->overflow = tot > cap ? 1u : 0u;
        cnt->energy = et;
        cnt->dropped = dt;
    }
}

// The streams: one wavefront per block at its scanned offset (blocks whose
// stream would pass cap write nothing).
__global__ __launch_bounds__(64 * PW) void qwrite_kernel(const float *__restrict__ coef, long long nblk, Aux ax,
                                                         uint32_t *ent, unsigned long long cap, uint32_t *boff,
                                                         uint16_t *bcnt)
{
    __shared__ uint32_t buf[PW][kMaxBlockWords];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long long b = (long long)blockIdx.x * PW + w;
    if (b >= nblk) return;   // wavefront-uniform; no barrier below
    const unsigned words = ax.words[b];
    const unsigned base = ax.gp[b / kScanGroup] + boff[b];
    if (l == 0) {
        boff[b] = base;
        bcnt[b] = (uint16_t)words;
    }
    if (!words || (unsigned long long)base + words > cap) return;
    float v[8];
    const float *c = coef + b * 512 + 8 * l;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = c[e];
    const float d = ax.d[b];
    const uint32_t h1 = ax.h1[b], h2 = ax.h2[b];
    const int mode = (int)(h1 >> 30);
    int q[8];
    unsigned mask;
    int width;
    quantize(v, d, mode, q, mask, width);
    unsigned vbits;
    const unsigned off = lane_bits(mask, width, mode, l, vbits);
    uint32_t *bw = buf[w];
    for (unsigned i = l; i < words; i += 64) bw[i] = 0u;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    if (l == 0) {
        bw[0] = __float_as_uint(d);
        bw[1] = h1;
        bw[2] = h2;
    }
    atomicOr(&bw[3 + (l >> 2)], mask << (8 * (l & 3)));
    unsigned pos = (unsigned)kHead * 32u + off;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        if (!((mask >> e) & 1u)) continue;
        if (mode) {
            bw[pos >> 5] = __float_as_uint(v[e]);
            pos += 32u;
            continue;
        }
        const unsigned m = ((unsigned)abs(q[e]) << 1) | (q[e] < 0 ? 1u : 0u);
        const unsigned sh = pos & 31u;
        atomicOr(&bw[pos >> 5], m << sh);
        if (sh + (unsigned)width > 32u) atomicOr(&bw[(pos >> 5) + 1], m >> (32u - sh));
        pos += (unsigned)width;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (unsigned i = l; i < words; i += 64) ent[base + i] = bw[i];
}

// Decode with 64 blocks per workgroup of 8 wavefronts, as the encode: the basis streamed once per 64 blocks, each lane's 8
// dequantized coefficients written to the LDS tile as one 16-byte row
// piece, the basis operands loaded one K step ahead.  Wavefront w parses
// blocks w, w + 8, .. and computes blocks 32 (w & 1) .. + 32 against voxels
// 128 (w >> 1) .. + 128.
constexpr int DB = bc::kTileBlocks;   // blocks per workgroup
constexpr int DW = bc::kTileWaves;    // wavefronts per workgroup

__device__ __forceinline__ us8 pack8(const uint16_t h[8])
{
    us8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] = h[q];
    return r;
}

__global__ __launch_bounds__(64 * DW) void qdecode64_kernel(bc::Basis V, const uint32_t *__restrict__ ent,
                                                           const uint32_t *__restrict__ boff,
                                                           const uint16_t *__restrict__ bcnt, int n, bc::TileSink sk)
{
    __shared__ __align__(16) uint16_t lds[2][DB][bc::LROW];   // coefficient tiles (hi, lo), then the output tile
    uint16_t(*ch)[bc::LROW] = lds[0];
    uint16_t(*cl)[bc::LROW] = lds[1];
    const int nb = n / 8;
    const long long nblk = (long long)nb * nb * nb;
    const long long b0 = (long long)blockIdx.x * DB;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int i = w; i < DB; i += DW) {
        const long long b = b0 + i;
        const int nw = b < nblk ? bcnt[b] : 0;
        uint16_t hh[8] = {}, ll[8] = {};
        if (nw > 0) {
            const uint32_t *e = ent + boff[b];
            const float d = __uint_as_float(e[0]);
            const uint32_t h1 = e[1], h2 = e[2];
            const int mode = (int)(h1 >> 30);
            const int g = l >> 3;
            const unsigned wdt = g < 6 ? (h1 >> (5 * g)) & 31u : (h2 >> (5 * (g - 6))) & 31u;
            const unsigned m = (e[3 + (l >> 2)] >> (8 * (l & 3))) & 255u;
            const unsigned mine = mode ? 32u * __popc(m) : __popc(m) * wdt;
            unsigned incl = mine;
#pragma unroll
            for (int s = 1; s < 64; s <<= 1) {
                const unsigned t = __shfl_up(incl, s, 64);
                if (l >= s) incl += t;
            }
            unsigned pos = (unsigned)kHead * 32u + incl - mine;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (!((m >> k) & 1u)) continue;
                float f;
                if (mode) {
                    f = __uint_as_float(e[pos >> 5]);
                    pos += 32u;
                } else {
                    const unsigned sh = pos & 31u;
                    unsigned long long x = e[pos >> 5];
                    if (sh + wdt > 32u) x |= (unsigned long long)e[(pos >> 5) + 1] << 32;
                    const unsigned code = (unsigned)(x >> sh) & ((1u << wdt) - 1u);
                    const int mag = (int)(code >> 1);
                    f = (float)((code & 1u) ? -mag : mag) * d;
                    pos += wdt;
                }
                hh[k] = to_bf(f);
                ll[k] = to_bf(f - from_bf(hh[k]));
            }
        }
        // every lane writes its 8 coefficients (zeros included): the tile needs no clearing
        *reinterpret_cast<us8 *>(&ch[i][8 * l]) = pack8(hh);
        *reinterpret_cast<us8 *>(&cl[i][8 * l]) = pack8(ll);
    }
    __syncthreads();
    bc::tile_f4 acc[4][4];
    bc::tile_mfma(ch, cl, V.v_h, V.v_l, w, l, acc);
    __syncthreads();
    bc::tile_store(&lds[0][0][0], acc, w, l, b0, n, sk);
}

}  // namespace

int encode(const bc::Basis &V, const float *P, long long off, long long sy, long long sx, int n,
           const bc::Mask &m, float tol, float flo, bc::Work &w, uint32_t *ent, long long cap, uint32_t *boff,
           uint16_t *bcnt, long long *kept, double *err, hipStream_t s)
{
    bc::Counters h;
    const int st = encode_start(V, P, off, sy, sx, n, m, tol, flo, w, ent, cap, boff, bcnt, nullptr, &h, s);
    if (st) return st;
    if (hipStreamSynchronize(s) != hipSuccess) return VTI_ERR_HIP;
    bc::encode_finish(h, cap, kept, err, nullptr);
    return VTI_OK;
}

int encode_start(const bc::Basis &V, const float *P, long long off, long long sy, long long sx, int n,
                 const bc::Mask &m, float tol, float flo, bc::Work &w, uint32_t *ent, long long cap, uint32_t *boff,
                 uint16_t *bcnt, hipEvent_t read, bc::Counters *hc, hipStream_t s)
{
    if (n < 8 || n % 8 || !(tol >= 0.0f) || !(flo >= 0.0f) || !hc) return VTI_ERR_ARG;
    const long long nb = n / 8, nblk = nb * nb * nb;
    if (nblk * kMaxBlockWords > 0xffffffffLL) return VTI_ERR_ARG;   // 32-bit word offsets
    int st = bc::work_reserve(w, n);
    if (st) return st;
    const long long need = aux_bytes(nblk);
    if (w.aux_cap < need) {
        (void)hipFree(w.aux);
        w.aux = nullptr;
        w.aux_cap = 0;
        if (hipMalloc(&w.aux, (size_t)need) != hipSuccess) {
            (void)hipGetLastError();
            return VTI_ERR_ALLOC;
        }
        w.aux_cap = need;
    }
    // the rounding of d q into bf16 hi + lo (2^-16 of each value) and the
    // two transforms (~2^-16 of the block norm each): 2^-26 of every
    // block's budget is set aside for them, as in codec 1 (whose
    // encode_finish turns the counters into the error bound)
    const double q2 = std::ldexp(1.0, -26);
    const double t2 = (double)tol * tol;
    const double tol2 = t2 > q2 ? t2 - q2 : 0.0;
    if (bc::launch_encode(V, P, off, sy, sx, n, m, w, s) != hipSuccess) return VTI_ERR_HIP;
    if (read && hipEventRecord(read, s) != hipSuccess) return VTI_ERR_HIP;
    const unsigned long long c = (unsigned long long)std::max(0LL, std::min(cap, nblk * (long long)kMaxBlockWords));
    const Aux ax = aux_view(w.aux, nblk);
    const long long ng = (nblk + kScanGroup - 1) / kScanGroup;
    const dim3 gb((unsigned)((nblk + PW - 1) / PW));
    hipLaunchKernelGGL(qsize_kernel, gb, dim3(64 * PW), 0, s, w.coef, nblk, tol2, flo, ax, w.cnt);
    hipLaunchKernelGGL(qscan_kernel, dim3((unsigned)ng), dim3(1024), 0, s, ax, nblk, boff);
    hipLaunchKernelGGL(qtop_kernel, dim3(1), dim3(1024), 0, s, ax, ng, c, w.cnt);
    hipLaunchKernelGGL(qwrite_kernel, gb, dim3(64 * PW), 0, s, w.coef, nblk, ax, ent, c, boff, bcnt);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hc, w.cnt, sizeof(bc::Counters), hipMemcpyDeviceToHost, s) != hipSuccess)
        return VTI_ERR_HIP;
    return VTI_OK;
}

hipError_t launch_decode(const bc::Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         const bc::TileSink &sk, hipStream_t s)
{
    const long long nb = n / 8, nblk = nb * nb * nb;
    hipLaunchKernelGGL(qdecode64_kernel, dim3((unsigned)((nblk + DB - 1) / DB)), dim3(64 * DW), 0, s, V, ent, boff,
                       bcnt, n, sk);
    return hipGetLastError();
}

hipError_t launch_decode(const bc::Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         float *out, hipStream_t s)
{
    bc::TileSink sk;
    sk.out = out;
    return bq::launch_decode(V, ent, boff, bcnt, n, sk, s);
}

}  // namespace bq
}  // namespace vti
