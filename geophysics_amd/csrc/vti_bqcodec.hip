// vti_bqcodec.hip — quantized sparse block-transform codec (codec 2 of the
// learned 3D RTM; gfx950).
//
// The projection on the fitted orthonormal 512 x 512 block basis is the
// sparse codec's (vti_bcodec.hip encode_kernel, MFMA).  Instead of keeping
// the largest coefficients at 15 significant bits each with a 9-bit index
// (codec 1: 32 bits per kept coefficient), every 8^3 block is quantized
// with its own uniform step: q_j = rint(c_j / d), dead zone included (a
// coefficient below d/2 becomes 0).  One wavefront per block bisects on the
// float bits of d for the largest step whose exact quantization error
// sum_j (c_j - d q_j)^2 fits the block's budget — the budget of codec 1,
// tol^2 x max(block energy, floor x the strongest block's energy), so the
// error bound and its Parseval argument are the same.  A block is stored
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
// group of 64 follow the coefficient sizes closely.  The words of a block
// are assembled in LDS with LDS atomics (values may straddle words) and
// written out coalesced; placement is one cursor atomic per 8-block
// workgroup, as in codec 1.  Decode: every lane finds its 8 coefficients'
// bit offsets with a wave prefix sum, dequantizes d q_j (an exact fp32
// product, split into bf16 hi + lo) into the LDS tile and the tile goes
// through codec 1's MFMA reconstruction.
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

constexpr int BB = bc::kBlocksPerWG;
constexpr int LROW = 512 + 8;
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

// quantization error of the lane's 8 coefficients at step d, wave-summed
__device__ __forceinline__ double qerr(const float v[8], float d)
{
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float r = v[e] - d * rintf(v[e] / d);
        s += (double)(r * r);
    }
    return wave_sum_d(s);
}

__global__ __launch_bounds__(64 * PW) void qpack_kernel(const float *__restrict__ coef, long long nblk, double tol2,
                                                        float flo, uint32_t *ent, unsigned cap, uint32_t *boff,
                                                        uint16_t *bcnt, bc::Counters *cnt)
{
    __shared__ uint32_t buf[PW][kMaxBlockWords];
    __shared__ double sd[PW], se[PW];
    __shared__ unsigned tot[PW], pre[PW];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long long b = (long long)blockIdx.x * PW + w;
    const bool live = b < nblk;   // wavefront-uniform
    double err = 0.0, eblk = 0.0;
    float v[8] = {};
    int q[8] = {};
    unsigned mask = 0u, words = 0u, off = 0u;
    int width = 0, mode = 0, nz = 0;
    float d = 0.0f;
    uint32_t h1 = 0u, h2 = 0u;
    if (live) {
        const float *c = coef + b * 512 + 8 * l;
        double e2 = 0.0;
        float amx = 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            v[e] = c[e];
            e2 += (double)(v[e] * v[e]);
            amx = fmaxf(amx, fabsf(v[e]));
        }
        eblk = wave_sum_d(e2);
        const float mx = wave_max(amx);
        const double emax = (double)__uint_as_float(cnt->emax_bits);
        const double budget = tol2 * fmax(eblk, (double)flo * emax);
        if (eblk <= budget) {
            err = eblk;                     // every coefficient quantizes to 0: no words
        } else {
            const float dmin = fmaxf(mx * 0x1p-22f, 1.17549435e-38f);
            if (qerr(v, dmin) > budget) {
                mode = 1;                   // raw fp32 values
                d = 0.0f;
                err = 0.0;
#pragma unroll
                for (int e = 0; e < 8; ++e) mask |= (v[e] != 0.0f ? 1u : 0u) << e;
            } else {
                unsigned lo = __float_as_uint(dmin), hi = __float_as_uint(mx * 4.0f);
                while (hi - lo > 1u) {
                    const unsigned mid = lo + (hi - lo) / 2u;
                    if (qerr(v, __uint_as_float(mid)) <= budget) lo = mid;
                    else hi = mid;
                }
                d = __uint_as_float(lo);
                err = qerr(v, d);
                int amq = 0;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    q[e] = (int)rintf(v[e] / d);
                    mask |= (q[e] != 0 ? 1u : 0u) << e;
                    amq = max(amq, abs(q[e]));
                }
                // the group of 8 lanes' largest |q|
#pragma unroll
                for (int s = 1; s < 8; s <<= 1) amq = max(amq, __shfl_xor(amq, s, 64));
                width = amq ? 33 - __clz(amq) : 0;   // magnitude bits + sign
            }
        }
        nz = __popc(mask);
        const unsigned mine = mode ? 32u * nz : (unsigned)(nz * width);
        unsigned incl = mine;
#pragma unroll
        for (int s = 1; s < 64; s <<= 1) {
            const unsigned t = __shfl_up(incl, s, 64);
            if (l >= s) incl += t;
        }
        off = incl - mine;
        const unsigned vbits = __shfl(incl, 63, 64);
        int nzt = nz;
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) nzt += __shfl_xor(nzt, s, 64);
        if (eblk > budget) {
            words = (unsigned)kHead + (vbits + 31u) / 32u;
            // header fields: every lane reads the widths of the 8 groups
            for (int g = 0; g < 8; ++g) {
                const unsigned wg = (unsigned)__shfl(width, 8 * g, 64);
                if (g < 6) h1 |= wg << (5 * g);
                else h2 |= wg << (5 * (g - 6));
            }
            h1 |= (unsigned)mode << 30;
            h2 |= (unsigned)nzt << 10;
        }
    }
    if (l == 0) {
        sd[w] = err;
        se[w] = eblk;
        tot[w] = words;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        double dsum = 0.0, esum = 0.0;
        for (int i = 0; i < PW; ++i) {
            pre[i] = run;
            run += tot[i];
            dsum += sd[i];
            esum += se[i];
        }
        const unsigned base = run > 0 ? atomicAdd(&cnt->cursor, run) : 0u;
        for (int i = 0; i < PW; ++i) pre[i] += base;
        atomicAdd(&cnt->dropped, dsum);
        atomicAdd(&cnt->energy, esum);
    }
    for (unsigned i = l; i < (unsigned)kMaxBlockWords; i += 64) buf[w][i] = 0u;
    __syncthreads();
    if (live && words) {
        if (l == 0) {
            buf[w][0] = __float_as_uint(d);
            buf[w][1] = h1;
            buf[w][2] = h2;
        }
        atomicOr(&buf[w][3 + (l >> 2)], mask << (8 * (l & 3)));
        unsigned pos = (unsigned)kHead * 32u + off;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if (!((mask >> e) & 1u)) continue;
            if (mode) {
                buf[w][pos >> 5] = __float_as_uint(v[e]);
                pos += 32u;
                continue;
            }
            const unsigned m = ((unsigned)abs(q[e]) << 1) | (q[e] < 0 ? 1u : 0u);
            const unsigned sh = pos & 31u;
            atomicOr(&buf[w][pos >> 5], m << sh);
            if (sh + (unsigned)width > 32u) atomicOr(&buf[w][(pos >> 5) + 1], m >> (32u - sh));
            pos += (unsigned)width;
        }
    }
    __syncthreads();
    if (live) {
        const unsigned base = pre[w];
        const bool fits = (unsigned long long)base + words <= cap;
        if (l == 0) {
            boff[b] = base;
            bcnt[b] = (uint16_t)words;
            if (!fits) atomicOr(&cnt->overflow, 1u);
        }
        if (fits)
            for (unsigned i = l; i < words; i += 64) ent[base + i] = buf[w][i];
    }
}

__global__ __launch_bounds__(256) void qdecode_kernel(bc::Basis V, const uint32_t *__restrict__ ent,
                                                      const uint32_t *__restrict__ boff,
                                                      const uint16_t *__restrict__ bcnt, int n,
                                                      float *__restrict__ out)
{
    __shared__ uint16_t ch[BB][LROW];
    __shared__ uint16_t cl[BB][LROW];
    const int nb = n / 8;
    const long long nblk = (long long)nb * nb * nb;
    const long long b0 = (long long)blockIdx.x * BB;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int i = w; i < BB; i += 4) {
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
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ch[i][8 * l + k] = hh[k];
            cl[i][8 * l + k] = ll[k];
        }
    }
    __syncthreads();
    f4 acc[2][8];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[r][c] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int kk = 0; kk < 512; kk += 32) {
        const int kc = kk + 8 * (l >> 4);
        us8 a[2], al[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            a[r] = *reinterpret_cast<const us8 *>(&ch[16 * r + (l & 15)][kc]);
            al[r] = *reinterpret_cast<const us8 *>(&cl[16 * r + (l & 15)][kc]);
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int k = 128 * w + 16 * c + (l & 15);
            const us8 bh = *reinterpret_cast<const us8 *>(V.v_h + (long long)k * 512 + kc);
            const us8 bl = *reinterpret_cast<const us8 *>(V.v_l + (long long)k * 512 + kc);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a[r]), as_bf(bh), acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(al[r]), as_bf(bh), acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(a[r]), as_bf(bl), acc[r][c], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long long b = b0 + 16 * r + 4 * (l >> 4) + e;
            if (b >= nblk) continue;
            const int bz = (int)(b % nb), bx = (int)((b / nb) % nb), by = (int)(b / ((long long)nb * nb));
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int k = 128 * w + 16 * c + (l & 15);
                const int y = 8 * by + (k >> 6), x = 8 * bx + ((k >> 3) & 7), z = 8 * bz + (k & 7);
                out[((long long)y * n + x) * n + z] = acc[r][c][e];
            }
        }
}

}  // namespace

int encode(const bc::Basis &V, const float *P, long long off, long long sy, long long sx, int n,
           const bc::Mask &m, float tol, float flo, bc::Work &w, uint32_t *ent, long long cap, uint32_t *boff,
           uint16_t *bcnt, long long *kept, double *err, hipStream_t s)
{
    if (n < 8 || n % 8 || !(tol >= 0.0f) || !(flo >= 0.0f)) return VTI_ERR_ARG;
    const long long nb = n / 8, nblk = nb * nb * nb;
    if (nblk * kMaxBlockWords > 0xffffffffLL) return VTI_ERR_ARG;   // 32-bit word offsets
    int st = bc::work_reserve(w, n);
    if (st) return st;
    // the rounding of d q into bf16 hi + lo (2^-16 of each value) and the
    // two transforms (~2^-16 of the block norm each): 2^-26 of every
    // block's budget is set aside for them, as in codec 1
    const double q2 = std::ldexp(1.0, -26);
    const double t2 = (double)tol * tol;
    const double tol2 = t2 > q2 ? t2 - q2 : 0.0;
    if (bc::launch_encode(V, P, off, sy, sx, n, m, w, s) != hipSuccess) return VTI_ERR_HIP;
    const unsigned c32 = (unsigned)std::max(0LL, std::min(cap, nblk * (long long)kMaxBlockWords));
    hipLaunchKernelGGL(qpack_kernel, dim3((unsigned)((nblk + PW - 1) / PW)), dim3(64 * PW), 0, s, w.coef, nblk, tol2,
                       flo, ent, c32, boff, bcnt, w.cnt);
    if (hipGetLastError() != hipSuccess) return VTI_ERR_HIP;
    bc::Counters h;
    if (hipMemcpyAsync(&h, w.cnt, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return VTI_ERR_HIP;
    const long long k = (long long)h.cursor;
    if (kept) *kept = (h.overflow && k <= cap) ? cap + 1 : k;
    if (err) *err = h.energy > 0.0 ? std::sqrt(std::min(1.0, (h.dropped + q2 * (h.energy - h.dropped)) / h.energy)) : 0.0;
    return VTI_OK;
}

hipError_t launch_decode(const bc::Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         float *out, hipStream_t s)
{
    const long long nb = n / 8, nblk = nb * nb * nb;
    hipLaunchKernelGGL(qdecode_kernel, dim3((unsigned)((nblk + BB - 1) / BB)), dim3(256), 0, s, V, ent, boff, bcnt, n,
                       out);
    return hipGetLastError();
}

}  // namespace bq
}  // namespace vti
