// vti_bcodec.hip — error-bounded sparse block-transform codec (gfx950).
//
// Encode (encode64_kernel): a workgroup of 8 wavefronts takes 64 blocks,
// stages them in LDS as bf16 hi/lo pairs (x = hi + lo to 2^-16), and
// computes the 64 x 512 coefficient tile on v_mfma_f32_16x16x32_bf16 as
// Xh Vh + Xl Vh + Xh Vl (fp32 accumulation; the dropped Xl Vl term is below
// 2^-16 relative); the staging also yields the strongest block energy.
// Pack: one wavefront per block finds, by bisection on the float bits of
// c^2 (31 wave reductions), the largest cut whose dropped energy fits the
// block's budget, keeps the coefficients above it, places them with a wave
// prefix sum, a prefix over the workgroup's 8 blocks and one atomic per
// workgroup.  Decode (decode64_kernel): the kept values (15 significant
// bits = bf16 hi + lo exactly) are scattered into zeroed LDS tiles of 64
// blocks and multiplied by V: Ch Vh + Cl Vh + Ch Vl.
#include "vti_bcodec.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/vti.h"

namespace vti {
namespace bc {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));


__device__ __forceinline__ bf16x8 as_bf(us8 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ uint16_t to_bf(float f)   // round to nearest even
{
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float from_bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// A kept coefficient as the top 23 bits of its fp32 pattern (sign, exponent,
// 14 mantissa bits), rounded to nearest even: relative error <= 2^-15.  The
// low 9 bits of the entry hold the coefficient's index.
__device__ __forceinline__ uint32_t to_v23(float f)
{
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return u & 0xfffffe00u;
    return (u + 0xffu + ((u >> 9) & 1u)) & 0xfffffe00u;
}

// Encode with 64 blocks per workgroup of 8 wavefronts: the 1 MB basis is
// streamed from L2 once per 64 blocks, the
// staging writes 16-byte rows of 8 voxels (one z line of one block per lane:
// the wavefront's 64 lanes read the same line of 64 consecutive blocks,
// adjacent in memory), block energies are summed per lane and then over the
// 8 wavefronts (no LDS atomics), and each wavefront's basis operands are
// loaded one K step ahead of the MFMAs that use them.  Wavefront w computes
// blocks 32 (w & 1) .. + 32 against components 128 (w >> 1) .. + 128 as
// Xh Vh + Xl Vh + Xh Vl.
constexpr int EB = kTileBlocks;   // blocks per workgroup
constexpr int EW = kTileWaves;    // wavefronts per workgroup

__device__ __forceinline__ us8 pack8(const uint16_t h[8])
{
    us8 r;
#pragma unroll
    for (int q = 0; q < 8; ++q) r[q] = h[q];
    return r;
}

__global__ __launch_bounds__(64 * EW) void encode64_kernel(Basis V, const float *__restrict__ P, long long off,
                                                          long long sy, long long sx, int n, Mask m,
                                                          float *__restrict__ coef, Counters *cnt)
{
    __shared__ __align__(16) uint16_t xh[EB][LROW];
    __shared__ __align__(16) uint16_t xl[EB][LROW];
    __shared__ float ep[EW][EB];
    const int nb = n / 8;
    const long long nblk = (long long)nb * nb * nb;
    const long long b0 = (long long)blockIdx.x * EB;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    {
        // lane l stages block b0 + l, z lines w, w + 8, .. (line = ty * 8 + tx)
        const long long b = b0 + l;
        float e2 = 0.0f;
        const bool live = b < nblk;
        int bz = 0, bx = 0, by = 0;
        if (live) {
            bz = (int)(b % nb);
            bx = (int)((b / nb) % nb);
            by = (int)(b / ((long long)nb * nb));
        }
        // two lines at a time, each line's loads issued before its conversion
        auto stage = [&](int line, const float *v) {   // the line's hi / lo parts into LDS, its energy into e2
            uint16_t hh[8], ll[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) hh[q] = ll[q] = 0;
            if (live) {
                const int y = 8 * by + (line >> 3), x = 8 * bx + (line & 7), z0 = 8 * bz;
                const bool yin = y >= m.y0 && y < m.y0 + m.by && x >= m.x0 && x < m.x0 + m.bx;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const bool masked = yin && z0 + q >= m.z0 && z0 + q < m.z0 + m.bz;
                    const float vv = masked ? 0.0f : v[q];
                    hh[q] = to_bf(vv);
                    ll[q] = to_bf(vv - from_bf(hh[q]));
                    e2 += vv * vv;
                }
            }
            *reinterpret_cast<us8 *>(&xh[l][line * 8]) = pack8(hh);
            *reinterpret_cast<us8 *>(&xl[l][line * 8]) = pack8(ll);
        };
        auto line_src = [&](int line) {
            return P + off + (long long)(8 * by + (line >> 3)) * sy + (long long)(8 * bx + (line & 7)) * sx + 8 * bz;
        };
#pragma unroll 2
        for (int i = 0; i < 8; ++i) {
            float v[8];
            if (live) {
                const float *src = line_src(w + 8 * i);
#pragma unroll
                for (int q = 0; q < 8; ++q) v[q] = src[q];
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) v[q] = 0.0f;
            }
            stage(w + 8 * i, v);
        }
        ep[w][l] = e2;
    }
    __syncthreads();
    if (threadIdx.x < 64) {   // wavefront 0: block energies, the workgroup's strongest
        float e = 0.0f;
#pragma unroll
        for (int k = 0; k < EW; ++k) e += ep[k][l];
        float mx = e;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) mx = fmaxf(mx, __shfl_xor(mx, d, 64));
        // non-negative floats order as their bits; the maximum only grows, so
        // a workgroup below the value already stored skips the atomic (the
        // same-address atomics of all workgroups would serialise)
        if (l == 0 && __float_as_uint(mx) > __hip_atomic_load(&cnt->emax_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&cnt->emax_bits, __float_as_uint(mx));
    }
    tile_f4 acc[4][4];
    tile_mfma(xh, xl, V.vt_h, V.vt_l, w, l, acc);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long long b = b0 + 16 * r + 4 * (l >> 4) + e;
            if (b >= nblk) continue;
#pragma unroll
            for (int c = 0; c < 4; ++c) coef[b * 512 + 64 * w + 16 * c + (l & 15)] = acc[r][c][e];
        }
}

// wave-uniform sum: lane 0's butterfly result broadcast (the lanes' own
// results may differ in the last bit)
__device__ __forceinline__ float wave_sum(float x)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return __shfl(x, 0, 64);
}

constexpr int PW = 8;   // blocks (wavefronts) per pack workgroup

// One wavefront per block, PW blocks per workgroup; the workgroup reserves
// its entries with one atomic on the cursor (same-address atomics serialise
// in L2, so one per block would dominate the kernel) and hands each block
// its range from a prefix over the workgroup's block totals.
__global__ __launch_bounds__(64 * PW) void pack_kernel(const float *__restrict__ coef, long long nblk, float tol2,
                                                       float flo, uint32_t *ent, unsigned cap, uint32_t *boff,
                                                       uint16_t *bcnt, Counters *cnt)
{
    __shared__ float sd[PW], se[PW];
    __shared__ unsigned tot[PW], pre[PW];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long long b = (long long)blockIdx.x * PW + w;
    const bool live = b < nblk;   // wavefront-uniform
    float dropped = 0.0f, eblk = 0.0f;
    float v[8];
    unsigned q[8];
    unsigned lo = 0u;
    int n_keep = 0, incl = 0, total = 0;
    if (live) {
        const float *c = coef + b * 512 + 8 * l;
        float e2 = 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            v[e] = c[e];
            q[e] = __float_as_uint(v[e] * v[e]);
            e2 += v[e] * v[e];
        }
        eblk = wave_sum(e2);
        const float emax = __uint_as_float(cnt->emax_bits);
        const float budget = tol2 * fmaxf(eblk, flo * emax);
        // largest cut t with sum of c^2 over bits(c^2) < t within budget
        unsigned hi = 0x7f800001u;
        while (hi - lo > 1u) {
            const unsigned mid = lo + (hi - lo) / 2u;
            float part = 0.0f;
#pragma unroll
            for (int e = 0; e < 8; ++e) part += q[e] < mid ? __uint_as_float(q[e]) : 0.0f;
            if (wave_sum(part) <= budget) lo = mid;
            else hi = mid;
        }
        float part = 0.0f;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            n_keep += q[e] >= lo;
            part += q[e] < lo ? __uint_as_float(q[e]) : 0.0f;
        }
        dropped = wave_sum(part);
        incl = n_keep;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (l >= d) incl += t;
        }
        total = __shfl(incl, 63, 64);
    }
    if (l == 0) {
        sd[w] = dropped;
        se[w] = eblk;
        tot[w] = (unsigned)total;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned run = 0;
        double dsum = 0.0, esum = 0.0;
        for (int i = 0; i < PW; ++i) {
            pre[i] = run;
            run += tot[i];
            dsum += (double)sd[i];
            esum += (double)se[i];
        }
        const unsigned base = run > 0 ? atomicAdd(&cnt->cursor, run) : 0u;
        for (int i = 0; i < PW; ++i) pre[i] += base;
        atomicAdd(&cnt->dropped, dsum);
        atomicAdd(&cnt->energy, esum);
    }
    __syncthreads();
    if (live) {
        const unsigned base = pre[w];
        const bool fits = (unsigned long long)base + (unsigned)total <= cap;
        if (l == 0) {
            boff[b] = base;
            bcnt[b] = (uint16_t)total;
            if (!fits) atomicOr(&cnt->overflow, 1u);
        }
        if (fits) {
            unsigned pos = base + (unsigned)(incl - n_keep);
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (q[e] >= lo) ent[pos++] = to_v23(v[e]) | (uint32_t)(8 * l + e);
        }
    }
}

// Decode with 64 blocks per workgroup of 8 wavefronts: the basis streamed
// once per 64 blocks, the basis operands loaded one K
// step ahead; wavefront w scatters blocks w, w + 8, .. into the zeroed LDS
// tile and computes blocks 32 (w & 1) .. + 32 against voxels
// 128 (w >> 1) .. + 128.
__global__ __launch_bounds__(64 * EW) void decode64_kernel(Basis V, const uint32_t *__restrict__ ent,
                                                          const uint32_t *__restrict__ boff,
                                                          const uint16_t *__restrict__ bcnt, int n, TileSink sk)
{
    __shared__ __align__(16) uint16_t lds[2][EB][LROW];   // coefficient tiles (hi, lo), then the output tile
    uint16_t(*ch)[LROW] = lds[0];
    uint16_t(*cl)[LROW] = lds[1];
    const int nb = n / 8;
    const long long nblk = (long long)nb * nb * nb;
    const long long b0 = (long long)blockIdx.x * EB;
    for (int i = threadIdx.x; i < EB * LROW; i += 64 * EW) reinterpret_cast<uint32_t *>(&lds[0][0][0])[i] = 0u;
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int i = w; i < EB; i += EW) {
        const long long b = b0 + i;
        if (b >= nblk) break;
        const int cnt = bcnt[b];
        const uint32_t *e = ent + boff[b];
        for (int q = l; q < cnt; q += 64) {
            const uint32_t x = e[q];
            const float f = __uint_as_float(x & 0xfffffe00u);   // = hi + lo exactly (15 significant bits)
            const uint16_t h = to_bf(f);
            ch[i][x & 511u] = h;
            cl[i][x & 511u] = to_bf(f - from_bf(h));
        }
    }
    __syncthreads();
    tile_f4 acc[4][4];
    tile_mfma(ch, cl, V.v_h, V.v_l, w, l, acc);
    __syncthreads();
    tile_store(&lds[0][0][0], acc, w, l, b0, n, sk);
}

__global__ void zero_kernel(uint32_t *p, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

uint16_t host_bf(float f)   // round to nearest even (finite inputs)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

float host_from_bf(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

int work_reserve(Work &w, int n)
{
    const long long nb = n / 8, nblk = nb * nb * nb;
    if (!w.cnt && hipMalloc(&w.cnt, sizeof(Counters)) != hipSuccess) return VTI_ERR_ALLOC;
    if (nblk > w.cap_blocks) {
        (void)hipFree(w.coef);
        w.coef = nullptr;
        w.cap_blocks = 0;
        if (hipMalloc(&w.coef, (size_t)nblk * 512 * sizeof(float)) != hipSuccess) return VTI_ERR_ALLOC;
        w.cap_blocks = nblk;
    }
    return VTI_OK;
}

void work_free(Work &w)
{
    (void)hipFree(w.coef);
    (void)hipFree(w.cnt);
    (void)hipFree(w.aux);
    w = Work{};
}

void basis_split(const float *V, uint16_t *out)
{
    uint16_t *vt_h = out, *vt_l = out + 512 * 512, *v_h = out + 2 * 512 * 512, *v_l = out + 3 * 512 * 512;
    for (int k = 0; k < 512; ++k)
        for (int j = 0; j < 512; ++j) {
            const float x = V[k * 512 + j];
            const uint16_t h = host_bf(x);
            const uint16_t lo = host_bf(x - host_from_bf(h));
            v_h[frag_off(k, j)] = h;
            v_l[frag_off(k, j)] = lo;
            vt_h[frag_off(j, k)] = h;
            vt_l[frag_off(j, k)] = lo;
        }
}

Basis basis_view(const uint16_t *dev)
{
    return Basis{dev, dev + 512 * 512, dev + 2 * 512 * 512, dev + 3 * 512 * 512};
}

hipError_t launch_encode(const Basis &V, const float *P, long long off, long long sy, long long sx, int n,
                         const Mask &m, const Work &w, hipStream_t s)
{
    const long long nb = n / 8, nblk = nb * nb * nb;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<uint32_t *>(w.cnt),
                       (long long)(sizeof(Counters) / 4));
    hipLaunchKernelGGL(encode64_kernel, dim3((unsigned)((nblk + EB - 1) / EB)), dim3(64 * EW), 0, s, V, P, off, sy, sx,
                       n, m, w.coef, w.cnt);
    return hipGetLastError();
}

hipError_t launch_pack(const Work &w, int n, float tol2, float flo, uint32_t *ent, long long cap, uint32_t *boff,
                       uint16_t *bcnt, hipStream_t s)
{
    const long long nb = n / 8, nblk = nb * nb * nb;
    // a snapshot has at most n^3 entries
    const unsigned c32 = (unsigned)std::max(0LL, std::min(cap, (long long)n * n * n));
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((nblk + PW - 1) / PW)), dim3(64 * PW), 0, s, w.coef, nblk, tol2,
                       flo, ent, c32, boff, bcnt, w.cnt);
    return hipGetLastError();
}

hipError_t launch_decode(const Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         const TileSink &sk, hipStream_t s)
{
    const long long nb = n / 8, nblk = nb * nb * nb;
    hipLaunchKernelGGL(decode64_kernel, dim3((unsigned)((nblk + EB - 1) / EB)), dim3(64 * EW), 0, s, V, ent, boff, bcnt,
                       n, sk);
    return hipGetLastError();
}

hipError_t launch_decode(const Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         float *out, hipStream_t s)
{
    TileSink sk;
    sk.out = out;
    return launch_decode(V, ent, boff, bcnt, n, sk, s);
}

int encode(const Basis &V, const float *P, long long off, long long sy, long long sx, int n, const Mask &m,
           float tol, float flo, Work &w, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt,
           long long *kept, double *err, double *energy, hipStream_t s)
{
    if (n < 8 || n % 8 || !(tol >= 0.0f) || !(flo >= 0.0f)) return VTI_ERR_ARG;
    const int st = work_reserve(w, n);
    if (st) return st;
    return encode_in(V, P, off, sy, sx, n, m, tol, flo, w, ent, cap, boff, bcnt, kept, err, energy, s);
}

int encode_in(const Basis &V, const float *P, long long off, long long sy, long long sx, int n, const Mask &m,
              float tol, float flo, const Work &w, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt,
              long long *kept, double *err, double *energy, hipStream_t s)
{
    Counters h;
    const int st = encode_start(V, P, off, sy, sx, n, m, tol, flo, w, ent, cap, boff, bcnt, nullptr, &h, s);
    if (st) return st;
    if (hipStreamSynchronize(s) != hipSuccess) return VTI_ERR_HIP;
    encode_finish(h, cap, kept, err, energy);
    return VTI_OK;
}

// a kept value is stored to 2^-15 of itself and the two transforms add
// ~2^-16 of the block norm each: 2^-26 of every block's budget is set aside
// for them
static double rounding_share() { return std::ldexp(1.0, -26); }

int encode_start(const Basis &V, const float *P, long long off, long long sy, long long sx, int n, const Mask &m,
                 float tol, float flo, const Work &w, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt,
                 hipEvent_t read, Counters *hc, hipStream_t s)
{
    const long long nb = n / 8;
    if (n < 8 || n % 8 || !(tol >= 0.0f) || !(flo >= 0.0f) || !w.cnt || !w.coef || nb * nb * nb > w.cap_blocks || !hc)
        return VTI_ERR_ARG;
    const double q2 = rounding_share();
    const double t2 = (double)tol * tol;
    const float tol2 = t2 > q2 ? (float)(t2 - q2) : 0.0f;
    if (launch_encode(V, P, off, sy, sx, n, m, w, s) != hipSuccess) return VTI_ERR_HIP;
    if (read && hipEventRecord(read, s) != hipSuccess) return VTI_ERR_HIP;
    if (launch_pack(w, n, tol2, flo, ent, cap, boff, bcnt, s) != hipSuccess ||
        hipMemcpyAsync(hc, w.cnt, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess)
        return VTI_ERR_HIP;
    return VTI_OK;
}

void encode_finish(const Counters &h, long long cap, long long *kept, double *err, double *energy)
{
    const double q2 = rounding_share();
    const long long k = (long long)h.cursor;
    if (kept) *kept = (h.overflow && k <= cap) ? cap + 1 : k;
    if (energy) *energy = h.energy;
    if (err) *err = h.energy > 0.0 ? std::sqrt(std::min(1.0, (h.dropped + q2 * (h.energy - h.dropped)) / h.energy)) : 0.0;
}

}  // namespace bc
}  // namespace vti
