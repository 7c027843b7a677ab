// vti_nn.hip — BF16-MFMA kernels of the snapshot autoencoder (gfx950).
//
// GEMM operand layout of v_mfma_f32_16x16x32_bf16 (one wavefront, lane l):
//   A (16 x 32): row l % 16, columns 8*(l/16) .. 8*(l/16)+7
//   B (32 x 16): rows 8*(l/16) .. +7, column l % 16
//   D (16 x 16): rows 4*(l/16) .. +3, column l % 16
// Forward / backward-data: A = weight rows n (16-byte loads of [Npad][Kpad]),
// B = the gathered source (8 consecutive channels of one tap: one 16-byte
// load), D = 4 output channels x 1 position per lane.  Weight gradient:
// A = the output gradient (rows n, 8 positions), B = the gathered source
// (columns k, 8 positions), reduced over positions in chunks and added to
// the fp32 gradient with global atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "vti_nn.h"

namespace vti {
namespace nn {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

// fp32 -> bf16 bits, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 as_bf(us8 v) { return __builtin_bit_cast(bf16x8, v); }

// source position of a tap; false when it falls in the zero padding
__device__ __forceinline__ bool src_pos(int kind, int tap, int Y, int X, int Z, int iy, int ix, int iz, int &sy,
                                        int &sx, int &sz)
{
    if (kind == DOWN) {
        sy = 2 * Y + (tap >> 2);
        sx = 2 * X + ((tap >> 1) & 1);
        sz = 2 * Z + (tap & 1);
        return true;
    }
    sy = Y + tap / 9 - 1;
    sx = X + (tap / 3) % 3 - 1;
    sz = Z + tap % 3 - 1;
    return sy >= 0 && sy < iy && sx >= 0 && sx < ix && sz >= 0 && sz < iz;
}

__device__ __forceinline__ void decode(long long p, int gy, int gx, int gz, int &b, int &Y, int &X, int &Z)
{
    Z = (int)(p % gz);
    p /= gz;
    X = (int)(p % gx);
    p /= gx;
    Y = (int)(p % gy);
    b = (int)(p / gy);
}

// one source element (ci == 1 path, or a single k of the weight gradient)
template <class G>
__device__ __forceinline__ uint16_t src_elem(const G &g, int b, int Y, int X, int Z, int k)
{
    if (g.kind == UP) return g.x[((((long long)b * g.iy + Y) * g.ix + X) * g.iz + Z) * g.ci + k];
    const int tap = k / g.ci, c = k - tap * g.ci;
    int sy, sx, sz;
    if (!src_pos(g.kind, tap, Y, X, Z, g.iy, g.ix, g.iz, sy, sx, sz)) return 0;
    if (g.xf) return f2bf(g.xf[g.xoff + b * g.xsb + sy * g.xsy + sx * g.xsx + sz] * g.xscale);
    return g.x[((((long long)b * g.iy + sy) * g.ix + sx) * g.iz + sz) * g.ci + c];
}

// B operand of the forward GEMM: 8 consecutive k of position (b,Y,X,Z)
__device__ __forceinline__ us8 load_b(const Gemm &g, int b, int Y, int X, int Z, int k)
{
    us8 r = {0, 0, 0, 0, 0, 0, 0, 0};
    if (k >= g.K) return r;
    if (g.kind == UP)
        return *reinterpret_cast<const us8 *>(g.x + ((((long long)b * g.iy + Y) * g.ix + X) * g.iz + Z) * g.ci + k);
    if (g.ci >= 8) {
        const int tap = k / g.ci, c = k - tap * g.ci;
        int sy, sx, sz;
        if (!src_pos(g.kind, tap, Y, X, Z, g.iy, g.ix, g.iz, sy, sx, sz)) return r;
        return *reinterpret_cast<const us8 *>(g.x + ((((long long)b * g.iy + sy) * g.ix + sx) * g.iz + sz) * g.ci + c);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e)
        if (k + e < g.K) r[e] = src_elem(g, b, Y, X, Z, k + e);
    return r;
}

template <int NT>
__global__ __launch_bounds__(256) void gemm_kernel(Gemm g)
{
    const int lane = threadIdx.x & 63;
    const long long P = (long long)g.nb * g.gy * g.gx * g.gz;
    const long long p0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (p0 >= P) return;   // wavefront-uniform
    const long long p = p0 + (lane & 15);
    const bool pv = p < P;
    int b = 0, Y = 0, X = 0, Z = 0;
    if (pv) decode(p, g.gy, g.gx, g.gz, b, Y, X, Z);
    const int kq = 8 * (lane >> 4);
    const int nbase = blockIdx.y * NT * 16;
    f4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (f4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.Kpad; k0 += 32) {
        const int k = k0 + kq;
        us8 bv = {0, 0, 0, 0, 0, 0, 0, 0};
        if (pv) bv = load_b(g, b, Y, X, Z, k);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const us8 av = *reinterpret_cast<const us8 *>(g.w + (long long)(nbase + t * 16 + (lane & 15)) * g.Kpad + k);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(av), as_bf(bv), acc[t], 0, 0, 0);
        }
    }
    if (!pv) return;
    const bool up = g.kind == UP;
    const int OY = up ? 2 * g.gy : g.gy, OX = up ? 2 * g.gx : g.gx, OZ = up ? 2 * g.gz : g.gz;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int n = nbase + t * 16 + 4 * (lane >> 4) + e;
            if (n >= g.N) continue;
            int c = n, oy = Y, ox = X, oz = Z;
            if (up) {
                const int tap = n / g.co;
                c = n - tap * g.co;
                oy = 2 * Y + (tap >> 2);
                ox = 2 * X + ((tap >> 1) & 1);
                oz = 2 * Z + (tap & 1);
            }
            const long long oa = ((((long long)b * OY + oy) * OX + ox) * OZ + oz) * g.co + c;
            float v = acc[t][e];
            if (g.gate) {
                if (!(bf2f(g.gate[oa]) > 0.0f)) v = kLeak * v;
            } else {
                if (g.bias) v = v + g.bias[c];
                if (g.act && !(v > 0.0f)) v = kLeak * v;
            }
            if (g.yf) g.yf[oa] = v * g.yscale;
            else g.y[oa] = f2bf(v);
        }
}

// gradient element of output column n at GEMM position pos
__device__ __forceinline__ uint16_t dval(const WGrad &g, long long pos, int n)
{
    if (g.kind != UP) return g.d[pos * g.co + n];
    int b, Y, X, Z;
    decode(pos, g.gy, g.gx, g.gz, b, Y, X, Z);
    const int tap = n / g.co, c = n - tap * g.co;
    const long long oa = ((((long long)b * 2 * g.gy + 2 * Y + (tap >> 2)) * 2 * g.gx + 2 * X + ((tap >> 1) & 1)) * 2 *
                              g.gz + 2 * Z + (tap & 1)) * g.co + c;
    return g.d[oa];
}

__global__ __launch_bounds__(256) void wgrad_kernel(WGrad g)
{
    const int lane = threadIdx.x & 63;
    const int ntk = g.Kpad / 16;
    const int n0 = (blockIdx.y / ntk) * 16, k0 = (blockIdx.y % ntk) * 16;
    const long long P = (long long)g.nb * g.gy * g.gx * g.gz;
    const long long ps = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * g.chunk;
    if (ps >= P) return;   // wavefront-uniform
    const long long pe = ps + g.chunk < P ? ps + g.chunk : P;
    const int n = n0 + (lane & 15), k = k0 + (lane & 15);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long p = ps; p < pe; p += 32) {
        us8 av = {0, 0, 0, 0, 0, 0, 0, 0}, bv = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long long pos = p + 8 * (lane >> 4) + e;
            if (pos >= pe) continue;
            if (n < g.N) av[e] = dval(g, pos, n);
            if (k < g.K) {
                int b, Y, X, Z;
                decode(pos, g.gy, g.gx, g.gz, b, Y, X, Z);
                bv[e] = src_elem(g, b, Y, X, Z, k);
            }
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(av), as_bf(bv), acc, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int nn = n0 + 4 * (lane >> 4) + e;
        if (nn < g.N && k < g.K) atomicAdd(g.dw + (long long)nn * g.Kpad + k, acc[e]);
    }
}

__global__ void mse_kernel(const float *__restrict__ out, const float *__restrict__ t, long long n, float *loss,
                           uint16_t *dout)
{
    __shared__ float red[256];
    float s = 0.0f;
    const float sc = 2.0f / (float)n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = out[i] - t[i];
        s += d * d;
        dout[i] = f2bf(sc * d);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)n);
}

__global__ void adam_kernel(float *w, float *m, float *v, const float *__restrict__ g, uint16_t *wb, long long n,
                            float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        wb[i] = f2bf(wi);
    }
}

// backward-data matrix of a layer (see vti_nn.h): DOWN and UP transpose,
// SAME transposes channels and reverses the 27 taps.
__global__ void wtrans_kernel(int kind, int ci, int co, const uint16_t *__restrict__ w, int Kpad, uint16_t *wt,
                              int Ktpad, int Ntpad)
{
    const long long total = (long long)Ntpad * Ktpad;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(i / Ktpad), c = (int)(i % Ktpad);
        uint16_t val = 0;
        if (kind == DOWN) {          // rows (tap, ci) = 8ci, cols co
            if (r < 8 * ci && c < co) val = w[(long long)c * Kpad + r];
        } else if (kind == UP) {     // rows ci, cols (tap, co) = 8co
            if (r < ci && c < 8 * co) val = w[(long long)c * Kpad + r];
        } else {                     // rows ci, cols (tap', co) = 27co
            if (r < ci && c < 27 * co) {
                const int tp = c / co, o = c - tp * co;
                val = w[(long long)o * Kpad + (26 - tp) * ci + r];
            }
        }
        wt[i] = val;
    }
}

__global__ void fill_u16_kernel(uint16_t *p, long long n, uint16_t v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_f32_kernel(float *p, long long n, float v)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void f2bf_kernel(const float *__restrict__ in, uint16_t *out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = f2bf(in[i]);
}
__global__ void bf2f_kernel(const uint16_t *__restrict__ in, float *out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = bf2f(in[i]);
}

__global__ void maxabs_kernel(const float *__restrict__ v, long long off, long long sy, long long sx, int ny, int nx,
                              int nz, unsigned *out)
{
    __shared__ unsigned red[256];
    unsigned m = 0;
    const long long n = (long long)ny * nx * nz;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % nz);
        const long long r = i / nz;
        const int x = (int)(r % nx), y = (int)(r / nx);
        const unsigned u = __float_as_uint(fabsf(v[off + y * sy + x * sx + z]));
        m = u > m ? u : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

__global__ void crops_kernel(const float *__restrict__ src, long long off, long long sy, long long sx,
                             const int *__restrict__ org, int ncrop, int n, float *dst)
{
    const long long per = (long long)n * n * n, total = per * ncrop;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const int cidx = (int)(i / per);
        long long r = i % per;
        const int z = (int)(r % n);
        r /= n;
        const int x = (int)(r % n), y = (int)(r / n);
        const int *o = org + 3 * cidx;
        dst[i] = src[off + (o[0] + y) * sy + (o[1] + x) * sx + o[2] + z];
    }
}

// one workgroup per crop: scale[c] = max |crop c|, then crop c /= scale[c]
// (an all-zero crop stays zero)
__global__ void crop_norm_kernel(float *dst, long long per, float *scale)
{
    __shared__ float red[256];
    float *p = dst + (long long)blockIdx.x * per;
    float m = 0.0f;
    for (long long i = threadIdx.x; i < per; i += blockDim.x) m = fmaxf(m, fabsf(p[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const float sc = red[0];
    if (threadIdx.x == 0) scale[blockIdx.x] = sc;
    if (!(sc > 0.0f)) return;
    for (long long i = threadIdx.x; i < per; i += blockDim.x) p[i] = p[i] / sc;
}

// B[b][k] = bf16(P[block b voxel k] / scale), k = (t2*8 + t1)*8 + t0 with
// the voxel offset (4 dy2 + 2 dy1 + dy0, ...) of an 8^3 block, tap bits
// (dy, dx, dz) = (t >> 2, (t >> 1) & 1, t & 1): the order in which the
// encoder's third layer sees a block.  blk[b] = index of the block in the
// (n/8)^3 grid of the box.
__global__ void gather_blocks_kernel(const float *__restrict__ P, long long off, long long sy, long long sx, int n,
                                     const int *__restrict__ blk, int nblk, const float *__restrict__ scale_ptr,
                                     uint16_t *B)
{
    const int nb8 = n / 8;
    const float inv = 1.0f / *scale_ptr;
    const long long total = (long long)nblk * 512;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(i / 512), k = (int)(i % 512);
        const int q = blk[b];
        const int bz = q % nb8, bx = (q / nb8) % nb8, by = q / (nb8 * nb8);
        const int t2 = k >> 6, t1 = (k >> 3) & 7, t0 = k & 7;
        const int y = 4 * (t2 >> 2) + 2 * (t1 >> 2) + (t0 >> 2);
        const int x = 4 * ((t2 >> 1) & 1) + 2 * ((t1 >> 1) & 1) + ((t0 >> 1) & 1);
        const int z = 4 * (t2 & 1) + 2 * (t1 & 1) + (t0 & 1);
        B[i] = f2bf(P[off + (8 * by + y) * sy + (8 * bx + x) * sx + 8 * bz + z] * inv);
    }
}

// G[a][b] += sum_p B[p][a] B[p][b] (K x K, K a multiple of 16) on MFMA:
// one 16x16 tile of G per wavefront and chunk of positions, fp32 atomics.
__global__ __launch_bounds__(256) void gram_kernel(const uint16_t *__restrict__ B, long long P, int K, float *G,
                                                   int chunk)
{
    const int lane = threadIdx.x & 63;
    const int nt = K / 16;
    const int a0 = (blockIdx.y / nt) * 16, b0 = (blockIdx.y % nt) * 16;
    const long long ps = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * chunk;
    if (ps >= P) return;   // wavefront-uniform
    const long long pe = ps + chunk < P ? ps + chunk : P;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long p = ps; p < pe; p += 32) {
        us8 av = {0, 0, 0, 0, 0, 0, 0, 0}, bv = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long long pos = p + 8 * (lane >> 4) + e;
            if (pos >= pe) continue;
            av[e] = B[pos * K + a0 + (lane & 15)];
            bv[e] = B[pos * K + b0 + (lane & 15)];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf(av), as_bf(bv), acc, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) atomicAdd(G + (long long)(a0 + 4 * (lane >> 4) + e) * K + b0 + (lane & 15), acc[e]);
}

unsigned blocks_for(long long n, int per = 256, unsigned cap = 8192)
{
    const long long b = (n + per - 1) / per;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, cap));
}

}  // namespace

hipError_t launch_gather_blocks(const float *P, long long off, long long sy, long long sx, int n, const int *blk,
                                int nblk, const float *scale_ptr, uint16_t *B, hipStream_t s)
{
    hipLaunchKernelGGL(gather_blocks_kernel, dim3(blocks_for((long long)nblk * 512)), dim3(256), 0, s, P, off, sy, sx, n,
                       blk, nblk, scale_ptr, B);
    return hipGetLastError();
}

hipError_t launch_gram(const uint16_t *B, long long P, int K, float *G, hipStream_t s)
{
    const int chunk = 256;
    const dim3 grd((unsigned)((P + 4LL * chunk - 1) / (4LL * chunk)), (unsigned)((K / 16) * (K / 16)));
    hipLaunchKernelGGL(gram_kernel, grd, dim3(256), 0, s, B, P, K, G, chunk);
    return hipGetLastError();
}

hipError_t launch_gemm(const Gemm &g, hipStream_t s)
{
    const long long P = (long long)g.nb * g.gy * g.gx * g.gz;
    const unsigned gx = (unsigned)((P + 63) / 64);
    const dim3 grd(gx, (unsigned)(g.Npad / (16 * g.NT)));
    switch (g.NT) {
    case 1: hipLaunchKernelGGL(gemm_kernel<1>, grd, dim3(256), 0, s, g); break;
    case 2: hipLaunchKernelGGL(gemm_kernel<2>, grd, dim3(256), 0, s, g); break;
    case 4: hipLaunchKernelGGL(gemm_kernel<4>, grd, dim3(256), 0, s, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_wgrad(const WGrad &g, hipStream_t s)
{
    const long long P = (long long)g.nb * g.gy * g.gx * g.gz;
    const long long per_block = 4LL * g.chunk;
    const dim3 grd((unsigned)((P + per_block - 1) / per_block), (unsigned)((g.Npad / 16) * (g.Kpad / 16)));
    hipLaunchKernelGGL(wgrad_kernel, grd, dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_mse(const float *out, const float *t, long long n, float *loss, uint16_t *dout, hipStream_t s)
{
    hipLaunchKernelGGL(mse_kernel, dim3(blocks_for(n, 256, 1024)), dim3(256), 0, s, out, t, n, loss, dout);
    return hipGetLastError();
}

hipError_t launch_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, long long n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t s)
{
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, s, w, m, v, g, wb, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_wtrans(int kind, int ci, int co, const uint16_t *w, int Kpad, uint16_t *wt, int Ktpad, int Ntpad,
                         hipStream_t s)
{
    hipLaunchKernelGGL(wtrans_kernel, dim3(blocks_for((long long)Ntpad * Ktpad)), dim3(256), 0, s, kind, ci, co, w, Kpad,
                       wt, Ktpad, Ntpad);
    return hipGetLastError();
}

hipError_t launch_fill_u16(uint16_t *p, long long n, uint16_t v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_u16_kernel, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks_for(n)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_f32_to_bf16(const float *in, uint16_t *out, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(f2bf_kernel, dim3(blocks_for(n)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_bf16_to_f32(const uint16_t *in, float *out, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(bf2f_kernel, dim3(blocks_for(n)), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_maxabs(const float *v, long long off, long long sy, long long sx, int ny, int nx, int nz,
                         unsigned *out_bits, hipStream_t s)
{
    hipLaunchKernelGGL(maxabs_kernel, dim3(blocks_for((long long)ny * nx * nz, 256, 2048)), dim3(256), 0, s, v, off, sy,
                       sx, ny, nx, nz, out_bits);
    return hipGetLastError();
}

hipError_t launch_crops(const float *src, long long off, long long sy, long long sx, const int *org, int ncrop, int n,
                        float *crop_scale, float *dst, hipStream_t s)
{
    const long long per = (long long)n * n * n;
    hipLaunchKernelGGL(crops_kernel, dim3(blocks_for(ncrop * per)), dim3(256), 0, s, src, off, sy, sx, org, ncrop, n, dst);
    hipLaunchKernelGGL(crop_norm_kernel, dim3(ncrop), dim3(256), 0, s, dst, per, crop_scale);
    return hipGetLastError();
}

}  // namespace nn
}  // namespace vti
