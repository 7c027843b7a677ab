// vti_nn.h — BF16-MFMA layers of the source-wavefield snapshot autoencoder.
//
// Volumes are channels-last, (nb, ny, nx, nz, c) with c fastest, stored as
// bf16 bits (uint16_t).  Every layer is one GEMM
//     D[n][p] = sum_k W[n][k] * B[k][p]
// on v_mfma_f32_16x16x32_bf16 (fp32 accumulation): n runs over output
// channels (or tap x channel for the transposed kind), p over positions of
// the GEMM grid, and B is gathered from the source volume on the fly:
//   kind 0 "down": kernel 2, stride 2, B[(tap, c)][p] = x[2p + tap][c]
//   kind 1 "same": kernel 3, stride 1, zero padding, B[(tap, c)][p] = x[p + tap - 1][c]
//   kind 2 "up":   kernel 2, stride 2 transposed, B[c][q] = x[q][c], D row
//                  n = (tap, c) scattered to output position 2q + tap.
// Weight matrices are stored [Npad][Kpad] (rows padded to 16, columns to 32,
// padding zero) so every A operand is one 16-byte load.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vti {
namespace nn {

enum Kind { DOWN = 0, SAME = 1, UP = 2 };
constexpr float kLeak = 0.125f;   // leaky ReLU slope (exact in bf16)

inline int ntap(int kind) { return kind == SAME ? 27 : 8; }
inline int pad16(int v) { return (v + 15) / 16 * 16; }
inline int pad32(int v) { return (v + 31) / 32 * 32; }

// One GEMM-form layer launch.
struct Gemm {
    int kind;
    int nb, gy, gx, gz;        // GEMM position grid (down/same: output grid; up: input grid)
    int iy, ix, iz;            // source grid (down/same gather; up: = GEMM grid)
    int ci;                    // source channels (1 or a multiple of 8)
    int K, Kpad, N, Npad, NT;  // GEMM sizes; NT = 16-row tiles per wavefront
    const uint16_t *x;         // bf16 source, or
    const float *xf;           // fp32 source read with strides and a scale (ci == 1)
    long long xsb, xsy, xsx, xoff;
    float xscale;
    const uint16_t *w;         // [Npad][Kpad]
    const float *bias;         // [co] or null
    int co;                    // output channels
    int act;                   // 1: leaky ReLU
    uint16_t *y;               // bf16 output, or
    float *yf;                 // fp32 output (dense channels-last), times yscale
    float yscale;
    const uint16_t *gate;      // backward: multiply by act'(gate) at the output element
};

// Weight gradient dW[n][k] += sum_p D[p][n] * B[k][p] (the layer's forward
// GEMM transposed), D = gradient at the pre-activation output.
struct WGrad {
    int kind;
    int nb, gy, gx, gz;        // forward GEMM grid
    int iy, ix, iz, ci;        // forward source grid / channels
    int K, Kpad, N, Npad, co;
    const uint16_t *x;         // forward source (bf16)
    const float *xf;
    long long xsb, xsy, xsx, xoff;
    float xscale;
    const uint16_t *d;         // gradient volume: down/same: GEMM grid x co; up: output grid x co
    float *dw;                 // [Npad][Kpad] fp32, accumulated
    int chunk;                 // positions per wavefront (multiple of 32)
};

hipError_t launch_gemm(const Gemm &g, hipStream_t s);
hipError_t launch_wgrad(const WGrad &g, hipStream_t s);
// loss += sum (out - t)^2 / n ; dout = bf16(2 (out - t) / n)
hipError_t launch_mse(const float *out, const float *t, long long n, float *loss, uint16_t *dout, hipStream_t s);
// Adam on n fp32 parameters; wb receives the bf16 copy
hipError_t launch_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, long long n, float lr,
                       float b1, float b2, float eps, float c1, float c2, hipStream_t s);
// Rebuild the backward-data matrices of one layer from its forward matrix.
hipError_t launch_wtrans(int kind, int ci, int co, const uint16_t *w, int Kpad, uint16_t *wt, int Ktpad,
                         int Ntpad, hipStream_t s);
// 8^3 blocks of an n^3 box as rows of 512 bf16 values (scaled by 1/scale),
// in the order the encoder's third layer sees them
hipError_t launch_gather_blocks(const float *P, long long off, long long sy, long long sx, int n, const int *blk,
                                int nblk, const float *scale_ptr, uint16_t *B, hipStream_t s);
// G (K x K fp32) += B^T B, B = P rows of K bf16 values (MFMA)
hipError_t launch_gram(const uint16_t *B, long long P, int K, float *G, hipStream_t s);
hipError_t launch_fill_u16(uint16_t *p, long long n, uint16_t v, hipStream_t s);
hipError_t launch_fill_f32(float *p, long long n, float v, hipStream_t s);
hipError_t launch_f32_to_bf16(const float *in, uint16_t *out, long long n, hipStream_t s);
hipError_t launch_bf16_to_f32(const uint16_t *in, float *out, long long n, hipStream_t s);
// max |v| over an (ny, nx, nz) box of a strided fp32 volume, as float bits
hipError_t launch_maxabs(const float *v, long long off, long long sy, long long sx, int ny, int nx, int nz,
                         unsigned *out_bits, hipStream_t s);
// crops: dst[c][y][x][z] = src[off + (oy_c + y) sy + (ox_c + x) sx + oz_c + z] / crop_scale[c],
// crop_scale[c] = max |crop c| (an all-zero crop stays zero)
hipError_t launch_crops(const float *src, long long off, long long sy, long long sx, const int *org, int ncrop,
                        int n, float *crop_scale, float *dst, hipStream_t s);

}  // namespace nn
}  // namespace vti
