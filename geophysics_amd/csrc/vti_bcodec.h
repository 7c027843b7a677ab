// vti_bcodec.h — error-bounded sparse block-transform codec for 3D source
// wavefield snapshots (the learned RTM's second codec).
//
// A snapshot box (n^3, n a multiple of 8) is cut into 8^3 blocks; every
// block is projected on an orthonormal 512 x 512 basis learned from earlier
// snapshots (the principal components of the autoencoder fit), and the
// smallest coefficients of each block are dropped while the block's error
// stays within tol of its own norm.  Because the basis is orthonormal the
// energy of the dropped coefficients is exactly the block's squared L2
// error (Parseval), so no decode is needed to bound it.  The bound is local
// (weak blocks far from the source keep their relative accuracy: the
// illumination-compensated image divides by the local source energy), with
// a floor: a block weaker than floor x the strongest block's energy is held
// to tol of that floor.
//
// Voxel order inside a block is natural: k = (ty*8 + tx)*8 + tz.
// A kept coefficient is one 32-bit entry: the top 23 bits of its fp32
// pattern (rounded; relative error <= 2^-15) | k in the low 9 bits.  Per block:
// the offset of its first entry (uint32, from the snapshot's base) and the
// number of entries (uint16).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vti {
namespace bc {

// Basis matrices on the device, bf16 hi/lo splits of the fp32 basis V
// (V[k][j]: voxel k, component j): vt_* = V^T (component rows), v_* = V,
// each 512 x 512 in fragment-major order (frag_off): one MFMA operand of a
// wavefront (16 rows x 32 columns; lane l: row l % 16, columns
// 8 (l / 16) .. + 8) is 1 KB contiguous, lane after lane.
struct Basis {
    const uint16_t *vt_h, *vt_l, *v_h, *v_l;
};

__host__ __device__ constexpr long long frag_off(int row, int col)
{
    return ((long long)(row >> 4) * 16 + (col >> 5)) * 512 + (((col >> 3) & 3) * 16 + (row & 15)) * 8 + (col & 7);
}

// Transform tiles of 64 blocks x 512 (the encode's coefficients, the
// decodes' voxels): LDS rows of LROW bf16 per block (hi and lo parts), a
// workgroup of 8 wavefronts, wavefront w owning output columns
// 64 w .. + 64 for all 64 blocks, so each basis element is loaded once per
// workgroup.
constexpr int kTileBlocks = 64;
constexpr int kTileWaves = 8;
constexpr int LROW = 512 + 8;   // LDS row (bf16), padded against bank conflicts

// Where a decoded tile goes: the dense n^3 box `out` (z fastest), or, when
// img is set, into an image of the same layout: img[o] += v * rp[off + y sy
// + x sx + z] for every voxel (y, x, z) (o its dense index), v the decoded
// value or, inside the box at (y0, x0, z0) of by x bx x bz voxels when box
// is set, box[((y - y0) bx + (x - x0)) bz + (z - z0)].
struct TileSink {
    float *out = nullptr;
    float *img = nullptr;
    const float *rp = nullptr;
    long long off = 0, sy = 0, sx = 0;
    const float *box = nullptr;
    int y0 = 0, x0 = 0, z0 = 0, by = 0, bx = 0, bz = 0;
};

#if defined(__HIPCC__)
typedef __bf16 tile_bf8 __attribute__((ext_vector_type(8)));
typedef float tile_f4 __attribute__((ext_vector_type(4)));
typedef unsigned short tile_us8 __attribute__((ext_vector_type(8)));

// acc[r][c] += X[16 r + .][k] M[64 w + 16 c + .][k] over k < 512, X = the
// LDS tile (hi, lo), M the fragment-major basis (hi, lo), as Xh Mh + Xl Mh +
// Xh Ml; the basis operands of the next K step are loaded before the MFMAs
// of the current one.  D layout of v_mfma_f32_16x16x32_bf16: acc[r][c][e]
// is row 16 r + 4 (l / 16) + e, column 64 w + 16 c + l % 16.
__device__ __forceinline__ void tile_mfma(const uint16_t (*xh)[LROW], const uint16_t (*xl)[LROW],
                                          const uint16_t *__restrict__ mh, const uint16_t *__restrict__ ml, int w,
                                          int l, tile_f4 acc[4][4])
{
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = tile_f4{0.0f, 0.0f, 0.0f, 0.0f};
    const int kl = 8 * (l >> 4);
    tile_us8 bh[4], bl[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = 64 * w + 16 * c + (l & 15);
        bh[c] = *reinterpret_cast<const tile_us8 *>(mh + frag_off(j, kl));
        bl[c] = *reinterpret_cast<const tile_us8 *>(ml + frag_off(j, kl));
    }
    for (int kk = 0; kk < 512; kk += 32) {
        const int kc = kk + kl;
        const int kn = (kk + 32 < 512 ? kk + 32 : kk) + kl;   // next K step (the last reloads its own)
        tile_us8 nh[4], nl[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int j = 64 * w + 16 * c + (l & 15);
            nh[c] = *reinterpret_cast<const tile_us8 *>(mh + frag_off(j, kn));
            nl[c] = *reinterpret_cast<const tile_us8 *>(ml + frag_off(j, kn));
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const tile_us8 ah = *reinterpret_cast<const tile_us8 *>(&xh[16 * r + (l & 15)][kc]);
            const tile_us8 al = *reinterpret_cast<const tile_us8 *>(&xl[16 * r + (l & 15)][kc]);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(tile_bf8, ah),
                                                                    __builtin_bit_cast(tile_bf8, bh[c]), acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(tile_bf8, al),
                                                                    __builtin_bit_cast(tile_bf8, bh[c]), acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(tile_bf8, ah),
                                                                    __builtin_bit_cast(tile_bf8, bl[c]), acc[r][c], 0, 0, 0);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            bh[c] = nh[c];
            bl[c] = nl[c];
        }
    }
}

// The decoded tile (acc as tile_mfma leaves it, blocks b0 .. b0 + 63 of the
// n^3 box, z fastest) to the sink.  Staged in LDS over the operand tiles
// (lds: their 2 x 64 x LROW bf16 = 64 x LROW fp32; every wavefront must be
// past its products, the caller's barrier), then stored as 16-byte pieces of
// 8-voxel z lines: one wavefront instruction writes both halves of one line
// of 32 consecutive blocks, contiguous along z within a row of blocks.
__device__ __forceinline__ void tile_store(uint16_t *lds, const tile_f4 acc[4][4], int w, int l, long long b0, int n,
                                           const TileSink &sk)
{
    float(*T)[LROW] = reinterpret_cast<float(*)[LROW]>(lds);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) T[16 * r + 4 * (l >> 4) + e][64 * w + 16 * c + (l & 15)] = acc[r][c][e];
    __syncthreads();
    const int nb = n / 8;
    const long long nblk = (long long)nb * nb * nb;
    float *const dst = sk.img ? sk.img : sk.out;
    const bool vec = (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
    // the receiver field in 16-byte pieces too when every z line's start is
    // 16-byte aligned (the dense snapshot copy; not a padded field)
    const bool rvec = sk.img && (reinterpret_cast<uintptr_t>(sk.rp + sk.off) & 15) == 0 && (sk.sy & 3) == 0 &&
                      (sk.sx & 3) == 0;
    for (int it = threadIdx.x; it < kTileBlocks * 128; it += 64 * kTileWaves) {
        const int half = it & 1, blk = (it >> 1) & (kTileBlocks - 1), line = it >> 7;
        const long long b = b0 + blk;
        if (b >= nblk) continue;
        const int bz = (int)(b % nb), bx = (int)((b / nb) % nb), by = (int)(b / ((long long)nb * nb));
        const int y = 8 * by + (line >> 3), x = 8 * bx + (line & 7), z = 8 * bz + 4 * half;
        tile_f4 v = *reinterpret_cast<const tile_f4 *>(&T[blk][8 * line + 4 * half]);
        float *q = dst + ((long long)y * n + x) * n + z;
        if (sk.img) {
            if (sk.box && y >= sk.y0 && y < sk.y0 + sk.by && x >= sk.x0 && x < sk.x0 + sk.bx) {
                const float *bl = sk.box + ((long long)(y - sk.y0) * sk.bx + (x - sk.x0)) * sk.bz - sk.z0;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (z + u >= sk.z0 && z + u < sk.z0 + sk.bz) v[u] = bl[z + u];
            }
            const float *rp = sk.rp + sk.off + y * sk.sy + x * sk.sx + z;
            const tile_f4 g = rvec ? *reinterpret_cast<const tile_f4 *>(rp) : tile_f4{rp[0], rp[1], rp[2], rp[3]};
            tile_f4 m = vec ? *reinterpret_cast<const tile_f4 *>(q) : tile_f4{q[0], q[1], q[2], q[3]};
#pragma unroll
            for (int u = 0; u < 4; ++u) m[u] = m[u] + v[u] * g[u];
            v = m;
        }
        if (vec) {
            *reinterpret_cast<tile_f4 *>(q) = v;
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = v[u];
        }
    }
}
#endif

// Box of the snapshot (box coordinates) that is zeroed before encoding.
struct Mask {
    int y0, x0, z0, by, bx, bz;
};

// Device scratch of one encoder: dense coefficients and the counters of the
// last encode (strongest block energy, entries, overflow flag, energy and
// dropped energy).
struct Counters {
    unsigned emax_bits;
    unsigned cursor;
    unsigned overflow;
    unsigned pad;
    double energy;
    double dropped;
};
struct Work {
    float *coef = nullptr;   // [nblk][512]
    Counters *cnt = nullptr;
    long long cap_blocks = 0;
    void *aux = nullptr;     // the quantized codec's per-block search results and scan sums
    long long aux_cap = 0;   // bytes
};

int work_reserve(Work &w, int n);
void work_free(Work &w);

// Host-side: bf16 hi/lo splits of V (512 x 512, row-major V[k][j]) into the
// 4 x 512 x 512 uint16 layout Basis points into (vt_h, vt_l, v_h, v_l).
void basis_split(const float *V, uint16_t *out);
Basis basis_view(const uint16_t *dev);

// Coefficients of the n^3 box at P[off + y*sy + x*sx + z] (mask zeroed) into
// w.coef; counters cleared, the strongest block's energy recorded.
hipError_t launch_encode(const Basis &V, const float *P, long long off, long long sy, long long sx, int n,
                         const Mask &m, const Work &w, hipStream_t s);
// Per block: drop the smallest coefficients while their energy stays within
// tol2 x max(block energy, flo x strongest block energy); entries of the
// rest at ent + the block's range (blocks whose range would pass cap write
// nothing and set the overflow flag).
hipError_t launch_pack(const Work &w, int n, float tol2, float flo, uint32_t *ent, long long cap, uint32_t *boff,
                       uint16_t *bcnt, hipStream_t s);
// Dense n^3 box (z fastest) from the entries, or the decoded box into the
// sink (TileSink: a dense box, or an image accumulation).
hipError_t launch_decode(const Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         float *out, hipStream_t s);
hipError_t launch_decode(const Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         const TileSink &sk, hipStream_t s);

// Whole encode of one snapshot: kernels, counter read-back (host sync).
// Every block's relative L2 error (to the floor above) is within tol: the
// dropped energy plus 2^-26 of the budget set aside for the value rounding
// and the transforms, so tol <= 2^-13 keeps every non-zero coefficient.  *kept =
// entries needed; when *kept > cap the entries are incomplete (the caller
// falls back).  *err = the bound on the box's relative L2 error; *energy =
// the box's energy.
int encode(const Basis &V, const float *P, long long off, long long sy, long long sx, int n, const Mask &m,
           float tol, float flo, Work &w, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt,
           long long *kept, double *err, double *energy, hipStream_t s);
// The same in a Work the caller has sized (no allocation): VTI_ERR_ARG when
// w holds fewer than (n/8)^3 blocks or no counters.
int encode_in(const Basis &V, const float *P, long long off, long long sy, long long sx, int n, const Mask &m,
              float tol, float flo, const Work &w, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt,
              long long *kept, double *err, double *energy, hipStream_t s);
// encode_in without the host sync: the kernels on s, `read` (when not null)
// recorded once the snapshot has been read (after the projection, before the
// pack), the counters copied into *hc (pinned host memory: the copy is
// asynchronous), to be turned into kept / err / energy by encode_finish once
// the stream has passed the copy.
int encode_start(const Basis &V, const float *P, long long off, long long sy, long long sx, int n, const Mask &m,
                 float tol, float flo, const Work &w, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt,
                 hipEvent_t read, Counters *hc, hipStream_t s);
void encode_finish(const Counters &h, long long cap, long long *kept, double *err, double *energy);

}  // namespace bc
}  // namespace vti
