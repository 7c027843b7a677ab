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
// Dense n^3 box (z fastest) from the entries.
hipError_t launch_decode(const Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         float *out, hipStream_t s);

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

}  // namespace bc
}  // namespace vti
