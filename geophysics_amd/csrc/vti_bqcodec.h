// vti_bqcodec.h — quantized sparse block-transform codec (codec 2 of the
// learned 3D RTM): the sparse codec's projection on the fitted block basis,
// then a per-block uniform quantizer whose step is the largest that keeps
// the block's exact quantization error within the same budget as codec 1,
// stored as a bit stream (step, group widths, bitmap, packed integers).
// Layout and kernels: vti_bqcodec.hip.
#pragma once

#include "vti_bcodec.h"

namespace vti {
namespace bq {

// words of one stored block at most: 19 header and bitmap words + 512 raw values
constexpr int kMaxBlockWords = 19 + 512;

// Encode of the n^3 box at P[off + y*sy + x*sx + z] (mask zeroed): block
// streams at ent + boff[b] (bcnt[b] words, 0 for a block within its budget
// at q = 0); *kept = words needed (> cap: incomplete, nothing written past
// cap); *err = the bound on the box's relative L2 error.  Host sync.
int encode(const bc::Basis &V, const float *P, long long off, long long sy, long long sx, int n,
           const bc::Mask &m, float tol, float flo, bc::Work &w, uint32_t *ent, long long cap, uint32_t *boff,
           uint16_t *bcnt, long long *kept, double *err, hipStream_t s);
// encode without the host sync, as bc::encode_start: `read` (when not null)
// recorded once the snapshot has been read, the counters copied into *hc
// (pinned host memory), bc::encode_finish (energy unused) once the stream
// has passed the copy.
int encode_start(const bc::Basis &V, const float *P, long long off, long long sy, long long sx, int n,
                 const bc::Mask &m, float tol, float flo, bc::Work &w, uint32_t *ent, long long cap, uint32_t *boff,
                 uint16_t *bcnt, hipEvent_t read, bc::Counters *hc, hipStream_t s);
// Dense n^3 box (z fastest) from the block streams, or the decoded box into
// the sink (a dense box, or an image accumulation).
hipError_t launch_decode(const bc::Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         float *out, hipStream_t s);
hipError_t launch_decode(const bc::Basis &V, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                         const bc::TileSink &sk, hipStream_t s);

}  // namespace bq
}  // namespace vti
