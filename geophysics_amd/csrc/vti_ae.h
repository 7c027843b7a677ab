// vti_ae.h — internal entry points of the snapshot autoencoder used by the
// learned-compression RTM (vti_rtm_learned.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/vti.h"
#include "vti_bcodec.h"
#include "vti_bqcodec.h"

namespace vti {

// Encode the n^3 box at P[off + y*sy + x*sx + z] (fp32) into latent
// ((n/8)^3 x 64 bf16); *scale_dev receives the box's largest |P|.
int ae_encode(vti_ae *A, const float *P, long long off, long long sy, long long sx, int n, uint16_t *latent,
              float *scale_dev, hipStream_t s);
// Decode into a dense n^3 fp32 box (z fastest), rescaled by *scale_dev.
int ae_decode(vti_ae *A, const uint16_t *latent, const float *scale_dev, int n, float *out, hipStream_t s);
// nsteps Adam steps on random crops of the (ny, nx, nz) box of P.
int ae_train_field(vti_ae *A, const float *P, long long off, long long sy, long long sx, int ny, int nx, int nz,
                   int nsteps, float *loss);
hipStream_t ae_stream(vti_ae *A);
// Principal-component fit: sample nblk random 8^3 blocks of the n^3 box
// (scaled by its max |P|) into the Gram matrix; ae_fit_end sets the weights.
int ae_fit_field(vti_ae *A, const float *P, long long off, long long sy, long long sx, int n, int nblk);
int ae_fit_end(vti_ae *A, double *explained);
// Sparse block-transform codec with the autoencoder's full basis (set by
// ae_fit_end or ae_set_basis; V[k][j] host, 512 x 512, natural voxel order).
int ae_set_basis(vti_ae *A, const float *V);
int ae_bc_encode(vti_ae *A, const float *P, long long off, long long sy, long long sx, int n, const bc::Mask &m,
                 float tol, float flo, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt, long long *kept, double *err,
                 double *energy, hipStream_t s);
int ae_bc_decode(vti_ae *A, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n, float *out,
                 hipStream_t s);
// Quantized codec (vti_bqcodec.h) with the same basis: block streams in
// 32-bit words, at most bq::kMaxBlockWords per block.
int ae_bq_encode(vti_ae *A, const float *P, long long off, long long sy, long long sx, int n, const bc::Mask &m,
                 float tol, float flo, uint32_t *ent, long long cap, uint32_t *boff, uint16_t *bcnt, long long *kept,
                 double *err, hipStream_t s);
int ae_bq_decode(vti_ae *A, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n, float *out,
                 hipStream_t s);
// Device bytes of the codecs' work buffers (coefficients, counters, the
// quantized codec's search results) as currently allocated.
long long ae_codec_work_bytes(const vti_ae *A);
// The two decodes into a sink (bc::TileSink: here the image accumulation
// of the decoded snapshot times the receiver field, the exact box included).
int ae_bc_decode_sink(vti_ae *A, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                      const bc::TileSink &sk, hipStream_t s);
int ae_bq_decode_sink(vti_ae *A, const uint32_t *ent, const uint32_t *boff, const uint16_t *bcnt, int n,
                      const bc::TileSink &sk, hipStream_t s);
// The two encodes without their host sync (bc::encode_start,
// bq::encode_start): `read` recorded once P has been read, the counters in
// *hc (pinned) for bc::encode_finish once s has passed their copy.
int ae_bc_encode_start(vti_ae *A, const float *P, long long off, long long sy, long long sx, int n,
                       const bc::Mask &m, float tol, float flo, uint32_t *ent, long long cap, uint32_t *boff,
                       uint16_t *bcnt, hipEvent_t read, bc::Counters *hc, hipStream_t s);
int ae_bq_encode_start(vti_ae *A, const float *P, long long off, long long sy, long long sx, int n,
                       const bc::Mask &m, float tol, float flo, uint32_t *ent, long long cap, uint32_t *boff,
                       uint16_t *bcnt, hipEvent_t read, bc::Counters *hc, hipStream_t s);

}  // namespace vti

extern "C" int vti_ae_set_params_padded_(vti_ae *A, const float *h);
