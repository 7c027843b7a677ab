/*
 * vti.h — C ABI of the MI355X (gfx950) first-order staggered-grid
 * pseudo-acoustic VTI finite-difference path (libvti.so).
 *
 * Plain C types only: ints, floats, pointers and sizes.  Streams are
 * passed as `void*` holding a hipStream_t (NULL = the context's own
 * stream).  Every call returns an int status (VTI_OK = 0, negative codes
 * below); nothing calls exit().
 *
 * Array layout everywhere is the reference's: z fastest, then x, then y
 * (2D: id = iz + ix*nz; 3D: id = iz + ix*nz + iy*nz*nx).  Shot records are
 * [irec][it] (t fastest), receivers ordered x fastest then y.
 *
 * Citations are path:line in the reference repository
 * (liaman/Geophysics).
 */
#ifndef VTI_H
#define VTI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define VTI_OK          0
#define VTI_ERR_ARG    -1   /* invalid argument or configuration        */
#define VTI_ERR_ALLOC  -2   /* device or host allocation failed         */
#define VTI_ERR_HIP    -3   /* HIP runtime error                        */
#define VTI_ERR_IO     -4   /* file open / read / write failed          */
#define VTI_ERR_CFL    -5   /* stability / dispersion check failed      */
#define VTI_ERR_STATE  -6   /* call out of order (no model, no shot...) */
#define VTI_ERR_NODEV  -7   /* no HIP device                            */

/* ---- semantics ------------------------------------------------------- */
/* The reference's GPU and CPU drivers compute different seismograms
 * (source footprint and phase, damping strength, stress-table form,
 * receiver scale).  The stencil kernels are the same for both; the mode
 * selects host-built tables and the fused source/receiver work. */
#define VTI_MODE_GPU_REF 0  /* GPU_vti2dfd_kernels.cu / GPU_vti3dfd_kernels.cu */
#define VTI_MODE_CPU_REF 1  /* CPU_vti2dfd_kernels.c                            */
/* 2D-vti-fd-omp/Toafd2domp_2.c (2D only): the CPU form's in-stress 7x7
 * Gaussian source, no absorbing layer (every damping factor 1), the first
 * time sample at t = 0 (:144), pfac 1000 (:70), its Ricker product order
 * (wtype 4, :154), the pz and qx updates in double precision (pow promotes
 * them, :176-180), P+Q recorded (:219). */
#define VTI_MODE_OMP_REF 2

#define VTI_REC_DEFAULT -1  /* mode default: P (GPU), P+P (CPU)                 */
#define VTI_REC_P        0  /* GPU_vti2dfd_kernels.cu:147                       */
#define VTI_REC_2P       1  /* CPU_vti2dfd_kernels.c:490                        */
#define VTI_REC_PQ       2  /* Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:331   */

typedef struct vti_cfg {
    int ndim;          /* 2 or 3                                            */
    int nx, ny, nz;    /* model sizes in cells (ny ignored in 2D)           */
    int npml;          /* absorbing-layer width in cells (npd / npml)       */
    float dx, dy, dz;  /* grid spacing, metres                              */
    float dt;          /* time step, seconds                                */
    int nt;            /* number of time steps                              */
    float favg;        /* wavelet peak frequency, Hz                        */
    float pfac;        /* source gain (reference: 10)                       */
    int wtype;         /* 1 Ricker, 2 / 3 Gaussian forms, 4 Ricker with the
                          product order of Toafd2domp_2.c:154               */
    int mode;          /* VTI_MODE_*                                        */
    int sv;            /* taper eps/delta around the source (SV flag)       */
    int sv_R, sv_r;    /* taper radii in cells; 0 = mode default            */
    int rec_kind;      /* VTI_REC_*                                         */
} vti_cfg;

/* One shot: source cell, taper centre and a regular receiver grid, all
 * as padded-grid indices (0-based, npml included). */
typedef struct vti_shot {
    int sx, sy, sz;            /* source cell                               */
    int sv_cx, sv_cy, sv_cz;   /* SV-taper centre                           */
    int rec_z;                 /* receiver depth row                        */
    int rec_x0, rec_dx, rec_nx;
    int rec_y0, rec_dy, rec_ny;/* ignored in 2D                             */
} vti_shot;

typedef struct vti_ctx vti_ctx;

/* Field selectors for vti_get_field / vti_set_field (padded volumes). */
enum {
    VTI_F_U = 0, VTI_F_V, VTI_F_W, VTI_F_PX, VTI_F_PY, VTI_F_PZ,
    VTI_F_QX, VTI_F_QY, VTI_F_QZ, VTI_F_P, VTI_F_Q,
    VTI_F_VP, VTI_F_EPS, VTI_F_DEL
};
/* Table selectors for vti_get_table (damping tables along one axis). */
enum {
    VTI_T_COFFX1 = 20, VTI_T_COFFX2, VTI_T_COFFY1, VTI_T_COFFY2,
    VTI_T_COFFZ1, VTI_T_COFFZ2, VTI_T_ACOFFX1, VTI_T_ACOFFX2,
    VTI_T_ACOFFY1, VTI_T_ACOFFY2, VTI_T_ACOFFZ1, VTI_T_ACOFFZ2,
    VTI_T_SRCAMP = 40, VTI_T_D0 = 41
};

/* ---- configuration helpers: the reference drivers' index conversions -- */
/* cfg for GPU_vti2dfd (GPU_vti2dfd_kernels.cu:301-352) or CPU_vti2dfd
 * (CPU_vti2dfd_kernels.c:366-390) arguments; dt_us in microseconds. */
int vti_cfg_ref2d(vti_cfg *cfg, int mode, int nx, int nz, int dx, int dz,
                  int npd, int SV, int favg, int nt, int dt_us);
/* cfg for GPU_vti3dfd (GPU_vti3dfd_kernels.cu:371-413). */
int vti_cfg_ref3d(vti_cfg *cfg, int nx, int ny, int nz, int dx, int dy,
                  int dz, int npml, int SV, int favg, int nt, int dt_us);
/* Shot `is` (1-based, as the 2D reference loop) of a 2D run: source
 * (GPU_vti2dfd_kernels.cu:185,202-203 / CPU_vti2dfd_kernels.c:61-66),
 * taper centre (:260-261 / :137-138), receivers at every x at 1-based
 * depth rec_depth (reference: 1 = row npd, :147). */
int vti_shot_ref2d(vti_shot *shot, const vti_cfg *cfg, int is, int fs,
                   int ds, int zs, int rec_depth);
/* Shot `is` (0-based) of a 3D run (GPU_vti3dfd_kernels.cu:63-65,82-84,
 * 155-157); receivers every rec_step cells in x and y at depth rec_depth
 * (reference: every cell at row npml, :325). */
int vti_shot_ref3d(vti_shot *shot, const vti_cfg *cfg, int is, int fsx,
                   int dsx, int fsy, int dsy, int fsz, int dsz,
                   int rec_depth, int rec_step);

/* Host-only table construction (no device needed): d0 and the 12 damping
 * tables (rows of max(nnx,nny,nnz) floats in the order of the VTI_T_COFF*
 * selectors) for an unpadded vp, exactly as vti_set_model builds them.
 * get_d02d GPU_vti2dfd_kernels.cu:102-106, initial_coffe2d :108-139,
 * get_constant / initial_coffe CPU_vti2dfd_kernels.c:179-332. */
int vti_host_tables(const vti_cfg *cfg, const float *vp, const float *eps,
                    float *tables, float *d0);
/* Host-only per-step source gains pfac*wavelet(t), nt floats
 * (add_source2d GPU_vti2dfd_kernels.cu:183-206). */
int vti_host_source_gains(const vti_cfg *cfg, float *amp);

/* ---- context lifecycle ------------------------------------------------ */
/* Replaces the allocation / table set-up blocks of GPU_vti2dfd
 * (GPU_vti2dfd_kernels.cu:354-398) and GPU_vti3dfd
 * (GPU_vti3dfd_kernels.cu:429-484) with a per-context, re-entrant form. */
int vti_create(const vti_cfg *cfg, vti_ctx **out);
int vti_destroy(vti_ctx *ctx);
int vti_set_stream(vti_ctx *ctx, void *hip_stream);
/* Model without padding (nx*nz or nx*ny*nz floats, z fastest).  on_device:
 * 0 = host pointers, 1 = device pointers.  Pads by edge replication
 * (pad_vv2d GPU_vti2dfd_kernels.cu:49-74, pad_vv3d GPU_vti3dfd_kernels.cu:208-255)
 * and builds d0 and the damping tables (get_d02d :102-106,
 * initial_coffe2d :108-139; CPU form CPU_vti2dfd_kernels.c:179-332). */
int vti_set_model(vti_ctx *ctx, const float *vp, const float *eps,
                  const float *del, int on_device);
/* Select the shot: source, taper centre, receivers; builds the per-step
 * source table (add_source2d :176-207 / ptsource CPU_vti2dfd_kernels.c:50-76). */
int vti_set_shot(vti_ctx *ctx, const vti_shot *shot);
/* Zero wavefields and the receiver buffer (the per-shot cudaMemset block,
 * GPU_vti2dfd_kernels.cu:407-417). */
int vti_reset(vti_ctx *ctx);

/* ---- per-time-step calls (GPU_vti2dfd_kernels.cu:423-429) -------------- */
/* Velocity update of step it; in GPU mode the point source of step it is
 * injected into P and Q first (add_source2d + update_vel2d). */
int vti_update_vel(vti_ctx *ctx, int it);
/* Stress / pressure update of step it with damping, the in-stress source
 * (CPU mode) and the receiver gather of row it (update_stress2d +
 * shot_record2d, :425-429). */
int vti_update_stress(vti_ctx *ctx, int it);
/* nsteps full steps starting at it0. */
int vti_step(vti_ctx *ctx, int it0, int nsteps);
/* reset + all nt steps. */
int vti_run(vti_ctx *ctx);

/* ---- outputs ------------------------------------------------------------ */
int vti_nrec(const vti_ctx *ctx);
/* Direct-wave mute of the receiver buffer: mute_directwave2d
 * (GPU_vti2dfd_kernels.cu:151-174, tt=100 at :439) / mute_directwave3d
 * (GPU_vti3dfd_kernels.cu:342-369).  fs,ds,zs,is as the shot helpers. */
int vti_mute(vti_ctx *ctx, int fs, int ds, int zs, int is, int tt);
int vti_mute3d(vti_ctx *ctx, int fsx, int dsx, int fsy, int dsy, int fsz,
               int is, int tt);
/* Shot record [irec][it] into host memory (nrec*nt floats). */
int vti_get_record(vti_ctx *ctx, float *dst);
/* Device pointer of the receiver buffer, layout [it][irec]. */
float *vti_record_device(vti_ctx *ctx);
/* Padded field / table copies (host memory).  Fields: nnx*nnz (2D) or
 * nnx*nny*nnz floats in the reference layout. */
int vti_get_field(vti_ctx *ctx, int which, float *dst);
int vti_set_field(vti_ctx *ctx, int which, const float *src);
int vti_get_table(vti_ctx *ctx, int which, float *dst);
/* Padded sizes (nnx, nny, nnz; nny = 1 in 2D). */
int vti_padded_dims(const vti_ctx *ctx, int *nnx, int *nny, int *nnz);

/* ---- reverse-time migration -------------------------------------------- */
/* Per-shot RTM on the context's model and shot: forward source wavefield
 * (record + illumination), receiver back-propagation of the observed record
 * and zero-lag cross-correlation imaging, following RTM()
 * 2D-vti-fd+rtm-cuda/Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1184-1566 and the 3D
 * main Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:686-1041.  The source
 * wavefield of every imaged step is the exact forward field, recomputed
 * from nsnap full-state checkpoints with the binomial (revolve) schedule;
 * the reference's boundary-ring reconstruction (:1423-1437) is replaced. */
typedef struct vti_rtm_opts {
    int sv_forward;    /* SV taper in the source propagation (reference: on)  */
    int sv_backward;   /* SV taper in the receiver propagation (reference: off) */
    int mute_tt;       /* windowed direct-wave mute of the modelled record when
                          no observed record is given; <0 = none (2D ref: 20) */
    int inject_z;      /* receiver injection row, padded; <0 = the record row  */
    int illum_mode;    /* 0: source energy, <=0 clamped to 1 while accumulating,
                          divide (2D ref :654-677, :700-715); 1: source +
                          receiver energy, divide where nonzero (3D ref)       */
    int image_every;   /* cross-correlate every k-th step                      */
    int nsnap;         /* checkpoint slots of the source state; 0 = one per
                          step (vti_rtm_checkpoint_bytes each)                 */
} vti_rtm_opts;
int vti_rtm_defaults(vti_rtm_opts *opts, int ndim);
/* Zero the accumulated image and illumination (interior nx*ny*nz). */
int vti_rtm_reset(vti_ctx *ctx);
/* Device bytes of one checkpoint slot of vti_rtm_shot under the current
 * settings: a full state, or, with the reduced split of vti_set_unsplit,
 * u, v, w, P, Q whole and the split fields of the absorbing layer's cells
 * only (the interior never reads them); 0 for a NULL context. */
long long vti_rtm_checkpoint_bytes(vti_ctx *ctx);
/* One shot; obs = observed record [irec][it] in host memory, or NULL to
 * migrate the modelled (muted) record.  *fwd_steps = forward steps taken.
 * A checkpointed state is imaged from its slot (not restored first). */
int vti_rtm_shot(vti_ctx *ctx, const vti_rtm_opts *opts, const float *obs,
                 long long *fwd_steps);
/* Image and illumination to host memory; finalize = divide by the
 * illumination and apply the adjoint Laplacian filter (:1486-1493). */
int vti_rtm_image(vti_ctx *ctx, int finalize, float *img, float *illum);
/* Windowed 3D direct-wave mute of the RTM main
 * (Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:353-381). */
int vti_mute3d_window(vti_ctx *ctx, int fsx, int dsx, int fsy, int dsy,
                      int fsz, int is, int tt);
/* Poynting-vector angle-domain common-image gathers (ADCIGs) of the RTM:
 * Poynting_Adcigs Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:718-766 (2D, every
 * cdp) and Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:618-660 (3D, every
 * dcdp-th cdp in x and y).  With nangle > 0, every imaging step of
 * vti_rtm_shot (and vti_adcig_accumulate) adds, per cdp cell, P_s P_g
 * cos^3(ia) to angle bin ia = (int)(half the angle between the source
 * Poynting vector (P u, [P v,] Q w) and the receiver one (P u, [P v,] Q w),
 * in degrees / dangle) of that cell's gather (the reference negates the
 * source vector of its backward-rebuilt source field, :746-747; the source
 * field here is the exact forward one), the cosine scaled by
 * 1/(1 - 0.1) as there (cells whose scaled cosine leaves [-1, 1] add
 * nothing).  Gathers [iya][ixa][ia][iz] (2D: [ix][ia][iz]), nx/dcdp x
 * ny/dcdp cdps in 3D; vti_rtm_reset zeroes them.  nangle = 0 switches
 * them off.  The learned RTM (vti_rtm_shot_learned) keeps only P of the
 * source field and refuses to run while gathers are configured. */
int vti_adcig_config(vti_ctx *ctx, int nangle, int dangle, int dcdp);
/* Number of gather samples (0 when off). */
long long vti_adcig_elems(const vti_ctx *ctx);
/* Host-only: the bin thresholds the kernel compares the scaled cosine a
 * with (thresholds[k-1]: the largest float a whose bin, by the reference's
 * arithmetic with the host's acosf, is at least k; -2 when no a reaches k)
 * and the cos^3 weight factors cosf(ia pi / 180), nangle each. */
int vti_adcig_tables(int nangle, int dangle, float *thresholds, float *weights);
int vti_adcig_reset(vti_ctx *ctx);
/* One accumulation from the current fields of a source-side and a
 * receiver-side context of the same grid into the receiver context's
 * gathers (ordered like vti_image_accumulate). */
int vti_adcig_accumulate(vti_ctx *src, vti_ctx *rcv);
/* The gathers into host memory; illum_divide: divided by the RTM's
 * illumination of each gather's cdp (adcigs_illum :768-786, 3D where
 * nonzero :663-684). */
int vti_adcig_get(vti_ctx *ctx, int illum_divide, float *adcigs);
/* Host post-processing (2D layout [ix][ia][iz]): the stack over angles then
 * the adjoint Laplacian (stk_adcigs :788-816), and the gathers of every
 * dAdcigs-th cdp, packed ((nx + dAdcigs - 1) / dAdcigs of them,
 * adcigs_smiled :818-847). */
int vti_adcig_stack(int nx, int nz, int nangle, const float *adcigs, float *migration);
int vti_adcig_thin(int nx, int nz, int nangle, int dAdcigs, const float *adcigs, float *out);

/* The checkpoint schedule of vti_rtm_shot / vti_batch_rtm / vti_grad_run
 * run on positions instead of wavefields (host only, no device needed):
 * VTI_OK when the reverse sweep receives X_nt .. X_1 in order with at most
 * nslot slots; *fwd_steps the forward steps it takes, *max_slots the slots
 * it touched.  Replaces the reference's boundary re-injection
 * (Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1423-1437), which has no schedule. */
int vti_sweep_check(int nt, int nslot, long long *fwd_steps, int *max_slots);

/* ---- batches of independent 2D shots ------------------------------------- */
/* Many 2D shots, each with its own grid, model and geometry (GPU form),
 * advanced together: one persistent work-list launch per phase covers every
 * tile of every shot; the time index lives on the device so a block of steps
 * is a replayable hipGraph; the tile shape and grid size come from an
 * autotuner.  Replaces the reference's serial shot loop with ~15 memsets per
 * shot (GPU_vti2dfd_kernels.cu:404-444) for the RTM of
 * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1350-1484.  Every shot's record and
 * image equal the single-shot path's bit for bit. */
typedef struct vti_batch vti_batch;
/* cfgs/shots/vp/eps/del: nshot entries; models unpadded, host memory. */
int vti_batch_create(int nshot, const vti_cfg *cfgs, const vti_shot *shots,
                     const float *const *vp, const float *const *eps,
                     const float *const *del, vti_batch **out);
int vti_batch_destroy(vti_batch *b);
/* tile_shape: 0..4 = 64x4, 64x2, 64x8, 128x2, 32x8 threads (z x x), one cell
 * each; 5..9 = 64x4, 64x4, 64x2, 32x8, 128x2 threads, each marching 16, 32,
 * 32, 16, 16 consecutive rows in x with the x-stencil in registers
 * (vti_batch_get_config's tile_x: the tile's extent in x); grid_mult:
 * persistent grid of grid_mult*256 workgroups (0 = one per tile);
 * graph_steps: steps per captured hipGraph block (0/1 = direct launches). */
int vti_batch_config(vti_batch *b, int tile_shape, int grid_mult, int graph_steps);
int vti_batch_get_config(vti_batch *b, int *tile_z, int *tile_x, int *grid_mult,
                         int *graph_steps);
/* Times `steps` forward steps for every shape and grid size, keeps the
 * fastest; returns the chosen shape index, *best_ms per step. */
int vti_batch_autotune(vti_batch *b, int steps, double *best_ms);
/* Updated cells per step summed over the shots. */
long long vti_batch_cells(const vti_batch *b);
/* Every shot's propagation with the unsplit interior, with the same rule as
 * the single-shot RTM (see vti_set_unsplit): both sides, the receiver
 * side's injection row in the split form.  Point-source form only; default
 * off. */
int vti_batch_set_unsplit(vti_batch *b, int on);
/* nsteps forward steps of every shot from rest; record: keep the records. */
int vti_batch_forward(vti_batch *b, int nsteps, int record);
/* RTM of all shots in lockstep (source illumination, illum_mode 0). */
int vti_batch_rtm(vti_batch *b, const vti_rtm_opts *opts, const float *const *obs,
                  long long *fwd_steps);
int vti_batch_image(vti_batch *b, int shot, int finalize, float *img, float *illum);
int vti_batch_record(vti_batch *b, int shot, float *dst);

/* ---- learned snapshot compression ---------------------------------------- */
/* A 4-level 3D convolutional autoencoder (BF16 weights and activations,
 * FP32 accumulation on MFMA, FP32 master weights and Adam moments) that
 * stores source-wavefield snapshots in 16x fewer bytes than FP32.  It has no
 * counterpart in the reference, whose RTM keeps no snapshots and re-runs the
 * forward operator from the final state
 * (Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1423-1437).  Layers: down-sampling
 * 2x2x2 convolutions 1->8->64->64, a 3x3x3 convolution 64->64 (latent),
 * a 3x3x3 convolution 64->64 and transposed 2x2x2 convolutions
 * 64->64->8->1; leaky ReLU (slope 1/8) between layers; no biases (the
 * network is positively homogeneous: zero in, zero out). */
typedef struct vti_ae vti_ae;
/* crop: side of the training crops (multiple of 8); batch: crops per step. */
int vti_ae_create(int crop, int batch, unsigned long long seed, vti_ae **out);
int vti_ae_destroy(vti_ae *ae);
/* Parameter count of the flat form: per layer its weight matrix (rows x
 * cols, row-major; down/same: rows = out channels, cols = tap*in + c;
 * up: rows = tap*out + c, cols = in channels). */
long long vti_ae_nparams(const vti_ae *ae);
int vti_ae_set_params(vti_ae *ae, const float *flat);
int vti_ae_get_params(vti_ae *ae, float *flat);
/* Gradients of the last training step, flat form. */
int vti_ae_get_grads(vti_ae *ae, float *flat);
int vti_ae_set_lr(vti_ae *ae, float lr);
/* One step on host crops (batch x crop^3 fp32): MSE loss, gradients and,
 * when apply != 0, an Adam update.  recon (optional, host) receives the
 * step's reconstruction. */
int vti_ae_train_batch(vti_ae *ae, const float *crops, int apply, float *loss, float *recon);
/* Stored output of layer i-1 (the input of layer i) of the last step, as
 * fp32 host values; i = 8: the reconstruction. */
int vti_ae_activation(vti_ae *ae, int i, float *dst);
/* nsteps Adam steps on random crops of the (ny, nx, nz) box of a device
 * field P[off + y*sy + x*sx + z], each crop scaled by its own max |P| (the
 * network is positively homogeneous, so the scale does not change what it
 * learns, only how much every crop weighs). */
int vti_ae_train_field(vti_ae *ae, const float *P, long long off, long long sy, long long sx,
                       int ny, int nx, int nz, int nsteps, float *loss);
/* Device-side encode of an n^3 box (n a multiple of 8) into latent
 * (vti_ae_latent_elems(n) bf16 values) and its scale (one float). */
int vti_ae_encode(vti_ae *ae, const float *P, long long off, long long sy, long long sx, int n,
                  void *latent, float *scale_dev);
int vti_ae_decode(vti_ae *ae, const void *latent, const float *scale_dev, int n, float *out);
long long vti_ae_latent_elems(int n);
/* Principal-component fit (the closed-form optimum of the linear
 * autoencoder): nblocks random 8^3 blocks of an n^3 box of a device field,
 * scaled by the box's max |P|, go into a Gram matrix (BF16 MFMA, FP32
 * sums); vti_ae_fit_end sets the encoder to the projection on the 64
 * leading components, the decoder to the reconstruction, the latent
 * convolutions to identity; *explained = captured energy fraction. */
int vti_ae_fit_field(vti_ae *ae, const float *P, long long off, long long sy, long long sx, int n,
                     int nblocks);
int vti_ae_fit_end(vti_ae *ae, double *explained);
/* The fit also keeps the full orthonormal 512 x 512 basis of 8^3 blocks
 * (V[k][j], k = (ty*8 + tx)*8 + tz the voxel, j the component, leading
 * components first) for the sparse codec below; set_basis replaces it
 * (any orthonormal V, host, row-major). */
int vti_ae_set_basis(vti_ae *ae, const float *V);
int vti_ae_get_basis(vti_ae *ae, float *V);
/* Sparse block-transform codec (device pointers; n^3 box at
 * P[off + y*sy + x*sx + z], n a multiple of 8).  Every 8^3 block is
 * projected on the basis and its smallest coefficients are dropped while
 * the block's L2 error stays <= tol x max(block norm, sqrt(floor) x the
 * strongest block's norm) (Parseval: the dropped energy is the error; 2^-26
 * of the budget covers the rounding of the kept values and the transforms,
 * so tol <= 2^-13 keeps every non-zero coefficient).  Kept coefficients are
 * entries ent[] = top 23 bits of the fp32 value (rounded, relative error
 * <= 2^-15) | 9-bit voxel-order index; per block (z fastest, then
 * x, then y over the (n/8)^3 blocks) boff = first entry, bcnt = count.
 * *kept = entries needed; when *kept > cap the entries are incomplete
 * (never written past cap).  *err = the bound on the box's relative L2
 * error.  Both calls run on the autoencoder's
 * stream and return after it has finished; work producing P must be
 * complete before the call.  No counterpart in the reference. */
int vti_bc_encode(vti_ae *ae, const float *P, long long off, long long sy, long long sx, int n, float tol,
                  float floor, unsigned *ent, long long cap, unsigned *boff, unsigned short *bcnt, long long *kept,
                  double *err);
/* vti_bc_encode without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and entries as vti_bc_encode (the entry order across
 * blocks may differ; every block's entries are the same); its scratch (the
 * dense coefficients and counters, 8 MiB + 256 bytes) is a static device
 * array that every GPU the library runs on holds until the process ends,
 * so n <= 128 (larger volumes: boxes of at most 128^3 through off / sy / sx;
 * VTI_ERR_ARG beyond).  Calls that overlap on one GPU share that scratch
 * and race; calls in sequence (this call returns after its work) do not. */
int vti_bc_encode_noalloc(vti_ae *ae, const float *P, long long off, long long sy, long long sx, int n,
                          float tol, float floor, unsigned *ent, long long cap, unsigned *boff,
                          unsigned short *bcnt, long long *kept, double *err);
/* Dense n^3 box (z fastest) from the entries of vti_bc_encode. */
int vti_bc_decode(vti_ae *ae, const unsigned *ent, const unsigned *boff, const unsigned short *bcnt, int n,
                  float *out);
/* Quantized block-transform codec (device pointers, the same box, basis
 * and error bound as vti_bc_encode: every 8^3 block's L2 error <=
 * tol x max(block norm, sqrt(floor) x the strongest block's norm)).  Every
 * coefficient of a block is quantized with the block's own uniform step d,
 * q = rint(c / d) (the largest d whose exact quantization error meets the
 * budget, found by bisection on d's float bits); a block is a stream of
 * 32-bit words at ent + boff[b], bcnt[b] words (0: the block is within its
 * budget at q = 0):
 *   word 0        d (fp32 bits; 0 in raw mode)
 *   word 1        bit widths of coefficient groups 0..5 (5 bits each,
 *                 group g = components 64g..64g+63) | mode << 30
 *   word 2        widths of groups 6, 7 | number of nonzero q << 10
 *   words 3..18   bitmap of the nonzero q (component j: bit j%32 of word 3+j/32)
 *   words 19..    the nonzero q in component order, |q| << 1 | sign, each in
 *                 its group's width, packed from bit 0 of word 19 upward
 *                 (raw mode 1, when even d = 2^-22 x the block's largest
 *                 |c| misses the budget: the nonzero c as fp32 words)
 * A block takes at most 531 words; vti_bq_max_words(n) = (n/8)^3 x 531.
 * *kept = words needed (> cap: incomplete, nothing written past cap);
 * *err = the bound on the box's relative L2 error.  Both calls return after
 * the autoencoder's stream has finished.  No counterpart in the reference. */
int vti_bq_encode(vti_ae *ae, const float *P, long long off, long long sy, long long sx, int n, float tol,
                  float floor, unsigned *ent, long long cap, unsigned *boff, unsigned short *bcnt, long long *kept,
                  double *err);
int vti_bq_decode(vti_ae *ae, const unsigned *ent, const unsigned *boff, const unsigned short *bcnt, int n,
                  float *out);
long long vti_bq_max_words(int n);
/* A stored snapshot decoded straight into an image, as the reverse sweep of
 * vti_rtm_shot_learned does (cal_migration's img += s * g,
 * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:679-698, with the source sample s
 * decoded): img[o] += v * rp[off + y*sy + x*sx + z] for every voxel (y, x,
 * z) of the n^3 box, o = (y*n + x)*n + z, v the decoded value or, when box
 * is not NULL, inside the by x bx x bz voxels at (y0, x0, z0) = box_geom[0..5]
 * = {y0, x0, z0, by, bx, bz} (host array), box[((y - y0)*bx + (x - x0))*bz +
 * (z - z0)] (an exactly kept box).  codec 1: entries of vti_bc_encode, 2:
 * streams of vti_bq_encode.  Device pointers except box_geom; returns after
 * the autoencoder's stream has finished. */
int vti_codec_decode_image(vti_ae *ae, int codec, const unsigned *ent, const unsigned *boff,
                           const unsigned short *bcnt, int n, float *img, const float *rp, long long off,
                           long long sy, long long sx, const float *box, const int *box_geom);
/* Symmetric eigen-decomposition used by the fit (cyclic Jacobi, host):
 * a n x n row-major, w eigenvalues, v eigenvectors as columns, unsorted. */
int vti_sym_eigen(int n, const double *a, double *w, double *v);

/* 3D RTM with the source wavefield stored as autoencoder latents (cubic
 * interiors, side a multiple of 8).  The first sweep is vti_rtm_shot's
 * (record, illumination); after every step it with it % save_every == 0 the
 * interior P is encoded.  The reverse sweep images those steps with the
 * decoded field (vti_rtm_shot with image_every = save_every images the same
 * steps with the exact field).  Fallback: every check_every-th snapshot is
 * decoded at once and compared with the true field; when its L2 error
 * relative to the largest snapshot energy so far exceeds tol it is kept
 * exactly (uncompressed interior; with
 * check_every == 1, up to max_exact snapshots) or, past that, it and the
 * next check_every - 1 are recomputed exactly from full-state checkpoints
 * kept every ck_every steps (plain checkpoint recomputation; ck_every = 0:
 * no checkpoints, the latent is used).  A box of half-width src_box cells
 * around the source is kept exactly in every snapshot and masked out of the
 * encoder's input. */
typedef struct {
    int save_every;    /* 4 */
    int check_every;   /* 8; 0: no checks */
    float tol;         /* 0.05 */
    int ck_every;      /* 0: no fallback checkpoints */
    int max_exact;     /* 0: no exactly kept snapshots */
    int src_box;       /* 8: half-width of the exactly kept box around the source (0: none) */
    int codec;         /* 0: autoencoder latents; 1: sparse block transform, every 8^3 block
                          of every stored snapshot within relative L2 error tol of
                          max(its energy, floor x the strongest block's energy)
                          (check_every, max_exact unused); 2: the same bound with the
                          quantized codec (vti_bq_encode) */
    float floor;       /* codecs 1, 2: 1e-8 */
    long long budget;  /* codecs 1, 2: bytes for the coded snapshots (0: nsnap * n^3, a quarter
                          of the fp32 snapshots); past it, snapshots are recomputed from the
                          ck_every checkpoints (or the call fails without them) */
    int serial_codec;  /* codecs 1, 2: where the codec runs (the results are the same):
                          0: encodes and decodes on a stream of their own, from snapshot
                          copies the steps' stress updates write, beside the steps;
                          1: both on the shot's stream between the steps; 2: encodes
                          between the steps, decodes beside them */
} vti_learned_opts;
int vti_learned_defaults(vti_learned_opts *o);
/* stats (optional, 9 doubles): encode ms, decode ms (time between HIP
 * events around the codec's launches on its stream, summed: beside the
 * steps these spans include the GPU time the stencil takes meanwhile; the
 * encode's up to the copy of its counters to the host), mean
 * checked relative error (codecs 1, 2: mean error bound of the stored
 * snapshots), snapshots above tol (codecs 1, 2: over budget), snapshots,
 * snapshots kept exactly, bytes of the compressed store (latents or
 * entries + block tables, exactly kept snapshots, source boxes), snapshots
 * recomputed from checkpoints, and the device bytes the call held for the
 * stored source field as allocated (the store, its sparse entries in 4 GB
 * chunks taken as needed up to the budget, source boxes, the work buffers:
 * the codec's coefficients and search results for codecs 1, 2 and, with
 * the codec beside the steps, the n^3 snapshot copy; two dense n^3 buffers
 * for the latents; fallback checkpoints): the memory a plain
 * checkpointed RTM would be given for a comparison at equal memory.  Codecs
 * 1, 2 decode each snapshot straight into the image. */
int vti_rtm_shot_learned(vti_ctx *ctx, vti_ae *ae, const vti_rtm_opts *opts,
                         const vti_learned_opts *lo, const float *obs, double *stats);
/* Forward modelling of the current shot; every save_every-th snapshot, the
 * box of half-width src_box around the source masked out (as
 * vti_rtm_shot_learned does), adds nblocks sampled 8^3 blocks to the
 * autoencoder's principal-component fit. */
int vti_ae_fit_shot(vti_ctx *ctx, vti_ae *ae, int save_every, int nblocks, int src_box);
/* Forward modelling of the current shot with nsteps Adam steps of the
 * autoencoder on random crops of every save_every-th snapshot. */
int vti_ae_train_shot(vti_ctx *ctx, vti_ae *ae, int save_every, int nsteps, float *mean_loss);

/* ---- adjoint and FWI gradient -------------------------------------------- */
/* No counterpart in the reference for VTI; the reference's FWI gradient is
 * the ISO one of 3D-iso-fd-rtm-fwi-cuda/fwi3d/Toa_gpu_3dFWI.cu:295-311,
 * 474-494.  A vti_grad context runs the forward step of the point-source
 * (GPU) form in FP32 (bit-identical to vti_step) or FP64, its exact
 * discrete adjoint, and the gradient of J = 1/2 sum (d - dobs)^2 with
 * respect to the unpadded vp (absorbing-layer tables held fixed).  Host
 * arrays at this boundary are double. */
typedef struct vti_grad vti_grad;
int vti_grad_create(const vti_cfg *cfg, int fp64, vti_grad **out);
int vti_grad_destroy(vti_grad *g);
/* keep_tables: reuse the damping tables of the previous model (so that a
 * perturbed model sees the same absorbing layer). */
int vti_grad_set_model(vti_grad *g, const double *vp, const double *eps,
                       const double *del, int keep_tables);
int vti_grad_set_shot(vti_grad *g, const vti_shot *shot);
/* Per-step source gain (nt values) replacing pfac*wavelet; NULL restores it. */
int vti_grad_set_source(vti_grad *g, const double *amp);
/* Forward only: record d [irec][it]. */
int vti_grad_forward(vti_grad *g, double *d);
/* Forward + adjoint sweep with nsnap checkpoint slots (0 = every step).
 * The residual is d - dobs, or resid when non-NULL.  Outputs (each may be
 * NULL): misfit, grad_vp (unpadded model size), grad_src (nt: gradient
 * with respect to the per-step source gain), d_out (record), fwd_steps. */
int vti_grad_run(vti_grad *g, const double *dobs, const double *resid, int nsnap,
                 double *misfit, double *grad_vp, double *grad_src, double *d_out,
                 long long *fwd_steps);

/* Wavefield state for caller-driven checkpointing (no reference
 * counterpart: the reference re-runs the forward operator,
 * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1423-1437): every allocated field of
 * u, v, w, px, py, pz, qx, qy, qz, P, Q (padded volumes, in that order) to
 * or from a device buffer of vti_state_bytes bytes, enqueued on the
 * context's stream (vti_sync waits). */
long long vti_state_bytes(const vti_ctx *ctx);
int vti_save_state(vti_ctx *ctx, void *dst);
int vti_load_state(vti_ctx *ctx, const void *src);
/* One imaging step of a caller-driven RTM between a source-side and a
 * receiver-side context on the same grid (cal_migration
 * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:679-698 / 3D
 * Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:508-521, cal_illumination
 * :654-677): img += P_src * P_rcv, and when illum is not NULL illum +=
 * P_src^2 + Q_src^2 (clamp: a sum <= 0 becomes 1), over the interior;
 * img and illum are device arrays (z fastest, then x, then y).  Ordered
 * after both contexts' queued work; later work of the source context
 * waits for it. */
int vti_image_accumulate(vti_ctx *src, vti_ctx *rcv, float *img, float *illum, int clamp);

/* Point-source (GPU) form, 2D and 3D: vti_step / vti_run /
 * vti_update_stress update P and Q directly wherever every damping factor
 * of the cell is exactly 1 (outside the absorbing layer), P -= tx + ty + tz
 * and Q likewise with the split form's per-direction increments (3D: 80 ->
 * 40, 2D: 60 -> 36 bytes per interior cell in the stress update).  In 3D a
 * cell of the absorbing layer keeps only the split fields of its damped
 * axes and carries the rest in P and Q (56 or 72 bytes; corners 80); in 2D
 * the absorbing layer keeps the split form.  Not the reference's
 * arithmetic (update_stress3d GPU_vti3dfd_kernels.cu:131-201,
 * update_stress2d GPU_vti2dfd_kernels.cu:240-299 keep every split
 * everywhere): the records agree with it to ~1e-6 rel-L2 and are as close
 * to an FP64 run.
 * Switch before a run (after vti_reset): the interior split fields are not
 * maintained while it is on, so a split-form step (vti_update_vel,
 * vti_update_stress, vti_step with it off) after unsplit steps returns
 * VTI_ERR_STATE until vti_reset; a state loaded with vti_load_state counts
 * as unsplit when the setting is on at the load.  Default off.  vti_rtm_shot and
 * vti_rtm_shot_learned use it for both propagations; on the receiver side
 * the row the records are written into keeps the split form, which
 * rebuilds P and Q so that the written values do not persist (as in the
 * reference).  vti_grad keeps the split form. */
int vti_set_unsplit(vti_ctx *ctx, int on);

/* 3D only: vti_step runs each step as one fused kernel (on = 1) or as the
 * velocity and stress kernels of vti_update_vel / vti_update_stress (on =
 * 0, the default and the faster one on MI355X: 4.47 against 7.3 ms per
 * step at 572^3); the results are bit-identical.  The fused step keeps
 * second copies of u, v, w, P, Q (5 more fields) and swaps them with the
 * context's fields every step.  No counterpart in the reference
 * (GPU_vti3dfd_kernels.cu:515-518 launches the two kernels). */
int vti_set_fused(vti_ctx *ctx, int on);

/* ---- measurement ------------------------------------------------------- */
/* When on, vti_step brackets every velocity and stress launch with HIP
 * events on the context stream; vti_prof_read sums them (ms).  A fused 3D
 * step counts entirely as stress time. */
int vti_prof_enable(vti_ctx *ctx, int on);
int vti_prof_read(vti_ctx *ctx, double *vel_ms, double *stress_ms, int *nlaunch);
/* Plain device copy of nbytes (bandwidth calibration kernel). */
int vti_calib_copy(void *dst, const void *src, size_t nbytes, void *hip_stream);
int vti_sync(vti_ctx *ctx);
const char *vti_strerror(int code);
const char *vti_last_error(void);

/* RTM of the current shot with the source wavefield rebuilt backward in
 * time instead of recomputed from checkpoints (effective-boundary
 * reconstruction; the reference's counterpart, wavefield_bndr
 * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:532-612 with :1394-1437, 3D
 * Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:405-506, saves a one-cell ring
 * and re-runs the forward operator, which is not a reversal).  The forward
 * sweep is vti_rtm_shot's (record, illumination, mute / obs) and saves after
 * every step the interior's outer 4-cell bands (x faces: u, P; y faces: v,
 * P; z faces: w, Q, P; vti_rtm_eb_bytes(ctx) bytes for the shot); the
 * reverse sweep steps the receiver side as vti_rtm_shot does and the
 * source side backward on the interior less the bands, with the saved bands
 * in place: two source-side propagations instead of the checkpointed
 * recomputation.  The rebuilt field differs from the forward one by
 * accumulated rounding.  Point-source (GPU) form only; opts.nsnap unused;
 * no angle gathers.  stats (optional, 2 doubles): band bytes, source-side
 * steps.  Leaves the main state's split fields stale (vti_reset before
 * split-form steps).  No counterpart in the reference beyond the above. */
int vti_rtm_shot_eb(vti_ctx *ctx, const vti_rtm_opts *opts, const float *obs, double *stats);
long long vti_rtm_eb_bytes(const vti_ctx *ctx);

/* ---- 2D TTI ----------------------------------------------------------- */
/* 2D-tti-fd+rtm-cuda/Toa_gpu_2dtti_fd_1orderfunction.cu: the GPU form with a
 * per-cell tilt theta.  The velocity update takes the x and z derivatives of
 * P (u) and Q (w) rotated by theta, the stress update those of u and w
 * (update_vel :81-114, update_stress :116-183), and every field is damped by
 * the product of its x and z factors.  vti_set_theta sets the tilt of the
 * interior (nx * nz, x-major, z fastest; radians, or degrees converted as
 * the reference reads them, theta * (pi/180.0) in double), padded by edge
 * replication; cos and sin of the float angle per cell are host-built
 * tables, so every later step, record and RTM of the context runs the TTI
 * update (theta NULL: back to VTI).  2D point-source (GPU) form, split
 * update only (vti_set_unsplit(0); VTI_ERR_STATE otherwise); not with
 * vti_rtm_shot_eb. */
int vti_set_theta(vti_ctx *ctx, const float *theta, int degrees);
/* main of Toa_gpu_2dtti_fd_1orderfunction.cu:313-499 as a function, its
 * hard-coded values (:329-359: 600 x 300, 5 m, npml 50, 40 Hz, pfac 10,
 * source (200, 140), 801 steps of 0.5 ms) as arguments: vp, eps, delta and
 * the tilt (degrees) read x-major, z fastest; the whole padded P every 100
 * steps after step 0 appended to FNsnap; the P record of row npml
 * ([ix][it]) to FNshot.  Returns the status (also in vti_dropin_status). */
int vti_tti2dfd(int nx, int nz, float dx, float dz, int npml, float favg, float pfac, int fs, int zs, int nt,
                float dt, const char FNv[], const char FNe[], const char FNd[], const char FNtheta[],
                const char FNshot[], const char FNsnap[]);

/* ---- shot records to SU / SEG-Y files (host) ----------------------------- */
/* vti_bin2su: utils-binary2segy/binary2segy.c:43-85 — each of ntrace output
 * traces is the 240-byte header of the template SU file's trace (kept as it
 * is) and ns samples from the raw float file, each sample's bytes reversed
 * when swap is non-zero (the reference always swaps, swap_float_4
 * :136-144).  Unlike the reference, a short template or raw file is an error
 * (VTI_ERR_IO), not garbage output. */
int vti_bin2su(const char *su_template, const char *raw, const char *out, long long ntrace, int ns, int swap);
/* Acquisition geometry of nshot shots of ntr traces of ns samples (the
 * trace-major records the library returns and the drop-ins' shot files
 * hold, shots one after another).  Metres; trace it of shot is has its
 * receiver at gx0 + is*dgx_shot + it*dgx (y alike) and its source at
 * sx0 + is*dsx; coordinates are stored as integers with scalco (SEG-Y:
 * negative divides, so -100 keeps centimetres); fldr = fldr0 + is. */
typedef struct {
    int nshot, ntr, ns, dt_us;
    int fldr0;
    int scalco;
    double sx0, dsx, sy0, dsy;
    double gx0, dgx, dgx_shot;
    double gy0, dgy, dgy_shot;
    double sz, gz;   /* source depth, receiver depth (gelev = -gz) */
} vti_su_geometry;
/* SU file: 240-byte trace headers (tracl, tracr, fldr, tracf, ep, trid 1,
 * offset, gelev, sdepth, scalel 1, scalco, sx, sy, gx, gy, counit 1, ns, dt)
 * and float samples, in the host's byte order (big_endian = 0) or
 * big-endian.  No counterpart in the reference (it converts existing SU
 * files, vti_bin2su). */
int vti_su_write(const char *path, const float *rec, const vti_su_geometry *g, int big_endian);
/* SEG-Y revision 1 file, big-endian: 3200-byte EBCDIC textual header (40
 * cards "C nn " + the lines of text, may be NULL), 400-byte binary header
 * (traces per record = ntr <= 32767, dt, ns, format 5 IEEE float, sorting 1,
 * metres, revision 0x0100, fixed length), then the SU trace headers and
 * samples big-endian.  No counterpart in the reference. */
int vti_segy_write(const char *path, const float *rec, const vti_su_geometry *g, const char *text);

/* ---- drop-in entry points with the reference signatures ------------------ */
/* GPU_vti2dfd_kernels.cu:301-304 (caller GTK_gpuvti2dfd.c:88). */
void GPU_vti2dfd(int nx, int nz, int dx, int dz, int npd, int SV,
                 const char FNv[], const char FNe[], const char FNd[],
                 int favg, int ns, int fs, int ds, int zs,
                 const char FNshot[], const char FNsnap[], int nt, int dt,
                 int run_count);
/* GPU_vti3dfd_kernels.cu:371-374 (caller GTK_gpuvti3dfd.c:67-68). */
void GPU_vti3dfd(int nx, int ny, int nz, int dx, int dy, int dz, int npml,
                 int SV, const char FNv[], const char FNe[], const char FNd[],
                 int favg, int ns, int fsx, int dsx, int fsy, int dsy,
                 int fsz, int dsz, const char FNshot[], const char FNsnap[],
                 int nt, int dt, int run_count);
/* CPU_vti2dfd_kernels.c:366-367 (caller GTK_cpuvti2dfd.c:57): the CPU
 * driver's semantics, computed on the GPU. */
void CPU_vti2dfd(int nx, int nz, int dx, int dz, int npd, int SV,
                 const char FNv[], const char FNe[], const char FNd[],
                 int favg, int ns, int fs, int ds, int zs,
                 const char FNshot[], const char FNsnap[], int nt, int dt,
                 int run_count);
/* main of 2D-vti-fd-omp/Toafd2domp_2.c:40-242, a program in the reference,
 * with its hard-coded values as arguments (there: 601, 301, 5.0, 5.0, 0,
 * 30, 1000, 100, 100, 4001, 0.0005, and the file names v_601_301.dat,
 * e_601_301.dat, d_601_301.dat, shot.dat, snap.dat): VTI_MODE_OMP_REF
 * semantics; models read x-major, z fastest (:91-104); the interior P every
 * 50 steps from step 0 into frame it/50 of FNsnap (:222-227); the P+Q
 * record of row npd, [ix][it], into FNshot (:219, :231-235; with npd = 0
 * that row is never updated and the record is zero).  Returns the status
 * (also kept for vti_dropin_status). */
int vti_omp2dfd(int nx, int nz, float dx, float dz, int npd, float favg,
                float pfac, int sx, int sz, int nt, float dt,
                const char FNv[], const char FNe[], const char FNd[],
                const char FNshot[], const char FNsnap[]);
/* Status of the last drop-in call (the reference entry points return void). */
int vti_dropin_status(void);

#ifdef __cplusplus
}
#endif
#endif /* VTI_H */
