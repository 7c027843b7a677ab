// vti_internal.h — structures shared by the host side (vti_api.hip,
// vti_tables.cpp, vti_dropin.cpp) and the gfx950 kernels (vti_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vti.h"

namespace vti {

// #define pi 3.141592653 (GPU_vti2dfd_kernels.cu:31)
constexpr double kPi = 3.141592653;

// Padded grid as laid out in HBM: z fastest with a pitch ldz >= nnz that
// keeps every (x,y) row 128-byte aligned, then x, then y.
struct Grid {
    int nnx, nny, nnz;
    int ldz;
    long long ldxy;   // ldz * nnx
};

// Velocity update: update_vel2d GPU_vti2dfd_kernels.cu:209-238,
// update_vel3d GPU_vti3dfd_kernels.cu:90-129.
struct VelArgs {
    Grid g;
    float *u, *v, *w;
    const float *P, *Q;
    const float *cx1, *cx2, *cy1, *cy2, *cz1, *cz2;
    float dtx, dty, dtz;
    long long src;   // cell that receives the point source before the update (-1: none)
    float amp;       // pfac * wavelet(t) of this step
    // 2D TTI (Toa_gpu_2dtti_fd_1orderfunction.cu:81-114): cos and sin of the
    // per-cell tilt (padded, pitched); nullptr: the VTI update
    const float *cth, *sth;
};

// Stress update: update_stress2d GPU_vti2dfd_kernels.cu:240-299,
// update_stress3d GPU_vti3dfd_kernels.cu:131-201, with the receiver gather
// (shot_record2d :141-149) and the CPU driver's in-stress source
// (CPU_vti2dfd_kernels.c:173-174) fused in.
struct StressArgs {
    Grid g;
    const float *u, *v, *w;
    float *px, *py, *pz, *qx, *qy, *qz, *P, *Q;
    const float *vp, *ep, *de;
    const float *ax1, *ax2, *ay1, *ay2, *az1, *az2;
    float dtx, dty, dtz;
    // SV taper: factor tap[r2] for integer r2 <= svR*svR around (cx,cy,cz)
    int sv, svR, cx, cy, cz;
    const double *tap;
    // in-stress 7x7 Gaussian footprint centred on (gx, gy, gz); gw[(dx+3)*7+(dz+3)]
    int gsrc, gx, gy, gz;
    float gamp;
    const double *gw;
    // receivers: row rz, x = rx0 + k*rdx (k < rnx), y = ry0 + j*rdy (j < rny)
    float *rec;      // row of this step, nrec floats (nullptr: no record)
    int rec_kind, rz, rx0, rdx, rnx, ry0, rdy, rny;
    // point-source form, reduced split (unsplit): in the interior P and Q
    // are updated directly; in 3D a cell of the absorbing layer keeps the
    // split fields only of its damped axes (all of them in the corners) and
    // carries the undamped rest in P, Q.  Row split_z (z index, -1: none)
    // keeps the full split form.
    int unsplit, split_z;
    // Toafd2domp_2.c form (2D): pz and qx are (vp*vp)*pow(1+2*del, 0.5)*dt*d
    // in double, subtracted in double and rounded once (:176-180)
    int dpq;
    // source illumination accumulated in the epilogue (RTM first sweep):
    // illum[interior cell] += P^2 + Q^2, a sum <= 0 set to 1 when illum_clamp;
    // the interior is inx x iny x inz cells from padded index inp (y only in 3D)
    float *illum;
    int illum_clamp, inp, inx, iny, inz;
    // 3D: the updated P of every interior cell (the illumination's interior)
    // also written to snap[z + x inz + y inz inx] (nullptr: none)
    float *snap;
    // 2D TTI (Toa_gpu_2dtti_fd_1orderfunction.cu:116-183): as VelArgs
    const float *cth, *sth;
};

// One fused 3D step (vti_fused3d.h): velocity update into new u, v, w from
// the previous P, Q (v.P, v.Q) and velocities (u0, v0, w0), then the stress
// update into new P, Q (s.P, s.Q), split fields in place.
struct FusedArgs {
    VelArgs v;
    const float *u0, *v0, *w0;
    StressArgs s;
};

// ---- kernel launchers (vti_kernels.hip) ---------------------------------
hipError_t launch_fused(const FusedArgs &a, hipStream_t s);
hipError_t launch_vel(int ndim, const VelArgs &a, hipStream_t s);
hipError_t launch_stress(int ndim, const StressArgs &a, hipStream_t s);
hipError_t launch_pad_model(const Grid &g, int ndim, int nx, int ny, int nz,
                            int npml, const float *src, float *dst, hipStream_t s);
hipError_t launch_zero(float *p, size_t n, hipStream_t s);
hipError_t launch_transpose(const float *in, float *out, int rows, int cols,
                            hipStream_t s);
hipError_t launch_mute2d(float *rec_t, int nrec, int nt, float dt, float favg,
                         float dx, float dz, int fs, int ds, int zs, int is,
                         float vp1, float eps1, int tt, int x0, int rdx, hipStream_t s);
hipError_t launch_mute3d(float *rec_t, int nrx, int nry, int nt, float dt,
                         float favg, float dx, float dy, float dz, int fsx,
                         int dsx, int fsy, int dsy, int fsz, int is,
                         float vp1, float eps1, int tt, int window,
                         int x0, int rdx, int y0, int rdy, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t nbytes, hipStream_t s);

// ---- host table construction (vti_tables.cpp) ---------------------------
void axis_tables(int nn_int, int npml, float dtd0, int cpu_form,
                 float *c1, float *c2, float *a1, float *a2);
float wavelet(float ts, float favg, int wtype);
void source_gains(const vti_cfg &c, std::vector<float> &amp);
void taper_table(int cpu_form, int R, int r, std::vector<double> &tap);
void gaussian_weights(double w[49]);

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

}  // namespace vti

// One set of wavefield volumes: u v w px py pz qx qy qz P Q (v/py/qy are
// null in 2D), each Grid-pitched.
struct VtiState {
    float *f[11] = {};
};

// Context: model, tables, the main wavefield state and per-shot buffers.
struct vti_ctx {
    vti_cfg cfg{};
    vti::Grid g{};
    size_t nvol = 0;            // floats per pitched volume
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;

    float *arena = nullptr;     // model + main state volumes
    float *f[11] = {};          // main state: u v w px py pz qx qy qz P Q
    float *vp = nullptr, *ep = nullptr, *de = nullptr;
    float *tab = nullptr;       // 12 damping tables, nnmax floats each
    int nnmax = 0;
    std::vector<float> htab;    // host copy of the tables
    float d0 = 0.0f;

    std::vector<float> amp;     // per-step source gain
    double *dtap = nullptr;     // SV taper factors
    int tapR = 0;
    double *dgw = nullptr;      // 7x7 Gaussian weights

    vti_shot shot{};
    int nrec = 0;
    float *rec = nullptr;       // [it][irec]
    float *rec_t = nullptr;     // [irec][it] scratch

    bool have_model = false, have_shot = false;
    float vp1 = 0, eps1 = 0;    // padded vp[1], eps[1] (mute_directwave2d :167)

    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    // ordering of vti_image_accumulate against the other context's stream
    hipEvent_t ev_acc[2] = {nullptr, nullptr};

    // second buffers of u, v, w, P, Q for the fused 3D step (vti_step), which
    // reads one set and writes the other; the pointers in f swap each step
    float *alt = nullptr;
    float *altf[11] = {};
    bool fused = false;
    bool unsplit = false;   // vti_step: unsplit interior (3D point-source form)
    // an unsplit stress update has run since the last vti_reset: the
    // interior split fields no longer add up to P and Q
    bool us_dirty = false;
    bool alt_valid = false;   // altf equals f outside the updated region

    // imaging (RTM): interior-sized image and illumination, accumulated over shots
    float *img = nullptr, *illum = nullptr;
    size_t nint = 0;            // nx*ny*nz
    int rtm_illum_mode = 0;
    // Poynting-vector angle gathers (vti_adcig_*): [iya][ixa][ia][iz] over
    // the cdps, host-built bin thresholds and cos^3 weights (nangle each)
    float *adcig = nullptr, *adcig_tab = nullptr;
    int adcig_na = 0, adcig_da = 1, adcig_dcdp = 1;
    size_t adcig_n = 0;
    // 2D TTI (vti_set_theta): cos and sin of the padded tilt, two pitched
    // volumes; nullptr: VTI
    float *tti = nullptr;
};

namespace vti {
// number of state volumes per state (8 in 2D, 11 in 3D)
int state_nfield(const vti_ctx *c);
// allocate / free an extra state (zeroed)
int state_alloc(vti_ctx *c, VtiState &s);
void state_free(VtiState &s);
hipError_t state_zero(vti_ctx *c, const VtiState &s);
hipError_t state_copy(vti_ctx *c, const VtiState &dst, const VtiState &src);
VtiState main_state(vti_ctx *c);
// kernel arguments for one step of state s; inject: apply the shot's source
// (point before the velocity update in GPU mode, footprint in the stress
// update in CPU mode); sv: SV taper; rec_row: receiver row of this step or null
VelArgs vel_args_for(vti_ctx *c, const VtiState &s, int it, bool inject);
StressArgs stress_args_for(vti_ctx *c, const VtiState &s, int it, bool inject, bool sv, float *rec_row);
// Source-side stress arguments (point source injected): the context's
// unsplit-interior setting applies (3D point-source form).
StressArgs src_stress_args(vti_ctx *c, const VtiState &s, int it, bool sv, float *rec_row);
// Receiver-side stress arguments of an RTM whose records are written into
// P and Q at row inj_z: the context's unsplit setting applies everywhere
// but on that row, which keeps the split form (it rebuilds P and Q, so the
// written values do not persist, as in the split form everywhere).
StressArgs rcv_stress_args(vti_ctx *c, const VtiState &s, int it, bool sv, int inj_z);
// a with the context's source illumination accumulated in its epilogue
StressArgs with_illum(vti_ctx *c, StressArgs a, int clamp);
// a 3D stress update that also copies the interior P into snap (dense, z fastest)
StressArgs with_snap(vti_ctx *c, StressArgs a, float *snap);
// one angle-gather accumulation (vti_adcig.hip) from a source and a receiver
// state of the context's grid, when the context has gathers configured
hipError_t launch_adcig(vti_ctx *c, const VtiState &src, const VtiState &rcv);
}  // namespace vti
