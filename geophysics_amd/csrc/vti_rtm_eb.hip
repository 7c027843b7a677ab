// vti_rtm_eb.hip — RTM whose source wavefield is rebuilt backward in time
// from the final forward state and saved boundary bands (effective-boundary
// reconstruction), a third source-state option beside checkpoint
// recomputation (vti_rtm_shot) and the compressed store
// (vti_rtm_shot_learned).
//
// The reference's form (wavefield_bndr Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:
// 532-612, its use :1394-1437; 3D :405-506) saves P and Q on a one-cell
// ring outside the interior and re-runs the forward operator from the final
// state with that ring overwritten, which is not a time reversal.  Here the
// reversal is exact in exact arithmetic.  The interior is split into R, the
// interior shrunk by 4 cells on every side, and the band B, its outer 4
// cells.  The stencils of R reach 4 cells, so no update of R reads outside
// R + B, and no damping acts there (the absorbing-layer factors are 1 in
// the interior).  With S_it = (U_it, P_it) the state after forward step it
// (the point source enters the velocity update of step it as P + amp at
// the source cell, it does not persist in P):
//   U_it = U_{it-1} - dt D+(P_{it-1} + amp_it)      P_it = P_{it-1} - K D-(U_it)
// so on R
//   P_{it-1} = P_it + K D-(U_it)                     U_{it-1} = U_it + dt D+(P_{it-1} + amp_it)
// needing U_it and P_{it-1} on B.  The forward sweep saves, after every
// step, the fields the reverse reads on B: the normal velocity and P on the
// x (y) faces, w, Q and P on the z faces (P on every face for the image).
// The reverse steps run the product's own velocity and stress kernels with
// negated dt and damping factors of exactly 1 (u - 1*(-dt)*d = u + dt*d and
// the unsplit update P - (K*(-dt)*d) = P + K*dt*d, bit for bit the formulas
// above), over the whole grid: what they write outside R is overwritten (B)
// or never read by R (the absorbing layer).  Floating-point rounding keeps
// the reversal from being exact; the rebuilt field drifts from the forward
// one by accumulated rounding, measured by the tests.
//
// Band storage per step: 2 * 4 * (ny*nz) cells on the x faces (u, P),
// 2 * 4 * (nx*nz) on the y faces (v, P; 3D), 2 * 4 * (nx*ny) on the z faces
// (w, Q, P), ny = 1 in 2D; nt steps of it in one device buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "vti_internal.h"
#include "vti_rtm.h"

using namespace vti;

namespace {

constexpr int kBand = 4;   // stencil half-width

struct BandGeom {
    Grid g;
    int np, nx, ny, nz, nd;
    long long cx, cy, cz;   // band cells on the x, y, z faces
};

BandGeom band_geom(const vti_ctx *c)
{
    BandGeom b{};
    b.g = c->g;
    b.nd = c->cfg.ndim;
    b.np = c->cfg.npml;
    b.nx = c->cfg.nx;
    b.nz = c->cfg.nz;
    b.ny = b.nd == 3 ? c->cfg.ny : 1;
    b.cx = 2LL * kBand * b.ny * b.nz;
    b.cy = b.nd == 3 ? 2LL * kBand * b.nx * b.nz : 0;
    b.cz = 2LL * kBand * b.nx * b.ny;
    return b;
}

long long band_floats(const BandGeom &b) { return 2 * b.cx + 2 * b.cy + 3 * b.cz; }

// mode 0: save every band field of state (P, Q, u, v, w) into slot;
// 1: load the velocities (u on x faces, v on y faces, w on z faces);
// 2: load the pressures (P on x and y faces, Q and P on z faces).
__global__ void band_kernel(BandGeom b, float *P, float *Q, float *u, float *v, float *w, float *slot, int mode)
{
    const long long tot = b.cx + b.cy + b.cz;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= tot) return;
    const long long ldz = b.g.ldz, ldxy = b.g.ldxy;
    const int np = b.np;
    const int yo = b.nd == 3 ? np : 0;   // y offset of the interior (2D: the single plane 0)
    long long cell;
    float *vel, *pv[2];
    float *sv, *sp[2];
    int npres;
    if (i < b.cx) {   // x faces: (side, layer, y, z), z fastest
        const long long k = i;
        const int z = (int)(k % b.nz);
        long long r = k / b.nz;
        const int y = (int)(r % b.ny);
        r /= b.ny;
        const int l = (int)(r % kBand), s = (int)(r / kBand);
        const int x = s ? np + b.nx - kBand + l : np + l;
        cell = (np + z) + (long long)x * ldz + (long long)(yo + y) * ldxy;
        vel = u; pv[0] = P; npres = 1;
        sv = slot + k; sp[0] = slot + b.cx + k;
    } else if (i < b.cx + b.cy) {   // y faces: (side, layer, x, z), z fastest
        const long long k = i - b.cx;
        const int z = (int)(k % b.nz);
        long long r = k / b.nz;
        const int x = (int)(r % b.nx);
        r /= b.nx;
        const int l = (int)(r % kBand), s = (int)(r / kBand);
        const int y = s ? np + b.ny - kBand + l : np + l;
        cell = (np + z) + (long long)(np + x) * ldz + (long long)y * ldxy;
        vel = v; pv[0] = P; npres = 1;
        sv = slot + 2 * b.cx + k; sp[0] = slot + 2 * b.cx + b.cy + k;
    } else {   // z faces: (side, y, x, layer), layer fastest (contiguous in z)
        const long long k = i - b.cx - b.cy;
        const int l = (int)(k % kBand);
        long long r = k / kBand;
        const int x = (int)(r % b.nx);
        r /= b.nx;
        const int y = (int)(r % b.ny), s = (int)(r / b.ny);
        const int z = s ? np + b.nz - kBand + l : np + l;
        cell = z + (long long)(np + x) * ldz + (long long)(yo + y) * ldxy;
        vel = w; pv[0] = Q; pv[1] = P; npres = 2;
        const long long base = 2 * b.cx + 2 * b.cy;
        sv = slot + base + k; sp[0] = slot + base + b.cz + k; sp[1] = slot + base + 2 * b.cz + k;
    }
    if (mode == 0) {
        *sv = vel[cell];
        for (int q = 0; q < npres; ++q) *sp[q] = pv[q][cell];
    } else if (mode == 1) {
        vel[cell] = *sv;
    } else {
        for (int q = 0; q < npres; ++q) pv[q][cell] = *sp[q];
    }
}

hipError_t launch_band(vti_ctx *c, const BandGeom &b, const VtiState &s, float *slot, int mode)
{
    const long long tot = b.cx + b.cy + b.cz;
    hipLaunchKernelGGL(band_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, b, s.f[VTI_F_P],
                       s.f[VTI_F_Q], s.f[VTI_F_U], s.f[VTI_F_V], s.f[VTI_F_W], slot, mode);
    return hipGetLastError();
}

struct Buf {
    void *p = nullptr;
    ~Buf() { (void)hipFree(p); }
};

}  // namespace

extern "C" {

long long vti_rtm_eb_bytes(const vti_ctx *c)
{
    if (!c) return 0;
    return band_floats(band_geom(c)) * (long long)c->cfg.nt * (long long)sizeof(float);
}

int vti_rtm_shot_eb(vti_ctx *c, const vti_rtm_opts *o, const float *obs, double *stats)
{
    if (!c || !o) return fail(VTI_ERR_ARG, "vti_rtm_shot_eb: null argument");
    if (!c->have_model || !c->have_shot) return fail(VTI_ERR_STATE, "vti_rtm_shot_eb: model and shot first");
    if (c->nrec <= 0) return fail(VTI_ERR_STATE, "vti_rtm_shot_eb: the shot has no receivers");
    if (o->image_every < 1) return fail(VTI_ERR_ARG, "vti_rtm_shot_eb: bad options");
    const vti_cfg &k = c->cfg;
    if (k.mode != VTI_MODE_GPU_REF)
        return fail(VTI_ERR_ARG, "vti_rtm_shot_eb: the point-source (GPU) form only (the in-stress source "
                                 "forms are not reversed)");
    if (c->tti) return fail(VTI_ERR_ARG, "vti_rtm_shot_eb: the VTI update only (the reversal runs the unsplit form)");
    if (k.nx < 2 * kBand + 1 || k.nz < 2 * kBand + 1 || (k.ndim == 3 && k.ny < 2 * kBand + 1))
        return fail(VTI_ERR_ARG, "vti_rtm_shot_eb: interior narrower than two 4-cell bands + 1");
    if (c->adcig_na > 0)
        return fail(VTI_ERR_STATE, "vti_rtm_shot_eb: angle gathers need the full source state "
                                   "(vti_rtm_shot); switch them off with vti_adcig_config(ctx, 0, 0, 0)");
    int st;
    if (!c->img && (st = vti_rtm_reset(c))) return st;
    const int nt = k.nt, nrec = c->nrec, nd = k.ndim;
    const int inj_z = o->inject_z >= 0 ? o->inject_z : c->shot.rec_z;
    if (inj_z < 4 || inj_z >= c->g.nnz - 4)
        return fail(VTI_ERR_ARG, "vti_rtm_shot_eb: injection row outside the updated region");
    const BandGeom bg = band_geom(c);
    const long long per = band_floats(bg);
    Buf band, ones;
    if (hipMalloc(&band.p, (size_t)per * nt * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(VTI_ERR_ALLOC, "vti_rtm_shot_eb: boundary bands (" +
                                       std::to_string((double)per * nt * 4 / 1e9) + " GB)");
    }
    float *BAND = static_cast<float *>(band.p);
    // damping factors of exactly 1 for the reverse steps
    const int M = c->nnmax;
    {
        std::vector<float> h(M, 1.0f);
        if (hipMalloc(&ones.p, (size_t)M * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(VTI_ERR_ALLOC, "vti_rtm_shot_eb: tables");
        }
        if (hipMemcpyAsync(ones.p, h.data(), (size_t)M * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return fail(VTI_ERR_HIP, "vti_rtm_shot_eb: tables");
    }
    const float *ONE = static_cast<const float *>(ones.p);
    VtiState G;
    if ((st = state_alloc(c, G))) return st;
    struct GGuard {
        VtiState &g;
        ~GGuard() { state_free(g); }
    } gg{G};
    const VtiState W = main_state(c);
    const bool svf = o->sv_forward != 0 && k.sv != 0;
    const bool svb = o->sv_backward != 0 && k.sv != 0;
    const int clamp = o->illum_mode == 0;
    hipStream_t s = c->stream;
    auto hip = [](hipError_t e, const char *w) { return e == hipSuccess ? VTI_OK : fail(VTI_ERR_HIP, w); };

    // ---- forward: vti_rtm_shot's first sweep, the bands saved after every step
    if ((st = hip(state_zero(c, W), "zero"))) return st;
    for (int it = 0; it < nt; ++it) {
        if (launch_vel(nd, vel_args_for(c, W, it, true), s) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
        float *row = c->rec + (size_t)it * nrec;
        if (launch_stress(nd, with_illum(c, src_stress_args(c, W, it, svf, row), clamp), s) != hipSuccess)
            return fail(VTI_ERR_HIP, "stress");
        if ((st = hip(launch_band(c, bg, W, BAND + (size_t)it * per, 0), "band save"))) return st;
    }
    // ---- the record: observed, or modelled + mute (as vti_rtm_shot)
    if (obs) {
        if (hipMemcpyAsync(c->rec_t, obs, (size_t)nrec * nt * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess)
            return fail(VTI_ERR_HIP, "obs upload");
        if ((st = hip(launch_transpose(c->rec_t, c->rec, nrec, nt, s), "transpose"))) return st;
    } else if (o->mute_tt >= 0) {
        const vti_shot &sh = c->shot;
        const int np = k.npml;
        st = nd == 2 ? vti_mute(c, sh.sx - np + 1, 0, sh.sz - np + 1, 1, o->mute_tt)
                     : vti_mute3d_window(c, sh.sx - np + 1, 0, sh.sy - np + 1, 0, sh.sz - np + 1, 0, o->mute_tt);
        if (st) return st;
    }
    // ---- reverse: receiver forward in reversed time, source rebuilt backward
    if ((st = hip(state_zero(c, G), "zero"))) return st;
    for (int it = nt - 1; it >= 0; --it) {
        if (launch_inject_rec(c, G, c->rec + (size_t)it * nrec, inj_z) != hipSuccess) return fail(VTI_ERR_HIP, "inject");
        if (launch_vel(nd, vel_args_for(c, G, it, false), s) != hipSuccess) return fail(VTI_ERR_HIP, "vel");
        StressArgs ra = rcv_stress_args(c, G, it, svb, inj_z);
        if (o->illum_mode == 1) ra = with_illum(c, ra, 0);
        if (launch_stress(nd, ra, s) != hipSuccess) return fail(VTI_ERR_HIP, "stress");
        // W holds S_it on R; P on B from the saved band of step it
        if (it % o->image_every == 0 &&
            (st = hip(launch_image(c, W.f[VTI_F_P], G.f[VTI_F_P], c->img), "image")))
            return st;
        if (it == 0) break;
        // S_it -> S_{it-1} on R
        if ((st = hip(launch_band(c, bg, W, BAND + (size_t)it * per, 1), "band load"))) return st;
        StressArgs sa = src_stress_args(c, W, it, svf, nullptr);
        sa.ax1 = sa.ax2 = sa.ay1 = sa.ay2 = sa.az1 = sa.az2 = ONE;
        sa.dtx = -sa.dtx; sa.dty = -sa.dty; sa.dtz = -sa.dtz;
        sa.unsplit = 1;
        sa.split_z = -1;
        sa.rec = nullptr;
        sa.illum = nullptr;
        sa.gsrc = 0;
        if (launch_stress(nd, sa, s) != hipSuccess) return fail(VTI_ERR_HIP, "reverse stress");
        if ((st = hip(launch_band(c, bg, W, BAND + (size_t)(it - 1) * per, 2), "band load"))) return st;
        VelArgs va = vel_args_for(c, W, it, true);
        va.cx1 = va.cx2 = va.cy1 = va.cy2 = va.cz1 = va.cz2 = ONE;
        va.dtx = -va.dtx; va.dty = -va.dty; va.dtz = -va.dtz;
        if (launch_vel(nd, va, s) != hipSuccess) return fail(VTI_ERR_HIP, "reverse vel");
    }
    c->rtm_illum_mode = o->illum_mode;
    c->us_dirty = 1;   // the main state's split fields no longer match P, Q
    if ((st = hip(hipStreamSynchronize(s), "sync"))) return st;
    if (stats) {
        stats[0] = (double)per * nt * sizeof(float);   // band bytes
        stats[1] = 2.0 * nt;                             // source-side steps (forward + reverse)
    }
    return VTI_OK;
}

}  // extern "C"
