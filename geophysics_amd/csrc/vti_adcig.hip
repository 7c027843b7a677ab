// vti_adcig.hip — Poynting-vector angle-domain common-image gathers of the
// RTM (Poynting_Adcigs Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:718-766 and
// Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:618-660; adcigs_illum :768-786
// / :663-684; stk_adcigs :788-816; adcigs_smiled :818-847).
//
// At every imaged step each cdp cell forms the source Poynting vector
// S_s = (P u, [P v,] Q w) and the receiver one S_g = (P u, [P v,] Q w)
// (the reference negates S_s, :746-747, because its source field is rebuilt
// backward in time, where u changes sign; the source field here is the exact
// forward field, whose P u points along the incident wave already),
// the (reference-scaled) cosine a = S_s.S_g / (|S_s||S_g| (1 - 0.1)), and,
// when |a| <= 1, adds P_s P_g cos^3(ia degrees) to bin ia = (int)(half the
// angle between the vectors, in degrees / dangle) of its own gather.  Each
// cell writes only its own gather column, so the accumulation is a plain
// read-modify-write (no atomics); one thread per (z, x) cdp cell, z on the
// wavefront lanes for coalesced field reads.
//
// Exactness.  The angle binning of the reference goes through acosf; here it
// is a comparison of a against nangle thresholds: T_k, the largest float a
// whose reference bin (the same float / double arithmetic with the host
// libm's acosf) is at least k, found by bisection over the ordered floats
// of [-1, 1] once per configuration.  The bin of a is the number of
// thresholds it does not exceed, which is the reference's bin for every a
// (the angle is monotone in a), so no device acosf / cosf enters and the
// gathers follow the host arithmetic bit for bit; the cos^3 weights are a
// host table of the reference's cosf values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "vti_internal.h"
#include "vti_rtm.h"

namespace vti {
namespace {

// Reference bin of a in [-1, 1] (Poynting_Adcigs :754-757): the angle in
// double rounded to float, then (int)(angle / dangle).
int ref_bin(float a, int dangle)
{
    const float ang = (float)(0.5 * (double)std::acos(a) * 180.0 / kPi);
    return (int)(ang / (dangle * 1.0));
}

// float <-> monotone integer key
int32_t fkey(float f)
{
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i >= 0 ? i : (int32_t)(0x80000000u - (uint32_t)i);
}
float kfloat(int32_t k)
{
    const int32_t i = k >= 0 ? k : (int32_t)(0x80000000u - (uint32_t)k);
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

// tab[0..na): T_1 .. T_na (largest a with ref_bin(a) >= k; -2 when none);
// tab[na..2na): cos^3 weight factors cosf(ia pi / 180) (:761-762)
void adcig_tables(int na, int da, std::vector<float> &tab)
{
    tab.assign(2 * (size_t)na, 0.0f);
    const int32_t lo0 = fkey(-1.0f), hi0 = fkey(1.0f);
    for (int k = 1; k <= na; ++k) {
        if (ref_bin(-1.0f, da) < k) {    // the largest angle (a = -1) does not reach bin k
            tab[k - 1] = -2.0f;
            continue;
        }
        int64_t lo = lo0, hi = hi0;      // invariant: ref_bin(lo) >= k
        if (ref_bin(1.0f, da) >= k) lo = hi;
        while (hi - lo > 1) {
            const int32_t mid = (int32_t)(lo + (hi - lo) / 2);
            if (ref_bin(kfloat(mid), da) >= k) lo = mid;
            else hi = mid;
        }
        tab[k - 1] = kfloat((int32_t)lo);
    }
    for (int ia = 0; ia < na; ++ia) tab[na + ia] = std::cos((float)(ia * kPi / 180.0));
}

template <bool THREE>
__global__ __launch_bounds__(256) void adcig_kernel(Grid g, int np, int nxa, int nz, int dcdp, int na,
                                                    const float *__restrict__ sP, const float *__restrict__ sQ,
                                                    const float *__restrict__ su, const float *__restrict__ sv,
                                                    const float *__restrict__ sw, const float *__restrict__ gP,
                                                    const float *__restrict__ gQ, const float *__restrict__ gu,
                                                    const float *__restrict__ gv, const float *__restrict__ gw,
                                                    const float *__restrict__ tab, float *adcigs)
{
    const int iz = blockIdx.x * blockDim.x + threadIdx.x;
    const int ixa = blockIdx.y * blockDim.y + threadIdx.y, iya = blockIdx.z;
    if (iz >= nz || ixa >= nxa) return;
    const long long c = (iz + np) + (long long)(ixa * dcdp + np) * g.ldz +
                        (THREE ? (long long)(iya * dcdp + np) * g.ldxy : 0LL);
    const float pS = sP[c], pG = gP[c];
    const float Ssx = pS * su[c], Ssz = sQ[c] * sw[c];
    const float Sgx = pG * gu[c], Sgz = gQ[c] * gw[c];
    float b1, b2, num;
    if (THREE) {
        const float Ssy = pS * sv[c], Sgy = pG * gv[c];
        b1 = Ssx * Ssx + Ssy * Ssy + Ssz * Ssz;
        b2 = Sgx * Sgx + Sgy * Sgy + Sgz * Sgz;
        num = Ssx * Sgx + Ssy * Sgy + Ssz * Sgz;
    } else {
        b1 = Ssx * Ssx + Ssz * Ssz;
        b2 = Sgx * Sgx + Sgz * Sgz;
        num = Ssx * Sgx + Ssz * Sgz;
    }
    const float a = (float)((double)num / ((double)sqrtf(b1 * b2) * (1 - 0.1)));
    if (!(a >= -1.0f && a <= 1.0f)) return;
    int ia = 0;
    while (ia < na && a <= tab[ia]) ++ia;
    if (ia >= na) return;
    const float w = tab[na + ia];
    const long long id = iz + (long long)nz * ia + (long long)nz * na * ixa + (long long)nz * na * nxa * iya;
    adcigs[id] += pS * pG * w * w * w;
}

// adcigs_illum: 2D divide (:783), 3D divide where nonzero (:679-680), by the
// illumination of the gather's cdp
__global__ void adcig_divide_kernel(float *adcigs, const float *__restrict__ illum, int nxa, int nya, int na, int nz,
                                    int nx, int dcdp, int guard)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)nxa * nya * na * nz;
    if (i >= n) return;
    const int iz = (int)(i % nz);
    const long long r = i / ((long long)nz * na);
    const int ixa = (int)(r % nxa), iya = (int)(r / nxa);
    const float d = illum[iz + (long long)ixa * dcdp * nz + (long long)iya * dcdp * nx * nz];
    if (!guard || d != 0.0f) adcigs[i] /= d;
}

void dims(const vti_ctx *c, int &nxa, int &nya)
{
    const bool three = c->cfg.ndim == 3;
    const int d = three ? c->adcig_dcdp : 1;
    nxa = c->cfg.nx / d;
    nya = three ? c->cfg.ny / d : 1;
}

}  // namespace

hipError_t launch_adcig(vti_ctx *c, const VtiState &s, const VtiState &r)
{
    if (c->adcig_na <= 0 || !c->adcig) return hipSuccess;
    int nxa, nya;
    dims(c, nxa, nya);
    if (nxa < 1 || nya < 1) return hipSuccess;
    const bool three = c->cfg.ndim == 3;
    const int d = three ? c->adcig_dcdp : 1;
    const dim3 blk(64, 4), grd((c->cfg.nz + 63) / 64, (nxa + 3) / 4, nya);
    if (three)
        hipLaunchKernelGGL(adcig_kernel<true>, grd, blk, 0, c->stream, c->g, c->cfg.npml, nxa, c->cfg.nz, d,
                           c->adcig_na, s.f[VTI_F_P], s.f[VTI_F_Q], s.f[VTI_F_U], s.f[VTI_F_V], s.f[VTI_F_W],
                           r.f[VTI_F_P], r.f[VTI_F_Q], r.f[VTI_F_U], r.f[VTI_F_V], r.f[VTI_F_W], c->adcig_tab,
                           c->adcig);
    else
        hipLaunchKernelGGL(adcig_kernel<false>, grd, blk, 0, c->stream, c->g, c->cfg.npml, nxa, c->cfg.nz, d,
                           c->adcig_na, s.f[VTI_F_P], s.f[VTI_F_Q], s.f[VTI_F_U], s.f[VTI_F_U], s.f[VTI_F_W],
                           r.f[VTI_F_P], r.f[VTI_F_Q], r.f[VTI_F_U], r.f[VTI_F_U], r.f[VTI_F_W], c->adcig_tab,
                           c->adcig);
    return hipGetLastError();
}

}  // namespace vti

using namespace vti;

extern "C" {

int vti_adcig_config(vti_ctx *c, int nangle, int dangle, int dcdp)
{
    if (!c || nangle < 0 || (nangle > 0 && (dangle < 1 || dcdp < 1)))
        return fail(VTI_ERR_ARG, "vti_adcig_config: bad arguments");
    hipFree(c->adcig);
    hipFree(c->adcig_tab);
    c->adcig = c->adcig_tab = nullptr;
    c->adcig_na = 0;
    c->adcig_n = 0;
    if (nangle == 0) return VTI_OK;
    c->adcig_da = dangle;
    c->adcig_dcdp = c->cfg.ndim == 3 ? dcdp : 1;
    c->adcig_na = nangle;
    int nxa, nya;
    dims(c, nxa, nya);
    if (nxa < 1 || nya < 1) {
        c->adcig_na = 0;
        return fail(VTI_ERR_ARG, "vti_adcig_config: no cdp (dcdp larger than the model)");
    }
    c->adcig_n = (size_t)nxa * nya * nangle * c->cfg.nz;
    std::vector<float> tab;
    adcig_tables(nangle, dangle, tab);
    if (hipMalloc(&c->adcig, c->adcig_n * sizeof(float)) != hipSuccess ||
        hipMalloc(&c->adcig_tab, tab.size() * sizeof(float)) != hipSuccess) {
        c->adcig_na = 0;
        return fail(VTI_ERR_ALLOC, "vti_adcig_config: gathers");
    }
    if (hipMemcpy(c->adcig_tab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        launch_zero(c->adcig, c->adcig_n, c->stream) != hipSuccess)
        return fail(VTI_ERR_HIP, "vti_adcig_config: set-up");
    return VTI_OK;
}

long long vti_adcig_elems(const vti_ctx *c) { return c ? (long long)c->adcig_n : 0; }

int vti_adcig_tables(int nangle, int dangle, float *thresholds, float *weights)
{
    if (nangle < 1 || dangle < 1 || !thresholds || !weights) return fail(VTI_ERR_ARG, "vti_adcig_tables: bad arguments");
    std::vector<float> tab;
    adcig_tables(nangle, dangle, tab);
    std::memcpy(thresholds, tab.data(), (size_t)nangle * sizeof(float));
    std::memcpy(weights, tab.data() + nangle, (size_t)nangle * sizeof(float));
    return VTI_OK;
}

int vti_adcig_reset(vti_ctx *c)
{
    if (!c) return fail(VTI_ERR_ARG, "vti_adcig_reset: null ctx");
    if (c->adcig && launch_zero(c->adcig, c->adcig_n, c->stream) != hipSuccess)
        return fail(VTI_ERR_HIP, "vti_adcig_reset: zero");
    return VTI_OK;
}

int vti_adcig_accumulate(vti_ctx *src, vti_ctx *rcv)
{
    if (!src || !rcv) return fail(VTI_ERR_ARG, "vti_adcig_accumulate: null argument");
    const vti_cfg &a = src->cfg, &b = rcv->cfg;
    if (a.ndim != b.ndim || a.nx != b.nx || a.nz != b.nz || (a.ndim == 3 && a.ny != b.ny) || a.npml != b.npml)
        return fail(VTI_ERR_ARG, "vti_adcig_accumulate: the contexts' grids differ");
    if (rcv->adcig_na <= 0) return fail(VTI_ERR_STATE, "vti_adcig_accumulate: no gathers configured (vti_adcig_config)");
    for (hipEvent_t &e : rcv->ev_acc)
        if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail(VTI_ERR_HIP, "vti_adcig_accumulate: event");
    if (src != rcv && (hipEventRecord(rcv->ev_acc[0], src->stream) != hipSuccess ||
                       hipStreamWaitEvent(rcv->stream, rcv->ev_acc[0], 0) != hipSuccess))
        return fail(VTI_ERR_HIP, "vti_adcig_accumulate: ordering after the source context");
    if (launch_adcig(rcv, main_state(src), main_state(rcv)) != hipSuccess)
        return fail(VTI_ERR_HIP, "vti_adcig_accumulate: launch");
    if (src != rcv && (hipEventRecord(rcv->ev_acc[1], rcv->stream) != hipSuccess ||
                       hipStreamWaitEvent(src->stream, rcv->ev_acc[1], 0) != hipSuccess))
        return fail(VTI_ERR_HIP, "vti_adcig_accumulate: ordering of the source context");
    return VTI_OK;
}

int vti_adcig_get(vti_ctx *c, int illum_divide, float *adcigs)
{
    if (!c || !adcigs) return fail(VTI_ERR_ARG, "vti_adcig_get: null argument");
    if (!c->adcig) return fail(VTI_ERR_STATE, "vti_adcig_get: no gathers configured");
    int nxa, nya;
    dims(c, nxa, nya);
    float *tmp = nullptr;
    if (illum_divide) {
        if (!c->illum) return fail(VTI_ERR_STATE, "vti_adcig_get: no illumination (run vti_rtm_shot first)");
        if (hipMalloc(&tmp, c->adcig_n * sizeof(float)) != hipSuccess)
            return fail(VTI_ERR_ALLOC, "vti_adcig_get: copy");
        const bool three = c->cfg.ndim == 3;
        if (hipMemcpyAsync(tmp, c->adcig, c->adcig_n * sizeof(float), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
            hipFree(tmp);
            return fail(VTI_ERR_HIP, "vti_adcig_get: copy");
        }
        hipLaunchKernelGGL(adcig_divide_kernel, dim3((unsigned)((c->adcig_n + 255) / 256)), dim3(256), 0, c->stream,
                           tmp, c->illum, nxa, nya, c->adcig_na, c->cfg.nz, c->cfg.nx, three ? c->adcig_dcdp : 1,
                           three ? 1 : 0);
    }
    const hipError_t e1 = hipStreamSynchronize(c->stream);
    const hipError_t e2 = hipMemcpy(adcigs, tmp ? tmp : c->adcig, c->adcig_n * sizeof(float), hipMemcpyDeviceToHost);
    hipFree(tmp);
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(VTI_ERR_HIP, "vti_adcig_get: copy");
    return VTI_OK;
}

int vti_adcig_stack(int nx, int nz, int nangle, const float *adcigs, float *migration)
{
    if (nx < 1 || nz < 1 || nangle < 1 || !adcigs || !migration) return fail(VTI_ERR_ARG, "vti_adcig_stack: bad arguments");
    std::vector<float> temp((size_t)nx * nz);
    for (int ix = 0; ix < nx; ++ix)
        for (int iz = 0; iz < nz; ++iz) {
            float stk = 0.0f;
            for (int ia = 0; ia < nangle; ++ia) stk += adcigs[(size_t)ix * nangle * nz + (size_t)ia * nz + iz];
            temp[(size_t)ix * nz + iz] = stk;
        }
    laplace_adj(nz, nx, 1, temp.data(), migration);
    return VTI_OK;
}

int vti_adcig_thin(int nx, int nz, int nangle, int dAdcigs, const float *adcigs, float *out)
{
    if (nx < 1 || nz < 1 || nangle < 1 || dAdcigs < 1 || !adcigs || !out)
        return fail(VTI_ERR_ARG, "vti_adcig_thin: bad arguments");
    for (int ix = 0; ix < nx; ix += dAdcigs)
        std::memcpy(out + (size_t)(ix / dAdcigs) * nangle * nz, adcigs + (size_t)ix * nangle * nz,
                    (size_t)nangle * nz * sizeof(float));
    return VTI_OK;
}

}  // extern "C"
