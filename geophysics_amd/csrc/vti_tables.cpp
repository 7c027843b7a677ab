// vti_tables.cpp — host-side tables of the VTI FD path: damping
// coefficients, source gains, SV-taper factors, Gaussian footprint.
//
// These run once per model / shot.  The reference computes them on the
// device in double precision and stores floats; the same double
// expressions are evaluated here, in the reference's order, and rounded
// to float where the reference stores a float.
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "vti_internal.h"

namespace vti {

// One axis of nn_int interior cells padded by npml on both sides.
// Velocity tables at half-grid positions xi = (npml-0.5-i)/npml, stress
// tables at integer positions (npml-i)/npml:
//   coff1 = 1/(1 + dt*d0*xi^2/2),  coff2 = coff1*(1 - dt*d0*xi^2/2)
// GPU form: initial_coffe2d GPU_vti2dfd_kernels.cu:108-139 (same as
// initial_coffe3d GPU_vti3dfd_kernels.cu:283-314).
// CPU form: initial_coffe CPU_vti2dfd_kernels.c:271-332, whose acoff2 is
// built from coff1 (:307,309,314,319) and whose right-side stress
// position is (1+i-nx-npd)/npd (:313).
void axis_tables(int nn_int, int npml, float dtd0, int cpu_form,
                 float *c1, float *c2, float *a1, float *a2)
{
    const int n = nn_int + 2 * npml;
    for (int i = 0; i < n; ++i) {
        if (i >= npml && i < npml + nn_int) {
            c1[i] = c2[i] = a1[i] = a2[i] = 1.0f;
            continue;
        }
        double xv, xs;
        if (i < npml) {
            xv = (npml - 0.5 - i) / npml;
            xs = ((npml - i) * 1.0) / npml;
        } else {
            xv = (0.5 + i - nn_int - npml) / npml;
            xs = cpu_form ? ((1 + i - nn_int - npml) * 1.0) / npml
                          : ((i - nn_int - npml) * 1.0) / npml;
        }
        const double hv = static_cast<double>(dtd0) * std::pow(xv, 2.0);
        const double hs = static_cast<double>(dtd0) * std::pow(xs, 2.0);
        c1[i] = static_cast<float>(1.0 / (1.0 + hv / 2.0));
        c2[i] = static_cast<float>(static_cast<double>(c1[i]) * (1.0 - hv / 2.0));
        a1[i] = static_cast<float>(1.0 / (1.0 + hs / 2.0));
        const float base = cpu_form ? c1[i] : a1[i];
        a2[i] = static_cast<float>(static_cast<double>(base) * (1.0 - hs / 2.0));
    }
}

// Wavelet of add_source2d GPU_vti2dfd_kernels.cu:187-198 (= get_wavelet
// CPU_vti2dfd_kernels.c:30-48): float time arguments, double formula.
float wavelet(float ts, float favg, int wtype)
{
    if (wtype == 1) {
        const float x = favg * ts;
        const float xx = x * x;
        return static_cast<float>((1 - 2 * kPi * kPi * xx) * std::exp(-(kPi * kPi * xx)));
    }
    if (wtype == 2) {
        const float x = static_cast<float>((-4) * favg * favg * kPi * kPi / std::log(0.1));
        const float e = -x * ts * ts;
        return static_cast<float>((-2) * kPi * kPi * ts * std::exp(static_cast<double>(e)));
    }
    if (wtype == 3) {
        const float x = static_cast<float>((-1) * favg * favg * kPi * kPi / std::log(0.1));
        const float e = -x * ts * ts;
        return static_cast<float>(std::exp(static_cast<double>(e)));
    }
    if (wtype == 4) {
        // Toafd2domp_2.c:154: (favg*ts*favg*ts) is a float product in the
        // first factor; inside exp, pi*pi*favg*ts*favg*ts is a double chain
        const float x4 = favg * ts * favg * ts;
        const double arg = kPi * kPi * favg * ts * favg * ts;
        return static_cast<float>((1 - 2 * kPi * kPi * x4) * std::exp(-arg));
    }
    return 0.0f;
}

// Per-step source gain pfac*wavelet(t - 1/favg), active while t <= 2/favg.
// The time is a float accumulated from dt (t=dt; t+=dt:
// GPU_vti2dfd_kernels.cu:419, CPU_vti2dfd_kernels.c:478), from 0 in the
// Toafd2domp_2.c form (t=0.0; t+=dt, :144).
void source_gains(const vti_cfg &c, std::vector<float> &amp)
{
    amp.assign(c.nt > 0 ? c.nt : 0, 0.0f);
    float t = c.mode == VTI_MODE_OMP_REF ? 0.0f : c.dt;
    const float tdelay = static_cast<float>(1.0 / c.favg);
    for (int it = 0; it < c.nt; ++it) {
        if (t <= 2 * tdelay) {
            const float ts = t - tdelay;
            amp[it] = c.pfac * wavelet(ts, c.favg, c.wtype);
        }
        t += c.dt;
    }
}

// SV-suppression taper factor for each integer r^2 in [0, R^2].
// GPU form (GPU_vti2dfd_kernels.cu:262-270, GPU_vti3dfd_kernels.cu:158-166):
//   r2 <= r*r -> 0, else 0.5*(1-cos(pi*((sqrtf(r2)-r)*4/(3R-1))))
// CPU form (CPU_vti2dfd_kernels.c:139-147), CR = R:
//   r2 <= CR*CR/16 -> 0, else 0.5*(1-cos(pi*((pow(r2,.5)-CR/4.0)*4/(3CR-1))))
void taper_table(int cpu_form, int R, int r, std::vector<double> &tap)
{
    tap.assign(static_cast<size_t>(R) * R + 1, 0.0);
    for (int r2 = 0; r2 <= R * R; ++r2) {
        double f = 0.0;
        if (!cpu_form) {
            if (r2 > r * r) {
                const float sr = std::sqrt(static_cast<float>(r2));
                f = 0.5 * (1 - std::cos(kPi * ((sr - r) * 4.0 / (R * 3.0 - 1))));
            }
        } else {
            if (r2 > (R * R / 16))
                f = 0.5 * (1 - std::cos(kPi * ((std::pow(static_cast<double>(r2), 0.5) - R / 4.0) *
                                               4.0 / (R * 3.0 - 1))));
        }
        tap[r2] = f;
    }
}

// exp(-z*z - x*x) over the 7x7 footprint of ptsource
// (CPU_vti2dfd_kernels.c:67-74), index (x+3)*7 + (z+3).
void gaussian_weights(double w[49])
{
    for (int x = -3; x <= 3; ++x)
        for (int z = -3; z <= 3; ++z)
            w[(x + 3) * 7 + (z + 3)] = std::exp(static_cast<double>(-z * z - x * x));
}

namespace {
std::mutex g_err_mu;
std::string g_err;
}

void set_error(const std::string &msg)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = msg;
}

const char *last_error_cstr()
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    thread_local std::string copy;
    copy = g_err;
    return copy.c_str();
}

}  // namespace vti
