// vti_dropin.cpp — the reference's three extern "C" entry points with
// their exact signatures and file semantics, computed on the GPU through
// the vti_* context API.
//
//   GPU_vti2dfd  GPU_vti2dfd_kernels.cu:301-484
//   GPU_vti3dfd  GPU_vti3dfd_kernels.cu:371-573
//   CPU_vti2dfd  CPU_vti2dfd_kernels.c:366-536 (the CPU driver's semantics)
//   vti_omp2dfd  main of 2D-vti-fd-omp/Toafd2domp_2.c:40-242 (a program in the
//                reference, with its constants as arguments here)
//
// Differences from the reference, all on error paths or defects: errors
// set vti_dropin_status() instead of exit(0) / continuing with NULL files;
// ISO models are padded completely (CPU_vti2dfd_kernels.c:244 stops
// padding at the first zero); the 3D y tables are sized nny, not nnx
// (GPU_vti3dfd_kernels.cu:471,474).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <string>
#include <vector>

#include "vti_internal.h"

namespace {

std::atomic<int> g_status{VTI_OK};

int read_model(const char *fn, std::vector<float> &dst, size_t n)
{
    FILE *fp = std::fopen(fn, "rb");
    if (!fp) {
        std::printf("error open <%s>!\n", fn);
        vti::set_error(std::string("cannot open ") + fn);
        return VTI_ERR_IO;
    }
    dst.resize(n);
    const size_t got = std::fread(dst.data(), sizeof(float), n, fp);
    std::fclose(fp);
    if (got != n) {
        vti::set_error(std::string("short read from ") + fn);
        return VTI_ERR_IO;
    }
    return VTI_OK;
}

struct Ctx {
    vti_ctx *c = nullptr;
    ~Ctx() { vti_destroy(c); }
};

// 2D driver shared by GPU_vti2dfd (mode GPU) and CPU_vti2dfd (mode CPU).
int run2d(int mode, int nx, int nz, int dx, int dz, int npd, int SV, const char *FNv,
          const char *FNe, const char *FNd, int favg, int ns, int fs, int ds, int zs,
          const char *FNshot, const char *FNsnap, int nt, int dt)
{
    vti_cfg cfg;
    int st = vti_cfg_ref2d(&cfg, mode, nx, nz, dx, dz, npd, SV, favg, nt, dt);
    if (st) return st;
    const size_t n = (size_t)nx * nz;
    std::vector<float> v, e, d;
    if ((st = read_model(FNv, v, n)) || (st = read_model(FNe, e, n)) || (st = read_model(FNd, d, n)))
        return st;
    Ctx x;
    if ((st = vti_create(&cfg, &x.c))) return st;
    if ((st = vti_set_model(x.c, v.data(), e.data(), d.data(), 0))) return st;
    FILE *fshot = std::fopen(FNshot, "wb");
    FILE *fsnap = std::fopen(FNsnap, "wb");
    if (!fshot || !fsnap) {
        if (fshot) std::fclose(fshot);
        if (fsnap) std::fclose(fsnap);
        vti::set_error("cannot open output files");
        return VTI_ERR_IO;
    }
    int nnx, nny, nnz;
    vti_padded_dims(x.c, &nnx, &nny, &nnz);
    std::vector<float> frame((size_t)nnx * nnz), shot((size_t)nx * nt);
    for (int is = 1; is <= ns && st == VTI_OK; ++is) {
        vti_shot sh;
        if ((st = vti_shot_ref2d(&sh, &cfg, is, fs, ds, zs, 1))) break;
        if ((st = vti_set_shot(x.c, &sh)) || (st = vti_reset(x.c))) break;
        // snapshots of P every 50 steps for shot 1 (GPU_vti2dfd_kernels.cu:432-437,
        // CPU_vti2dfd_kernels.c:502-506)
        int it = 0;
        while (it < nt && st == VTI_OK) {
            int last = is == 1 ? (it + 49) / 50 * 50 : nt - 1;   // next snapshot step >= it
            if (last > nt - 1) last = nt - 1;
            if ((st = vti_step(x.c, it, last - it + 1))) break;
            it = last + 1;
            if (is == 1 && last % 50 == 0) {
                if ((st = vti_get_field(x.c, VTI_F_P, frame.data()))) break;
                std::fseek(fsnap, (long)(last / 50) * nnx * nnz * 4L, SEEK_SET);
                std::fwrite(frame.data(), 4, frame.size(), fsnap);
            }
        }
        if (st) break;
        if (mode == VTI_MODE_GPU_REF && (st = vti_mute(x.c, fs, ds, zs, is, 100))) break;   // :439
        if ((st = vti_get_record(x.c, shot.data()))) break;
        std::fseek(fshot, (long)(is - 1) * nt * nx * (long)sizeof(float), SEEK_SET);
        std::fwrite(shot.data(), sizeof(float), shot.size(), fshot);
    }
    std::fclose(fshot);
    std::fclose(fsnap);
    return st;
}

int run3d(int nx, int ny, int nz, int dx, int dy, int dz, int npml, int SV, const char *FNv,
          const char *FNe, const char *FNd, int favg, int ns, int fsx, int dsx, int fsy, int dsy,
          int fsz, int dsz, const char *FNshot, const char *FNsnap, int nt, int dt)
{
    vti_cfg cfg;
    int st = vti_cfg_ref3d(&cfg, nx, ny, nz, dx, dy, dz, npml, SV, favg, nt, dt);
    if (st) return st;
    const size_t n = (size_t)nx * ny * nz;
    std::vector<float> v, e, d;
    if ((st = read_model(FNv, v, n)) || (st = read_model(FNe, e, n)) || (st = read_model(FNd, d, n)))
        return st;
    Ctx x;
    if ((st = vti_create(&cfg, &x.c))) return st;
    if ((st = vti_set_model(x.c, v.data(), e.data(), d.data(), 0))) return st;
    FILE *fshot = std::fopen(FNshot, "wb");
    FILE *fsnap = std::fopen(FNsnap, "wb");
    if (!fshot || !fsnap) {
        if (fshot) std::fclose(fshot);
        if (fsnap) std::fclose(fsnap);
        vti::set_error("cannot open output files");
        return VTI_ERR_IO;
    }
    int nnx, nny, nnz;
    vti_padded_dims(x.c, &nnx, &nny, &nnz);
    std::vector<float> shot((size_t)nx * ny * nt), full, win;
    for (int is = 0; is < ns && st == VTI_OK; ++is) {
        vti_shot sh;
        if ((st = vti_shot_ref3d(&sh, &cfg, is, fsx, dsx, fsy, dsy, fsz, dsz, 1, 1))) break;
        if ((st = vti_set_shot(x.c, &sh)) || (st = vti_reset(x.c))) break;
        // windowed snapshots every 300 steps of shot 0 (GPU_vti3dfd_kernels.cu:526-532)
        int it = 0;
        while (it < nt && st == VTI_OK) {
            int last = nt - 1;
            if (is == 0) last = std::max(300, (it + 299) / 300 * 300);   // next snapshot step
            if (last > nt - 1) last = nt - 1;
            if ((st = vti_step(x.c, it, last - it + 1))) break;
            it = last + 1;
            if (is == 0 && last != 0 && last % 300 == 0) {
                full.resize((size_t)nnx * nny * nnz);
                win.resize(n);
                if ((st = vti_get_field(x.c, VTI_F_P, full.data()))) break;
                for (int iy = 0; iy < ny; ++iy)          // window3d :329-340
                    for (int ix = 0; ix < nx; ++ix)
                        for (int iz = 0; iz < nz; ++iz)
                            win[iz + (size_t)nz * ix + (size_t)nz * nx * iy] =
                                full[(iz + npml) + (size_t)nnz * (ix + npml) + (size_t)nnz * nnx * (iy + npml)];
                std::fseek(fsnap, (long)(last / 300 - 1) * (long)n * 4L, SEEK_SET);
                std::fwrite(win.data(), 4, n, fsnap);
            }
        }
        if (st) break;
        if ((st = vti_mute3d(x.c, fsx, dsx, fsy, dsy, fsz, is, 100))) break;   // :534
        if ((st = vti_get_record(x.c, shot.data()))) break;
        std::fseek(fshot, (long)is * nt * nx * ny * (long)sizeof(float), SEEK_SET);
        std::fwrite(shot.data(), sizeof(float), shot.size(), fshot);
    }
    std::fclose(fshot);
    std::fclose(fsnap);
    return st;
}

// main of Toafd2domp_2.c (2D-vti-fd-omp/Toafd2domp_2.c:40-242), its
// hard-coded values (:55-76) as arguments: models read x-major, z fastest
// (:91-104); P of the interior every 50 steps from step 0 into frame it/50
// (:222-227); the P+Q record at row npd, [ix][it] (:219, :231-235).
int run_omp2d(int nx, int nz, float dx, float dz, int npd, float favg, float pfac, int sx, int sz, int nt,
              float dt, const char *FNv, const char *FNe, const char *FNd, const char *FNshot, const char *FNsnap)
{
    vti_cfg cfg{};
    cfg.ndim = 2;
    cfg.nx = nx; cfg.ny = 1; cfg.nz = nz; cfg.npml = npd;
    cfg.dx = dx; cfg.dy = dx; cfg.dz = dz;
    cfg.dt = dt; cfg.nt = nt;
    cfg.favg = favg; cfg.pfac = pfac;
    cfg.wtype = 4;
    cfg.mode = VTI_MODE_OMP_REF;
    cfg.sv = 0;
    cfg.rec_kind = VTI_REC_PQ;
    const size_t n = (size_t)nx * nz;
    std::vector<float> v, e, d;
    int st;
    if ((st = read_model(FNv, v, n)) || (st = read_model(FNe, e, n)) || (st = read_model(FNd, d, n)))
        return st;
    Ctx x;
    if ((st = vti_create(&cfg, &x.c))) return st;
    if ((st = vti_set_model(x.c, v.data(), e.data(), d.data(), 0))) return st;
    vti_shot sh{};
    sh.sx = (int)(sx + 0.5) + npd - 1;                           // :150
    sh.sz = (int)(sz + 0.5) + npd - 1;                           // :151
    sh.sv_cx = sh.sx; sh.sv_cz = sh.sz;
    sh.rec_z = npd;                                              // :219
    sh.rec_x0 = npd; sh.rec_dx = 1; sh.rec_nx = nx;
    sh.rec_dy = 1; sh.rec_ny = 1;
    if ((st = vti_set_shot(x.c, &sh)) || (st = vti_reset(x.c))) return st;
    FILE *fshot = std::fopen(FNshot, "wb");
    FILE *fsnap = std::fopen(FNsnap, "wb");
    if (!fshot || !fsnap) {
        if (fshot) std::fclose(fshot);
        if (fsnap) std::fclose(fsnap);
        vti::set_error("cannot open output files");
        return VTI_ERR_IO;
    }
    int nnx, nny, nnz;
    vti_padded_dims(x.c, &nnx, &nny, &nnz);
    std::vector<float> full((size_t)nnx * nnz), win(n), shot((size_t)nx * nt);
    for (int it = 0; it < nt && st == VTI_OK;) {
        int last = (it + 49) / 50 * 50;                          // next snapshot step >= it
        if (last > nt - 1) last = nt - 1;
        if ((st = vti_step(x.c, it, last - it + 1))) break;
        it = last + 1;
        if (last % 50 == 0) {
            if ((st = vti_get_field(x.c, VTI_F_P, full.data()))) break;
            for (int ix = 0; ix < nx; ++ix)
                for (int iz = 0; iz < nz; ++iz)
                    win[iz + (size_t)nz * ix] = full[(iz + npd) + (size_t)nnz * (ix + npd)];
            std::fseek(fsnap, (long)(last / 50) * (long)n * 4L, SEEK_SET);
            std::fwrite(win.data(), 4, n, fsnap);
        }
    }
    if (st == VTI_OK && (st = vti_get_record(x.c, shot.data())) == VTI_OK)
        std::fwrite(shot.data(), sizeof(float), shot.size(), fshot);
    std::fclose(fshot);
    std::fclose(fsnap);
    return st;
}

// main of 2D-tti-fd+rtm-cuda/Toa_gpu_2dtti_fd_1orderfunction.cu:313-499, its
// hard-coded values (:329-359) as arguments: the four models read x-major,
// z fastest, the tilt in degrees (read_file :217-242); one shot; the whole
// padded P every 100 steps after step 0, frames appended (:448-454); the P
// record of row npml, [ix][it] (shot_record :283, :457-459).
int run_tti2d(int nx, int nz, float dx, float dz, int npml, float favg, float pfac, int fs, int zs, int nt,
              float dt, const char *FNv, const char *FNe, const char *FNd, const char *FNth, const char *FNshot,
              const char *FNsnap)
{
    vti_cfg cfg{};
    cfg.ndim = 2;
    cfg.nx = nx; cfg.ny = 1; cfg.nz = nz; cfg.npml = npml;
    cfg.dx = dx; cfg.dy = dx; cfg.dz = dz;
    cfg.dt = dt; cfg.nt = nt;
    cfg.favg = favg; cfg.pfac = pfac;
    cfg.wtype = 1;                                               // :329
    cfg.mode = VTI_MODE_GPU_REF;
    cfg.sv = 0;                                                  // update_stress(..., false) :444
    cfg.rec_kind = VTI_REC_P;
    const size_t n = (size_t)nx * nz;
    std::vector<float> v, e, d, th;
    int st;
    if ((st = read_model(FNv, v, n)) || (st = read_model(FNe, e, n)) || (st = read_model(FNd, d, n)) ||
        (st = read_model(FNth, th, n)))
        return st;
    Ctx x;
    if ((st = vti_create(&cfg, &x.c))) return st;
    if ((st = vti_set_model(x.c, v.data(), e.data(), d.data(), 0)) || (st = vti_set_theta(x.c, th.data(), 1)))
        return st;
    vti_shot sh{};
    sh.sx = (int)((float)fs + 0.5) + npml - 1;                   // add_source :74
    sh.sz = (int)((float)zs + 0.5) + npml - 1;                   // :75
    sh.sv_cx = fs + npml; sh.sv_cz = zs + npml;                  // :136-137
    sh.rec_z = npml;                                             // :283
    sh.rec_x0 = npml; sh.rec_dx = 1; sh.rec_nx = nx;
    sh.rec_dy = 1; sh.rec_ny = 1;
    if ((st = vti_set_shot(x.c, &sh)) || (st = vti_reset(x.c))) return st;
    FILE *fshot = std::fopen(FNshot, "wb");
    FILE *fsnap = std::fopen(FNsnap, "wb");
    if (!fshot || !fsnap) {
        if (fshot) std::fclose(fshot);
        if (fsnap) std::fclose(fsnap);
        vti::set_error("cannot open output files");
        return VTI_ERR_IO;
    }
    int nnx, nny, nnz;
    vti_padded_dims(x.c, &nnx, &nny, &nnz);
    std::vector<float> full((size_t)nnx * nnz), shot((size_t)nx * nt);
    for (int it = 0; it < nt && st == VTI_OK;) {
        int last = std::max(100, (it + 99) / 100 * 100);         // next frame step >= it (not step 0)
        if (last > nt - 1) last = nt - 1;
        if ((st = vti_step(x.c, it, last - it + 1))) break;
        it = last + 1;
        if (last % 100 == 0) {
            if ((st = vti_get_field(x.c, VTI_F_P, full.data()))) break;
            std::fwrite(full.data(), 4, full.size(), fsnap);
        }
    }
    if (st == VTI_OK && (st = vti_get_record(x.c, shot.data())) == VTI_OK)
        std::fwrite(shot.data(), sizeof(float), shot.size(), fshot);
    std::fclose(fshot);
    std::fclose(fsnap);
    return st;
}

}  // namespace

extern "C" {

void GPU_vti2dfd(int nx, int nz, int dx, int dz, int npd, int SV, const char FNv[],
                 const char FNe[], const char FNd[], int favg, int ns, int fs, int ds, int zs,
                 const char FNshot[], const char FNsnap[], int nt, int dt, int run_count)
{
    (void)run_count;
    g_status = run2d(VTI_MODE_GPU_REF, nx, nz, dx, dz, npd, SV, FNv, FNe, FNd, favg, ns, fs, ds, zs,
                     FNshot, FNsnap, nt, dt);
}

void CPU_vti2dfd(int nx, int nz, int dx, int dz, int npd, int SV, const char FNv[],
                 const char FNe[], const char FNd[], int favg, int ns, int fs, int ds, int zs,
                 const char FNshot[], const char FNsnap[], int nt, int dt, int run_count)
{
    (void)run_count;
    g_status = run2d(VTI_MODE_CPU_REF, nx, nz, dx, dz, npd, SV, FNv, FNe, FNd, favg, ns, fs, ds, zs,
                     FNshot, FNsnap, nt, dt);
}

void GPU_vti3dfd(int nx, int ny, int nz, int dx, int dy, int dz, int npml, int SV,
                 const char FNv[], const char FNe[], const char FNd[], int favg, int ns, int fsx,
                 int dsx, int fsy, int dsy, int fsz, int dsz, const char FNshot[],
                 const char FNsnap[], int nt, int dt, int run_count)
{
    (void)run_count;
    g_status = run3d(nx, ny, nz, dx, dy, dz, npml, SV, FNv, FNe, FNd, favg, ns, fsx, dsx, fsy, dsy,
                     fsz, dsz, FNshot, FNsnap, nt, dt);
}

int vti_omp2dfd(int nx, int nz, float dx, float dz, int npd, float favg, float pfac, int sx, int sz, int nt,
                float dt, const char FNv[], const char FNe[], const char FNd[], const char FNshot[],
                const char FNsnap[])
{
    const int st = run_omp2d(nx, nz, dx, dz, npd, favg, pfac, sx, sz, nt, dt, FNv, FNe, FNd, FNshot, FNsnap);
    g_status = st;
    return st;
}

int vti_tti2dfd(int nx, int nz, float dx, float dz, int npml, float favg, float pfac, int fs, int zs, int nt,
                float dt, const char FNv[], const char FNe[], const char FNd[], const char FNtheta[],
                const char FNshot[], const char FNsnap[])
{
    const int st = run_tti2d(nx, nz, dx, dz, npml, favg, pfac, fs, zs, nt, dt, FNv, FNe, FNd, FNtheta, FNshot,
                             FNsnap);
    g_status = st;
    return st;
}

int vti_dropin_status(void) { return g_status.load(); }

}  // extern "C"
