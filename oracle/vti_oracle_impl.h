/*
 * vti_oracle_impl.h — body of the CPU restatement, included twice by
 * vti_oracle.c with REAL = float (SFX _f32) and REAL = double (SFX _f64).
 * TEST INFRASTRUCTURE ONLY (see vti_oracle.h).
 *
 * The float instantiation keeps the reference's left-to-right evaluation
 * order and its float/double promotions; vti_oracle.c is compiled with
 * -ffp-contract=off so no multiply-add is fused.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define SIM FN(orc_sim)

typedef struct SIM {
    orc_params p;
    int nnx, nny, nnz;
    long n;
    REAL *vp, *ep, *de;     /* padded model */
    REAL *tab[12];          /* coffx1 coffx2 coffy1 coffy2 coffz1 coffz2 acoffx1 .. acoffz2 */
    REAL *f[11];            /* u v w px py pz qx qy qz P Q */
    REAL d0;
    REAL c[4];
    REAL dtx, dty, dtz;
    REAL *amp;              /* per-step source gain pfac*source (0 when inactive) */
    int ng;                 /* Gaussian footprint (CPU source form) */
    long gidx[49];
    double gw[49];
    int tap_n;              /* SV taper factor per integer r^2, 0..tap_n-1 */
    double *tap;
    int split_z;            /* unsplit mode: row kept in the split form (-1: none) */
    REAL *cth, *sth;        /* TTI: cos and sin of the padded tilt per cell (NULL otherwise) */
} SIM;

/* Ricker / Gaussian wavelet.
 * GPU_vti2dfd_kernels.cu:183-198, CPU_vti2dfd_kernels.c:30-48. */
static REAL FN(wavelet)(REAL ts, REAL favg, int wtype)
{
    REAL src = 0;
    if (wtype == 1) {
        REAL x = favg * ts;
        REAL xx = x * x;
        src = (REAL)((1 - 2 * ORC_PI * ORC_PI * (xx)) * exp(-(ORC_PI * ORC_PI * xx)));
    } else if (wtype == 2) {
        REAL x = (REAL)((-4) * favg * favg * ORC_PI * ORC_PI / log(0.1));
        src = (REAL)((-2) * ORC_PI * ORC_PI * ts * exp(-x * ts * ts));
    } else if (wtype == 3) {
        REAL x = (REAL)((-1) * favg * favg * ORC_PI * ORC_PI / log(0.1));
        src = (REAL)exp(-x * ts * ts);
    } else if (wtype == 4) {
        /* Toafd2domp_2.c:154: (favg*ts*favg*ts) in float inside the first
         * factor, a double product chain inside exp */
        REAL x4 = favg * ts * favg * ts;
        src = (REAL)((1 - 2 * ORC_PI * ORC_PI * (x4)) *
                     exp(-(ORC_PI * ORC_PI * (double)favg * (double)ts * (double)favg * (double)ts)));
    }
    return src;
}

/* 1-D damping tables for one axis of nn_int interior cells.
 * GPU form: GPU_vti2dfd_kernels.cu:108-139 (= GPU_vti3dfd_kernels.cu:283-314).
 * CPU form: CPU_vti2dfd_kernels.c:271-332 — acoff2 is built from coff1
 * (:307,309,314,319) and the right-side stress position is
 * (1+i-nx-npd)/npd (:313). */
static void FN(axis_tables)(int nn_int, int npml, REAL dtd0, int kind,
                            REAL *c1, REAL *c2, REAL *a1, REAL *a2)
{
    int N = nn_int + 2 * npml;
    for (int id = 0; id < N; id++) {
        double xv, xs;
        if (id >= npml && id < npml + nn_int) {
            c1[id] = 1; c2[id] = 1; a1[id] = 1; a2[id] = 1;
            continue;
        }
        if (id < npml) {
            xv = (npml - 0.5 - id) / npml;
            xs = ((npml - id) * 1.0) / npml;
        } else {
            xv = (0.5 + id - nn_int - npml) / npml;
            xs = kind == 0 ? ((id - nn_int - npml) * 1.0) / npml
                           : ((1 + id - nn_int - npml) * 1.0) / npml;
        }
        double ev = (double)dtd0 * pow(xv, 2.0);
        double es = (double)dtd0 * pow(xs, 2.0);
        c1[id] = (REAL)(1.0 / (1.0 + ev / 2.0));
        c2[id] = (REAL)((double)c1[id] * (1.0 - ev / 2.0));
        a1[id] = (REAL)(1.0 / (1.0 + es / 2.0));
        a2[id] = kind == 0 ? (REAL)((double)a1[id] * (1.0 - es / 2.0))
                           : (REAL)((double)c1[id] * (1.0 - es / 2.0));
    }
}

void *FN(orc_create)(const orc_params *p, const float *vp, const float *eps,
                     const float *del)
{
    SIM *s = (SIM *)calloc(1, sizeof(SIM));
    if (!s) return NULL;
    s->p = *p;
    s->split_z = -1;
    int ny = p->ndim == 3 ? p->ny : 1;
    int np = p->npml;
    s->nnx = p->nx + 2 * np;
    s->nnz = p->nz + 2 * np;
    s->nny = p->ndim == 3 ? ny + 2 * np : 1;
    s->n = (long)s->nnx * s->nny * s->nnz;
    long n = s->n;
    s->vp = (REAL *)malloc(n * sizeof(REAL));
    s->ep = (REAL *)malloc(n * sizeof(REAL));
    s->de = (REAL *)malloc(n * sizeof(REAL));
    for (int k = 0; k < 11; k++) s->f[k] = (REAL *)calloc(n, sizeof(REAL));
    int nnmax = s->nnx > s->nnz ? s->nnx : s->nnz;
    if (s->nny > nnmax) nnmax = s->nny;
    for (int k = 0; k < 12; k++) s->tab[k] = (REAL *)calloc(nnmax, sizeof(REAL));
    s->amp = (REAL *)calloc(p->nt > 0 ? p->nt : 1, sizeof(REAL));
    if (p->tti) {   /* tilt 0 until orc_set_theta */
        s->cth = (REAL *)malloc(n * sizeof(REAL));
        s->sth = (REAL *)calloc(n, sizeof(REAL));
        for (long k = 0; k < n; k++) s->cth[k] = 1;
    }

    /* Edge-replicating pad: pad_vv2d GPU_vti2dfd_kernels.cu:49-74 and
     * pad_vv3d GPU_vti3dfd_kernels.cu:208-255 copy x, then y, then z;
     * the result equals clamping each coordinate into the interior.
     * (CPU_vti2dfd_kernels.c:244 returns early on a zero value and leaves
     * pad cells uninitialised; that defect is not restated.) */
    for (int iy = 0; iy < s->nny; iy++)
        for (int ix = 0; ix < s->nnx; ix++)
            for (int iz = 0; iz < s->nnz; iz++) {
                int jx = ix < np ? 0 : (ix >= np + p->nx ? p->nx - 1 : ix - np);
                int jz = iz < np ? 0 : (iz >= np + p->nz ? p->nz - 1 : iz - np);
                int jy = 0;
                if (p->ndim == 3) jy = iy < np ? 0 : (iy >= np + ny ? ny - 1 : iy - np);
                long src = (long)jz + (long)jx * p->nz + (long)jy * p->nz * p->nx;
                long id = (long)iz + (long)ix * s->nnz + (long)iy * s->nnz * s->nnx;
                s->vp[id] = vp[src];
                s->ep[id] = eps[src];
                s->de[id] = del[src];
            }

    /* c[] GPU_vti2dfd_kernels.cu:37, cal_c CPU_vti2dfd_kernels.c:334-343 */
    s->c[0] = (REAL)1.196289f;
    s->c[1] = (REAL)-0.0797526f;
    s->c[2] = (REAL)0.009570313f;
    s->c[3] = (REAL)-0.0006975447f;
    if (sizeof(REAL) == sizeof(double)) {
        s->c[0] = 1.196289; s->c[1] = -0.0797526;
        s->c[2] = 0.009570313; s->c[3] = -0.0006975447;
    }
    REAL dt = p->dt;
    s->dtx = dt / (REAL)p->dx;
    s->dty = dt / (REAL)p->dy;
    s->dtz = dt / (REAL)p->dz;

    /* d0: get_d02d GPU_vti2dfd_kernels.cu:105, get_d03d GPU_vti3dfd_kernels.cu:205,
     * get_constant CPU_vti2dfd_kernels.c:185-192 */
    if (p->d0_kind == 0) {
        if (p->ndim == 2) {
            REAL sdx = (REAL)p->dx + (REAL)p->dz;
            s->d0 = (REAL)(10.0 * s->vp[(long)s->nnx * s->nnz / 2] * log(100000.0) /
                           (2.0 * np * (sdx / 2.0)));
        } else {
            REAL sdx = (REAL)p->dx + (REAL)p->dy + (REAL)p->dz;
            s->d0 = (REAL)(10.0 * s->vp[(long)s->nny * s->nnx * s->nnz / 2] * log(100000.0) /
                           (2.0 * np * (sdx / 3.0)));
        }
    } else if (p->d0_kind == 1) {
        REAL vmax = s->vp[np];
        for (long id = np; id < n; id++)
            if (vmax < s->vp[id]) vmax = s->vp[id];
        s->d0 = (REAL)(3.0 * vmax * log(100000.0) / (2.0 * np * (REAL)p->dx));
    } else {
        s->d0 = 0;
    }
    REAL dtd0 = dt * s->d0;
    FN(axis_tables)(p->nx, np, dtd0, p->acoff_kind, s->tab[0], s->tab[1], s->tab[6], s->tab[7]);
    FN(axis_tables)(p->nz, np, dtd0, p->acoff_kind, s->tab[4], s->tab[5], s->tab[10], s->tab[11]);
    if (p->ndim == 3)
        FN(axis_tables)(ny, np, dtd0, p->acoff_kind, s->tab[2], s->tab[3], s->tab[8], s->tab[9]);

    /* Per-step source gain: add_source2d GPU_vti2dfd_kernels.cu:176-207 with
     * the float time t (t=dt, t+=dt: :419), ptsource CPU_vti2dfd_kernels.c:50-76. */
    {
        REAL t = p->t0_kind == 0 ? dt : (REAL)0;
        REAL tdelay = (REAL)(1.0 / p->favg);
        for (int it = 0; it < p->nt; it++) {
            REAL a = 0;
            if (t <= 2 * tdelay) {
                REAL ts = t - tdelay;
                REAL src = FN(wavelet)(ts, (REAL)p->favg, p->wtype);
                a = (REAL)p->pfac * src;
            }
            s->amp[it] = a;
            t += dt;
        }
    }
    /* 7x7 Gaussian footprint exp(-z*z-x*x), CPU_vti2dfd_kernels.c:67-74 */
    s->ng = 0;
    if (p->src_kind == 1) {
        for (int j = p->sz - 3; j <= p->sz + 3; j++)
            for (int i = p->sx - 3; i <= p->sx + 3; i++) {
                int x = i - p->sx, z = j - p->sz;
                s->gidx[s->ng] = (long)i * s->nnz + j +
                                 (long)p->sy * s->nnz * s->nnx * (p->ndim == 3);
                s->gw[s->ng] = exp(-z * z - x * x);
                s->ng++;
            }
    }
    /* SV taper factor as a function of the integer r^2.
     * GPU form GPU_vti2dfd_kernels.cu:260-278 (R=15, r=4) and
     * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:342-382 (R=18, r=7);
     * CPU form CPU_vti2dfd_kernels.c:124,137-155 (CR=15, inner CR*CR/16). */
    {
        int R = p->sv_R;
        s->tap_n = R * R + 1;
        s->tap = (double *)calloc(s->tap_n, sizeof(double));
        for (int r2 = 0; r2 <= R * R; r2++) {
            double f;
            if (p->sv_kind == 0) {
                int r = p->sv_r;
                if (r2 <= r * r) f = 0.0;
                else {
                    REAL sr = (REAL)sqrt((REAL)r2);
                    if (sizeof(REAL) == sizeof(float)) sr = sqrtf((float)r2);
                    f = 0.5 * (1 - cos(ORC_PI * ((sr - r) * 4.0 / (R * 3.0 - 1))));
                }
            } else {
                int CR = R;
                if (r2 <= (CR * CR / 16)) f = 0.0;
                else f = 0.5 * (1 - cos(ORC_PI * ((pow((r2), 0.5) - CR / 4.0) * 4.0 / (CR * 3.0 - 1))));
            }
            s->tap[r2] = f;
        }
    }
    return s;
}

void FN(orc_destroy)(void *h)
{
    SIM *s = (SIM *)h;
    if (!s) return;
    free(s->vp); free(s->ep); free(s->de);
    for (int k = 0; k < 11; k++) free(s->f[k]);
    for (int k = 0; k < 12; k++) free(s->tab[k]);
    free(s->amp); free(s->tap);
    free(s->cth); free(s->sth);
    free(s);
}

/* Tilt of the interior (radians, x-major, z fastest), padded by edge
 * replication as the model (pad_vv Toa_gpu_2dtti_fd_1orderfunction.cu:
 * 190-215 on theta after read_file :217-242 converted it to radians, :235);
 * the kernels take cos(theta[id]) and sin(theta[id]) of the float angle
 * (:110-111, :170-177): cosf / sinf here. */
int FN(orc_set_theta)(void *h, const float *theta)
{
    SIM *s = (SIM *)h;
    if (!s || !s->cth || !theta) return -1;
    const int np = s->p.npml, nx = s->p.nx, nz = s->p.nz;
    for (int ix = 0; ix < s->nnx; ix++)
        for (int iz = 0; iz < s->nnz; iz++) {
            int jx = ix < np ? 0 : (ix >= np + nx ? nx - 1 : ix - np);
            int jz = iz < np ? 0 : (iz >= np + nz ? nz - 1 : iz - np);
            const float th = theta[(long)jz + (long)jx * nz];
            const long id = (long)iz + (long)ix * s->nnz;
            if (sizeof(REAL) == sizeof(float)) {
                s->cth[id] = (REAL)cosf(th);
                s->sth[id] = (REAL)sinf(th);
            } else {
                s->cth[id] = (REAL)cos((double)th);
                s->sth[id] = (REAL)sin((double)th);
            }
        }
    return 0;
}

/* Effective eps/delta after the SV taper. */
static inline void FN(taper)(const SIM *s, int sv, int ix, int iy, int iz, long id, REAL *ee, REAL *dd)
{
    *ee = s->ep[id];
    *dd = s->de[id];
    if (!sv) return;
    int rx = ix - s->p.sv_cx, rz = iz - s->p.sv_cz, ry = 0;
    if (s->p.ndim == 3) ry = iy - s->p.sv_cy;
    int r2 = rx * rx + ry * ry + rz * rz;
    if (r2 <= s->p.sv_R * s->p.sv_R) {
        *ee = (REAL)(s->tap[r2] * (double)s->ep[id]);
        *dd = (REAL)(s->tap[r2] * (double)s->de[id]);
    }
}

#define SQRT_R(x) (sizeof(REAL) == sizeof(float) ? (REAL)sqrtf((float)(x)) : (REAL)sqrt((double)(x)))

/* Velocity update.  update_vel2d GPU_vti2dfd_kernels.cu:209-238
 * (= CPU_vti2dfd_kernels.c:79-106); update_vel3d GPU_vti3dfd_kernels.cu:90-129. */
static void FN(vel)(SIM *s)
{
    const int nnx = s->nnx, nny = s->nny, nnz = s->nnz;
    const long sx = nnz, sy = (long)nnz * nnx;
    REAL *u = s->f[0], *v = s->f[1], *w = s->f[2];
    const REAL *P = s->f[9], *Q = s->f[10];
    const REAL *c = s->c;
    const REAL *cx1 = s->tab[0], *cx2 = s->tab[1], *cy1 = s->tab[2], *cy2 = s->tab[3];
    const REAL *cz1 = s->tab[4], *cz2 = s->tab[5];
    const REAL dtx = s->dtx, dty = s->dty, dtz = s->dtz;
    if (s->p.tti) {
        /* update_vel Toa_gpu_2dtti_fd_1orderfunction.cu:81-114: the x and z
         * derivatives of P and Q rotated by the tilt, both components
         * damped by the product of the x and z factors */
        const REAL *cth = s->cth, *sth = s->sth;
#pragma omp parallel for schedule(static)
        for (int ix = 4; ix < nnx - 4; ix++)
            for (int iz = 4; iz < nnz - 4; iz++) {
                long id = (long)ix * nnz + iz;
                REAL Px = 0, Pz = 0, Qz = 0, Qx = 0;
                for (int im = 0; im < 4; im++) {
                    Px += c[im] * (P[id + (im + 1) * sx] - P[id - im * sx]);
                    Pz += c[im] * (P[id + im + 1] - P[id - im]);
                    Qz += c[im] * (Q[id + im + 1] - Q[id - im]);
                    Qx += c[im] * (Q[id + (im + 1) * sx] - Q[id - im * sx]);
                }
                const REAL cs = cth[id], sn = sth[id];
                const REAL g2 = ORC_PLANT == 9 ? cx2[ix] : cx2[ix] * cz2[iz];
                const REAL g1 = ORC_PLANT == 9 ? cx1[ix] : cx1[ix] * cz1[iz];
                const REAL su = ORC_PLANT == 7 ? -sn : sn;
                u[id] = g2 * u[id] - g1 * (cs * dtx * Px - su * dtz * Pz);
                w[id] = cx2[ix] * cz2[iz] * w[id] - cx1[ix] * cz1[iz] * (sn * dtx * Qx + cs * dtz * Qz);
            }
        return;
    }
    if (s->p.ndim == 2) {
#pragma omp parallel for schedule(static)
        for (int ix = 4; ix < nnx - 4; ix++)
            for (int iz = 4; iz < nnz - 4; iz++) {
                long id = (long)ix * nnz + iz;
                REAL xx = 0, zz = 0;
                for (int im = 0; im < 4; im++) {
                    xx += c[im] * (P[id + (im + 1) * sx] - P[id - im * sx]);
                    zz += c[im] * (Q[id + im + 1] - Q[id - im]);
                }
                u[id] = cx2[ix] * u[id] - cx1[ix] * dtx * xx;
                w[id] = cz2[iz] * w[id] - cz1[iz] * dtz * zz;
            }
    } else {
#pragma omp parallel for schedule(static)
        for (int iy = 4; iy < nny - 4; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + (long)ix * nnz + (long)iy * sy;
                    REAL xx = 0, yy = 0, zz = 0;
                    const int pyo = ORC_PLANT == 1, ty = ORC_PLANT == 2 ? ix : iy;
                    for (int im = 0; im < 4; im++) {
                        yy += c[im] * (P[id + (im + 1 - pyo) * sy] - P[id - (im + pyo) * sy]);
                        xx += c[im] * (P[id + (im + 1) * sx] - P[id - im * sx]);
                        zz += c[im] * (Q[id + im + 1] - Q[id - im]);
                    }
                    u[id] = cx2[ix] * u[id] - cx1[ix] * dtx * xx;
                    v[id] = cy2[ty] * v[id] - cy1[ty] * dty * yy;
                    w[id] = cz2[iz] * w[id] - cz1[iz] * dtz * zz;
                }
    }
}

/* Stress update.  update_stress2d GPU_vti2dfd_kernels.cu:240-299
 * (= CPU_vti2dfd_kernels.c:108-177 with P=px+pz+s at :173-174);
 * update_stress3d GPU_vti3dfd_kernels.cu:131-201. */
static void FN(stress)(SIM *s, int it, int inject, int sv)
{
    const int nnx = s->nnx, nny = s->nny, nnz = s->nnz;
    const long sx = nnz, sy = (long)nnz * nnx;
    const REAL *u = s->f[0], *v = s->f[1], *w = s->f[2];
    REAL *px = s->f[3], *py = s->f[4], *pz = s->f[5];
    REAL *qx = s->f[6], *qy = s->f[7], *qz = s->f[8];
    REAL *P = s->f[9], *Q = s->f[10];
    const REAL *c = s->c, *vp = s->vp;
    const REAL *ax1 = s->tab[6], *ax2 = s->tab[7], *ay1 = s->tab[8], *ay2 = s->tab[9];
    const REAL *az1 = s->tab[10], *az2 = s->tab[11];
    const REAL dtx = s->dtx, dty = s->dty, dtz = s->dtz;
    if (s->p.tti) {
        /* update_stress Toa_gpu_2dtti_fd_1orderfunction.cu:116-183 */
        const REAL *cth = s->cth, *sth = s->sth;
#pragma omp parallel for schedule(static)
        for (int ix = 4; ix < nnx - 4; ix++)
            for (int iz = 4; iz < nnz - 4; iz++) {
                long id = (long)ix * nnz + iz;
                REAL ee, dd;
                FN(taper)(s, sv, ix, 0, iz, id, &ee, &dd);
                REAL ux = 0, uz = 0, wz = 0, wx = 0;
                for (int im = 0; im < 4; im++) {
                    ux += c[im] * (u[id + im * sx] - u[id - (im + 1) * sx]);
                    uz += c[im] * (u[id + im] - u[id - im - 1]);
                    wz += c[im] * (w[id + im] - w[id - im - 1]);
                    wx += c[im] * (w[id + im * sx] - w[id - (im + 1) * sx]);
                }
                const REAL vv = vp[id], cs = cth[id], sn = sth[id];
                const REAL a2 = ax2[ix] * az2[iz], a1 = ax1[ix] * az1[iz];
                const REAL e2 = 1 + 2 * ee, sd = SQRT_R(1 + 2 * dd);
                const REAL cw = ORC_PLANT == 8 ? sn : cs, sw = ORC_PLANT == 8 ? cs : sn;
                px[id] = a2 * px[id] - a1 * vv * vv * e2 * (cs * dtx * ux - sn * dtz * uz);
                pz[id] = a2 * pz[id] - a1 * vv * vv * sd * (sw * dtx * wx + cw * dtz * wz);
                qx[id] = a2 * qx[id] - a1 * vv * vv * sd * (cs * dtx * ux - sn * dtz * uz);
                qz[id] = a2 * qz[id] - a1 * vv * vv * (sn * dtx * wx + cs * dtz * wz);
                P[id] = px[id] + pz[id];
                Q[id] = qx[id] + qz[id];
            }
        return;
    }
    if (s->p.ndim == 2) {
#pragma omp parallel for schedule(static)
        for (int ix = 4; ix < nnx - 4; ix++)
            for (int iz = 4; iz < nnz - 4; iz++) {
                long id = (long)ix * nnz + iz;
                REAL ee, dd;
                FN(taper)(s, sv, ix, 0, iz, id, &ee, &dd);
                REAL xx = 0, zz = 0;
                for (int im = 0; im < 4; im++) {
                    xx += c[im] * (u[id + im * sx] - u[id - (im + 1) * sx]);
                    zz += c[im] * (w[id + im] - w[id - im - 1]);
                }
                REAL vv = vp[id];
                REAL e2 = 1 + 2 * ee;
                REAL sd = SQRT_R(1 + 2 * dd);
                if (s->p.unsplit && s->p.src_kind == 0 && iz != s->split_z && ax1[ix] == 1 && ax2[ix] == 1 &&
                    az1[iz] == 1 && az2[iz] == 1) {
                    /* unsplit interior (product option, not the reference) */
                    REAL a = vv * vv;
                    P[id] = P[id] - (a * e2 * dtx * xx + a * sd * dtz * zz);
                    Q[id] = Q[id] - (a * sd * dtx * xx + a * dtz * zz);
                    continue;
                }
                px[id] = ax2[ix] * px[id] - ax1[ix] * vv * vv * e2 * dtx * xx;
                if (s->p.omp_form) {
                    /* Toafd2domp_2.c:176-180: (vp*vp)*pow(1+2*del,0.5)*dt*xx
                     * is a double product, subtracted in double */
                    const double sq = pow((double)(1 + 2 * dd), 0.5);
                    pz[id] = (REAL)((double)(az2[iz] * pz[id]) -
                                    (double)(az1[iz] * vv * vv) * sq * (double)dtz * (double)zz);
                    qx[id] = (REAL)((double)(ax2[ix] * qx[id]) -
                                    (double)(ax1[ix] * vv * vv) * sq * (double)dtx * (double)xx);
                } else {
                    pz[id] = az2[iz] * pz[id] - az1[iz] * vv * vv * sd * dtz * zz;
                    qx[id] = ax2[ix] * qx[id] - ax1[ix] * vv * vv * sd * dtx * xx;
                }
                qz[id] = az2[iz] * qz[id] - az1[iz] * vv * vv * dtz * zz;
                P[id] = px[id] + pz[id];
                Q[id] = qx[id] + qz[id];
            }
    } else {
#pragma omp parallel for schedule(static)
        for (int iy = 4; iy < nny - 4; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + (long)ix * nnz + (long)iy * sy;
                    REAL ee, dd;
                    FN(taper)(s, sv, ix, iy, iz, id, &ee, &dd);
                    REAL xx = 0, yy = 0, zz = 0;
                    const int pxo = ORC_PLANT == 3;
                    for (int im = 0; im < 4; im++) {
                        yy += c[im] * (v[id + im * sy] - v[id - (im + 1) * sy]);
                        xx += c[im] * (u[id + (im + pxo) * sx] - u[id - (im + 1 - pxo) * sx]);
                        zz += c[im] * (w[id + im] - w[id - im - 1]);
                    }
                    REAL vv = vp[id];
                    REAL e2 = 1 + 2 * ee;
                    REAL sd = SQRT_R(1 + 2 * dd);
                    const REAL dtyp = ORC_PLANT == 4 ? dtx : dty;
                    const REAL sdy = ORC_PLANT == 6 ? e2 : sd;
                    const int dX = !(ax1[ix] == 1 && ax2[ix] == 1), dY = !(ay1[iy] == 1 && ay2[iy] == 1);
                    const int dZ = !(az1[iz] == 1 && az2[iz] == 1);
                    const int fs = iz == s->split_z;
                    if (s->p.unsplit && !fs && dX + dY + dZ == 0) {
                        /* unsplit interior (product option, not the reference):
                         * P -= sum of the three split increments, Q likewise */
                        REAL a = vv * vv;
                        P[id] = P[id] - (a * e2 * dtx * xx + a * e2 * dtyp * yy + a * sd * dtz * zz);
                        Q[id] = Q[id] - (a * sd * dtx * xx + a * sdy * dtyp * yy + a * dtz * zz);
                        continue;
                    }
                    if (s->p.unsplit && !fs && dX + dY + dZ < 3) {
                        /* one or two damped axes: their split components are
                         * kept, the undamped rest is P minus them */
                        REAL a = vv * vv, z0 = 0;
                        REAL npx = dX ? ax2[ix] * px[id] - ax1[ix] * vv * vv * e2 * dtx * xx : z0;
                        REAL npy = dY ? ay2[iy] * py[id] - ay1[iy] * vv * vv * e2 * dtyp * yy : z0;
                        REAL npz = dZ ? az2[iz] * pz[id] - az1[iz] * vv * vv * sd * dtz * zz : z0;
                        REAL nqx = dX ? ax2[ix] * qx[id] - ax1[ix] * vv * vv * sd * dtx * xx : z0;
                        REAL nqy = dY ? ay2[iy] * qy[id] - ay1[iy] * vv * vv * sdy * dtyp * yy : z0;
                        REAL nqz = dZ ? az2[iz] * qz[id] - az1[iz] * vv * vv * dtz * zz : z0;
                        REAL rp = ((P[id] - (dX ? px[id] : z0)) - (dY ? py[id] : z0)) - (dZ ? pz[id] : z0);
                        REAL rq = ((Q[id] - (dX ? qx[id] : z0)) - (dY ? qy[id] : z0)) - (dZ ? qz[id] : z0);
                        REAL tp = ((dX ? z0 : a * e2 * dtx * xx) + (dY ? z0 : a * e2 * dtyp * yy)) + (dZ ? z0 : a * sd * dtz * zz);
                        REAL tq = ((dX ? z0 : a * sd * dtx * xx) + (dY ? z0 : a * sdy * dtyp * yy)) + (dZ ? z0 : a * dtz * zz);
                        if (dX) px[id] = npx, qx[id] = nqx;
                        if (dY) py[id] = npy, qy[id] = nqy;
                        if (dZ) pz[id] = npz, qz[id] = nqz;
                        P[id] = ((npx + npy) + npz) + (rp - tp);
                        Q[id] = ((nqx + nqy) + nqz) + (rq - tq);
                        continue;
                    }
                    px[id] = ax2[ix] * px[id] - ax1[ix] * vv * vv * e2 * dtx * xx;
                    py[id] = ay2[iy] * py[id] - ay1[iy] * vv * vv * e2 * dtyp * yy;
                    pz[id] = az2[iz] * pz[id] - az1[iz] * vv * vv * sd * dtz * zz;
                    qx[id] = ax2[ix] * qx[id] - ax1[ix] * vv * vv * sd * dtx * xx;
                    qy[id] = ay2[iy] * qy[id] - ay1[iy] * vv * vv * sdy * dtyp * yy;
                    qz[id] = az2[iz] * qz[id] - az1[iz] * vv * vv * dtz * zz;
                    P[id] = px[id] + py[id] + pz[id];
                    Q[id] = qx[id] + qy[id] + qz[id];
                }
    }
    /* In-stress Gaussian source: P = px+pz+s, Q = qx+qz+s (only inside the
     * updated region, where s is nonzero only on the footprint). */
    if (inject && s->p.src_kind == 1 && s->amp[it] != 0) {
        for (int k = 0; k < s->ng; k++) {
            long id = s->gidx[k];
            REAL sv = (REAL)((double)s->amp[it] * s->gw[k]);
            if (s->p.ndim == 2)
                P[id] = px[id] + pz[id] + sv, Q[id] = qx[id] + qz[id] + sv;
            else
                P[id] = px[id] + py[id] + pz[id] + sv, Q[id] = qx[id] + qy[id] + qz[id] + sv;
        }
    }
}

/* Receiver gather after the stress update: shot_record2d
 * GPU_vti2dfd_kernels.cu:141-149 (P), CPU_vti2dfd_kernels.c:488-491 (P+P),
 * 3D RTM shot_record Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:331-332 (P+Q). */
static void FN(record)(const SIM *s, REAL *rec)
{
    const orc_params *p = &s->p;
    const REAL *P = s->f[9], *Q = s->f[10];
    int nry = p->ndim == 3 ? p->rec_ny : 1;
    for (int iry = 0; iry < nry; iry++)
        for (int irx = 0; irx < p->rec_nx; irx++) {
            int ix = p->rec_x0 + irx * p->rec_dx;
            int iy = p->ndim == 3 ? p->rec_y0 + iry * p->rec_dy : 0;
            long id = p->rec_z + (long)ix * s->nnz + (long)iy * s->nnz * s->nnx;
            REAL val = P[id];
            if (p->rec_kind == 1) val = P[id] + P[id];
            else if (p->rec_kind == 2) val = P[id] + Q[id];
            rec[irx + (long)p->rec_nx * iry] = val;
        }
}

static int FN(step_ex)(SIM *s, int it, int inject, int sv, REAL *rec)
{
    if (!s || it < 0 || it >= s->p.nt) return -1;
    if (inject && s->p.src_kind == 0) {
        /* add_source2d GPU_vti2dfd_kernels.cu:200-206: before update_vel */
        long id = s->p.sz + (long)s->p.sx * s->nnz + (long)s->p.sy * s->nnz * s->nnx * (s->p.ndim == 3);
        REAL a = s->amp[it];
        if (a != 0 && s->p.unsplit) {
            /* unsplit interior: the injected value must not persist in P, Q
             * (in the split form the stress update rebuilds them) */
            const REAL p0 = s->f[9][id], q0 = s->f[10][id];
            s->f[9][id] += a;
            s->f[10][id] += a;
            FN(vel)(s);
            s->f[9][id] = p0;
            s->f[10][id] = q0;
            FN(stress)(s, it, inject, sv);
            if (rec) FN(record)(s, rec);
            return 0;
        }
        if (a != 0 && ORC_PLANT == 5) {
            FN(vel)(s);
            s->f[9][id] += a;
            s->f[10][id] += a;
            FN(stress)(s, it, inject, sv);
            if (rec) FN(record)(s, rec);
            return 0;
        }
        if (a != 0) {
            s->f[9][id] += a;
            s->f[10][id] += a;
        }
    }
    FN(vel)(s);
    FN(stress)(s, it, inject, sv);
    if (rec) FN(record)(s, rec);
    return 0;
}

int FN(orc_step)(void *h, int it, REAL *rec)
{
    SIM *s = (SIM *)h;
    if (!s) return -1;
    return FN(step_ex)(s, it, 1, s->p.sv, rec);
}

int FN(orc_run)(void *h, REAL *shot)
{
    SIM *s = (SIM *)h;
    if (!s) return -1;
    const orc_params *p = &s->p;
    long nrec = (long)p->rec_nx * (p->ndim == 3 ? p->rec_ny : 1);
    REAL *row = (REAL *)malloc((nrec > 0 ? nrec : 1) * sizeof(REAL));
    for (int it = 0; it < p->nt; it++) {
        FN(orc_step)(h, it, row);
        if (shot)
            for (long r = 0; r < nrec; r++) shot[it + (long)p->nt * r] = row[r];
    }
    free(row);
    return 0;
}

static REAL *FN(select)(SIM *s, int which, long *len)
{
    if (which >= ORC_U && which <= ORC_Q) { *len = s->n; return s->f[which]; }
    if (which == ORC_VP) { *len = s->n; return s->vp; }
    if (which == ORC_EPS) { *len = s->n; return s->ep; }
    if (which == ORC_DEL) { *len = s->n; return s->de; }
    if (which >= ORC_COFFX1 && which <= ORC_ACOFFZ2) {
        int k = which - ORC_COFFX1;
        int axis = (k % 6) / 2;   /* 0 x, 1 y, 2 z */
        *len = axis == 0 ? s->nnx : (axis == 1 ? s->nny : s->nnz);
        return s->tab[k];
    }
    if (which == ORC_SRCAMP) { *len = s->p.nt; return s->amp; }
    if (which == ORC_D0) { *len = 1; return &s->d0; }
    if (which == ORC_CTH && s->cth) { *len = s->n; return s->cth; }
    if (which == ORC_STH && s->sth) { *len = s->n; return s->sth; }
    *len = 0;
    return NULL;
}

int FN(orc_get)(void *h, int which, REAL *out)
{
    long len;
    REAL *src = FN(select)((SIM *)h, which, &len);
    if (!src) return -1;
    memcpy(out, src, len * sizeof(REAL));
    return (int)(len > 0x7fffffff ? 0x7fffffff : len);
}

int FN(orc_set)(void *h, int which, const REAL *in)
{
    long len;
    REAL *dst = FN(select)((SIM *)h, which, &len);
    if (!dst) return -1;
    memcpy(dst, in, len * sizeof(REAL));
    return 0;
}

/* Poynting-vector angle gathers, one accumulation: Poynting_Adcigs
 * Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:718-766 (2D, every cdp) and
 * Toa_gpu_3dvti_rtm_adcigs_1orderfunction.cu:618-660 (3D, every dcdp-th cdp
 * in x and y).  Source Poynting vector (P u, [P v,] Q w), receiver
 * (P u, [P v,] Q w).  The reference negates the source vector (:746-747)
 * because its source field is rebuilt backward in time, where u changes
 * sign; the source field here is the exact forward field, whose own P u
 * already points along the incident wave (with the negation the bins would
 * hold 90 degrees minus the reflection angle).
 * a = S_s.S_g / (|S_s||S_g| (1 - 0.1)); cells with
 * |a| <= 1 add P_s P_g cos^3(ia degrees) to angle bin ia = (int)(0.5 acos(a)
 * in degrees / dangle) of their own gather when ia < nangle.  Fields are
 * padded volumes; adcigs [iy][ix][ia][iz] over the cdps.  Float arithmetic
 * as the reference's: the quotient in double (the 1 - 0.1 literal), acosf,
 * the angle in double rounded to float, cosf of a double argument. */
static void FN(adcig_core)(const SIM *s, int nangle, int dangle, int dcdp,
                           const REAL *sP, const REAL *sQ, const REAL *su, const REAL *sv, const REAL *sw,
                           const REAL *gP, const REAL *gQ, const REAL *gu, const REAL *gv, const REAL *gw,
                           REAL *adcigs)
{
    const orc_params *p = &s->p;
    const int three = p->ndim == 3, np = p->npml, nz = p->nz;
    const int d = three ? dcdp : 1;
    const int nxa = p->nx / d, nya = three ? p->ny / d : 1;
    for (int iya = 0; iya < nya; iya++)
        for (int ixa = 0; ixa < nxa; ixa++)
            for (int iz = 0; iz < nz; iz++) {
                long c = (iz + np) + (long)(ixa * d + np) * s->nnz + (three ? (long)(iya * d + np) * s->nnz * s->nnx : 0);
                REAL Ssx = sP[c] * su[c], Ssz = sQ[c] * sw[c];
                REAL Sgx = gP[c] * gu[c], Sgz = gQ[c] * gw[c];
                REAL b1, b2, num;
                if (three) {
                    REAL Ssy = sP[c] * sv[c], Sgy = gP[c] * gv[c];
                    b1 = Ssx * Ssx + Ssy * Ssy + Ssz * Ssz;
                    b2 = Sgx * Sgx + Sgy * Sgy + Sgz * Sgz;
                    num = Ssx * Sgx + Ssy * Sgy + Ssz * Sgz;
                } else {
                    b1 = Ssx * Ssx + Ssz * Ssz;
                    b2 = Sgx * Sgx + Sgz * Sgz;
                    num = Ssx * Sgx + Ssz * Sgz;
                }
                REAL a = (REAL)(num / (SQRT_R(b1 * b2) * (1 - 0.1)));
                if (!(a >= -1 && a <= 1)) continue;
                REAL ang = (REAL)(0.5 * (sizeof(REAL) == sizeof(float) ? (double)acosf((float)a) : acos((double)a)) *
                                  180.0 / ORC_PI);
                int ia = (int)(ang / (dangle * 1.0));
                if (ia >= nangle) continue;
                REAL cw = sizeof(REAL) == sizeof(float) ? (REAL)cosf((float)(ia * ORC_PI / 180.0))
                                                        : (REAL)cos(ia * ORC_PI / 180.0);
                long id = iz + (long)nz * ia + (long)nz * nangle * ixa + (long)nz * nangle * nxa * iya;
                adcigs[id] += sP[c] * gP[c] * cw * cw * cw;
            }
}

/* One accumulation from the current fields of a source and a receiver
 * simulation of the same grid. */
int FN(orc_adcig)(void *hs, void *hr, int nangle, int dangle, int dcdp, REAL *adcigs)
{
    SIM *s = (SIM *)hs, *g = (SIM *)hr;
    if (!s || !g || nangle < 1 || dangle < 1 || dcdp < 1) return -1;
    FN(adcig_core)(s, nangle, dangle, dcdp, s->f[9], s->f[10], s->f[0], s->f[1], s->f[2],
                   g->f[9], g->f[10], g->f[0], g->f[1], g->f[2], adcigs);
    return 0;
}

/* RTM of one shot, restating RTM() Toa_gpu2DvtiFdRtmAdcigsLaplace.cu:1378-1484
 * with the exact forward field of every step (all P snapshots kept in
 * memory).  hs: the shot's simulation (source side), fresh; hr: a second
 * simulation of the same model and shot (receiver side), fresh.
 * obs: [irec][it] record or NULL (use the modelled record, muted by the
 * caller-provided callback values already applied: pass the muted record
 * as obs).  img, illum: interior nx*ny*nz, accumulated. */
int FN(orc_rtm)(void *hs, void *hr, int sv_fwd, int sv_bwd, int inject_z, int illum_mode,
                int image_every, const REAL *obs, REAL *modelled, REAL *img, REAL *illum)
{
    return FN(orc_rtm_ex)(hs, hr, sv_fwd, sv_bwd, inject_z, illum_mode, image_every, obs, modelled, img, illum,
                          0, 1, 1, NULL);
}

/* orc_rtm with the angle gathers of every imaged step (nangle > 0): the
 * source fields P, Q, u, v, w of every step are kept (padded). */
int FN(orc_rtm_ex)(void *hs, void *hr, int sv_fwd, int sv_bwd, int inject_z, int illum_mode,
                   int image_every, const REAL *obs, REAL *modelled, REAL *img, REAL *illum,
                   int nangle, int dangle, int dcdp, REAL *adcigs)
{
    SIM *s = (SIM *)hs, *g = (SIM *)hr;
    if (!s || !g || image_every < 1) return -1;
    const orc_params *p = &s->p;
    const int np = p->npml, nt = p->nt;
    const int ny = p->ndim == 3 ? p->ny : 1;
    const long nint = (long)p->nx * ny * p->nz;
    const long nrec = (long)p->rec_nx * (p->ndim == 3 ? p->rec_ny : 1);
    const int gat = nangle > 0 && adcigs != NULL;
    REAL *snap = (REAL *)malloc((size_t)nt * nint * sizeof(REAL));
    REAL *row = (REAL *)malloc((nrec > 0 ? nrec : 1) * sizeof(REAL));
    REAL *full = gat ? (REAL *)malloc((size_t)nt * 5 * s->n * sizeof(REAL)) : NULL;
    if (!snap || !row || (gat && !full)) { free(snap); free(row); free(full); return -2; }
    const int yo = p->ndim == 3 ? np : 0;
    /* the records replace P, Q of row inject_z before each receiver step: in
     * the unsplit mode that row keeps the split form, which rebuilds P, Q
     * from the split fields, so the loaded values do not persist */
    g->split_z = inject_z;
    for (int it = 0; it < nt; it++) {
        FN(step_ex)(s, it, 1, sv_fwd, row);
        for (long r = 0; r < nrec; r++) modelled[it + (long)nt * r] = row[r];
        if (gat) {
            const int sel[5] = {9, 10, 0, 1, 2};   /* P Q u v w */
            for (int k = 0; k < 5; k++)
                memcpy(full + ((size_t)it * 5 + k) * s->n, s->f[sel[k]], s->n * sizeof(REAL));
        }
        REAL *P = s->f[9], *Q = s->f[10];
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < p->nx; ix++)
                for (int iz = 0; iz < p->nz; iz++) {
                    long c = (iz + np) + (long)(ix + np) * s->nnz + (long)(iy + yo) * s->nnz * s->nnx;
                    long o = iz + (long)ix * p->nz + (long)iy * p->nz * p->nx;
                    snap[(size_t)it * nint + o] = P[c];
                    /* cal_illumination :671-675 */
                    REAL v = illum[o] + (P[c] * P[c] + Q[c] * Q[c]);
                    if (illum_mode == 0 && v <= 0) v = 1;
                    illum[o] = v;
                }
    }
    const REAL *data = obs ? obs : modelled;
    for (int it = nt - 1; it >= 0; it--) {
        /* shot_record(load) :525-528 into g_P and g_Q */
        for (long r = 0; r < nrec; r++) {
            int irx = (int)(r % p->rec_nx), iry = (int)(r / p->rec_nx);
            int ix = p->rec_x0 + irx * p->rec_dx;
            int iy = p->ndim == 3 ? p->rec_y0 + iry * p->rec_dy : 0;
            long c = inject_z + (long)ix * g->nnz + (long)iy * g->nnz * g->nnx;
            g->f[9][c] = data[it + (long)nt * r];
            g->f[10][c] = data[it + (long)nt * r];
        }
        FN(step_ex)(g, it, 0, sv_bwd, NULL);
        REAL *gP = g->f[9], *gQ = g->f[10];
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < p->nx; ix++)
                for (int iz = 0; iz < p->nz; iz++) {
                    long c = (iz + np) + (long)(ix + np) * g->nnz + (long)(iy + yo) * g->nnz * g->nnx;
                    long o = iz + (long)ix * p->nz + (long)iy * p->nz * p->nx;
                    if (illum_mode == 1) illum[o] = illum[o] + (gP[c] * gP[c] + gQ[c] * gQ[c]);
                    if (it % image_every == 0) img[o] += snap[(size_t)it * nint + o] * gP[c];
                }
        if (gat && it % image_every == 0) {
            const REAL *f = full + (size_t)it * 5 * s->n;
            FN(adcig_core)(s, nangle, dangle, dcdp, f, f + s->n, f + 2 * s->n, f + 3 * s->n, f + 4 * s->n,
                           g->f[9], g->f[10], g->f[0], g->f[1], g->f[2], adcigs);
        }
    }
    free(snap);
    free(row);
    free(full);
    return 0;
}

/* Misfit gradient J = 1/2 sum (d - dobs)^2 with respect to vp, by the
 * transpose of every line of the forward step written in scatter form
 * (the forward: vel then stress then record, GPU or CPU source form).
 * The absorbing-layer tables are held fixed.  All forward velocity states
 * are kept in memory (small grids).  resid: if non-NULL it replaces
 * d - dobs (the adjoint of the linear source-to-data map then follows).
 * Outputs: *misfit, grad_int (interior nx*ny*nz), grad_src (nt: adjoint
 * of the per-step source gain, point source form only), d_out ([irec][it]). */
int FN(orc_fwi_grad)(void *h, const REAL *dobs, const REAL *resid, REAL *misfit, REAL *grad_int,
                     REAL *grad_src, REAL *d_out)
{
    SIM *s = (SIM *)h;
    if (!s) return -1;
    const orc_params *p = &s->p;
    const int nt = p->nt, np = p->npml, three = p->ndim == 3;
    const int nnx = s->nnx, nny = s->nny, nnz = s->nnz;
    const long n = s->n, sx = nnz, sy = (long)nnz * nnx;
    const long nrec = (long)p->rec_nx * (three ? p->rec_ny : 1);
    const int nvf = three ? 3 : 2;
    REAL *hist = (REAL *)malloc((size_t)nt * nvf * n * sizeof(REAL));
    REAL *row = (REAL *)malloc((nrec > 0 ? nrec : 1) * sizeof(REAL));
    REAL *d = (REAL *)malloc((size_t)(nrec > 0 ? nrec : 1) * nt * sizeof(REAL));
    if (!hist || !row || !d) { free(hist); free(row); free(d); return -2; }
    /* forward: fresh state, keep u (v) w after every step */
    for (int k = 0; k < 11; k++) memset(s->f[k], 0, n * sizeof(REAL));
    for (int it = 0; it < nt; it++) {
        FN(step_ex)(s, it, 1, p->sv, row);
        for (long r = 0; r < nrec; r++) d[it + (long)nt * r] = row[r];
        memcpy(hist + ((size_t)it * nvf + 0) * n, s->f[0], n * sizeof(REAL));
        memcpy(hist + ((size_t)it * nvf + 1) * n, s->f[2], n * sizeof(REAL));
        if (three) memcpy(hist + ((size_t)it * nvf + 2) * n, s->f[1], n * sizeof(REAL));
    }
    double J = 0;
    for (long i = 0; i < nrec * nt; i++) {
        REAL rr = resid ? resid[i] : d[i] - dobs[i];
        J += 0.5 * (double)rr * (double)rr;
    }
    if (misfit) *misfit = (REAL)J;
    if (d_out) memcpy(d_out, d, (size_t)nrec * nt * sizeof(REAL));
    /* per-cell stress coefficients (taper applied) and their vp-derivatives */
    REAL *C[6], *dC[6];   /* px py pz qx qy qz */
    for (int k = 0; k < 6; k++) { C[k] = (REAL *)calloc(n, sizeof(REAL)); dC[k] = (REAL *)calloc(n, sizeof(REAL)); }
    const REAL *ax1 = s->tab[6], *ay1 = s->tab[8], *az1 = s->tab[10];
    for (int iy = 0; iy < nny; iy++)
        for (int ix = 0; ix < nnx; ix++)
            for (int iz = 0; iz < nnz; iz++) {
                long id = iz + ix * sx + iy * sy;
                REAL ee, dd;
                FN(taper)(s, p->sv, ix, iy, iz, id, &ee, &dd);
                REAL vv = s->vp[id], e2 = 1 + 2 * ee, sd = SQRT_R(1 + 2 * dd);
                REAL a1y = three ? ay1[iy] : 0;
                C[0][id] = ax1[ix] * vv * vv * e2 * s->dtx;  dC[0][id] = 2 * ax1[ix] * vv * e2 * s->dtx;
                C[1][id] = a1y * vv * vv * e2 * s->dty;      dC[1][id] = 2 * a1y * vv * e2 * s->dty;
                C[2][id] = az1[iz] * vv * vv * sd * s->dtz;  dC[2][id] = 2 * az1[iz] * vv * sd * s->dtz;
                C[3][id] = ax1[ix] * vv * vv * sd * s->dtx;  dC[3][id] = 2 * ax1[ix] * vv * sd * s->dtx;
                C[4][id] = a1y * vv * vv * sd * s->dty;      dC[4][id] = 2 * a1y * vv * sd * s->dty;
                C[5][id] = az1[iz] * vv * vv * s->dtz;       dC[5][id] = 2 * az1[iz] * vv * s->dtz;
            }
    /* adjoint state (zero outside the updated region R) */
    REAL *A[11];   /* u v w px py pz qx qy qz P Q */
    for (int k = 0; k < 11; k++) A[k] = (REAL *)calloc(n, sizeof(REAL));
    REAL *g = (REAL *)calloc(n, sizeof(REAL));
    REAL *ua = A[0], *va = A[1], *wa = A[2], *Pa = A[9], *Qa = A[10];
    REAL *sa[6] = {A[3], A[4], A[5], A[6], A[7], A[8]};
    const REAL *c = s->c;
    const REAL *cx1 = s->tab[0], *cx2 = s->tab[1], *cy1 = s->tab[2], *cy2 = s->tab[3];
    const REAL *cz1 = s->tab[4], *cz2 = s->tab[5];
    const REAL *ax2 = s->tab[7], *ay2 = s->tab[9], *az2 = s->tab[11];
    const int y0 = three ? 4 : 0, y1 = three ? nny - 4 : 1;
#define INR(ix_, iy_, iz_) ((ix_) >= 4 && (ix_) < nnx - 4 && (iz_) >= 4 && (iz_) < nnz - 4 && \
                            (!three || ((iy_) >= 4 && (iy_) < nny - 4)))
    for (int it = nt - 1; it >= 0; it--) {
        const REAL *uh = hist + ((size_t)it * nvf + 0) * n, *wh = hist + ((size_t)it * nvf + 1) * n;
        const REAL *vh = three ? hist + ((size_t)it * nvf + 2) * n : NULL;
        /* record^T */
        for (long r = 0; r < nrec; r++) {
            int irx = (int)(r % p->rec_nx), iry = (int)(r / p->rec_nx);
            long cc = p->rec_z + (long)(p->rec_x0 + irx * p->rec_dx) * sx +
                      (three ? (long)(p->rec_y0 + iry * p->rec_dy) * sy : 0);
            REAL rr = resid ? resid[it + (long)nt * r] : d[it + (long)nt * r] - dobs[it + (long)nt * r];
            if (p->rec_kind == 0) Pa[cc] += rr;
            else if (p->rec_kind == 1) Pa[cc] += rr + rr;
            else { Pa[cc] += rr; Qa[cc] += rr; }
        }
        /* stress^T: P = sum of p-splits, Q = sum of q-splits, then the splits */
        for (int iy = y0; iy < y1; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + ix * sx + iy * sy;
                    sa[0][id] += Pa[id]; sa[2][id] += Pa[id]; if (three) sa[1][id] += Pa[id];
                    sa[3][id] += Qa[id]; sa[5][id] += Qa[id]; if (three) sa[4][id] += Qa[id];
                    Pa[id] = 0; Qa[id] = 0;
                }
        for (int iy = y0; iy < y1; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + ix * sx + iy * sy;
                    REAL xx = 0, yy = 0, zz = 0;
                    for (int m = 0; m < 4; m++) {
                        xx += c[m] * (uh[id + m * sx] - uh[id - (m + 1) * sx]);
                        zz += c[m] * (wh[id + m] - wh[id - m - 1]);
                        if (three) yy += c[m] * (vh[id + m * sy] - vh[id - (m + 1) * sy]);
                    }
                    g[id] -= sa[0][id] * dC[0][id] * xx + sa[3][id] * dC[3][id] * xx +
                             sa[2][id] * dC[2][id] * zz + sa[5][id] * dC[5][id] * zz;
                    if (three) g[id] -= sa[1][id] * dC[1][id] * yy + sa[4][id] * dC[4][id] * yy;
                    REAL fx = sa[0][id] * C[0][id] + sa[3][id] * C[3][id];
                    REAL fz = sa[2][id] * C[2][id] + sa[5][id] * C[5][id];
                    REAL fy = three ? sa[1][id] * C[1][id] + sa[4][id] * C[4][id] : 0;
                    for (int m = 0; m < 4; m++) {
                        /* px[x] -= C c_m (u[x+m] - u[x-m-1]) */
                        if (INR(ix + m, iy, iz)) ua[id + m * sx] -= c[m] * fx;
                        if (INR(ix - m - 1, iy, iz)) ua[id - (m + 1) * sx] += c[m] * fx;
                        if (INR(ix, iy, iz + m)) wa[id + m] -= c[m] * fz;
                        if (INR(ix, iy, iz - m - 1)) wa[id - m - 1] += c[m] * fz;
                        if (three) {
                            if (INR(ix, iy + m, iz)) va[id + m * sy] -= c[m] * fy;
                            if (INR(ix, iy - m - 1, iz)) va[id - (m + 1) * sy] += c[m] * fy;
                        }
                    }
                }
        for (int iy = y0; iy < y1; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + ix * sx + iy * sy;
                    sa[0][id] *= ax2[ix]; sa[3][id] *= ax2[ix];
                    sa[2][id] *= az2[iz]; sa[5][id] *= az2[iz];
                    if (three) { sa[1][id] *= ay2[iy]; sa[4][id] *= ay2[iy]; }
                }
        /* vel^T: u[x] -= cx1 dtx sum c_m (P[x+m+1] - P[x-m]) */
        for (int iy = y0; iy < y1; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + ix * sx + iy * sy;
                    REAL hx = cx1[ix] * s->dtx * ua[id];
                    REAL hz = cz1[iz] * s->dtz * wa[id];
                    REAL hy = three ? cy1[iy] * s->dty * va[id] : 0;
                    for (int m = 0; m < 4; m++) {
                        if (INR(ix + m + 1, iy, iz)) Pa[id + (m + 1) * sx] -= c[m] * hx;
                        if (INR(ix - m, iy, iz)) Pa[id - m * sx] += c[m] * hx;
                        if (INR(ix, iy, iz + m + 1)) Qa[id + m + 1] -= c[m] * hz;
                        if (INR(ix, iy, iz - m)) Qa[id - m] += c[m] * hz;
                        if (three) {
                            if (INR(ix, iy + m + 1, iz)) Pa[id + (m + 1) * sy] -= c[m] * hy;
                            if (INR(ix, iy - m, iz)) Pa[id - m * sy] += c[m] * hy;
                        }
                    }
                }
        for (int iy = y0; iy < y1; iy++)
            for (int ix = 4; ix < nnx - 4; ix++)
                for (int iz = 4; iz < nnz - 4; iz++) {
                    long id = iz + ix * sx + iy * sy;
                    ua[id] *= cx2[ix]; wa[id] *= cz2[iz];
                    if (three) va[id] *= cy2[iy];
                }
        /* source^T (point form: P[src] += a, Q[src] += a before the velocity update) */
        if (grad_src) {
            long sid = p->sz + (long)p->sx * sx + (three ? (long)p->sy * sy : 0);
            grad_src[it] = p->src_kind == 0 ? Pa[sid] + Qa[sid] : 0;
        }
    }
#undef INR
    /* fold the padded-grid gradient onto the interior cells it copies */
    const int ny = three ? p->ny : 1;
    for (long j = 0; j < (long)p->nx * ny * p->nz; j++) grad_int[j] = 0;
    for (int iy = 0; iy < nny; iy++)
        for (int ix = 0; ix < nnx; ix++)
            for (int iz = 0; iz < nnz; iz++) {
                int jx = ix < np ? 0 : (ix >= np + p->nx ? p->nx - 1 : ix - np);
                int jz = iz < np ? 0 : (iz >= np + p->nz ? p->nz - 1 : iz - np);
                int jy = 0;
                if (three) jy = iy < np ? 0 : (iy >= np + ny ? ny - 1 : iy - np);
                grad_int[jz + (long)jx * p->nz + (long)jy * p->nz * p->nx] += g[iz + ix * sx + iy * sy];
            }
    for (int k = 0; k < 6; k++) { free(C[k]); free(dC[k]); }
    for (int k = 0; k < 11; k++) free(A[k]);
    free(g); free(hist); free(row); free(d);
    return 0;
}

/* Forward with a caller-given source gain per step (linear source map). */
int FN(orc_set_source)(void *h, const REAL *amp)
{
    SIM *s = (SIM *)h;
    if (!s) return -1;
    memcpy(s->amp, amp, s->p.nt * sizeof(REAL));
    return 0;
}

#undef SIM
#undef FN
#undef SQRT_R
