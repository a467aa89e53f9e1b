/*
 * CPU oracle for the advection hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or the CPU baseline.  The product path
 * (offline_particles_amd) never calls it.
 *
 * This is a plain-C FP64 restatement of the reference's Cython kernels, written
 * to keep their operation order so that results are reproducible bit for bit on
 * x86-64 (compile with -ffp-contract=off; the reference is built without FMA).
 * Each function cites the reference file:line it restates (paths relative to
 * src/offline_particles/ in the reference tree).
 *
 * Like the reference, kernels read *blended* FP64 field arrays: the caller first
 * performs the full-grid time blend (orc_blend_*), exactly as
 * TimeDependentField.get_field_data does (fields.py:394-430).
 *
 * Particle loops are OpenMP `schedule(static)` loops like the reference's
 * `prange(..., schedule='static')`; the RK2 kernels use chunks of 3000
 * particles (rk2_w_advection.pyx:81-91).
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* status codes: kernels/status.pxd:1-14 */
enum {
    ORC_INACTIVE = 1 << 7,
    ORC_NORMAL = 0,
    ORC_NONFINITE = 1 | ORC_INACTIVE,
    ORC_OUT_OF_DOMAIN = 2 | ORC_INACTIVE,
    ORC_MULTISTEP_1 = 10,
    ORC_MULTISTEP_2 = 11,
};

/* A contiguous FP64 array of 1, 2 or 3 active dimensions plus the offsets that
 * map rho-centred particle indices into it (kernels/_core.pyx:5-24). */
typedef struct {
    const double *a;
    int64_t n0, n1, n2;
    double o0, o1, o2;
} orc_field;

typedef struct {
    orc_field u, v, w, dx, dy, h, C, zeta;
} orc_fields;

/* SoA particle arrays; names follow the reference particle fields
 * (ab3_w_advection.pyx:245-258, rk2_w_advection.pyx:372-380). */
typedef struct {
    int64_t n;
    uint8_t *status;
    double *zidx, *yidx, *xidx, *z;
    double *dz0, *dz1, *dz2;
    double *dyidx0, *dyidx1, *dyidx2;
    double *dxidx0, *dxidx1, *dxidx2;
} orc_particles;

typedef struct {
    double dt, alpha, hc;
    int32_t NZ, pad_;
    double zmin, zmax, ymin, ymax, xmin, xmax;
} orc_scalars;

/* ------------------------------------------------------------------ */
/* L0: interpolation (kernels/_interpolation/linear.pxd)               */
/* ------------------------------------------------------------------ */

/* truncate_index (linear.pxd:3-9): `<Py_ssize_t>idx` then clamp to
 * [0, max_idx].  The cast truncates toward zero; on x86-64 an out-of-range or
 * NaN double converts to INT64_MIN, which the clamp then maps to 0.  That is
 * written out here so the C below is well defined. */
static inline int64_t orc_truncate_index(double idx, int64_t max_idx)
{
    if (!(idx > -9223372036854775808.0 && idx < 9223372036854775808.0))
        return 0; /* NaN / out of int64 range -> INT64_MIN -> clamped to 0 */
    int64_t i = (int64_t)idx;
    if (i < 0)
        i = 0;
    else if (i > max_idx)
        i = max_idx;
    return i;
}

/* linear_interpolation (linear.pxd:11-17) */
double orc_linear(const double *a, int64_t n0, double idx)
{
    int64_t I0 = orc_truncate_index(idx, n0 - 2);
    double f0 = idx - (double)I0;
    double g0 = 1.0 - f0;
    return g0 * a[I0] + f0 * a[I0 + 1];
}

/* bilinear_interpolation (linear.pxd:19-33) */
double orc_bilinear(const double *a, int64_t n0, int64_t n1, double idx0, double idx1)
{
    int64_t I0 = orc_truncate_index(idx0, n0 - 2);
    int64_t I1 = orc_truncate_index(idx1, n1 - 2);
    double f0 = idx0 - (double)I0;
    double f1 = idx1 - (double)I1;
    double g0 = 1.0 - f0;
    double g1 = 1.0 - f1;
    double v00 = a[I0 * n1 + I1];
    double v01 = a[I0 * n1 + I1 + 1];
    double v10 = a[(I0 + 1) * n1 + I1];
    double v11 = a[(I0 + 1) * n1 + I1 + 1];
    return g0 * g1 * v00 + g0 * f1 * v01 + f0 * g1 * v10 + f0 * f1 * v11;
}

/* trilinear_interpolation (linear.pxd:35-65), sum order of :58-65 */
double orc_trilinear(const double *a, int64_t n0, int64_t n1, int64_t n2,
                     double idx0, double idx1, double idx2)
{
    int64_t I0 = orc_truncate_index(idx0, n0 - 2);
    int64_t I1 = orc_truncate_index(idx1, n1 - 2);
    int64_t I2 = orc_truncate_index(idx2, n2 - 2);
    double f0 = idx0 - (double)I0;
    double f1 = idx1 - (double)I1;
    double f2 = idx2 - (double)I2;
    double g0 = 1.0 - f0;
    double g1 = 1.0 - f1;
    double g2 = 1.0 - f2;
    const double *p00 = a + (I0 * n1 + I1) * n2 + I2;
    const double *p01 = a + (I0 * n1 + I1 + 1) * n2 + I2;
    const double *p10 = a + ((I0 + 1) * n1 + I1) * n2 + I2;
    const double *p11 = a + ((I0 + 1) * n1 + I1 + 1) * n2 + I2;
    double v000 = p00[0], v001 = p00[1];
    double v010 = p01[0], v011 = p01[1];
    double v100 = p10[0], v101 = p10[1];
    double v110 = p11[0], v111 = p11[1];
    return (g0 * g1 * g2 * v000 +
            g0 * g1 * f2 * v001 +
            g0 * f1 * g2 * v010 +
            g0 * f1 * f2 * v011 +
            f0 * g1 * g2 * v100 +
            f0 * g1 * f2 * v101 +
            f0 * f1 * g2 * v110 +
            f0 * f1 * f2 * v111);
}

static inline double bil(const orc_field *F, double y, double x)
{
    return orc_bilinear(F->a, F->n0, F->n1, y + F->o0, x + F->o1);
}

static inline double tril(const orc_field *F, double z, double y, double x)
{
    return orc_trilinear(F->a, F->n0, F->n1, F->n2, z + F->o0, y + F->o1, x + F->o2);
}

/* ------------------------------------------------------------------ */
/* L0: ROMS vertical coordinate (kernels/roms/_vertical_coordinate.pxd) */
/* ------------------------------------------------------------------ */

/* _sigma_coordinate (:19-21) */
static inline double sigma_coordinate(double zidx, int NZ) { return (zidx + 0.5) / NZ - 1.0; }
/* _S_coordinate (:23-25) */
static inline double S_coordinate(double hc, double sigma, double h, double C)
{
    return (hc * sigma + h * C) / (hc + h);
}
/* _z_coordinate (:27-29) */
static inline double z_coordinate(double S, double h, double zeta) { return zeta + (zeta + h) * S; }

/* compute_z (:31-35) */
double orc_compute_z(double zidx, int NZ, double hc, double h, double C, double zeta)
{
    double sigma = sigma_coordinate(zidx, NZ);
    double S = S_coordinate(hc, sigma, h, C);
    return z_coordinate(S, h, zeta);
}

/* _S_from_z (:41-43) */
static inline double S_from_z(double z, double h, double zeta) { return (z - zeta) / (zeta + h); }
/* _sigma_from_Cidx (:45-48) */
static inline double sigma_from_Cidx(int64_t Cidx, double C_offset, int NZ)
{
    double zidx = (double)Cidx - C_offset;
    return sigma_coordinate(zidx, NZ);
}

/* _compute_Cidx_from_S (:50-84): edge returns then binary search */
static int64_t compute_Cidx_from_S(double S, double hc, int NZ, double h, const double *C,
                                   int64_t C_size, double C_offset)
{
    int64_t lo = 0, hi = C_size - 2, mid;
    double S_mid;
    if (S <= S_coordinate(hc, sigma_from_Cidx(0, C_offset, NZ), h, C[0]))
        return 0;
    else if (S >= S_coordinate(hc, sigma_from_Cidx(C_size - 2, C_offset, NZ), h, C[C_size - 2]))
        return C_size - 2;
    while (lo < hi - 1) {
        mid = (lo + hi) / 2; /* lo, hi >= 0 so C division == Python floor division */
        S_mid = S_coordinate(hc, sigma_from_Cidx(mid, C_offset, NZ), h, C[mid]);
        if (S_mid <= S)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

/* _zidx_from_S (:86-96) */
static double zidx_from_S(double S, double hc, int NZ, double h, const double *C,
                          int64_t C_size, double C_offset)
{
    int64_t Cidx = compute_Cidx_from_S(S, hc, NZ, h, C, C_size, C_offset);
    double S_low = S_coordinate(hc, sigma_from_Cidx(Cidx, C_offset, NZ), h, C[Cidx]);
    double S_high = S_coordinate(hc, sigma_from_Cidx(Cidx + 1, C_offset, NZ), h, C[Cidx + 1]);
    double f = (S - S_low) / (S_high - S_low);
    return (double)Cidx - C_offset + f;
}

/* compute_zidx (:98-103) */
double orc_compute_zidx(double z, double h, double zeta, double hc, int NZ, const double *C,
                        int64_t C_size, double C_offset)
{
    double S = S_from_z(z, h, zeta);
    return zidx_from_S(S, hc, NZ, h, C, C_size, C_offset);
}

/* ------------------------------------------------------------------ */
/* Time blend (fields.py:409-428) and FP64 unpacking (_core.pyx:5-24)  */
/* ------------------------------------------------------------------ */

/* FP64 data: out = prev*gt + next*ft, each op rounded in FP64 (numpy
 * multiply, multiply, add with separate output buffers). */
void orc_blend_f64(const double *prev, const double *next, int64_t n, double ft, double gt,
                   double *out)
{
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        double a = prev[i] * gt;
        double b = next[i] * ft;
        out[i] = a + b;
    }
}

/* FP32 data: ft and gt are float32 scalars (fields.py:423-424), both
 * products and the sum are rounded in FP32, then cast to the FP64 output
 * (np.add(..., out=float64, casting="unsafe") runs the float32 loop). */
void orc_blend_f32(const float *prev, const float *next, int64_t n, float ft, float gt, double *out)
{
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        float a = prev[i] * gt;
        float b = next[i] * ft;
        float s = a + b;
        out[i] = (double)s;
    }
}

/* np.ascontiguousarray(fp32_array, dtype=float64) (_core.pyx:5-24) */
void orc_f32_to_f64(const float *src, int64_t n, double *out)
{
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i)
        out[i] = (double)src[i];
}

/* ------------------------------------------------------------------ */
/* L1: validation (kernels/validation.pyx)                             */
/* ------------------------------------------------------------------ */

/* _finite_indices (validation.pyx:21-43) */
void orc_finite_indices(orc_particles *p)
{
    int64_t i, n = p->n;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        if (!isfinite(p->zidx[i]) || !isfinite(p->yidx[i]) || !isfinite(p->xidx[i]))
            p->status[i] = ORC_NONFINITE;
    }
}

/* _domain_bounds (validation.pyx:45-77) */
void orc_domain_bounds(orc_particles *p, const orc_scalars *s)
{
    int64_t i, n = p->n;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        if (!(s->zmin <= p->zidx[i] && p->zidx[i] <= s->zmax &&
              s->ymin <= p->yidx[i] && p->yidx[i] <= s->ymax &&
              s->xmin <= p->xidx[i] && p->xidx[i] <= s->xmax))
            p->status[i] = ORC_OUT_OF_DOMAIN;
    }
}

/* ------------------------------------------------------------------ */
/* L1: RK2 with w advection (kernels/roms/rk2_w_advection.pyx)         */
/* ------------------------------------------------------------------ */

#define RK2_CHUNK 3000

/* _rk2_step_1 (rk2_w_advection.pyx:31-153) */
void orc_rk2_step_1(orc_particles *p, const orc_scalars *s, const orc_fields *F)
{
    int64_t chunk, n = p->n;
    const double hc = s->hc;
    const int NZ = s->NZ;
#pragma omp parallel for schedule(static)
    for (chunk = 0; chunk < n; chunk += RK2_CHUNK) {
        int64_t end = chunk + RK2_CHUNK < n ? chunk + RK2_CHUNK : n;
        for (int64_t i = chunk; i < end; ++i) {
            if (p->status[i] & ORC_INACTIVE)
                continue;
            double h_value = bil(&F->h, p->yidx[i], p->xidx[i]);
            double zeta_value = bil(&F->zeta, p->yidx[i], p->xidx[i]);
            double C_value = orc_linear(F->C.a, F->C.n0, p->zidx[i] + F->C.o0);
            p->z[i] = orc_compute_z(p->zidx[i], NZ, hc, h_value, C_value, zeta_value);
            double u_value = tril(&F->u, p->zidx[i], p->yidx[i], p->xidx[i]);
            double v_value = tril(&F->v, p->zidx[i], p->yidx[i], p->xidx[i]);
            double dx_value = bil(&F->dx, p->yidx[i], p->xidx[i]);
            double dy_value = bil(&F->dy, p->yidx[i], p->xidx[i]);
            p->dxidx1[i] = u_value / dx_value;
            p->dyidx1[i] = v_value / dy_value;
            p->dz1[i] = tril(&F->w, p->zidx[i], p->yidx[i], p->xidx[i]);
        }
    }
}

/* _rk2_step_2 (rk2_w_advection.pyx:155-275) */
void orc_rk2_step_2(orc_particles *p, const orc_scalars *s, const orc_fields *F)
{
    int64_t chunk, n = p->n;
    const double dt = s->dt, alpha = s->alpha, hc = s->hc;
    const int NZ = s->NZ;
#pragma omp parallel for schedule(static)
    for (chunk = 0; chunk < n; chunk += RK2_CHUNK) {
        int64_t end = chunk + RK2_CHUNK < n ? chunk + RK2_CHUNK : n;
        for (int64_t i = chunk; i < end; ++i) {
            if (p->status[i] & ORC_INACTIVE)
                continue;
            double z_int = p->z[i] + alpha * dt * p->dz1[i];
            double yidx_int = p->yidx[i] + alpha * dt * p->dyidx1[i];
            double xidx_int = p->xidx[i] + alpha * dt * p->dxidx1[i];
            double h_value = bil(&F->h, yidx_int, xidx_int);
            double zeta_value = bil(&F->zeta, yidx_int, xidx_int);
            double zidx_int = orc_compute_zidx(z_int, h_value, zeta_value, hc, NZ, F->C.a,
                                               F->C.n0, F->C.o0);
            double u_value = tril(&F->u, zidx_int, yidx_int, xidx_int);
            double v_value = tril(&F->v, zidx_int, yidx_int, xidx_int);
            double dx_value = bil(&F->dx, yidx_int, xidx_int);
            double dy_value = bil(&F->dy, yidx_int, xidx_int);
            p->dxidx2[i] = u_value / dx_value;
            p->dyidx2[i] = v_value / dy_value;
            p->dz2[i] = tril(&F->w, zidx_int, yidx_int, xidx_int);
        }
    }
}

/* _rk2_update (rk2_w_advection.pyx:278-354); b1 = 1/(2 alpha) weights
 * stage 1 and b2 = 1 - b1 stage 2, as the code does (:316-341). */
void orc_rk2_update(orc_particles *p, const orc_scalars *s, const orc_fields *F)
{
    int64_t chunk, n = p->n;
    const double dt = s->dt, alpha = s->alpha, hc = s->hc;
    const int NZ = s->NZ;
    const double b1 = 1.0 / (2.0 * alpha);
    const double b2 = 1.0 - b1;
#pragma omp parallel for schedule(static)
    for (chunk = 0; chunk < n; chunk += RK2_CHUNK) {
        int64_t end = chunk + RK2_CHUNK < n ? chunk + RK2_CHUNK : n;
        for (int64_t i = chunk; i < end; ++i) {
            if (p->status[i] & ORC_INACTIVE)
                continue;
            p->z[i] = p->z[i] + b1 * dt * p->dz1[i] + b2 * dt * p->dz2[i];
            p->yidx[i] = p->yidx[i] + b1 * dt * p->dyidx1[i] + b2 * dt * p->dyidx2[i];
            p->xidx[i] = p->xidx[i] + b1 * dt * p->dxidx1[i] + b2 * dt * p->dxidx2[i];
            double h_value = bil(&F->h, p->yidx[i], p->xidx[i]);
            double zeta_value = bil(&F->zeta, p->yidx[i], p->xidx[i]);
            p->zidx[i] = orc_compute_zidx(p->z[i], h_value, zeta_value, hc, NZ, F->C.a, F->C.n0,
                                          F->C.o0);
        }
    }
}

/* ------------------------------------------------------------------ */
/* L1: AB3 with w advection (kernels/roms/ab3_w_advection.pyx)         */
/* ------------------------------------------------------------------ */

/* _ab3_w_advection (ab3_w_advection.pyx:20-178) */
void orc_ab3(orc_particles *p, const orc_scalars *s, const orc_fields *F)
{
    int64_t i, n = p->n;
    const double dt = s->dt, hc = s->hc;
    const int NZ = s->NZ;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        double u_value = tril(&F->u, p->zidx[i], p->yidx[i], p->xidx[i]);
        double dx_value = bil(&F->dx, p->yidx[i], p->xidx[i]);
        p->dxidx0[i] = u_value / dx_value;
        double v_value = tril(&F->v, p->zidx[i], p->yidx[i], p->xidx[i]);
        double dy_value = bil(&F->dy, p->yidx[i], p->xidx[i]);
        p->dyidx0[i] = v_value / dy_value;
        p->dz0[i] = tril(&F->w, p->zidx[i], p->yidx[i], p->xidx[i]);

        /* start-up (:121-138) */
        if (p->status[i] == ORC_MULTISTEP_1) {
            p->dxidx1[i] = p->dxidx0[i];
            p->dyidx1[i] = p->dyidx0[i];
            p->dz1[i] = p->dz0[i];
            p->dxidx2[i] = p->dxidx0[i];
            p->dyidx2[i] = p->dyidx0[i];
            p->dz2[i] = p->dz0[i];
            p->status[i] = ORC_MULTISTEP_2;
        } else if (p->status[i] == ORC_MULTISTEP_2) {
            p->dxidx2[i] = -p->dxidx0[i] + 2.0 * p->dxidx1[i];
            p->dyidx2[i] = -p->dyidx0[i] + 2.0 * p->dyidx1[i];
            p->dz2[i] = -p->dz0[i] + 2.0 * p->dz1[i];
            p->status[i] = ORC_NORMAL;
        }

        /* z at the current position (:144-165) */
        double h_value = bil(&F->h, p->yidx[i], p->xidx[i]);
        double zeta_value = bil(&F->zeta, p->yidx[i], p->xidx[i]);
        double C_value = orc_linear(F->C.a, F->C.n0, p->zidx[i] + F->C.o0);
        p->z[i] = orc_compute_z(p->zidx[i], NZ, hc, h_value, C_value, zeta_value);

        /* AB3 update (:168-170), exact operation order */
        p->z[i] += dt * (23.0 * p->dz0[i] - 16.0 * p->dz1[i] + 5.0 * p->dz2[i]) / 12.0;
        p->yidx[i] += dt * (23.0 * p->dyidx0[i] - 16.0 * p->dyidx1[i] + 5.0 * p->dyidx2[i]) / 12.0;
        p->xidx[i] += dt * (23.0 * p->dxidx0[i] - 16.0 * p->dxidx1[i] + 5.0 * p->dxidx2[i]) / 12.0;

        /* history shift (:173-178) */
        p->dz2[i] = p->dz1[i];
        p->dxidx2[i] = p->dxidx1[i];
        p->dyidx2[i] = p->dyidx1[i];
        p->dz1[i] = p->dz0[i];
        p->dxidx1[i] = p->dxidx0[i];
        p->dyidx1[i] = p->dyidx0[i];
    }
}

/* _ab3_post_w_advection (ab3_w_advection.pyx:180-229) */
void orc_ab3_post(orc_particles *p, const orc_scalars *s, const orc_fields *F)
{
    int64_t i, n = p->n;
    const double hc = s->hc;
    const int NZ = s->NZ;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        double h_value = bil(&F->h, p->yidx[i], p->xidx[i]);
        double zeta_value = bil(&F->zeta, p->yidx[i], p->xidx[i]);
        p->zidx[i] = orc_compute_zidx(p->z[i], h_value, zeta_value, hc, NZ, F->C.a, F->C.n0,
                                      F->C.o0);
    }
}

/* ------------------------------------------------------------------ */
/* L1: AB3 isopycnal following (kernels/roms/ab3_isopycnal_following.pyx) */
/* ------------------------------------------------------------------ */

/* _ab3_isopycnal_following (ab3_isopycnal_following.pyx:26-191): the AB3
 * step of orc_ab3 with the vertical rate dz0 = w_max * drho / (rho_scale +
 * |drho|), drho = rho(x) - rho_t (:125-132), in place of w.  The post step is
 * orc_ab3_post (:193-242 is the same computation). */
void orc_ab3_iso(orc_particles *p, const orc_scalars *s, const orc_fields *F, const orc_field *rho,
                 const double *rho_t, double w_max, double rho_scale)
{
    int64_t i, n = p->n;
    const double dt = s->dt, hc = s->hc;
    const int NZ = s->NZ;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        double u_value = tril(&F->u, p->zidx[i], p->yidx[i], p->xidx[i]);
        double dx_value = bil(&F->dx, p->yidx[i], p->xidx[i]);
        p->dxidx0[i] = u_value / dx_value;
        double v_value = tril(&F->v, p->zidx[i], p->yidx[i], p->xidx[i]);
        double dy_value = bil(&F->dy, p->yidx[i], p->xidx[i]);
        p->dyidx0[i] = v_value / dy_value;
        double rho_value = tril(rho, p->zidx[i], p->yidx[i], p->xidx[i]);
        double delta_rho = rho_value - rho_t[i];
        p->dz0[i] = w_max * delta_rho / (rho_scale + fabs(delta_rho));

        if (p->status[i] == ORC_MULTISTEP_1) {
            p->dxidx1[i] = p->dxidx0[i];
            p->dyidx1[i] = p->dyidx0[i];
            p->dz1[i] = p->dz0[i];
            p->dxidx2[i] = p->dxidx0[i];
            p->dyidx2[i] = p->dyidx0[i];
            p->dz2[i] = p->dz0[i];
            p->status[i] = ORC_MULTISTEP_2;
        } else if (p->status[i] == ORC_MULTISTEP_2) {
            p->dxidx2[i] = -p->dxidx0[i] + 2.0 * p->dxidx1[i];
            p->dyidx2[i] = -p->dyidx0[i] + 2.0 * p->dyidx1[i];
            p->dz2[i] = -p->dz0[i] + 2.0 * p->dz1[i];
            p->status[i] = ORC_NORMAL;
        }

        double h_value = bil(&F->h, p->yidx[i], p->xidx[i]);
        double zeta_value = bil(&F->zeta, p->yidx[i], p->xidx[i]);
        double C_value = orc_linear(F->C.a, F->C.n0, p->zidx[i] + F->C.o0);
        p->z[i] = orc_compute_z(p->zidx[i], NZ, hc, h_value, C_value, zeta_value);

        p->z[i] += dt * (23.0 * p->dz0[i] - 16.0 * p->dz1[i] + 5.0 * p->dz2[i]) / 12.0;
        p->yidx[i] += dt * (23.0 * p->dyidx0[i] - 16.0 * p->dyidx1[i] + 5.0 * p->dyidx2[i]) / 12.0;
        p->xidx[i] += dt * (23.0 * p->dxidx0[i] - 16.0 * p->dxidx1[i] + 5.0 * p->dxidx2[i]) / 12.0;

        p->dz2[i] = p->dz1[i];
        p->dxidx2[i] = p->dxidx1[i];
        p->dyidx2[i] = p->dyidx1[i];
        p->dz1[i] = p->dz0[i];
        p->dxidx1[i] = p->dxidx0[i];
        p->dyidx1[i] = p->dyidx0[i];
    }
}

/* ------------------------------------------------------------------ */
/* L1: output events (kernels/output/linear_interpolation.pyx,          */
/*     kernels/roms/_roms_core.pyx)                                     */
/* ------------------------------------------------------------------ */

/* _trilinear_interpolation_kernel_function (linear_interpolation.pyx:107-136):
 * out[i] = trilinear value of the blended field at the particle's indices plus
 * the field's offsets; inactive particles keep their previous out[i]. */
void orc_output_trilinear(const orc_particles *p, const orc_field *F, double *out)
{
    int64_t i, n = p->n;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        out[i] = tril(F, p->zidx[i], p->yidx[i], p->xidx[i]);
    }
}

/* _compute_z (_roms_core.pyx:17-73): z of each active particle from h, zeta
 * (bilinear) and C (linear at zidx); inactive particles keep out[i]. */
void orc_output_compute_z(const orc_particles *p, const orc_scalars *s, const orc_fields *F, double *out)
{
    int64_t i, n = p->n;
    const double hc = s->hc;
    const int NZ = s->NZ;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
        if (p->status[i] & ORC_INACTIVE)
            continue;
        double h_value = bil(&F->h, p->yidx[i], p->xidx[i]);
        double zeta_value = bil(&F->zeta, p->yidx[i], p->xidx[i]);
        double C_value = orc_linear(F->C.a, F->C.n0, p->zidx[i] + F->C.o0);
        out[i] = orc_compute_z(p->zidx[i], NZ, hc, h_value, C_value, zeta_value);
    }
}

/* bounding box reductions (launcher.py:122-148); min/max of each index. */
void orc_bbox(const orc_particles *p, double *out6)
{
    double zmin = INFINITY, zmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY,
           xmin = INFINITY, xmax = -INFINITY;
    int64_t i, n = p->n;
#pragma omp parallel for schedule(static) reduction(min : zmin, ymin, xmin) reduction(max : zmax, ymax, xmax)
    for (i = 0; i < n; ++i) {
        zmin = fmin(zmin, p->zidx[i]);
        zmax = fmax(zmax, p->zidx[i]);
        ymin = fmin(ymin, p->yidx[i]);
        ymax = fmax(ymax, p->yidx[i]);
        xmin = fmin(xmin, p->xidx[i]);
        xmax = fmax(xmax, p->xidx[i]);
    }
    out6[0] = zmin; out6[1] = zmax; out6[2] = ymin; out6[3] = ymax; out6[4] = xmin; out6[5] = xmax;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
#else
    (void)n;
#endif
}
