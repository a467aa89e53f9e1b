// Plain drop-in kernels: one HIP kernel per reference Cython kernel, FP64
// positions and FP64 interpolation, one thread per particle.  Compiled with
// -ffp-contract=off so results match the reference's x86-64 FP64 arithmetic
// bit for bit (the reference is built without FMA).
//
// Reference kernels (paths relative to src/offline_particles/):
//   _finite_indices / _domain_bounds   kernels/validation.pyx:21-77
//   _rk2_step_1 / _rk2_step_2 / _update kernels/roms/rk2_w_advection.pyx:31-354
//   _ab3_w_advection / _ab3_post        kernels/roms/ab3_w_advection.pyx:20-229
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "opx.h"
#include "opx_numerics.hpp"
#include "opx_runtime.hpp"

namespace opx {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool is_finite(double v) { return __builtin_isfinite(v); }

// ---------------- validation (validation.pyx) ----------------

__global__ void __launch_bounds__(kBlock) k_finite_indices(opx_particles p)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || (p.status[i] & OPX_INACTIVE))
        return;
    if (!is_finite(p.zidx[i]) || !is_finite(p.yidx[i]) || !is_finite(p.xidx[i]))
        p.status[i] = OPX_NONFINITE;
}

__device__ __forceinline__ bool in_bounds(const opx_scalars &s, double z, double y, double x)
{
    return s.zidx_min <= z && z <= s.zidx_max && s.yidx_min <= y && y <= s.yidx_max &&
           s.xidx_min <= x && x <= s.xidx_max;
}

__global__ void __launch_bounds__(kBlock) k_domain_bounds(opx_particles p, opx_scalars s)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || (p.status[i] & OPX_INACTIVE))
        return;
    if (!in_bounds(s, p.zidx[i], p.yidx[i], p.xidx[i]))
        p.status[i] = OPX_OUT_OF_DOMAIN;
}

// finite_indices then domain_bounds in one pass: a particle flagged NONFINITE
// is INACTIVE and domain_bounds skips it, so the result is the same.
__global__ void __launch_bounds__(kBlock) k_validate(opx_particles p, opx_scalars s)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n)
        return;
    uint8_t st = p.status[i];
    if (st & OPX_INACTIVE)
        return;
    double z = p.zidx[i], y = p.yidx[i], x = p.xidx[i];
    if (!is_finite(z) || !is_finite(y) || !is_finite(x))
        p.status[i] = OPX_NONFINITE;
    else if (!in_bounds(s, z, y, x))
        p.status[i] = OPX_OUT_OF_DOMAIN;
}

// ---------------- RK2 (rk2_w_advection.pyx) ----------------

__global__ void __launch_bounds__(kBlock) k_rk2_step_1(opx_particles p, opx_scalars s, opx_roms_fields f)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || (p.status[i] & OPX_INACTIVE))
        return;
    const double zi = p.zidx[i], yi = p.yidx[i], xi = p.xidx[i];
    double h_value = bilinear_f64(f.h, yi, xi);
    double zeta_value = bilinear_f64(f.zeta, yi, xi);
    double C_value = linear_f64(f.C, zi);
    p.z[i] = compute_z(zi, s.NZ, s.hc, h_value, C_value, zeta_value);
    double u_value = trilinear_f64(f.u, zi, yi, xi);
    double v_value = trilinear_f64(f.v, zi, yi, xi);
    double dx_value = bilinear_f64(f.dx, yi, xi);
    double dy_value = bilinear_f64(f.dy, yi, xi);
    p.dxidx1[i] = u_value / dx_value;
    p.dyidx1[i] = v_value / dy_value;
    p.dz1[i] = trilinear_f64(f.w, zi, yi, xi);
}

__global__ void __launch_bounds__(kBlock) k_rk2_step_2(opx_particles p, opx_scalars s, opx_roms_fields f)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || (p.status[i] & OPX_INACTIVE))
        return;
    const double dt = s.dt, alpha = s.alpha;
    double z_int = p.z[i] + alpha * dt * p.dz1[i];
    double yidx_int = p.yidx[i] + alpha * dt * p.dyidx1[i];
    double xidx_int = p.xidx[i] + alpha * dt * p.dxidx1[i];
    double h_value = bilinear_f64(f.h, yidx_int, xidx_int);
    double zeta_value = bilinear_f64(f.zeta, yidx_int, xidx_int);
    double zidx_int = compute_zidx(z_int, h_value, zeta_value, s.hc, s.NZ, f.C);
    double u_value = trilinear_f64(f.u, zidx_int, yidx_int, xidx_int);
    double v_value = trilinear_f64(f.v, zidx_int, yidx_int, xidx_int);
    double dx_value = bilinear_f64(f.dx, yidx_int, xidx_int);
    double dy_value = bilinear_f64(f.dy, yidx_int, xidx_int);
    p.dxidx2[i] = u_value / dx_value;
    p.dyidx2[i] = v_value / dy_value;
    p.dz2[i] = trilinear_f64(f.w, zidx_int, yidx_int, xidx_int);
}

// b1 = 1/(2 alpha) weights stage 1 and b2 = 1 - b1 stage 2, as the reference
// code does (rk2_w_advection.pyx:316-341).
__global__ void __launch_bounds__(kBlock) k_rk2_update(opx_particles p, opx_scalars s, opx_roms_fields f)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || (p.status[i] & OPX_INACTIVE))
        return;
    const double dt = s.dt;
    const double b1 = 1.0 / (2.0 * s.alpha);
    const double b2 = 1.0 - b1;
    double z = p.z[i] + b1 * dt * p.dz1[i] + b2 * dt * p.dz2[i];
    double y = p.yidx[i] + b1 * dt * p.dyidx1[i] + b2 * dt * p.dyidx2[i];
    double x = p.xidx[i] + b1 * dt * p.dxidx1[i] + b2 * dt * p.dxidx2[i];
    p.z[i] = z;
    p.yidx[i] = y;
    p.xidx[i] = x;
    double h_value = bilinear_f64(f.h, y, x);
    double zeta_value = bilinear_f64(f.zeta, y, x);
    p.zidx[i] = compute_zidx(z, h_value, zeta_value, s.hc, s.NZ, f.C);
}

// ---------------- AB3 (ab3_w_advection.pyx) ----------------

// ISO: the vertical rate is dz0 = w_max * drho / (rho_scale + |drho|) with
// drho = rho(x) - rho_t (ab3_isopycnal_following.pyx:125-132) instead of w.
struct IsoArgs {
    opx_field rho;
    const double *rho_t;
    double w_max, rho_scale;
};

template <bool ISO>
__device__ __forceinline__ void ab3_particle(int64_t i, const opx_particles &p, const opx_scalars &s,
                                             const opx_roms_fields &f, const IsoArgs &iso)
{
    if (i >= p.n)
        return;
    uint8_t st = p.status[i];
    if (st & OPX_INACTIVE)
        return;
    const double zi = p.zidx[i], yi = p.yidx[i], xi = p.xidx[i];
    const double dt = s.dt;

    double u_value = trilinear_f64(f.u, zi, yi, xi);
    double dx_value = bilinear_f64(f.dx, yi, xi);
    double dx0 = u_value / dx_value;
    double v_value = trilinear_f64(f.v, zi, yi, xi);
    double dy_value = bilinear_f64(f.dy, yi, xi);
    double dy0 = v_value / dy_value;
    double dz0;
    if constexpr (ISO) {
        double rho_value = trilinear_f64(iso.rho, zi, yi, xi);
        double delta_rho = rho_value - iso.rho_t[i];
        dz0 = iso.w_max * delta_rho / (iso.rho_scale + fabs(delta_rho));
    } else {
        dz0 = trilinear_f64(f.w, zi, yi, xi);
    }

    double dx1, dy1, dz1, dx2, dy2, dz2;
    if (st == OPX_MULTISTEP_1) { // :122-131
        dx1 = dx0; dy1 = dy0; dz1 = dz0;
        dx2 = dx0; dy2 = dy0; dz2 = dz0;
        p.status[i] = OPX_MULTISTEP_2;
    } else if (st == OPX_MULTISTEP_2) { // :132-138
        dx1 = p.dxidx1[i]; dy1 = p.dyidx1[i]; dz1 = p.dz1[i];
        dx2 = -dx0 + 2.0 * dx1;
        dy2 = -dy0 + 2.0 * dy1;
        dz2 = -dz0 + 2.0 * dz1;
        p.status[i] = OPX_NORMAL;
    } else {
        dx1 = p.dxidx1[i]; dy1 = p.dyidx1[i]; dz1 = p.dz1[i];
        dx2 = p.dxidx2[i]; dy2 = p.dyidx2[i]; dz2 = p.dz2[i];
    }

    double h_value = bilinear_f64(f.h, yi, xi);
    double zeta_value = bilinear_f64(f.zeta, yi, xi);
    double C_value = linear_f64(f.C, zi);
    double z = compute_z(zi, s.NZ, s.hc, h_value, C_value, zeta_value);

    // :168-170, exact operation order
    z += dt * (23.0 * dz0 - 16.0 * dz1 + 5.0 * dz2) / 12.0;
    double y = yi + dt * (23.0 * dy0 - 16.0 * dy1 + 5.0 * dy2) / 12.0;
    double x = xi + dt * (23.0 * dx0 - 16.0 * dx1 + 5.0 * dx2) / 12.0;
    p.z[i] = z;
    p.yidx[i] = y;
    p.xidx[i] = x;
    p.dxidx0[i] = dx0;
    p.dyidx0[i] = dy0;
    p.dz0[i] = dz0;
    // history shift (:173-178)
    p.dz2[i] = dz1;
    p.dxidx2[i] = dx1;
    p.dyidx2[i] = dy1;
    p.dz1[i] = dz0;
    p.dxidx1[i] = dx0;
    p.dyidx1[i] = dy0;
}

__global__ void __launch_bounds__(kBlock) k_ab3(opx_particles p, opx_scalars s, opx_roms_fields f)
{
    ab3_particle<false>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, p, s, f, IsoArgs{});
}

__global__ void __launch_bounds__(kBlock) k_ab3_iso(opx_particles p, opx_scalars s, opx_roms_fields f, IsoArgs iso)
{
    ab3_particle<true>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, p, s, f, iso);
}

__global__ void __launch_bounds__(kBlock) k_ab3_post(opx_particles p, opx_scalars s, opx_roms_fields f)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n || (p.status[i] & OPX_INACTIVE))
        return;
    const double y = p.yidx[i], x = p.xidx[i];
    double h_value = bilinear_f64(f.h, y, x);
    double zeta_value = bilinear_f64(f.zeta, y, x);
    p.zidx[i] = compute_zidx(p.z[i], h_value, zeta_value, s.hc, s.NZ, f.C);
}

// ---------------- diagnostics: compute_z and output interpolation ----------------

// compute_z_kernel (kernels/roms/_roms_core.pyx:17-75)
__global__ void __launch_bounds__(kBlock)
    k_compute_z(int64_t n, const uint8_t *status, const double *zidx, const double *yidx, const double *xidx,
                opx_field h, opx_field zeta, opx_field C, double hc, int NZ, double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (status[i] & OPX_INACTIVE))
        return;
    double h_value = bilinear_f64(h, yidx[i], xidx[i]);
    double zeta_value = bilinear_f64(zeta, yidx[i], xidx[i]);
    double C_value = linear_f64(C, zidx[i]);
    out[i] = compute_z(zidx[i], NZ, hc, h_value, C_value, zeta_value);
}

// _linear/_bilinear/_trilinear_interpolation_kernel_function
// (kernels/output/linear_interpolation.pyx:34-145)
__global__ void __launch_bounds__(kBlock)
    k_interp_linear(int64_t n, const uint8_t *status, const double *idx, opx_field f, double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (status[i] & OPX_INACTIVE))
        return;
    out[i] = linear_f64(f, idx[i]);
}

__global__ void __launch_bounds__(kBlock)
    k_interp_bilinear(int64_t n, const uint8_t *status, const double *idx0, const double *idx1, opx_field f,
                      double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (status[i] & OPX_INACTIVE))
        return;
    out[i] = bilinear_f64(f, idx0[i], idx1[i]);
}

__global__ void __launch_bounds__(kBlock)
    k_interp_trilinear(int64_t n, const uint8_t *status, const double *zidx, const double *yidx, const double *xidx,
                       opx_field f, double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (status[i] & OPX_INACTIVE))
        return;
    out[i] = trilinear_f64(f, zidx[i], yidx[i], xidx[i]);
}

// ---------------- argument checks ----------------

int check_particles(const opx_particles *p, bool positions_only, bool ab3, bool rk2_stage2)
{
    OPX_REQUIRE(p != nullptr, "null particles");
    OPX_REQUIRE(p->n >= 0, "negative particle count %lld", (long long)p->n);
    if (p->n == 0)
        return OPX_OK;
    OPX_REQUIRE(p->status && p->zidx && p->yidx && p->xidx, "null status/zidx/yidx/xidx");
    if (positions_only)
        return OPX_OK;
    OPX_REQUIRE(p->z, "null z");
    OPX_REQUIRE(p->dz1 && p->dyidx1 && p->dxidx1, "null _d*1 arrays");
    if (ab3)
        OPX_REQUIRE(p->dz0 && p->dyidx0 && p->dxidx0 && p->dz2 && p->dyidx2 && p->dxidx2,
                    "null AB3 history arrays");
    if (rk2_stage2)
        OPX_REQUIRE(p->dz2 && p->dyidx2 && p->dxidx2, "null _d*2 arrays");
    return OPX_OK;
}

int check_fields(const opx_roms_fields *f, bool full, bool need_w = true)
{
    OPX_REQUIRE(f != nullptr, "null fields");
    int rc;
    if (full) {
        if ((rc = check_field(f->u, 3, "u", true))) return rc;
        if ((rc = check_field(f->v, 3, "v", true))) return rc;
        if (need_w && (rc = check_field(f->w, 3, "w", true))) return rc;
        if ((rc = check_field(f->dx, 2, "dx", true))) return rc;
        if ((rc = check_field(f->dy, 2, "dy", true))) return rc;
    }
    if ((rc = check_field(f->h, 2, "h", true))) return rc;
    if ((rc = check_field(f->zeta, 2, "zeta", true))) return rc;
    if ((rc = check_field(f->C, 1, "C", true))) return rc;
    return OPX_OK;
}

int check_scalars(const opx_scalars *s, bool need_nz)
{
    OPX_REQUIRE(s != nullptr, "null scalars");
    if (need_nz)
        OPX_REQUIRE(s->NZ > 0, "NZ must be positive, got %d", s->NZ);
    return OPX_OK;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

} // namespace
} // namespace opx

using namespace opx;

extern "C" {

const char *opx_last_error(void) { return opx::g_err; }

int opx_abi_version(void) { return OPX_ABI_VERSION; }

int opx_finite_indices(const opx_particles *p, void *stream)
{
    int rc = check_particles(p, true, false, false);
    if (rc) return rc;
    if (p->n == 0) return OPX_OK;
    hipLaunchKernelGGL(k_finite_indices, dim3(grid_for(p->n)), dim3(kBlock), 0, as_stream(stream), *p);
    return check_launch("opx_finite_indices");
}

int opx_domain_bounds(const opx_particles *p, const opx_scalars *s, void *stream)
{
    int rc = check_particles(p, true, false, false);
    if (rc) return rc;
    if ((rc = check_scalars(s, false))) return rc;
    if (p->n == 0) return OPX_OK;
    hipLaunchKernelGGL(k_domain_bounds, dim3(grid_for(p->n)), dim3(kBlock), 0, as_stream(stream), *p, *s);
    return check_launch("opx_domain_bounds");
}

int opx_validate(const opx_particles *p, const opx_scalars *s, void *stream)
{
    int rc = check_particles(p, true, false, false);
    if (rc) return rc;
    if ((rc = check_scalars(s, false))) return rc;
    if (p->n == 0) return OPX_OK;
    hipLaunchKernelGGL(k_validate, dim3(grid_for(p->n)), dim3(kBlock), 0, as_stream(stream), *p, *s);
    return check_launch("opx_validate");
}

#define OPX_STAGE(NAME, KERNEL, FULL, AB3, S2)                                                     \
    int NAME(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f, void *stream) \
    {                                                                                              \
        int rc = check_particles(p, false, AB3, S2);                                               \
        if (rc) return rc;                                                                         \
        if ((rc = check_scalars(s, true))) return rc;                                              \
        if ((rc = check_fields(f, FULL))) return rc;                                               \
        if (p->n == 0) return OPX_OK;                                                              \
        hipLaunchKernelGGL(KERNEL, dim3(grid_for(p->n)), dim3(kBlock), 0, as_stream(stream), *p,   \
                           *s, *f);                                                                \
        return check_launch(#NAME);                                                                \
    }

OPX_STAGE(opx_rk2_step_1, k_rk2_step_1, true, false, false)
OPX_STAGE(opx_rk2_step_2, k_rk2_step_2, true, false, true)
OPX_STAGE(opx_rk2_update, k_rk2_update, false, false, true)
OPX_STAGE(opx_ab3_w_advection, k_ab3, true, true, false)

int opx_ab3_post_w_advection(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f,
                             void *stream)
{
    int rc = check_particles(p, true, false, false);
    if (rc) return rc;
    OPX_REQUIRE(p->n == 0 || p->z != nullptr, "null z");
    if ((rc = check_scalars(s, true))) return rc;
    if ((rc = check_fields(f, false))) return rc;
    if (p->n == 0) return OPX_OK;
    hipLaunchKernelGGL(k_ab3_post, dim3(grid_for(p->n)), dim3(kBlock), 0, as_stream(stream), *p, *s, *f);
    return check_launch("opx_ab3_post_w_advection");
}

int opx_ab3_isopycnal_following(const opx_particles *p, const double *rho_t, const opx_scalars *s,
                                const opx_roms_fields *f, const opx_field *rho, double w_max, double rho_scale,
                                void *stream)
{
    int rc = check_particles(p, false, true, false);
    if (rc) return rc;
    if ((rc = check_scalars(s, true))) return rc;
    if ((rc = check_fields(f, true, false))) return rc;
    OPX_REQUIRE(rho != nullptr, "null rho field");
    if ((rc = check_field(*rho, 3, "rho", true))) return rc;
    if (p->n == 0) return OPX_OK;
    OPX_REQUIRE(rho_t != nullptr, "null rho_t");
    IsoArgs iso{*rho, rho_t, w_max, rho_scale};
    hipLaunchKernelGGL(k_ab3_iso, dim3(grid_for(p->n)), dim3(kBlock), 0, as_stream(stream), *p, *s, *f, iso);
    return check_launch("opx_ab3_isopycnal_following");
}

int opx_compute_z(int64_t n, const uint8_t *status, const double *zidx, const double *yidx, const double *xidx,
                  const opx_field *h, const opx_field *zeta, const opx_field *C, double hc, int32_t NZ, double *out,
                  void *stream)
{
    OPX_REQUIRE(n >= 0, "negative particle count %lld", (long long)n);
    OPX_REQUIRE(h && zeta && C, "null field");
    int rc;
    if ((rc = check_field(*h, 2, "h", true))) return rc;
    if ((rc = check_field(*zeta, 2, "zeta", true))) return rc;
    if ((rc = check_field(*C, 1, "C", true))) return rc;
    OPX_REQUIRE(NZ > 0, "NZ must be positive");
    if (n == 0) return OPX_OK;
    OPX_REQUIRE(status && zidx && yidx && xidx && out, "null particle array");
    hipLaunchKernelGGL(k_compute_z, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, status, zidx, yidx,
                       xidx, *h, *zeta, *C, hc, (int)NZ, out);
    return check_launch("opx_compute_z");
}

int opx_interp_linear(int64_t n, const uint8_t *status, const double *idx, const opx_field *f, double *out,
                      void *stream)
{
    OPX_REQUIRE(n >= 0, "negative particle count %lld", (long long)n);
    OPX_REQUIRE(f != nullptr, "null field");
    int rc;
    if ((rc = check_field(*f, 1, "field", true))) return rc;
    if (n == 0) return OPX_OK;
    OPX_REQUIRE(status && idx && out, "null particle array");
    hipLaunchKernelGGL(k_interp_linear, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, status, idx, *f,
                       out);
    return check_launch("opx_interp_linear");
}

int opx_interp_bilinear(int64_t n, const uint8_t *status, const double *idx0, const double *idx1,
                        const opx_field *f, double *out, void *stream)
{
    OPX_REQUIRE(n >= 0, "negative particle count %lld", (long long)n);
    OPX_REQUIRE(f != nullptr, "null field");
    int rc;
    if ((rc = check_field(*f, 2, "field", true))) return rc;
    if (n == 0) return OPX_OK;
    OPX_REQUIRE(status && idx0 && idx1 && out, "null particle array");
    hipLaunchKernelGGL(k_interp_bilinear, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, status, idx0,
                       idx1, *f, out);
    return check_launch("opx_interp_bilinear");
}

int opx_interp_trilinear(int64_t n, const uint8_t *status, const double *zidx, const double *yidx,
                         const double *xidx, const opx_field *f, double *out, void *stream)
{
    OPX_REQUIRE(n >= 0, "negative particle count %lld", (long long)n);
    OPX_REQUIRE(f != nullptr, "null field");
    int rc;
    if ((rc = check_field(*f, 3, "field", true))) return rc;
    if (n == 0) return OPX_OK;
    OPX_REQUIRE(status && zidx && yidx && xidx && out, "null particle array");
    hipLaunchKernelGGL(k_interp_trilinear, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, status, zidx,
                       yidx, xidx, *f, out);
    return check_launch("opx_interp_trilinear");
}

} // extern "C"
