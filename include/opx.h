/*
 * opx.h - C-ABI of libopx.so, the MI355X (gfx950) advection hot path.
 *
 * Plain C: pointers, sizes and PODs only.  All device pointers are HIP device
 * memory owned by the caller; `stream` is a hipStream_t (NULL = default
 * stream).  Every entry point launches asynchronously on `stream`, returns
 * OPX_OK (0) or a negative error code, and never throws; opx_last_error()
 * returns the calling thread's last error message.
 *
 * Reference interface replaced (paths relative to src/offline_particles/ of
 * Jamie-Hilditch/offline_particles):
 *
 *   KernelFunction(particles, scalars, fielddata) -> None     kernels/_kernels.py:12
 *   with the per-stage Cython kernels
 *     finite_indices / domain_bounds                           kernels/validation.pyx:21-77
 *     rk2_w_advection_step_1 / _step_2 / _update                kernels/roms/rk2_w_advection.pyx:31-354
 *     ab3_w_advection / ab3_post_w_advection                    kernels/roms/ab3_w_advection.pyx:20-229
 *   and FieldData(array, offsets) (fields.py:12) as unpacked by
 *   unpack_fielddata_{1,2,3}d (kernels/_core.pyx:5-24).
 */
#ifndef OPX_H
#define OPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPX_ABI_VERSION 1

/* ---- storage types of field data ---- */
enum opx_dtype {
    OPX_F64 = 0,
    OPX_F32 = 1,
    OPX_BF16 = 2, /* raw bfloat16 bits (uint16), widened to FP32 for arithmetic */
};

/* ---- particle status codes: identical to kernels/status.pxd:1-14 ---- */
enum opx_status {
    OPX_INACTIVE = 128, /* bit 7 */
    OPX_NORMAL = 0,
    OPX_NONFINITE = 129,
    OPX_OUT_OF_DOMAIN = 130,
    OPX_MULTISTEP_1 = 10,
    OPX_MULTISTEP_2 = 11,
};

/* ---- error codes ---- */
enum opx_error {
    OPX_OK = 0,
    OPX_ERR_INVALID = -1,     /* bad argument (null pointer, shape, dtype, size) */
    OPX_ERR_HIP = -2,         /* a HIP runtime call failed */
    OPX_ERR_UNSUPPORTED = -3, /* valid but not supported combination */
};

/*
 * A field as a kernel sees it: the reference's FieldData (fields.py:12) for one
 * launch, except that the time blend of a time-dependent field is not
 * materialised.  `prev`/`next` are snapshots It and It+1 and the kernel blends
 * each stencil value it reads, value = prev*gt + next*ft, rounded in the
 * storage type exactly like TimeDependentField.get_field_data
 * (fields.py:409-428).  A static field sets next = NULL (ft, gt unused).
 * Data are C-contiguous over `ndim` active dimensions (1, 2 or 3) of size
 * shape[0..ndim-1] >= 2 each; `offsets` are the stagger offsets that map
 * rho-centred particle indices to array indices (spatial_arrays.py:29-40).
 * ft/gt carry the storage type's rounded weights (FP32 values for F32/BF16).
 */
typedef struct opx_field {
    const void *prev;
    const void *next;
    int32_t dtype;
    int32_t ndim;
    int64_t shape[3];
    double offsets[3];
    double ft;
    double gt;
} opx_field;

/* The eight simulation fields the ROMS w-advection kernels declare
 * (rk2_w_advection.pyx:386-395, ab3_w_advection.pyx:265-274). */
typedef struct opx_roms_fields {
    opx_field u, v, w, dx, dy, h, C, zeta;
} opx_roms_fields;

/* Scalars the kernels read (rk2_w_advection.pyx:171-174, ab3_w_advection.pyx:43-45,
 * validation.pyx:51-56).  `alpha` is _RK2_alpha; bounds are inclusive. */
typedef struct opx_scalars {
    double dt;
    double alpha;
    double hc;
    int32_t NZ;
    int32_t pad_;
    double zidx_min, zidx_max, yidx_min, yidx_max, xidx_min, xidx_max;
} opx_scalars;

/*
 * SoA particle arrays of the plain drop-in, named after the reference's particle
 * fields (ab3_w_advection.pyx:245-258; the RK2 kernels use _d*1 and _d*2 of
 * the same names, rk2_w_advection.pyx:372-410).  A pointer a kernel does not
 * declare may be NULL.
 */
typedef struct opx_particles {
    int64_t n;
    uint8_t *status;
    double *zidx, *yidx, *xidx, *z;
    double *dz0, *dz1, *dz2;
    double *dyidx0, *dyidx1, *dyidx2;
    double *dxidx0, *dxidx1, *dxidx2;
} opx_particles;

const char *opx_last_error(void);
int opx_abi_version(void);

/* ------------------------------------------------------------------ */
/* Plain drop-in: one entry point per reference kernel, FP64 positions */
/* and FP64 interpolation (FP32/BF16 storage is widened after the      */
/* blend, as the reference does).                                      */
/* ------------------------------------------------------------------ */

/* kernels/validation.pyx:21-43 (_finite_indices) */
int opx_finite_indices(const opx_particles *p, void *stream);
/* kernels/validation.pyx:45-77 (_domain_bounds) */
int opx_domain_bounds(const opx_particles *p, const opx_scalars *s, void *stream);
/* kernels/validation.pyx:124 (validation_kernel = finite_indices -> domain_bounds), one pass */
int opx_validate(const opx_particles *p, const opx_scalars *s, void *stream);

/* kernels/roms/rk2_w_advection.pyx:31-153 (_rk2_step_1) */
int opx_rk2_step_1(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f, void *stream);
/* kernels/roms/rk2_w_advection.pyx:155-275 (_rk2_step_2) */
int opx_rk2_step_2(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f, void *stream);
/* kernels/roms/rk2_w_advection.pyx:278-354 (_rk2_update); uses only h, C, zeta */
int opx_rk2_update(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f, void *stream);

/* kernels/roms/ab3_w_advection.pyx:20-178 (_ab3_w_advection) */
int opx_ab3_w_advection(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f,
                        void *stream);
/* kernels/roms/ab3_w_advection.pyx:180-229 (_ab3_post_w_advection); uses only h, C, zeta */
int opx_ab3_post_w_advection(const opx_particles *p, const opx_scalars *s, const opx_roms_fields *f,
                             void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OPX_H */
