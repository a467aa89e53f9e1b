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
 *
 * `packed` (optional, NULL for none) holds the same values arranged so one
 * load brings a node's stencil corners in (y, x), built by opx_pack_field
 * from prev (and next): OPX_PACK_QUAD for a static field, the corners
 * (y, x), (y, x+1), (y+1, x), (y+1, x+1) of every node; OPX_PACK_OCT for a
 * time-dependent field, those corners of prev then of next.  F32 and BF16
 * fields of 2 or 3 dimensions only.  The interpolants read the corners from
 * it and blend them with the same arithmetic, so results do not change; a
 * stencil then touches 2 cache lines of a 3-D field instead of 8.
 * OPX_PACK_TPAIR (time-dependent fields) interleaves the snapshots, prev[i]
 * then next[i] for every node: no value is stored twice, and one load brings
 * a stencil row's x-neighbours of both snapshots.  It suits particles sorted
 * by cell, whose neighbours in a wavefront share rows: the 4 copies of every
 * node in an OCT copy would then only add footprint.  OPX_PACK_HEX
 * (time-dependent 3-D fields) holds the 8 corners (z, z+1) x (y, y+1) x
 * (x, x+1) of both snapshots of every node, prev's then next's, corner
 * (dz, dy, dx) at (dz * 2 + dy) * 2 + dx: 16 values, 64 bytes (FP32) or 32
 * (BF16), so a trilinear stencil is one cache line.
 *
 * `packed_stride` (QUAD and OCT; 0 for the others) is the distance in bytes
 * between consecutive nodes' elements; 0 means the element size (QUAD 16
 * bytes FP32, 8 BF16; OCT 32 and 16).  It lets fields of one shape share one
 * buffer, their elements of a node side by side: the drop-in's host code puts
 * dx, dy and h's QUADs in one 64-byte (FP32) element per node, and zeta's OCT
 * with h's QUAD in another, so the static stencils of a step are one cache
 * line and h with zeta another.  A multiple of the QUAD size (QUAD) or of 16
 * bytes (OCT), at least the element size, at most 256.
 */
enum opx_pack_kind { OPX_PACK_NONE = 0, OPX_PACK_QUAD = 1, OPX_PACK_OCT = 2, OPX_PACK_TPAIR = 3, OPX_PACK_HEX = 4 };

typedef struct opx_field {
    const void *prev;
    const void *next;
    int32_t dtype;
    int32_t ndim;
    int64_t shape[3];
    double offsets[3];
    double ft;
    double gt;
    const void *packed;
    int32_t packed_kind;   /* opx_pack_kind */
    int32_t packed_stride; /* bytes between nodes' QUAD / OCT elements; 0: the element size */
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

/* kernels/roms/ab3_isopycnal_following.pyx:26-191 (_ab3_isopycnal_following):
 * the AB3 step of opx_ab3_w_advection with the vertical rate
 * dz0 = w_max * drho / (rho_scale + |drho|), drho = rho(x) - rho_t[i], where
 * `rho` is a 3-D (blended) density field and rho_t the per-particle target
 * density (the particle field "rho_t"); f->w is not read.  Its post step
 * (_ab3_post_isopycnal_following, :175-229) is the same computation as
 * opx_ab3_post_w_advection. */
int opx_ab3_isopycnal_following(const opx_particles *p, const double *rho_t, const opx_scalars *s,
                                const opx_roms_fields *f, const opx_field *rho, double w_max, double rho_scale,
                                void *stream);

/* ------------------------------------------------------------------ */
/* Diagnostics at particle positions (FP64, blended field data as in the */
/* entry points above; particles with the INACTIVE bit are skipped).   */
/* ------------------------------------------------------------------ */

/* compute_z_kernel (kernels/roms/_roms_core.pyx:17-112): out[i] = depth z of
 * (zidx, yidx, xidx)[i] from h and zeta (2-D, y/x) and C (1-D, z). */
int opx_compute_z(int64_t n, const uint8_t *status, const double *zidx, const double *yidx, const double *xidx,
                  const opx_field *h, const opx_field *zeta, const opx_field *C, double hc, int32_t NZ, double *out,
                  void *stream);

/* Output interpolation kernels (kernels/output/linear_interpolation.pyx:34-145):
 * linear over one particle index, bilinear over two (the field's active
 * dims in order), trilinear over (zidx, yidx, xidx). */
int opx_interp_linear(int64_t n, const uint8_t *status, const double *idx, const opx_field *f, double *out,
                      void *stream);
int opx_interp_bilinear(int64_t n, const uint8_t *status, const double *idx0, const double *idx1,
                        const opx_field *f, double *out, void *stream);
int opx_interp_trilinear(int64_t n, const uint8_t *status, const double *zidx, const double *yidx,
                         const double *xidx, const opx_field *f, double *out, void *stream);

/* Spatial ordering of the drop-in's particle arrays (opt-in; no counterpart
 * in the reference, which processes particles in the caller's order,
 * ab3_w_advection.pyx:80).  opx_spatial_order writes perm[j] = the particle
 * placed at j: particles sorted by a Morton key of their cell floor(index)
 * (clamped to the grid; the low `xlow` x bits kept contiguous), those with the
 * INACTIVE bit or a non-finite index last; stable.  `work` holds at least
 * opx_spatial_order_workspace_bytes(n) bytes.  opx_gather applies a
 * permutation to one array: dst[j] = src[perm[j]] for elements of 1, 2, 4 or 8
 * bytes (src != dst).  Each particle's results are independent of the order. */
int opx_spatial_order_workspace_bytes(int64_t n, uint64_t *bytes);
int opx_spatial_order(int64_t n, const uint8_t *status, const double *zidx, const double *yidx, const double *xidx,
                      int32_t NZ, int32_t NY, int32_t NX, int32_t xlow, void *work, uint64_t work_bytes,
                      int32_t *perm, void *stream);
int opx_gather(int64_t n, const int32_t *perm, const void *src, void *dst, int32_t elem_bytes, void *stream);

/* ------------------------------------------------------------------ */
/* Fused form: one kernel per AB3 step over split positions.           */
/*                                                                     */
/* The reference's AB3 step is three launches, each with a full-grid   */
/* time blend (timesteppers.py:296-305: validation_kernel,             */
/* ab3_w_advection, advance_time, ab3_post_w_advection).  Here one     */
/* kernel does the pre-step validation, the AB3 stage and update and   */
/* the post-step zidx at t_{n+1} per particle, with the time blend at  */
/* the stencil values.  Positions are split into a packed int32 cell   */
/* and 32-bit fixed-point offsets; u, v, w are FP32 or BF16 snapshot   */
/* rings; h, dx, dy, zeta FP32; the vertical coordinate, the position  */
/* update and the clock run in FP64.                                   */
/* The staggering is the ROMS C-grid's: u (NZ, NY, NX-1), v (NZ, NY-1, */
/* NX), w (NZ+1, NY, NX), zeta/h/dx/dy (NY, NX), C = Cs_w (NZ+1).      */
/* ------------------------------------------------------------------ */

/* Split-position particle state, SoA.  Per axis, index = c + q * 2^-32 with
 * c = floor(index) and q a 32-bit fixed-point offset, so a position converts
 * to FP64 exactly.  `cell` packs the three axes as
 * (c_z + 2) << (xbits + ybits) | (c_y + 2) << xbits | (c_x + 2); an axis field
 * of all ones marks a non-finite index (NaN/inf: flagged NONFINITE by the next
 * validation, exported as NaN).  hz1..hx2 are the FP32 AB3 derivative history
 * f_{n-1}, f_{n-2} (the reference's _dz1.._dxidx2).  `z` is each particle's
 * depth z (m, FP64) after its last fused step, which the next step starts from
 * instead of recomputing it from zidx (the two differ by the rounding of that
 * round trip); NaN where not known yet: opx_split_from_f64 writes NaN and the
 * first step computes z from zidx (compute_z).  The step entry points require
 * it; opx_rebin permutes it.  `rel` (release id per particle) may be NULL (all
 * release 0); `pid` is permuted with the state by opx_rebin and may be NULL
 * when rebinning is not used.  `rho_t` (FP64 target density of the isopycnal-following step,
 * the reference's particle field rho_t) is NULL for w advection and is
 * permuted with the state by opx_rebin. */
typedef struct opx_split_state {
    int64_t n;
    uint8_t *status;
    int32_t *cell;
    uint32_t *fz, *fy, *fx;
    float *hz1, *hy1, *hx1, *hz2, *hy2, *hx2;
    int32_t *pid;
    uint16_t *rel;
    double *z;
    double *rho_t;
} opx_split_state;

/* Grid sizes (rho points), cell-field widths with (1 << bits) >= N + 5 and
 * xbits + ybits + zbits <= 31, hc and the inclusive index bounds. */
typedef struct opx_grid {
    int32_t NZ, NY, NX;
    int32_t xbits, ybits, zbits;
    double hc;
    double zidx_min, zidx_max, yidx_min, yidx_max, xidx_min, xidx_max;
} opx_grid;

/* Layouts of u, v, w stencil data. */
enum opx_layout {
    OPX_LAYOUT_RING = 0, /* read snapshots s0 and s1 of the rings: 16 loads per field */
    OPX_LAYOUT_PAIR = 1, /* packed window, node i = (s0[i], s1[i]): 8 loads per field */
    OPX_LAYOUT_QUAD = 2, /* packed window, node i = (s0[i], s0[i+1], s1[i], s1[i+1]): 4 loads */
    OPX_LAYOUT_PAIR_LDS = 3, /* PAIR windows; the AB3 step stages each wave's stencil box in LDS
                                (64-thread workgroups; `block` is ignored) */
    OPX_LAYOUT_OCT = 4,      /* packed window, node i = the (y, x) corners i, i+1, i+n2, i+n2+1 of s0
                                and of s1: one load per z-plane (16 B for BF16, 32 B for FP32) */
    OPX_LAYOUT_YPAIR = 5,    /* packed window, node i = (s0[i], s1[i], s0[i+n2], s1[i+n2]); nodes i, i+1
                                hold a z-plane's four (y, x) corners: one 16-byte load per z-plane
                                for BF16, two for FP32 */
};

/* A ring of `nslots` snapshots per time-dependent field, contiguous with the
 * given element strides between slots; `hdd` packs the static (h, dx, dy, 0)
 * per rho point as float4.  With npacked == nslots the window also has the
 * pair ring: for every slot s a packed stencil window (built by
 * opx_pack_window) of the slot pair (s, (s + 1) mod nslots) in `layout`,
 * window s at byte offset s * p*_stride from pu, pv, pw and p2.  A step that
 * reads the pair with first slot s0 reads window s0; the caller keeps every
 * window a launch reads packed from the current contents of its two slots.
 * With npacked == 0 every step reads the rings.  The third 3-D field `w` has
 * `wlevels` z-levels: NZ + 1 (0 means the same) for w at w-points, NZ for the
 * density at rho points that the isopycnal-following step reads in its place. */
typedef struct opx_window {
    const void *u, *v, *w;  /* dtype3 storage */
    const float *zeta;
    int64_t u_stride, v_stride, w_stride, zeta_stride;
    int32_t dtype3; /* OPX_F32 or OPX_BF16 */
    int32_t nslots;
    const float *hdd; /* (NY, NX, 4) */
    const double *C;  /* (NZ + 1) */
    int32_t layout;   /* opx_layout of the pair ring */
    int32_t npacked;  /* 0 (no pair ring) or nslots */
    const void *pu, *pv, *pw;
    const void *p2; /* (NY, NX) float4 {h, zeta[s0], zeta[s1], dx} per window */
    uint64_t pu_stride, pv_stride, pw_stride, p2_stride; /* bytes from one window to the next */
    const float *dy; /* (NY, NX) static dy, read with packed windows */
    int32_t wlevels; /* z-levels of w: 0 or NZ + 1 (w), NZ (density, isopycnal step) */
    int32_t reserved;
} opx_window;

/* Bytes of one packed window of u, v and w for a layout (PAIR, QUAD, OCT or
 * YPAIR; PAIR_LDS packs as PAIR) and of its 2-D part. */
int opx_packed_bytes(const opx_grid *g, int32_t dtype3, int32_t layout, uint64_t *bytes_u, uint64_t *bytes_v,
                     uint64_t *bytes_w, uint64_t *bytes_2d);
/* Build a packed window of ring slots (s0, s1) into caller-owned pu, pv, pw
 * and p2. */
int opx_pack_window(const opx_window *win, const opx_grid *g, int32_t layout, int32_t s0, int32_t s1, void *pu,
                    void *pv, void *pw, void *p2, void *stream);
/* Packed stencil copies of a drop-in field (opx_field.packed): OPX_PACK_QUAD
 * of `prev` when next is NULL, else OPX_PACK_OCT of (prev, next), for a
 * C-contiguous (nz, ny, nx) F32 or BF16 array (nz = 1 for a 2-D field; ny,
 * nx >= 2).  `out` is device memory of opx_packed_field_bytes bytes (QUAD:
 * 4 values per node, OCT: 8, TPAIR: 2, HEX: 16).  No counterpart in the reference: its kernels
 * read the blended FP64 array (fields.py:394-430); the values here are the
 * same, stored so that each stencil row is not its own cache line. */
int opx_packed_field_bytes(int32_t dtype, int32_t kind, int64_t nz, int64_t ny, int64_t nx, uint64_t *bytes);
int opx_pack_field(const void *prev, const void *next, int32_t dtype, int64_t nz, int64_t ny, int64_t nx, void *out,
                   void *stream);

/* opx_pack_field with the kind named: OPX_PACK_QUAD (next NULL),
 * OPX_PACK_OCT, OPX_PACK_TPAIR or OPX_PACK_HEX (next given; TPAIR: 2 values
 * per node, HEX: 16, nz >= 2).  No counterpart in the reference, as
 * opx_pack_field. */
int opx_pack_field_as(const void *prev, const void *next, int32_t dtype, int32_t kind, int64_t nz, int64_t ny,
                      int64_t nx, void *out, void *stream);

/* opx_pack_window for the rho rows [y0, y1) only (each field's rows in that
 * range, every z-level and column; 0 <= y0 < y1 <= NY): refreshes the part of
 * a packed window whose slots received new data there. */
int opx_pack_window_rows(const opx_window *win, const opx_grid *g, int32_t layout, int32_t s0, int32_t s1,
                         int32_t y0, int32_t y1, void *pu, void *pv, void *pw, void *p2, void *stream);

/* opx_pack_window for the nodes of `boxes` only (nboxes x (y0, y1, x0, x1)
 * in rho rows and columns, all levels; each box widened by the row above and
 * the column to its left, whose packed elements hold its nodes as y + 1 and
 * x + 1 neighbours, and clipped to each field's shape), and the 2-D window
 * whole: the repack of a pair whose slots hold only the tiles the particles
 * can reach.  No counterpart in the reference; the tiled window is the
 * counterpart of its bounding-box chunk loading (spatial_arrays.py:260-295). */
int opx_pack_window_boxes(const opx_window *win, const opx_grid *g, int32_t layout, int32_t s0, int32_t s1,
                          const int32_t *boxes, int32_t nboxes, void *pu, void *pv, void *pw, void *p2, void *stream);

/* Sparse snapshot streaming: where the particles are and which parts of a
 * snapshot to copy to the device.  Counterpart of the reference's lazy
 * chunked field (ChunkedDaskArray.get_data_subset, spatial_arrays.py:260-295,
 * which loads only the chunks inside the particles' padded bounding box from
 * Launcher.construct_bbox, launcher.py:122-148); here the region is a set of
 * horizontal tiles rather than one box, so separate clouds of particles do not
 * pull in the space between them.
 *
 * opx_tile_occupancy: map[ty * ntx + tx] = 1 (else 0) for every tile of
 * tile x tile rho columns, nty = ceil(NY / tile) by ntx = ceil(NX / tile),
 * that holds a particle i < s->n whose status is not INACTIVE and whose y and
 * x cells are finite (a cell outside the grid counts in the nearest edge
 * tile).  `map` is device memory of nty * ntx bytes; the call clears it with a
 * kernel (no memset), so it may be captured in a graph. */
int opx_tile_occupancy(const opx_split_state *s, const opx_grid *g, int32_t tile, uint8_t *map, int32_t nty,
                       int32_t ntx, void *stream);
/* Copy boxes of a C-contiguous (nz, ny, nx) host array of elem_bytes elements
 * to the same places of a device array of the same shape: box b = boxes[4b ..
 * 4b + 3] = (y0, y1, x0, x1) covers rows [y0, y1) and columns [x0, x1) of
 * every z-level, one rectangular DMA copy each (hipMemcpy3DAsync) on `stream`.
 * `boxes` is host memory; `src` should be pinned for the copies to be
 * asynchronous. */
int opx_copy_boxes_h2d(void *dst, const void *src, int64_t nz, int64_t ny, int64_t nx, int32_t elem_bytes,
                       const int32_t *boxes, int32_t nboxes, void *stream);
/* The same boxes copied by a kernel on `stream` that reads the host array
 * through its device mapping (hipHostGetDevicePointer): `src` must be pinned
 * host memory mapped to the device over its whole extent (else
 * OPX_ERR_INVALID, nothing launched).  For 1- and 2-byte elements, whose
 * rectangular DMA copies run far below the link rate (rows not 4-byte aligned,
 * or narrow). */
int opx_gather_boxes_h2d(void *dst, const void *src, int64_t nz, int64_t ny, int64_t nx, int32_t elem_bytes,
                         const int32_t *boxes, int32_t nboxes, void *stream);

/* Per-release FP64 clock kept on the device.  A release advances at global
 * steps >= start_step while steps_left > 0. */
typedef struct opx_release_clock {
    double t;
    int32_t start_step;
    int32_t steps_left;
} opx_release_clock;

/* Values written to the device error flags (err_flag arguments). */
enum opx_err_flag {
    OPX_FLAG_TIME_OUT_OF_RANGE = 1, /* a release's time left the time array */
    OPX_FLAG_UNPACKED_PAIR = 2,     /* a several-step launch of a packed layout found no pair ring */
};

/* Per-release weights of one step, written by opx_clock_advance (16 bytes,
 * one vector load per particle): u/v/w/zeta blend at t_n with (ft, 1 - ft),
 * zeta at t_{n+1} with (ft1, 1 - ft1), FP32-rounded like fields.py:423-424.
 * The ring holds at most 1023 slots. */
typedef struct opx_release_step {
    float ft;  /* at t (AB3 step, RK2 stage 1); gt = 1.0f - ft */
    float ft1; /* at t + dt (AB3 post step, RK2 update) */
    float ftm; /* at t + alpha dt (RK2 stage 2; 0 when alpha is 0) */
    /* ring slots of the snapshot pairs and the active flag: slot s0 of t in
     * bits 0-9, of t + dt in bits 10-19, of t + alpha dt in bits 20-29, active
     * in bit 30; the pair's second slot is (s0 + 1) mod nslots */
    uint32_t slots;
} opx_release_step;

/* Advance `nrel` release clocks by one step of `dt` and write their step
 * weights (with alpha > 0 also those of the RK2 intermediate time t + alpha dt,
 * timesteppers.py:250-251).  The time index restates Timestepper.get_time_index
 * (timesteppers.py:64-85) on the device: searchsorted over time_array, then
 * idx + (t - t0)/(t1 - t0), and divmod(tidx, 1) (fields.py:409).  Snapshot It
 * lives in slot It mod nslots.  A release whose time leaves the time array is
 * stopped and *err_flag is set to 1 (the reference raises ValueError /
 * IndexError).  Graph-capturable form: with nrel_dev != NULL only the first
 * *nrel_dev (<= nrel) releases are advanced; with step_dev != NULL the global
 * step is read from *step_dev and incremented on the device after the clocks. */
int opx_clock_advance(opx_release_clock *clocks, opx_release_step *table, int32_t nrel, const int32_t *nrel_dev,
                      const double *time_array, int32_t ntime, int32_t nslots, int32_t global_step,
                      int32_t *step_dev, double dt, double alpha, int32_t *err_flag, void *stream);

/* opx_clock_advance over `nsteps` (1..128) consecutive global steps: step k's
 * entries go to table[k * table_stride + r] (table_stride >= nrel), and the
 * device step counter advances by nsteps.  The same per-step arithmetic as
 * nsteps calls of opx_clock_advance. */
int opx_clock_advance_steps(opx_release_clock *clocks, opx_release_step *table, int32_t table_stride,
                            int32_t nsteps, int32_t nrel, const int32_t *nrel_dev, const double *time_array,
                            int32_t ntime, int32_t nslots, int32_t global_step, int32_t *step_dev, double dt,
                            double alpha, int32_t *err_flag, void *stream);

/* One fused AB3 step (validation -> ab3_w_advection -> ab3_post_w_advection)
 * for particles [0, n) (or [0, *n_active) when n_active is not NULL), one
 * thread per particle in workgroups of `block` threads (128 or 256; 0 = 128).
 * NZ must be <= 127. */
int opx_ab3_fused(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                  const opx_release_step *table, double dt, const int32_t *n_active, int32_t block, void *stream);

/* opx_ab3_fused with the window descriptor in device memory (`win_dev`, of
 * storage type dtype3 and stencil layout `layout`), so that a captured graph
 * stays valid while the host repacks windows between replays. */
int opx_ab3_fused_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                      int32_t layout, const opx_release_step *table, double dt, const int32_t *n_active, int32_t block,
                      void *stream);

/* `nsteps` (1..128) fused AB3 steps in one launch, with step k's release
 * weights at table[k * table_stride + release] (from opx_clock_advance_steps).
 * Each particle's state stays in registers between the steps, and the u, v, w
 * and 2-D stencil nodes it read are blended again at the next step's weights
 * while its cells and snapshot pair are unchanged: the results are those of
 * nsteps one-step launches (timesteppers.py:296-305 applied nsteps times),
 * bit for bit.  With nsteps > 1 and a packed layout the steps read only the
 * pair ring (npacked > 0), whose windows of every pair the steps read must be
 * current; without a pair ring a particle stops, unchanged from that step on,
 * and *err_flag (if not NULL) is set to OPX_FLAG_UNPACKED_PAIR.  The RING
 * layout reads the rings.  No counterpart in the reference, which launches
 * per step. */
int opx_ab3_fused_steps(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                        const opx_release_step *table, int32_t table_stride, int32_t nsteps, double dt,
                        const int32_t *n_active, int32_t block, int32_t *err_flag, void *stream);

/* opx_ab3_fused_steps with the window descriptor in device memory. */
int opx_ab3_fused_steps_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                            int32_t layout, const opx_release_step *table, int32_t table_stride, int32_t nsteps,
                            double dt, const int32_t *n_active, int32_t block, int32_t *err_flag, void *stream);

/* `nsteps` (1..128) fused isopycnal-following AB3 steps in one launch: the
 * steps of opx_ab3_fused_steps with the vertical rate
 * w = w_max * drho / (rho_scale + |drho|), drho = rho(x_n, t_n) - rho_t[i]
 * (ab3_isopycnal_following.pyx:26-191, post step :193-242; launch order of
 * ab3_isopycnal_following_timestepper, models/roms/__init__.py:109-140), rho
 * read from the window's third field (wlevels = NZ, FP32 storage, rho points)
 * and trilinearly interpolated in FP64.  `s->rho_t` is required.  Replaces
 * the reference's per-step launches of validation_kernel,
 * ab3_isopycnal_following_kernel and its post kernel.  PAIR_LDS reads PAIR
 * windows. */
int opx_ab3_iso_fused_steps(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                            const opx_release_step *table, int32_t table_stride, int32_t nsteps, double dt,
                            double w_max, double rho_scale, const int32_t *n_active, int32_t block,
                            int32_t *err_flag, void *stream);

/* opx_ab3_iso_fused_steps with the window descriptor in device memory. */
int opx_ab3_iso_fused_steps_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev,
                                int32_t layout, const opx_release_step *table, int32_t table_stride, int32_t nsteps,
                                double dt, double w_max, double rho_scale, const int32_t *n_active, int32_t block,
                                int32_t *err_flag, void *stream);

/* One fused RK2 step (validation -> _rk2_step_1 at t -> _rk2_step_2 at
 * t + alpha dt -> _rk2_update at t + dt, rk2_w_advection.pyx:31-354 in the
 * order of timesteppers.py:241-257) per particle, with the weights b1 =
 * 1/(2 alpha) on stage 1 and b2 = 1 - b1 on stage 2 as the reference code has
 * them.  Positions, z and the stage derivatives are FP64 inside the step;
 * u/dx and v/dy are FP32 (fields stored in FP32 or BF16), w FP64.  The AB3
 * history arrays of `s` are not used (may be NULL).  Same launch rules as
 * opx_ab3_fused; `table` must come from opx_clock_advance with this alpha. */
int opx_rk2_fused(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                  const opx_release_step *table, double dt, double alpha, const int32_t *n_active, int32_t block,
                  void *stream);

/* `nsteps` (1..128) fused RK2 steps in one launch, step k's release weights at
 * table[k * table_stride + release] (from opx_clock_advance_steps with this
 * alpha); the state stays in registers and stage 2 and the update blend the
 * stencil nodes stage 1 read when they fall in the same cells and pair.  The
 * same bits as nsteps one-step launches; with nsteps > 1 and a packed layout
 * the pairs rule and *err_flag of opx_ab3_fused_steps apply.  No counterpart
 * in the reference, which launches per stage. */
int opx_rk2_fused_steps(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                        const opx_release_step *table, int32_t table_stride, int32_t nsteps, double dt, double alpha,
                        const int32_t *n_active, int32_t block, int32_t *err_flag, void *stream);

/* opx_rk2_fused_steps with the window descriptor in device memory. */
int opx_rk2_fused_steps_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                            int32_t layout, const opx_release_step *table, int32_t table_stride, int32_t nsteps,
                            double dt, double alpha, const int32_t *n_active, int32_t block, int32_t *err_flag,
                            void *stream);
int opx_rk2_fused_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                      int32_t layout, const opx_release_step *table, double dt, double alpha,
                      const int32_t *n_active, int32_t block, void *stream);

/* Output events of the fused form.  Each particle's samples form one record
 * of `out_stride` FP64 slots at out + o * out_stride, o = pid[i] (or i when
 * pid is NULL), the order of opx_split_to_f64: slot k holds sample k of an
 * active particle (status without the INACTIVE bit), every other slot and
 * every slot of an inactive particle holds NaN.  out_stride is 1, 2, 4 or 8
 * (>= nsamples; `out` 16-byte aligned above 1): a record of 4 or 8 slots is
 * one 32- or 64-byte piece of memory written whole, so the scattered records
 * cost no partial-line writes. */
enum opx_sample_kind {
    OPX_SAMPLE_LINEAR = 1,    /* 1-D field over zidx (linear_interpolation.pyx:34-68) */
    OPX_SAMPLE_BILINEAR = 2,  /* 2-D field over (yidx, xidx) (linear_interpolation.pyx:70-105) */
    OPX_SAMPLE_TRILINEAR = 3, /* 3-D field (linear_interpolation.pyx:107-136) */
    OPX_SAMPLE_COMPUTE_Z = 4, /* z from h, zeta (bilinear) and C (linear) (_roms_core.pyx:17-73) */
};

typedef struct opx_sample {
    int32_t kind;   /* opx_sample_kind */
    int32_t field;  /* index of the field in a release's row; h for OPX_SAMPLE_COMPUTE_Z */
    int32_t field2; /* zeta for OPX_SAMPLE_COMPUTE_Z */
    int32_t field3; /* C for OPX_SAMPLE_COMPUTE_Z */
} opx_sample;

/* Evaluate `nsamples` samples at the split positions of `s` in one launch.
 * `fields_dev` (device memory) holds one row of `nfields` opx_field per
 * release: row r is the fields' snapshot pairs and blend weights at release
 * r's time index (a static field has next == NULL); `samples_dev` (device
 * memory) lists the samples (at most 8).  The FP64 index of a split position is exact, so
 * each value is bit for bit what opx_interp_* / opx_compute_z give at the
 * indices opx_split_to_f64 exports; the gathers follow the state's (re-binned)
 * order instead of the caller's.  Replaces, for the fused form, the
 * reference's output kernels launched per field and compute_z_kernel
 * (kernels/output/linear_interpolation.pyx:34-136, kernels/roms/_roms_core.pyx:17-73). */
int opx_sample_split(const opx_split_state *s, const opx_grid *g, const opx_field *fields_dev, int32_t nfields,
                     int32_t nrel, const opx_sample *samples_dev, int32_t nsamples, double hc, double *out,
                     int32_t out_stride, void *stream);

/* Split <-> FP64 index conversion.  opx_split_from_f64 writes the state of
 * particle i from idx arrays (and status/rel untouched); opx_split_to_f64
 * writes zidx/yidx/xidx[pid[i]] (or [i] if pid is NULL). */
int opx_split_from_f64(const opx_split_state *s, const opx_grid *g, const double *zidx, const double *yidx,
                       const double *xidx, void *stream);
int opx_split_to_f64(const opx_split_state *s, const opx_grid *g, double *zidx, double *yidx, double *xidx,
                     void *stream);

/* Re-bin: sort particles by a spatial key and permute the whole state from
 * `src` into `dst` (same n).  Three classes, in this order:
 *   active  status without the INACTIVE bit and a release that still has steps
 *           (or not yet started): Morton order of the cell with the low `xlow`
 *           x bits kept contiguous;
 *   parked  INACTIVE status, or a release whose steps are done (steps_left 0);
 *   free    particles of retired releases (steps_left < 0).
 * The radix sort is stable, so each class keeps its input order outside the
 * spatial key.  Writes the active count to *n_active and, if n_live is not
 * NULL, the active + parked count to *n_live (device ints).  `clocks` (nrel
 * entries, indexed by rel) may be NULL: then only the status decides.  `work`
 * is caller-owned device memory of at least opx_rebin_workspace_bytes(n). */
int opx_rebin_workspace_bytes(int64_t n, uint64_t *bytes);
int opx_rebin(const opx_split_state *src, const opx_split_state *dst, const opx_grid *g, int32_t xlow,
              const opx_release_clock *clocks, int32_t nrel, void *work, uint64_t work_bytes, int32_t *n_active,
              int32_t *n_live, void *stream);

/* opx_rebin without the `work` argument: the keys, values and sort storage live
 * in a static 64 MiB device buffer of the library (one per device, held until
 * the process ends), so the call allocates nothing.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Refuses, with OPX_ERR_INVALID, a
 * state whose opx_rebin_workspace_bytes(n) exceeds 64 MiB (16 bytes per
 * particle plus the radix sort's temporary storage: about 3.9 million
 * particles).  Calls that overlap on one device share the buffer and race. */
int opx_rebin_noalloc(const opx_split_state *src, const opx_split_state *dst, const opx_grid *g, int32_t xlow,
                      const opx_release_clock *clocks, int32_t nrel, int32_t *n_active, int32_t *n_live,
                      void *stream);

/* Results of one release: every particle i < s->n with rel[i] == r writes its
 * FP64 indices (NaN for a non-finite index) and status at position
 * pid[i] - pid_base (uint32 arithmetic), i.e. the particles of a release
 * admitted with pids pid_base .. pid_base + m - 1 come out in admission order.
 * Particles whose offset is not below m are skipped.  No counterpart in the
 * reference (its particle arrays are the results); the caller of a served
 * release reads them from here (`status` may be NULL). */
int opx_extract_release(const opx_split_state *s, const opx_grid *g, int32_t r, uint32_t pid_base, int64_t m,
                        double *zidx, double *yidx, double *xidx, uint8_t *status, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OPX_H */
