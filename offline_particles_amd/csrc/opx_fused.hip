// Fused AB3 step over split positions, the on-device release clock, and
// Morton re-binning.
//
// One thread per particle does, in registers, what the reference does in
// three launches per step (timesteppers.py:296-305, paths relative to
// src/offline_particles/ of the reference):
//   validation_kernel       kernels/validation.pyx:21-77
//   ab3_w_advection         kernels/roms/ab3_w_advection.pyx:20-178
//   ab3_post_w_advection    kernels/roms/ab3_w_advection.pyx:180-229 (at t_{n+1})
// with the time blend done on the stencil values (fields.py:409-428) instead
// of over the whole grid.
//
// Precision: u, v, w are stored as FP32 or BF16; u and v are interpolated in
// FP32, w, h and zeta with FP64 weights, and the ROMS vertical transform
// (kernels/roms/_vertical_coordinate.pxd) runs in FP64.  A position is a
// packed int32 cell (floor of the index, per axis) plus a 32-bit fixed-point
// offset (units of 2^-32 cell), so it converts to FP64 exactly and the AB3
// update is applied in FP64 and rounded once to 2^-32 per step.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "opx.h"
#include "opx_numerics.hpp"
#include "opx_runtime.hpp"

namespace opx {
namespace {

constexpr int kBlock = 256;
constexpr int kXcds = 8; // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kMaxC = 128; // largest NZ + 1 whose C is staged in LDS

constexpr int kLevels = 2 * kMaxC + 1; // doubles of the staged level tables
constexpr int kInvNZ = 2 * kMaxC;      // index of 1 / NZ in them
constexpr double kInv12 = 1.0 / 12.0;  // the AB3 update's 1/12 (one multiply, not a divide)

// The w-level tables every particle reads, staged in LDS by a workgroup:
// L[k] = C[k] (Cs_w), L[kMaxC + k] = k / NZ - 1, the w-level sigma of the
// S-coordinate (_vertical_coordinate.pxd:19-35), the same FP64 values the
// formula gives inline, and L[kInvNZ] = 1 / NZ.
__device__ __forceinline__ void stage_levels(double *L, const double *C, int NZ, int t, int nt)
{
    for (int k = t; k <= NZ; k += nt) {
        L[k] = C[k];
        L[kMaxC + k] = (double)k / NZ - 1.0;
    }
    if (t == 0)
        L[kInvNZ] = 1.0 / NZ;
}

// L[kInvNZ] read where it is used: the index passes through an empty asm
// statement, so the LDS read is not hoisted out of a step loop (where the
// value would hold two VGPRs for the whole launch, for a branch that is
// rarely taken).
__device__ __forceinline__ double inv_nz_here(const double *L)
{
    int k = kInvNZ;
    asm volatile("" : "+v"(k));
    return L[k];
}

// ---------------------------------------------------------------- helpers

// Pointers read from the window descriptor are generic; loads through them
// compile to FLAT instructions, which also count against the LDS counter.
// Field data always lives in global memory: load through global pointers.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gbl(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// p[i] of a 4, 8, 16 or 32-byte element type through a global pointer.
template <typename V>
__device__ __forceinline__ V gload(const V *p, uint32_t i)
{
    static_assert(sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16 || sizeof(V) == 32, "element size");
    if constexpr (sizeof(V) == 4) {
        return __builtin_bit_cast(V, gbl(reinterpret_cast<const uint32_t *>(p))[i]);
    } else if constexpr (sizeof(V) == 8) {
        return __builtin_bit_cast(V, gbl(reinterpret_cast<const u32x2 *>(p))[i]);
    } else if constexpr (sizeof(V) == 16) {
        return __builtin_bit_cast(V, gbl(reinterpret_cast<const u32x4 *>(p))[i]);
    } else {
        struct Two {
            u32x4 a, b;
        };
        const auto *q = gbl(reinterpret_cast<const u32x4 *>(p));
        return __builtin_bit_cast(V, Two{q[2 * i], q[2 * i + 1]});
    }
}

// p[i] and p[i + 1] of a 4 or 8-byte element type as ONE 8 or 16-byte load.
// Multi-dword global loads need only dword alignment (the two low address
// bits are ignored), so the pair needs no more than the element's alignment.
// The compiler does not merge the two loads by itself: with a 32-bit index
// i + 1 may wrap, so it cannot prove the second address is the first + size.
// The casts are spelled out: passed through a template such as gbl(), the
// typedefs would lose their reduced alignment.
typedef uint32_t u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t u32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef const __attribute__((address_space(1))) u32x2_a4 *gptr_u32x2_a4;
typedef const __attribute__((address_space(1))) u32x4_a8 *gptr_u32x4_a8;

template <typename V>
__device__ __forceinline__ void gload_pair(const V *p, uint32_t i, V &a, V &b)
{
    static_assert(sizeof(V) == 4 || sizeof(V) == 8, "element size");
    // Elements are copied out before the bit casts: a bit cast of a vector
    // element expression (t.y) reads element 0.
    if constexpr (sizeof(V) == 4) {
        const u32x2_a4 t = *(gptr_u32x2_a4)(reinterpret_cast<const uint32_t *>(p) + i);
        const uint32_t x = t.x, y = t.y;
        a = __builtin_bit_cast(V, x);
        b = __builtin_bit_cast(V, y);
    } else {
        const u32x4_a8 t = *(gptr_u32x4_a8)(reinterpret_cast<const u32x2 *>(p) + i);
        const u32x2 x = {t.x, t.y}, y = {t.z, t.w};
        a = __builtin_bit_cast(V, x);
        b = __builtin_bit_cast(V, y);
    }
}

// The element (or pair of elements) of a packed window at byte offset `off`.
// A packed window is at most 4 GiB (opx_packed_bytes refuses larger ones), so
// the offset fits 32 bits: where the window base is the same in every lane it
// stays in scalar registers and each load takes the 32-bit offset alone (the
// scalar-base form of global loads) instead of a 64-bit address per lane.
template <typename V>
__device__ __forceinline__ V gload_at(const void *base, uint32_t off)
{
    return gload(reinterpret_cast<const V *>(static_cast<const char *>(base) + off), 0u);
}

template <typename V>
__device__ __forceinline__ void gload_pair_at(const void *base, uint32_t off, V &a, V &b)
{
    gload_pair(reinterpret_cast<const V *>(static_cast<const char *>(base) + off), 0u, a, b);
}

__device__ __forceinline__ float ld3(const float *__restrict__ p, uint32_t i) { return gbl(p)[i]; }
__device__ __forceinline__ float ld3(const uint16_t *__restrict__ p, uint32_t i)
{
    return __uint_as_float((uint32_t)gbl(p)[i] << 16);
}

constexpr double kQ = 4294967296.0; // 2^32: one cell in fixed-point offset units
constexpr int kBias = 2;            // stored cell = floor(index) + 2, range [-2, N+1]

// One axis of a particle position: index = c + q * 2^-32 with c = floor(index)
// (unclamped, in [-2, N+1]); `ok` is false for a non-finite index.
struct Axis {
    int c;
    uint32_t q;
    bool ok;
    __device__ __forceinline__ double pos() const { return (double)c + (double)q * (1.0 / kQ); }
    __device__ __forceinline__ float frac() const { return (float)q * (1.0f / 4294967296.0f); }
    __device__ __forceinline__ bool upper_half() const { return q >= 0x80000000u; }
};

__device__ __forceinline__ Axis decode_axis(uint32_t cell, int shift, int bits)
{
    const uint32_t mask = (1u << bits) - 1u;
    const uint32_t v = (cell >> shift) & mask;
    Axis a;
    a.ok = v != mask;
    a.c = (int)v - kBias;
    a.q = 0;
    return a;
}

// Round an FP64 index to the fixed-point form; non-finite -> marker.  Cells
// beyond [-2, N+1] are clamped (such a particle is out of the domain either way
// and the next validation flags it).
__device__ __forceinline__ Axis encode_axis(double v, int N)
{
    Axis a;
    if (!__builtin_isfinite(v)) {
        a.ok = false;
        a.c = 0;
        a.q = 0;
        return a;
    }
    a.ok = true;
    double fl = fmin(fmax(floor(v), (double)-kBias), (double)(N + 1));
    double r = rint((v - fl) * kQ);
    if (r >= kQ) {
        if (fl < (double)(N + 1)) {
            fl += 1.0;
            r -= kQ;
        } else {
            r = kQ - 1.0;
        }
    }
    r = fmax(r, 0.0);
    a.c = (int)fl;
    a.q = (uint32_t)r;
    return a;
}

__device__ __forceinline__ uint32_t axis_field(const Axis &a, int bits)
{
    return a.ok ? (uint32_t)(a.c + kBias) : (1u << bits) - 1u;
}

__device__ __forceinline__ uint32_t pack_cell(const Axis &z, const Axis &y, const Axis &x, const opx_grid &g)
{
    return (axis_field(z, g.zbits) << (g.xbits + g.ybits)) | (axis_field(y, g.ybits) << g.xbits) |
           axis_field(x, g.xbits);
}

// An index rounded to a multiple of 2^-32 the way encode_axis rounds it, kept
// in FP64 between the steps of one launch.  For v in [-2, N+2) it equals the
// decoded stored form of v exactly: encode_axis rounds (v - floor v) 2^32 to an
// integer (ties to even), v 2^32 differs from that by the even integer
// floor(v) 2^32, so rint lands on the same multiple; both scalings by 2^32 are
// exact.  The steps therefore need no encode/decode until the launch stores the
// particle.  Outside that range the particle is out of the domain and the value
// is left unclamped (the store clamps it); non-finite stays non-finite.
__device__ __forceinline__ double round32(double v) { return ldexp(rint(ldexp(v, 32)), -32); }

// One coordinate of a particle about to be sampled: c = floor(v), the exact
// offset d = v - c (d * 2^32 is the stored offset q) and its FP32 rounding f
// (the same value as (float)q * 2^-32).
struct Coord {
    int c;
    double d;
    float f;
};

__device__ __forceinline__ Coord coord(double v)
{
    const double fl = floor(v);
    const double d = v - fl;
    return {(int)fl, d, (float)d};
}

__device__ __forceinline__ Coord coord(const Axis &a)
{
    const double d = (double)a.q * (1.0 / kQ);
    return {a.c, d, a.frac()};
}

// FP64 weight of the upper node for base b and stagger offset o at index v:
// v - b + o, exact for a decoded stored axis or a round32 value in the grid.
__device__ __forceinline__ double w64(double v, int b, double o) { return (v - (double)b) + o; }
// w64 of a rho-centred axis (offset 0): v - b (the add of +0.0 that w64 would
// keep, for a -0.0 it cannot rule out, changes the sign of a zero weight only)
__device__ __forceinline__ double w64(double v, int b) { return v - (double)b; }

// The base of the 2-D corners at a position just stepped to: floor(v) clamped
// to [0, N-2] in FP64, so that any finite v converts safely.
__device__ __forceinline__ int base2(double v, int N) { return (int)fmin(fmax(floor(v), 0.0), (double)(N - 2)); }

// One interpolation axis: base index and weight of the upper node.
struct Ax {
    int b;
    float w;
};

// CENTER stagger (offset 0, size N): truncate(idx) clamped to [0, N-2]
// (kernels/_interpolation/linear.pxd:3-9), weight unclamped (extrapolation).
__device__ __forceinline__ Ax ax_center(const Coord &a, int N)
{
    int b = max(0, min(a.c, N - 2));
    return {b, a.f + (float)(a.c - b)};
}

// INNER stagger (offset -0.5, size N-1): truncate(idx - 0.5) clamped to [0, N-3].
__device__ __forceinline__ Ax ax_inner(const Coord &a, int N)
{
    int b = a.c - 1 + (a.d >= 0.5 ? 1 : 0);
    b = max(0, min(b, N - 3));
    return {b, a.f + ((float)(a.c - b) - 0.5f)};
}

// OUTER stagger (offset +0.5, size N+1): truncate(idx + 0.5) clamped to [0, N-1].
__device__ __forceinline__ Ax ax_outer(const Coord &a, int N)
{
    int b = a.c + (a.d >= 0.5 ? 1 : 0);
    b = max(0, min(b, N - 1));
    return {b, a.f + ((float)(a.c - b) + 0.5f)};
}

__device__ __forceinline__ Ax ax_center(const Axis &a, int N) { return ax_center(coord(a), N); }
__device__ __forceinline__ Ax ax_inner(const Axis &a, int N) { return ax_inner(coord(a), N); }
__device__ __forceinline__ Ax ax_outer(const Axis &a, int N) { return ax_outer(coord(a), N); }

// The time blend of one stencil value, p*gt + q*ft with each operation rounded
// in FP32 (no contraction): the same arithmetic as the reference's FP32
// full-grid blend (fields.py:426-428).  The fused forms keep it for the
// isopycnal-following step's density, whose rate takes the difference of the
// density and a target close to it and so magnifies any change in its last
// bits.
__device__ __forceinline__ float blend_rn(float p, float q, float gt, float ft)
{
#pragma clang fp contract(off)
    return p * gt + q * ft;
}

// a * b + c as the three-operand v_fma_f32, the instruction fmaf compiles to
// (the same result).  Left to itself the compiler picks the two-operand
// v_fmac_f32, which overwrites c; where c is a stencil node cached across the
// steps of a launch it then copies every cached node once per step to keep it.
__device__ __forceinline__ float fma3(float a, float b, float c)
{
    float r;
    asm("v_fma_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// The fused forms' time blend of velocities and the free surface: p + (q - p)*ft
// as one difference and one fused multiply-add (two instructions instead of
// three; within an FP32 rounding or two of blend_rn).  RN selects blend_rn.
// Every stencil layout and launch form blends with the same operations.
template <bool RN = false>
__device__ __forceinline__ float blend_t(float p, float q, float gt, float ft)
{
    if constexpr (RN)
        return blend_rn(p, q, gt, ft);
    else
        return fma3(q - p, ft, p);
}

template <typename T, bool RN = false>
__device__ __forceinline__ float blend(const T *__restrict__ p, const T *__restrict__ q, uint32_t i, float gt,
                                       float ft)
{
    return blend_t<RN>(ld3(p, i), ld3(q, i), gt, ft);
}

// ---------------------------------------------------------------- stencil corners

// Packed window elements: PAIR node i = (s0[i], s1[i]); QUAD node i =
// (s0[i], s0[i+1], s1[i], s1[i+1]) (x-neighbour pair of both snapshots).
template <typename T>
struct Packed;
// OCT node i = the 2 x 2 (y, x) corners i, i+1, i+n2, i+n2+1 of both snapshots.
struct float8 {
    float4 a, b; // a: s0 corners, b: s1 corners (order i, i+1, i+n2, i+n2+1)
};
template <>
struct Packed<float> {
    using pair = float2;
    using quad = float4;
    using oct = float8;
    using ypair = float4; // s0[i], s1[i], s0[i+n2], s1[i+n2]
};
template <>
struct Packed<uint16_t> {
    using pair = uint32_t; // lo: s0, hi: s1
    using quad = uint2;    // x: s0[i] | s0[i+1] << 16, y: s1[i] | s1[i+1] << 16
    using oct = uint4;     // x: s0[i] | s0[i+1] << 16, y: s0[i+n2] | s0[i+n2+1] << 16, z, w: the same of s1
    using ypair = uint2;   // x: s0[i] | s1[i] << 16, y: s0[i+n2] | s1[i+n2] << 16
};

__device__ __forceinline__ float bf_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }
__device__ __forceinline__ float2 unpair(float2 v) { return v; }
__device__ __forceinline__ float2 unpair(uint32_t v) { return make_float2(bf_lo(v), bf_hi(v)); }
__device__ __forceinline__ float4 unquad(float4 v) { return v; }
__device__ __forceinline__ float4 unquad(uint2 v)
{
    return make_float4(bf_lo(v.x), bf_hi(v.x), bf_lo(v.y), bf_hi(v.y));
}

// Raw stencil nodes of one field at one base: both snapshots, before the time
// blend.  A particle that stays in its cells keeps them and blends them again
// at the next step's weights instead of reloading them.
template <typename T, int LAYOUT>
struct Raw3;
template <typename T>
struct Raw3<T, OPX_LAYOUT_PAIR> {
    typename Packed<T>::pair n[8]; // the 8 corners in stencil order
};
template <typename T>
struct Raw3<T, OPX_LAYOUT_QUAD> {
    typename Packed<T>::quad n[4]; // the 4 rows (axis0, axis1) of x-pairs
};
template <typename T>
struct Raw3<T, OPX_LAYOUT_OCT> {
    typename Packed<T>::oct n[2]; // the 2 z-planes
};
template <typename T>
struct Raw3<T, OPX_LAYOUT_YPAIR> {
    typename Packed<T>::ypair n[4]; // nodes i, i+1 of the 2 z-planes
};

// Raw nodes come in units loaded and blended together: a row of x-neighbours
// (PAIR, QUAD: 4 units) or a z-plane (OCT, YPAIR: 2 units).
template <int LAYOUT>
constexpr int kRawUnits = (LAYOUT == OPX_LAYOUT_OCT || LAYOUT == OPX_LAYOUT_YPAIR) ? 2 : 4;

// Bytes of one packed node of a layout.
template <typename T, int LAYOUT>
constexpr uint32_t kNodeBytes =
    LAYOUT == OPX_LAYOUT_PAIR   ? sizeof(typename Packed<T>::pair)
    : LAYOUT == OPX_LAYOUT_QUAD ? sizeof(typename Packed<T>::quad)
    : LAYOUT == OPX_LAYOUT_OCT  ? sizeof(typename Packed<T>::oct)
                                : sizeof(typename Packed<T>::ypair);

// Loads unit k of the raw nodes of a field whose stencil rows (axis0, axis1)
// start at byte offsets rows[0..3] of packed window pp.
template <typename T, int LAYOUT>
__device__ __forceinline__ void load_unit(const void *__restrict__ pp, const uint32_t rows[4], int k,
                                          Raw3<T, LAYOUT> &r)
{
    if constexpr (LAYOUT == OPX_LAYOUT_PAIR) {
        // the x-neighbours (i, i+1) in one load
        gload_pair_at(pp, rows[k], r.n[2 * k], r.n[2 * k + 1]);
    } else if constexpr (LAYOUT == OPX_LAYOUT_QUAD) {
        r.n[k] = gload_at<typename Packed<T>::quad>(pp, rows[k]);
    } else if constexpr (LAYOUT == OPX_LAYOUT_OCT) {
        // one node per z-plane holds that plane's four (y, x) corners of both snapshots
        r.n[k] = gload_at<typename Packed<T>::oct>(pp, rows[2 * k]);
    } else {
        // YPAIR: nodes i, i+1 of a z-plane: (y, x), (y+1, x) and (y, x+1), (y+1, x+1)
        using E = typename Packed<T>::ypair;
        if constexpr (sizeof(T) == 4) {
            r.n[2 * k] = gload_at<E>(pp, rows[2 * k]);
            r.n[2 * k + 1] = gload_at<E>(pp, rows[2 * k] + (uint32_t)sizeof(E));
        } else {
            gload_pair_at(pp, rows[2 * k], r.n[2 * k], r.n[2 * k + 1]); // one 16-byte load per z-plane
        }
    }
}

// Blends unit k of raw nodes into stencil values v (order 000, 001, 010, 011,
// 100, 101, 110, 111 over (axis0, axis1, axis2)): p*gt + q*ft in FP32, the
// same operations in every layout.
template <typename T, int LAYOUT, bool RN = false>
__device__ __forceinline__ void blend_unit(const Raw3<T, LAYOUT> &r, int k, float gt, float ft, float v[8])
{
    if constexpr (LAYOUT == OPX_LAYOUT_PAIR) {
        const float2 e0 = unpair(r.n[2 * k]), e1 = unpair(r.n[2 * k + 1]);
        v[2 * k] = blend_t<RN>(e0.x, e0.y, gt, ft);
        v[2 * k + 1] = blend_t<RN>(e1.x, e1.y, gt, ft);
    } else if constexpr (LAYOUT == OPX_LAYOUT_QUAD) {
        const float4 e = unquad(r.n[k]);
        v[2 * k] = blend_t<RN>(e.x, e.z, gt, ft);
        v[2 * k + 1] = blend_t<RN>(e.y, e.w, gt, ft);
    } else if constexpr (LAYOUT == OPX_LAYOUT_OCT) {
        if constexpr (sizeof(T) == 4) {
            const float8 &e = r.n[k];
            v[4 * k + 0] = blend_t<RN>(e.a.x, e.b.x, gt, ft);
            v[4 * k + 1] = blend_t<RN>(e.a.y, e.b.y, gt, ft);
            v[4 * k + 2] = blend_t<RN>(e.a.z, e.b.z, gt, ft);
            v[4 * k + 3] = blend_t<RN>(e.a.w, e.b.w, gt, ft);
        } else {
            const uint4 &e = r.n[k];
            v[4 * k + 0] = blend_t<RN>(bf_lo(e.x), bf_lo(e.z), gt, ft);
            v[4 * k + 1] = blend_t<RN>(bf_hi(e.x), bf_hi(e.z), gt, ft);
            v[4 * k + 2] = blend_t<RN>(bf_lo(e.y), bf_lo(e.w), gt, ft);
            v[4 * k + 3] = blend_t<RN>(bf_hi(e.y), bf_hi(e.w), gt, ft);
        }
    } else {
        if constexpr (sizeof(T) == 4) {
            const float4 &a = r.n[2 * k], &c = r.n[2 * k + 1];
            v[4 * k + 0] = blend_t<RN>(a.x, a.y, gt, ft);
            v[4 * k + 1] = blend_t<RN>(c.x, c.y, gt, ft);
            v[4 * k + 2] = blend_t<RN>(a.z, a.w, gt, ft);
            v[4 * k + 3] = blend_t<RN>(c.z, c.w, gt, ft);
        } else {
            const uint2 &a = r.n[2 * k], &c = r.n[2 * k + 1];
            v[4 * k + 0] = blend_t<RN>(bf_lo(a.x), bf_hi(a.x), gt, ft);
            v[4 * k + 1] = blend_t<RN>(bf_lo(c.x), bf_hi(c.x), gt, ft);
            v[4 * k + 2] = blend_t<RN>(bf_lo(a.y), bf_hi(a.y), gt, ft);
            v[4 * k + 3] = blend_t<RN>(bf_lo(c.y), bf_hi(c.y), gt, ft);
        }
    }
}

// First node of each stencil row (axis0, axis1) of a field with dims
// (., n1, n2) at base (b0, b1, b2), as element indices times `esize` (1: the
// index; a packed node's bytes: its byte offset in the window).
__device__ __forceinline__ void stencil_rows(int n1, int n2, int b0, int b1, int b2, uint32_t rows[4],
                                             uint32_t esize = 1)
{
    const uint32_t s1 = (uint32_t)n2 * esize, s0 = (uint32_t)n1 * s1;
    // 24-bit multiplies (full rate): check_grid keeps (NZ + 1) NY and NX below
    // 2^24, and the node index itself below 2^31
    const uint32_t b = (__umul24(__umul24(b0, n1) + b1, n2) + b2) * esize;
    rows[0] = b;
    rows[1] = b + s1;
    rows[2] = b + s0;
    rows[3] = b + s0 + s1;
}

// All raw nodes of a field at base (b0, b1, b2) from packed window pp.
template <typename T, int LAYOUT>
__device__ __forceinline__ void load_raw(const void *__restrict__ pp, int n1, int n2, int b0, int b1, int b2,
                                         Raw3<T, LAYOUT> &r)
{
    uint32_t rows[4];
    stencil_rows(n1, n2, b0, b1, b2, rows, kNodeBytes<T, LAYOUT>);
#pragma unroll
    for (int k = 0; k < kRawUnits<LAYOUT>; ++k)
        load_unit<T, LAYOUT>(pp, rows, k, r);
}

template <typename T, int LAYOUT, bool RN = false>
__device__ __forceinline__ void blend_raw(const Raw3<T, LAYOUT> &r, float gt, float ft, float v[8])
{
#pragma unroll
    for (int k = 0; k < kRawUnits<LAYOUT>; ++k)
        blend_unit<T, LAYOUT, RN>(r, k, gt, ft, v);
}

// The 8 blended stencil values of a field with dims (., n1, n2) and base
// (b0, b1, b2).  RING reads slot pointers p and q; the packed layouts read
// window p, blending each unit as soon as it is loaded.
template <typename T, int LAYOUT, bool RN = false>
__device__ __forceinline__ void corners(const void *__restrict__ pp, const void *__restrict__ qq, float gt, float ft,
                                        int n1, int n2, int b0, int b1, int b2, float v[8])
{
    uint32_t rows[4];
    if constexpr (LAYOUT == OPX_LAYOUT_RING) {
        stencil_rows(n1, n2, b0, b1, b2, rows);
        const T *p = static_cast<const T *>(pp), *q = static_cast<const T *>(qq);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            v[2 * r] = blend<T, RN>(p, q, rows[r], gt, ft);
            v[2 * r + 1] = blend<T, RN>(p, q, rows[r] + 1, gt, ft);
        }
    } else {
        stencil_rows(n1, n2, b0, b1, b2, rows, kNodeBytes<T, LAYOUT>);
        Raw3<T, LAYOUT> r;
#pragma unroll
        for (int k = 0; k < kRawUnits<LAYOUT>; ++k) {
            load_unit<T, LAYOUT>(pp, rows, k, r);
            blend_unit<T, LAYOUT, RN>(r, k, gt, ft, v);
        }
    }
}

// Linear blends a * g + b * f (g = 1 - f) as a product and one fused
// multiply-add: one rounding fewer than two products and a sum, one
// instruction fewer, and the same result wherever the helper is inlined.
__device__ __forceinline__ float lerp32(float a, float b, float g, float f) { return fmaf(b, f, a * g); }

// a / b for the fused steps' index rates (velocity over grid spacing): the
// hardware reciprocal estimate of b, the quotient, and one Newton-Raphson
// correction of the quotient from its fused residual (Markstein's method),
// four instructions instead of the ten of the IEEE divide sequence.  b is a
// positive normal grid spacing; the result is the correctly rounded quotient
// but for rare last-bit cases.
__device__ __forceinline__ float div_nr(float a, float b)
{
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return fmaf(fmaf(-b, q, a), r, q);
}

__device__ __forceinline__ float tri32(const float v[8], float f0, float f1, float f2)
{
    const float g0 = 1.0f - f0, g1 = 1.0f - f1, g2 = 1.0f - f2;
    const float c00 = lerp32(v[0], v[1], g2, f2), c01 = lerp32(v[2], v[3], g2, f2);
    const float c10 = lerp32(v[4], v[5], g2, f2), c11 = lerp32(v[6], v[7], g2, f2);
    return lerp32(lerp32(c00, c01, g1, f1), lerp32(c10, c11, g1, f1), g0, f0);
}

// The FP64 trilinear of FP32 stencil values with every product and sum rounded
// on its own (no fused multiply-add), for the density of the isopycnal rate: the rate takes the difference of this
// density and a target close to it, which magnifies any change in the last
// bits of the density, so the density keeps this rounding.
__device__ __forceinline__ double tri64_rn(const float v[8], double f0, double f1, double f2)
{
#pragma clang fp contract(off)
    const double g0 = 1.0 - f0, g1 = 1.0 - f1, g2 = 1.0 - f2;
    double c00 = (double)v[0] * g2 + (double)v[1] * f2, c01 = (double)v[2] * g2 + (double)v[3] * f2;
    double c10 = (double)v[4] * g2 + (double)v[5] * f2, c11 = (double)v[6] * g2 + (double)v[7] * f2;
    return (c00 * g1 + c01 * f1) * g0 + (c10 * g1 + c11 * f1) * f0;
}

// FP64 bilinear of a 2-D field whose corner values are given.
__device__ __forceinline__ double bil64(double v00, double v01, double v10, double v11, double fy, double fx)
{
    const double gy = 1.0 - fy, gx = 1.0 - fx;
    return fma(fy * fx, v11, fma(fy * gx, v10, fma(gy * fx, v01, gy * gx * v00)));
}

// The four rho-point corners of the 2-D fields around a particle: h, zeta of
// the two snapshots, dx, dy, and the base they were read at.
struct Corners2 {
    float h[4], z0[4], z1[4], dx[4], dy[4];
    int by, bx;
};

__device__ __forceinline__ void load2_packed(const float4 *__restrict__ P, const float *__restrict__ dy_, int NX,
                                             int by, int bx, bool want_dxdy, Corners2 &c)
{
    // byte offsets: 16-byte nodes of the packed window, 4-byte dy values
    const uint32_t b = __umul24(by, NX) + bx;
    const uint32_t o[4] = {16 * b, 16 * (b + 1), 16 * (b + NX), 16 * (b + NX + 1)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float4 m = gload_at<float4>(P, o[k]);
        c.h[k] = m.x;
        c.z0[k] = m.y;
        c.z1[k] = m.z;
        c.dx[k] = m.w;
    }
    if (want_dxdy) {
        gload_pair_at(dy_, 4 * b, c.dy[0], c.dy[1]);
        gload_pair_at(dy_, 4 * (b + NX), c.dy[2], c.dy[3]);
    }
    c.by = by;
    c.bx = bx;
}

__device__ __forceinline__ void load2_ring(const float4 *__restrict__ hdd, const float *__restrict__ zp_,
                                           const float *__restrict__ zq_, int NX, int by, int bx, bool want_dxdy,
                                           Corners2 &c)
{
    const auto *zp = gbl(zp_), *zq = gbl(zq_);
    const uint32_t b = (uint32_t)by * (uint32_t)NX + (uint32_t)bx;
    const uint32_t o[4] = {b, b + 1, b + NX, b + NX + 1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        float4 m = gload(hdd, o[k]);
        c.h[k] = m.x;
        c.dx[k] = m.y;
        c.dy[k] = m.z;
        c.z0[k] = zp[o[k]];
        c.z1[k] = zq[o[k]];
    }
    c.by = by;
    c.bx = bx;
    (void)want_dxdy;
}

// h and zeta (blended in FP32, then widened) bilinear with FP64 weights.
__device__ __forceinline__ void eval2(const Corners2 &c, float gt, float ft, double wy, double wx, double &h,
                                      double &zeta)
{
    h = bil64(c.h[0], c.h[1], c.h[2], c.h[3], wy, wx);
    zeta = bil64(blend_t(c.z0[0], c.z1[0], gt, ft), blend_t(c.z0[1], c.z1[1], gt, ft),
                 blend_t(c.z0[2], c.z1[2], gt, ft), blend_t(c.z0[3], c.z1[3], gt, ft), wy, wx);
}

__device__ __forceinline__ float bil32(const float v[4], float fy, float fx)
{
    const float gy = 1.0f - fy, gx = 1.0f - fx;
    return lerp32(lerp32(v[0], v[1], gx, fx), lerp32(v[2], v[3], gx, fx), gy, fy);
}

// Packed window j of a pair ring at `base` (windows `stride` bytes apart).
__device__ __forceinline__ const void *window(const void *base, uint64_t stride, int j)
{
    return static_cast<const char *>(base) + (uint64_t)(uint32_t)j * stride;
}

// A release's step weights and ring slots, unpacked from its 16-byte table entry.
struct Step {
    float ft, gt, ft1, gt1, ftm, gtm;
    int s0, s1, s0p, s1p, s0m, s1m;
    bool active;
};

// Constants of the isopycnal-following vertical rate (the reference's scalars
// w_max and rho_scale, ab3_isopycnal_following.pyx:54-55).
struct IsoRate {
    double w_max, rho_scale;
};

constexpr int kMaxSlots = 1023; // slot fields of opx_release_step are 10 bits
constexpr int kMaxLaunchSteps = 128; // steps one fused launch may take
constexpr uint64_t kMaxWindowBytes = uint64_t(1) << 32; // one packed window (32-bit byte offsets)

__device__ __forceinline__ int next_slot(int s, int nslots) { return s + 1 == nslots ? 0 : s + 1; }

__device__ __forceinline__ Step load_release(const opx_release_step *__restrict__ table, int r, int nslots)
{
    const opx_release_step e = gload(table, (uint32_t)r);
    Step t;
    t.ft = e.ft;
    t.gt = 1.0f - e.ft;
    t.ft1 = e.ft1;
    t.gt1 = 1.0f - e.ft1;
    t.ftm = e.ftm;
    t.gtm = 1.0f - e.ftm;
    t.s0 = (int)(e.slots & 1023u);
    t.s0p = (int)((e.slots >> 10) & 1023u);
    t.s0m = (int)((e.slots >> 20) & 1023u);
    t.s1 = next_slot(t.s0, nslots);
    t.s1p = next_slot(t.s0p, nslots);
    t.s1m = next_slot(t.s0m, nslots);
    t.active = (e.slots >> 30) & 1u;
    return t;
}

// The packed window of the slot pair with first slot s0 (-1: the rings): the
// pair ring holds one window per slot.
template <int LAYOUT>
__device__ __forceinline__ int packed_index(const opx_window &win, int s0, int s1)
{
    (void)s1; // the pair's second slot is next_slot(s0)
    if constexpr (LAYOUT == OPX_LAYOUT_RING)
        return -1;
    else
        return win.npacked > 0 ? s0 : -1;
}

// ---------------------------------------------------------------- clock

// a + b * c with both operations rounded (no FMA): the reference's
// intermediate time `time + alpha * dt` (timesteppers.py:250).
__device__ __forceinline__ double add_mul_rn(double a, double b, double c)
{
#pragma clang fp contract(off)
    return a + b * c;
}


// Timestepper.get_time_index (timesteppers.py:64-85): false where the
// reference raises (time outside the array, or t == t_last).
__device__ bool time_index(const double *__restrict__ ta, int n, double t, double &tidx)
{
    if (n < 2 || !(t >= ta[0] && t <= ta[n - 1]))
        return false;
    // searchsorted(side="right"): the count of entries <= t.  First the entry g
    // where a uniformly spaced array would hold t: if ta[g] <= t < ta[g + 1]
    // the count is g + 1 (two loads, where each step's time lies for a
    // uniform array); otherwise the binary search runs on the side of g that
    // holds the count, so the result is the search's over any sorted array
    int g = 0;
    if (n > 2 && ta[n - 1] > ta[0])
        g = max(0, min(n - 2, (int)((t - ta[0]) / (ta[n - 1] - ta[0]) * (double)(n - 1))));
    int lo, hi;
    if (ta[g] <= t) {
        lo = g + 1;
        hi = ta[g + 1] <= t ? n : g + 1;
    } else {
        lo = 0;
        hi = g;
    }
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (ta[mid] <= t)
            lo = mid + 1;
        else
            hi = mid;
    }
    int idx = lo - 1;
    if (idx + 1 >= n)
        return false;
    double t0 = ta[idx], t1 = ta[idx + 1];
    double fraction = (t - t0) / (t1 - t0);
    tidx = (double)idx + fraction;
    return true;
}

// divmod(tidx, 1) and FP32 weights (fields.py:409-424), slots in the ring.
__device__ __forceinline__ void weights(double tidx, int nslots, float &ft, float &gt, int &s0, int &s1)
{
    double It = floor(tidx);
    double f = tidx - It;
    ft = (float)f;
    gt = 1.0f - ft;
    long long I = (long long)It;
    s0 = (int)(I % nslots);
    s1 = (int)((I + 1) % nslots);
}

// One release's clock at global step `step`: its table entry (weights and
// slots at t, t + dt and the RK2 time t + alpha dt), then the advance.
__device__ __forceinline__ void clock_step(opx_release_clock &c, opx_release_step *entry, const double *ta,
                                           int ntime, int nslots, int step, double dt, double alpha, int32_t *err)
{
    opx_release_step e{};
    float gt_unused;
    int s1_unused, s0 = 0, s0p = 0, s0m = 0, active = 0;
    if (c.steps_left > 0 && step >= c.start_step) {
        double tidx0, tidx1, tidxm = 0.0;
        double t1 = c.t + dt; // advance_time: time += dt (timesteppers.py:159)
        bool ok = time_index(ta, ntime, c.t, tidx0) && time_index(ta, ntime, t1, tidx1);
        if (alpha > 0.0) // RK2 intermediate time (timesteppers.py:250-251)
            ok = ok && time_index(ta, ntime, add_mul_rn(c.t, alpha, dt), tidxm);
        if (ok) {
            weights(tidx0, nslots, e.ft, gt_unused, s0, s1_unused);
            weights(tidx1, nslots, e.ft1, gt_unused, s0p, s1_unused);
            if (alpha > 0.0)
                weights(tidxm, nslots, e.ftm, gt_unused, s0m, s1_unused);
            active = 1;
            c.t = t1;
            c.steps_left -= 1;
        } else {
            c.steps_left = 0;
            if (err)
                *err = OPX_FLAG_TIME_OUT_OF_RANGE;
        }
    }
    e.slots = (uint32_t)s0 | ((uint32_t)s0p << 10) | ((uint32_t)s0m << 20) | ((uint32_t)active << 30);
    *entry = e;
}

// Advances each release's clock over `nsteps` global steps from `step`,
// writing step k's entry to table[k * tstride + r].
__global__ void k_clock(opx_release_clock *clocks, opx_release_step *table, int tstride, int nsteps, int nrel,
                        const int32_t *nrel_dev, const double *ta, int ntime, int nslots, int step,
                        const int32_t *step_dev, double dt, double alpha, int32_t *err)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (nrel_dev)
        nrel = min(nrel, *nrel_dev);
    if (step_dev)
        step = *step_dev;
    if (r >= nrel)
        return;
    opx_release_clock c = clocks[r];
    for (int k = 0; k < nsteps; ++k)
        clock_step(c, table + (int64_t)k * tstride + r, ta, ntime, nslots, step + k, dt, alpha, err);
    clocks[r] = c;
}

__global__ void k_step_incr(int32_t *step, int nsteps) { *step += nsteps; }

// ---------------------------------------------------------------- arithmetic shared by the fused steps

// a / b for the fused forms' vertical transform: the hardware reciprocal of
// b refined by two Newton-Raphson steps, the quotient, and one correction of
// the quotient from its fused residual (eight instructions instead of the IEEE
// divide's eleven; within an ulp or so of it).  b is a positive normal number
// well inside the FP64 range (a depth, a layer-level difference).
__device__ __forceinline__ double div64(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    r = fma(r, fma(-b, r, 1.0), r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

// zidx of an FP64 z with h, zeta at the particle (_compute_zidx,
// _vertical_coordinate.pxd:41-103): S = (z - zeta) / (zeta + h), the bracket k
// with S_k <= S < S_{k+1} on the w-levels clamped to [0, NZ-1] (S_k increases
// with k, so a walk from the hint k0 finds the reference's binary-search k),
// and the linear inverse k - 1/2 + (S - S_k) / (S_{k+1} - S_k).  The fused
// forms assume a water column (zeta + h > 0).
__device__ __forceinline__ double zidx_from_z(double z, double h, double zeta, double hc, int NZ, const double *Cw,
                                              int k0)
{
    // S and the level values S_k = (hc sigma_k + h C_k) / (hc + h), all scaled
    // by (hc + h)(zeta + h) > 0 (a water column: order and ratios unchanged):
    // the bracket search compares without dividing and the inverse takes one
    // division instead of three
    const double A = (z - zeta) * (hc + h), D = zeta + h;
    auto T_at = [&](int k) { return (hc * Cw[kMaxC + k] + h * Cw[k]) * D; };
    int k = max(0, min(k0, NZ - 1));
    double Tk = T_at(k);
    while (k > 0 && A < Tk) {
        --k;
        Tk = T_at(k);
    }
    double Tk1 = T_at(k + 1);
    while (k < NZ - 1 && A >= Tk1) {
        ++k;
        Tk = Tk1;
        Tk1 = T_at(k + 1);
    }
    return ((double)k - 0.5) + div64(A - Tk, Tk1 - Tk);
}

// Euler -> AB2 -> AB3 start-up of one coordinate (ab3_w_advection.pyx:121-138):
// the derivatives f_{n-1}, f_{n-2} the update uses, from this step's f0 and the
// stored history, read only where the status needs it (h1(), h2()).
template <typename H1, typename H2>
__device__ __forceinline__ void ab3_startup(uint8_t st, float f0, H1 h1, H2 h2, float &f1, float &f2)
{
    if (st == OPX_MULTISTEP_1) {
        f1 = f2 = f0;
    } else {
        f1 = h1();
        f2 = st == OPX_MULTISTEP_2 ? -f0 + 2.0f * f1 : h2();
    }
}

__device__ __forceinline__ uint8_t ab3_next_status(uint8_t st)
{
    return st == OPX_MULTISTEP_1 ? (uint8_t)OPX_MULTISTEP_2 : (st == OPX_MULTISTEP_2 ? (uint8_t)OPX_NORMAL : st);
}

// z at the particle (ab3_w_advection.pyx:144-165; Vtransform 2,
// _vertical_coordinate.pxd:19-35) in FP64, C = Cs_w at w-levels (offset +0.5).
__device__ __forceinline__ double z_at(double zi, int ob, double h, double zeta, double hc, const double *Cw)
{
    const double wC = w64(zi, ob, 0.5);
    const double Cv = (1.0 - wC) * Cw[ob] + wC * Cw[ob + 1];
    const double sigma = (zi + 0.5) * Cw[kInvNZ] - 1.0;
    const double S0 = div64(hc * sigma + h * Cv, hc + h);
    return zeta + (zeta + h) * S0;
}

// The AB3 update of one coordinate (ab3_w_advection.pyx:168-170) in FP64 from
// the FP32 rates (this step's and the FP32 history): the combination
// 23 f0 - 16 f1 + 5 f2 with fused multiply-adds (formed in FP32 it would lose
// bits to the cancellation between its terms), added to the coordinate times
// dt / 12 with one more.
__device__ __forceinline__ double ab3_advance(double x, double dt12, float f0, float f1, float f2)
{
    const double c = fma(5.0, (double)f2, fma(-16.0, (double)f1, 23.0 * (double)f0));
    return fma(c, dt12, x);
}

// ---------------------------------------------------------------- fused AB3

// A particle that passed the step's validation: decoded position, status and
// its release's weights.
struct Pre {
    Axis az, ay, ax;
    Step rs;
    uint8_t st;
    bool sample; // stepped (and sampled) this step
};

// Validation prologue of the fused step (validation.pyx:21-77, inclusive
// bounds in FP64); flags NONFINITE / OUT_OF_DOMAIN particles.
__device__ __forceinline__ Pre ab3_pre(int64_t i, int64_t n, const opx_split_state &s, const opx_grid &g,
                                       const opx_release_step *__restrict__ table, int nslots)
{
    Pre p;
    p.sample = false;
    if (i >= n)
        return p;
    p.st = s.status[i];
    if (p.st & OPX_INACTIVE)
        return p;
    const int r = s.rel ? (int)s.rel[i] : 0;
    p.rs = load_release(table, r, nslots);
    if (!p.rs.active)
        return p;
    const uint32_t cell = (uint32_t)s.cell[i];
    p.az = decode_axis(cell, g.xbits + g.ybits, g.zbits);
    p.ay = decode_axis(cell, g.xbits, g.ybits);
    p.ax = decode_axis(cell, 0, g.xbits);
    if (!(p.az.ok && p.ay.ok && p.ax.ok)) {
        s.status[i] = OPX_NONFINITE;
        return p;
    }
    p.az.q = s.fz[i];
    p.ay.q = s.fy[i];
    p.ax.q = s.fx[i];
    const double zi = p.az.pos(), yi = p.ay.pos(), xi = p.ax.pos();
    if (!(g.zidx_min <= zi && zi <= g.zidx_max && g.yidx_min <= yi && yi <= g.yidx_max && g.xidx_min <= xi &&
          xi <= g.xidx_max)) {
        s.status[i] = OPX_OUT_OF_DOMAIN;
        return p;
    }
    p.sample = true;
    return p;
}

// ---------------------------------------------------------------- wave tiles in LDS

// Nodes [lo0, lo0 + d0) x [lo1, lo1 + d1) x [lo2, lo2 + d2) of a field, row-major
// (x fastest) at dword `off` of the LDS pool.
struct Box {
    int lo0, lo1, lo2, d1, d2, off;
};

// The stencil boxes of one wave's particles: u, v, w of packed window j and
// the 2-D (h, zeta, dx | dy) box t2 (lo1/lo2: y/x origin, d1/d2: extents;
// its dy values follow the float4 values at dword off + 4 * d1 * d2).
struct Tile {
    Box u, v, w, t2;
    int j;
};

template <typename T>
__device__ __forceinline__ void corners_tile(const uint32_t *pool, const Box &b, int b0, int b1, int b2, float gt,
                                             float ft, float v[8])
{
    const int base = ((b0 - b.lo0) * b.d1 + (b1 - b.lo1)) * b.d2 + (b2 - b.lo2);
    const int plane = b.d1 * b.d2;
    const int rows[4] = {base, base + b.d2, base + plane, base + plane + b.d2};
    if constexpr (sizeof(T) == 4) {
        const float2 *P = reinterpret_cast<const float2 *>(pool + b.off);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float2 e0 = P[rows[r]], e1 = P[rows[r] + 1];
            v[2 * r] = blend_t(e0.x, e0.y, gt, ft);
            v[2 * r + 1] = blend_t(e1.x, e1.y, gt, ft);
        }
    } else {
        const uint32_t *P = pool + b.off;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float2 e0 = unpair(P[rows[r]]), e1 = unpair(P[rows[r] + 1]);
            v[2 * r] = blend_t(e0.x, e0.y, gt, ft);
            v[2 * r + 1] = blend_t(e1.x, e1.y, gt, ft);
        }
    }
}

__device__ __forceinline__ void load2_tile(const uint32_t *pool, const Box &b, int by, int bx, bool want_dxdy,
                                           Corners2 &c)
{
    const int base = (by - b.lo1) * b.d2 + (bx - b.lo2);
    const int o[4] = {base, base + 1, base + b.d2, base + b.d2 + 1};
    const float4 *P = reinterpret_cast<const float4 *>(pool + b.off);
    const float *D = reinterpret_cast<const float *>(pool + b.off + 4 * b.d1 * b.d2);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float4 m = P[o[k]];
        c.h[k] = m.x;
        c.z0[k] = m.y;
        c.z1[k] = m.z;
        c.dx[k] = m.w;
        if (want_dxdy)
            c.dy[k] = D[o[k]];
    }
    c.by = by;
    c.bx = bx;
}

__device__ __forceinline__ bool in_box2(const Box &b, int by, int bx)
{
    return by >= b.lo1 && by + 1 < b.lo1 + b.d1 && bx >= b.lo2 && bx + 1 < b.lo2 + b.d2;
}

// One particle's fused AB3 step after ab3_pre.  Cw is the stretching curve
// staged in LDS; with TILE the packed window tile.j and the 2-D fields at the
// particle are read from the wave's LDS tile.
template <typename T, int LAYOUT, bool TILE>
__device__ __forceinline__ void ab3_body(int64_t i, const Pre &p, const opx_split_state &s, const opx_grid &g,
                                         const opx_window &win, double dt, const double *Cw, const uint32_t *pool,
                                         const Tile &tile)
{
    const int NZ = g.NZ, NY = g.NY, NX = g.NX;
    const Step &rs = p.rs;
    const uint8_t st = p.st;
    const Axis az = p.az, ay = p.ay, ax = p.ax;
    const double zi = az.pos(), yi = ay.pos(), xi = ax.pos();

    // ---- f0 = (u/dx, v/dy, w) at (x_n, t_n)  (ab3_w_advection.pyx:86-119)
    const float gt = rs.gt, ft = rs.ft;
    const Ax cz = ax_center(az, NZ), cy = ax_center(ay, NY), cx = ax_center(ax, NX);
    const Ax ixu = ax_inner(ax, NX), iyv = ax_inner(ay, NY), ow = ax_outer(az, NZ);
    float vu[8], vv[8], vw[8];
    const int j = packed_index<LAYOUT>(win, rs.s0, rs.s1);
    const bool tiled = TILE && j == tile.j;
    if (tiled) {
        corners_tile<T>(pool, tile.u, cz.b, cy.b, ixu.b, gt, ft, vu);
        corners_tile<T>(pool, tile.v, cz.b, iyv.b, cx.b, gt, ft, vv);
        corners_tile<T>(pool, tile.w, ow.b, cy.b, cx.b, gt, ft, vw);
    } else if (LAYOUT != OPX_LAYOUT_RING && j >= 0) {
        const void *pu = window(win.pu, win.pu_stride, j);
        const void *pv = window(win.pv, win.pv_stride, j);
        const void *pw = window(win.pw, win.pw_stride, j);
        corners<T, LAYOUT>(pu, nullptr, gt, ft, NY, NX - 1, cz.b, cy.b, ixu.b, vu);
        corners<T, LAYOUT>(pv, nullptr, gt, ft, NY - 1, NX, cz.b, iyv.b, cx.b, vv);
        corners<T, LAYOUT>(pw, nullptr, gt, ft, NY, NX, ow.b, cy.b, cx.b, vw);
    } else {
        const T *U = static_cast<const T *>(win.u), *V = static_cast<const T *>(win.v);
        const T *W = static_cast<const T *>(win.w);
        corners<T, OPX_LAYOUT_RING>(U + rs.s0 * win.u_stride, U + rs.s1 * win.u_stride, gt, ft, NY, NX - 1, cz.b,
                                    cy.b, ixu.b, vu);
        corners<T, OPX_LAYOUT_RING>(V + rs.s0 * win.v_stride, V + rs.s1 * win.v_stride, gt, ft, NY - 1, NX, cz.b,
                                    iyv.b, cx.b, vv);
        corners<T, OPX_LAYOUT_RING>(W + rs.s0 * win.w_stride, W + rs.s1 * win.w_stride, gt, ft, NY, NX, ow.b, cy.b,
                                    cx.b, vw);
    }
    const float u = tri32(vu, cz.w, cy.w, ixu.w);
    const float v = tri32(vv, cz.w, iyv.w, cx.w);
    const double w = (double)tri32(vw, ow.w, cy.w, cx.w);
    const float4 *hdd = reinterpret_cast<const float4 *>(win.hdd);
    Corners2 c2;
    const bool packed0 = LAYOUT != OPX_LAYOUT_RING && j >= 0;
    if (tiled)
        load2_tile(pool, tile.t2, cy.b, cx.b, true, c2);
    else if (packed0)
        load2_packed(static_cast<const float4 *>(window(win.p2, win.p2_stride, j)), win.dy, NX, cy.b, cx.b, true, c2);
    else
        load2_ring(hdd, win.zeta + rs.s0 * win.zeta_stride, win.zeta + rs.s1 * win.zeta_stride, NX, cy.b, cx.b, true,
                   c2);
    const double wy0 = w64(yi, cy.b), wx0 = w64(xi, cx.b);
    const float dxv = bil32(c2.dx, (float)wy0, (float)wx0), dyv = bil32(c2.dy, (float)wy0, (float)wx0);
    const float dx0 = div_nr(u, dxv), dy0 = div_nr(v, dyv), dz0 = (float)w;

    // ---- Euler -> AB2 -> AB3 start-up (history read only where needed)
    float dx1, dy1, dz1, dx2, dy2, dz2;
    ab3_startup(st, dx0, [&] { return s.hx1[i]; }, [&] { return s.hx2[i]; }, dx1, dx2);
    ab3_startup(st, dy0, [&] { return s.hy1[i]; }, [&] { return s.hy2[i]; }, dy1, dy2);
    ab3_startup(st, dz0, [&] { return s.hz1[i]; }, [&] { return s.hz2[i]; }, dz1, dz2);
    const uint8_t st_new = ab3_next_status(st);

    // ---- z at x_n, t_n (kept from the last step, else compute_z), then the
    // AB3 update from the FP32 rates; the new position rounded as stored (the
    // several-step kernels keep this form)
    const double hc = g.hc, dt12 = dt * kInv12;
    double z0 = s.z[i];
    if (__builtin_isnan(z0)) {
        double h0, zeta0;
        eval2(c2, gt, ft, wy0, wx0, h0, zeta0);
        z0 = z_at(zi, ow.b, h0, zeta0, hc, Cw);
    }
    const double z = ab3_advance(z0, dt12, dz0, dz1, dz2);
    const double yn = round32(ab3_advance(yi, dt12, dy0, dy1, dy2));
    const double xn = round32(ab3_advance(xi, dt12, dx0, dx1, dx2));

    // ---- post step: zidx from z with h, zeta at t_{n+1} (ab3_w_advection.pyx:214-229)
    double zn = __builtin_nan(""); // zidx of a non-finite position is non-finite too
    if (__builtin_isfinite(yn) && __builtin_isfinite(xn)) {
        // h, zeta at t_{n+1} at the new position; most particles stay in their
        // cell and the zeta snapshot pair does not change, so the corners read
        // above are reused
        const int by1 = base2(yn, NY), bx1 = base2(xn, NX);
        const int j1 = packed_index<LAYOUT>(win, rs.s0p, rs.s1p);
        const bool same_pair = packed0 ? (j1 == j) : (j1 < 0 && rs.s0p == rs.s0 && rs.s1p == rs.s1);
        if (!(same_pair && by1 == c2.by && bx1 == c2.bx)) {
            if (tiled && j1 == j && in_box2(tile.t2, by1, bx1))
                load2_tile(pool, tile.t2, by1, bx1, false, c2);
            else if (LAYOUT != OPX_LAYOUT_RING && j1 >= 0)
                load2_packed(static_cast<const float4 *>(window(win.p2, win.p2_stride, j1)), win.dy, NX, by1, bx1, false,
                             c2);
            else
                load2_ring(hdd, win.zeta + rs.s0p * win.zeta_stride, win.zeta + rs.s1p * win.zeta_stride, NX, by1,
                           bx1, false, c2);
        }
        double h1, zeta1;
        eval2(c2, rs.gt1, rs.ft1, w64(yn, by1), w64(xn, bx1), h1, zeta1);
        zn = round32(zidx_from_z(z, h1, zeta1, hc, NZ, Cw, ow.b));
    }

    // ---- store: position, history shift (ab3_w_advection.pyx:173-178), status
    const Axis azn = encode_axis(zn, NZ), ayn = encode_axis(yn, NY), axn = encode_axis(xn, NX);
    s.cell[i] = (int32_t)pack_cell(azn, ayn, axn, g);
    s.fz[i] = azn.q;
    s.fy[i] = ayn.q;
    s.fx[i] = axn.q;
    s.hz2[i] = dz1;
    s.hy2[i] = dy1;
    s.hx2[i] = dx1;
    s.hz1[i] = dz0;
    s.hy1[i] = dy0;
    s.hx1[i] = dx0;
    if (st_new != st)
        s.status[i] = st_new;
    s.z[i] = z;
}

template <typename T, int LAYOUT>
__device__ __forceinline__ void ab3_one(int64_t i, int64_t n, const opx_split_state &s, const opx_grid &g,
                                        const opx_window &win, const opx_release_step *__restrict__ table, double dt,
                                        const double *Cw)
{
    const Pre p = ab3_pre(i, n, s, g, table, win.nslots);
    if (p.sample)
        ab3_body<T, LAYOUT, false>(i, p, s, g, win, dt, Cw, nullptr, Tile{});
}

// A particle's state while one launch steps it in registers.
struct PState {
    double zi, yi, xi; // the indices in the form round32 gives (non-finite where the stored axis is)
    double z;          // z (m) after the last step; NaN: not known yet (compute_z from zi)
    float hz1, hy1, hx1, hz2, hy2, hx2; // derivative history f_{n-1}, f_{n-2}
    uint8_t st;
};

template <typename T>
struct Raw3<T, OPX_LAYOUT_RING> {}; // the ring path caches no u, v, w nodes

// A particle's stencils across the steps of one launch: the raw u, v, w nodes
// of its last sample, valid while the snapshot pair (first slot s0) and their
// bases are unchanged, and the 2-D corners c2 at (c2.by, c2.bx) of the pair
// with first slot s0_2d.  Particles move ~0.04 cells per step, so most steps
// blend cached nodes at the new time weights instead of gathering them.
template <typename T, int LAYOUT>
struct StencilCache {
    Raw3<T, LAYOUT> u, v, w;
    uint32_t ku, kv, kw; // bases of the nodes, packed as a cell (base_key)
    int s0;              // -1: no u, v, w nodes cached
    int s0_2d;           // -1: no 2-D corners cached
    Corners2 c2;
};

// A field's stencil base (b0, b1, b2) packed like a cell: one compare per field.
__device__ __forceinline__ uint32_t base_key(int b0, int b1, int b2, const opx_grid &g)
{
    return ((uint32_t)b0 << (g.xbits + g.ybits)) | ((uint32_t)b1 << g.xbits) | (uint32_t)b2;
}

// The 2-D corners (h, zeta of pair (s0, s1), dx, dy) at base (by, bx), from
// packed window j or the rings; kept if already cached.
// RING_OK: a pair that is not packed (j < 0) reads the rings; without it the
// caller has checked that j >= 0.
template <int LAYOUT, bool RING_OK>
__device__ __forceinline__ void sample2(const opx_window &win, int j, int s0, int s1, int NX, int by, int bx,
                                        Corners2 &c2, int &s0_2d)
{
    if (s0_2d == s0 && c2.by == by && c2.bx == bx)
        return;
    if (LAYOUT == OPX_LAYOUT_RING || (RING_OK && j < 0))
        load2_ring(reinterpret_cast<const float4 *>(win.hdd), win.zeta + s0 * win.zeta_stride,
                   win.zeta + s1 * win.zeta_stride, NX, by, bx, true, c2);
    else
        load2_packed(static_cast<const float4 *>(window(win.p2, win.p2_stride, j)), win.dy, NX, by, bx, true, c2);
    s0_2d = s0;
}

// The blended u, v, w stencil values of pair (s0, s1) at weights (gt, ft) and
// stencil bases bz, by, bx (center), bxu, byv (inner x, y) and bzw (outer z).
// A packed LAYOUT blends its cached raw nodes again and reloads a field only
// when its bases or the pair changed; with RING_OK a pair that is not packed
// (j < 0) reads the rings and leaves nothing cached, without it the caller has
// checked that j >= 0.  RN3: the third field (the isopycnal-following
// step's density) is blended with blend_rn.
template <typename T, int LAYOUT, bool RING_OK, bool RN3 = false>
__device__ __forceinline__ void blend_uvw(const opx_window &win, const opx_grid &g, int j, int s0, int s1, float gt,
                                          float ft, int bz, int by, int bx, int bxu, int byv, int bzw,
                                          StencilCache<T, LAYOUT> &c, float vu[8], float vv[8], float vw[8])
{
    const int NY = g.NY, NX = g.NX;
    bool ring = LAYOUT == OPX_LAYOUT_RING;
    if constexpr (LAYOUT != OPX_LAYOUT_RING) {
        if (RING_OK && j < 0) {
            ring = true;
        } else {
            const bool same = c.s0 == s0;
            const uint32_t ku = base_key(bz, by, bxu, g), kv = base_key(bz, byv, bx, g), kw = base_key(bzw, by, bx, g);
            if (!(same && c.ku == ku))
                load_raw<T, LAYOUT>(window(win.pu, win.pu_stride, j), NY, NX - 1, bz, by, bxu, c.u);
            if (!(same && c.kv == kv))
                load_raw<T, LAYOUT>(window(win.pv, win.pv_stride, j), NY - 1, NX, bz, byv, bx, c.v);
            if (!(same && c.kw == kw))
                load_raw<T, LAYOUT>(window(win.pw, win.pw_stride, j), NY, NX, bzw, by, bx, c.w);
            c.s0 = s0;
            c.ku = ku;
            c.kv = kv;
            c.kw = kw;
            blend_raw<T, LAYOUT>(c.u, gt, ft, vu);
            blend_raw<T, LAYOUT>(c.v, gt, ft, vv);
            blend_raw<T, LAYOUT, RN3>(c.w, gt, ft, vw);
        }
    }
    if constexpr (LAYOUT == OPX_LAYOUT_RING || RING_OK) {
        if (ring) {
            const T *U = static_cast<const T *>(win.u), *V = static_cast<const T *>(win.v);
            const T *W = static_cast<const T *>(win.w);
            corners<T, OPX_LAYOUT_RING>(U + s0 * win.u_stride, U + s1 * win.u_stride, gt, ft, NY, NX - 1, bz, by,
                                        bxu, vu);
            corners<T, OPX_LAYOUT_RING>(V + s0 * win.v_stride, V + s1 * win.v_stride, gt, ft, NY - 1, NX, bz, byv,
                                        bx, vv);
            corners<T, OPX_LAYOUT_RING, RN3>(W + s0 * win.w_stride, W + s1 * win.w_stride, gt, ft, NY, NX, bzw, by,
                                             bx, vw);
            c.s0 = -1;
        }
    }
}

// One AB3 step of a validated particle in registers (ab3_w_advection.pyx:20-229,
// the same operations as ab3_body).  A packed LAYOUT reads only its packed
// windows: it returns false, leaving the particle unchanged, if a snapshot pair
// the step needs is not packed (the caller runs such steps with RING).
// ISO: the isopycnal-following step (ab3_isopycnal_following.pyx:26-242): the
// window's third field is the density at rho points and the vertical rate is
// iso.w_max * drho / (iso.rho_scale + |drho|), drho = rho - rho_t.
template <typename T, int LAYOUT, bool ISO = false>
__device__ __forceinline__ bool ab3_step(PState &p, StencilCache<T, LAYOUT> &c, const Step &rs, const opx_grid &g,
                                         const opx_window &win, double dt, const double *Cw,
                                         IsoRate iso = IsoRate{}, double rho_t = 0.0)
{
    const int NZ = g.NZ, NY = g.NY, NX = g.NX;
    const double zi = p.zi, yi = p.yi, xi = p.xi;
    const Coord kz = coord(zi), ky = coord(yi), kx = coord(xi);

    // ---- f0 = (u/dx, v/dy, w) at (x_n, t_n)  (ab3_w_advection.pyx:86-119)
    const float gt = rs.gt, ft = rs.ft;
    const Ax cz = ax_center(kz, NZ), cy = ax_center(ky, NY), cx = ax_center(kx, NX);
    const Ax ixu = ax_inner(kx, NX), iyv = ax_inner(ky, NY), ow = ax_outer(kz, NZ);
    float vu[8], vv[8], vw[8];
    const int j = packed_index<LAYOUT>(win, rs.s0, rs.s1);
    const int j1 = packed_index<LAYOUT>(win, rs.s0p, rs.s1p);
    if (LAYOUT != OPX_LAYOUT_RING && (j < 0 || j1 < 0))
        return false;
    // the third field's z base: w-points (outer) for w, rho points for density
    const int bz3 = ISO ? cz.b : ow.b;
    blend_uvw<T, LAYOUT, false, ISO>(win, g, j, rs.s0, rs.s1, gt, ft, cz.b, cy.b, cx.b, ixu.b, iyv.b, bz3, c, vu,
                                     vv, vw);
    const float u = tri32(vu, cz.w, cy.w, ixu.w);
    const float v = tri32(vv, cz.w, iyv.w, cx.w);
    // w in FP32 like u and v; the density of the isopycnal rate with FP64
    // weights and separately rounded terms (the rate magnifies its last bits)
    double w;
    if constexpr (ISO) {
        const double rho = tri64_rn(vw, w64(zi, cz.b), w64(yi, cy.b), w64(xi, cx.b));
        const double drho = rho - rho_t;
        w = iso.w_max * drho / (iso.rho_scale + fabs(drho));
    } else {
        w = (double)tri32(vw, ow.w, cy.w, cx.w);
    }
    Corners2 &c2 = c.c2;
    sample2<LAYOUT, false>(win, j, rs.s0, rs.s1, NX, cy.b, cx.b, c2, c.s0_2d);
    const double wy0 = w64(yi, cy.b), wx0 = w64(xi, cx.b);
    const float dxv = bil32(c2.dx, (float)wy0, (float)wx0), dyv = bil32(c2.dy, (float)wy0, (float)wx0);
    const float dx0 = div_nr(u, dxv), dy0 = div_nr(v, dyv), dz0 = (float)w;

    // ---- Euler -> AB2 -> AB3 start-up, history from registers; a wave whose
    // particles are all past the start-up (every step after a release's second)
    // takes the history as it is, without the selects
    const uint8_t st = p.st;
    float dx1, dy1, dz1, dx2, dy2, dz2;
    if (__all(st == OPX_NORMAL)) {
        dx1 = p.hx1;
        dx2 = p.hx2;
        dy1 = p.hy1;
        dy2 = p.hy2;
        dz1 = p.hz1;
        dz2 = p.hz2;
    } else {
        ab3_startup(st, dx0, [&] { return p.hx1; }, [&] { return p.hx2; }, dx1, dx2);
        ab3_startup(st, dy0, [&] { return p.hy1; }, [&] { return p.hy2; }, dy1, dy2);
        ab3_startup(st, dz0, [&] { return p.hz1; }, [&] { return p.hz2; }, dz1, dz2);
        p.st = ab3_next_status(st);
    }

    // ---- z at x_n, t_n (kept from the last step, else compute_z), then the
    // AB3 update (as ab3_body)
    const double hc = g.hc, dt12 = dt * kInv12;
    double z0 = p.z;
    if (__builtin_isnan(z0)) {
        double h0, zeta0;
        eval2(c2, gt, ft, wy0, wx0, h0, zeta0);
        z0 = z_at(zi, ow.b, h0, zeta0, hc, Cw);
    }
    const double z = ab3_advance(z0, dt12, dz0, dz1, dz2);
    const double yn = round32(ab3_advance(yi, dt12, dy0, dy1, dy2));
    const double xn = round32(ab3_advance(xi, dt12, dx0, dx1, dx2));

    // ---- post step: zidx from z with h, zeta at t_{n+1} (ab3_w_advection.pyx:214-229)
    double zn = __builtin_nan(""); // zidx of a non-finite position is non-finite too
    if (__builtin_isfinite(yn) && __builtin_isfinite(xn)) {
        // h, zeta at t_{n+1} at the new position: most particles stay in their
        // cell and the zeta snapshot pair does not change, so the corners read
        // above are reused
        const int by1 = base2(yn, NY), bx1 = base2(xn, NX);
        sample2<LAYOUT, false>(win, j1, rs.s0p, rs.s1p, NX, by1, bx1, c2, c.s0_2d);
        double h1, zeta1;
        eval2(c2, rs.gt1, rs.ft1, w64(yn, by1), w64(xn, bx1), h1, zeta1);
        zn = round32(zidx_from_z(z, h1, zeta1, hc, NZ, Cw, ow.b));
    }

    // ---- history shift (ab3_w_advection.pyx:173-178)
    p.zi = zn;
    p.yi = yn;
    p.xi = xn;
    p.z = z;
    p.hz2 = dz1;
    p.hy2 = dy1;
    p.hx2 = dx1;
    p.hz1 = dz0;
    p.hy1 = dy0;
    p.hx1 = dx0;
    return true;
}

enum Scheme { kAB3 = 0, kRK2 = 1, kAB3ISO = 2 };

template <typename T, int LAYOUT, bool RING_OK>
__device__ __forceinline__ bool rk2_step(PState &p, StencilCache<T, LAYOUT> &c, const Step &rs, const opx_grid &g,
                                         const opx_window &win, double dt, double alpha, const double *Cw);

// `nsteps` fused steps of particle i (AB3 or RK2); step k reads its release's
// weights from table[k * tstride + release].  The particle's state is loaded
// once, stepped in registers and stored once: the same values as nsteps
// one-step launches.  Each step validates first (validation.pyx:21-77,
// inclusive bounds in FP64); a flagged particle stops.  A step that meets a
// snapshot pair its packed window lacks (only without RING_OK) stops the
// particle and sets *err.
template <typename T, int LAYOUT, int SCHEME, bool RING_OK>
__device__ __forceinline__ void fused_loop(int64_t i, const opx_split_state &s, const opx_grid &g,
                                          const opx_window &win, const opx_release_step *__restrict__ table,
                                          int tstride, int nsteps, double dt, double alpha, const double *Cw,
                                          int32_t *err, IsoRate iso, uint8_t st0, int r)
{
    PState p;
    p.st = st0;
    StencilCache<T, LAYOUT> c;
    c.s0 = -1;
    c.s0_2d = -1;
    // the target density, read once per launch (ab3_isopycnal_following.pyx:131)
    double rho_t = 0.0;
    if constexpr (SCHEME == kAB3ISO)
        rho_t = s.rho_t[i];
    // the state, loaded before the steps (the loop then needs none of the
    // state's pointers)
    {
        const uint32_t cell = (uint32_t)s.cell[i];
        Axis az = decode_axis(cell, g.xbits + g.ybits, g.zbits);
        Axis ay = decode_axis(cell, g.xbits, g.ybits);
        Axis ax = decode_axis(cell, 0, g.xbits);
        if (az.ok && ay.ok && ax.ok) {
            az.q = s.fz[i];
            ay.q = s.fy[i];
            ax.q = s.fx[i];
        }
        const double nan = __builtin_nan("");
        p.zi = az.ok ? az.pos() : nan;
        p.yi = ay.ok ? ay.pos() : nan;
        p.xi = ax.ok ? ax.pos() : nan;
        p.z = s.z[i];
        if constexpr (SCHEME != kRK2) {
            if (p.st != OPX_MULTISTEP_1) {
                p.hz1 = s.hz1[i];
                p.hy1 = s.hy1[i];
                p.hx1 = s.hx1[i];
                if (p.st != OPX_MULTISTEP_2) {
                    p.hz2 = s.hz2[i];
                    p.hy2 = s.hy2[i];
                    p.hx2 = s.hx2[i];
                }
            }
        }
    }
    bool moved = false;
#pragma unroll 1
    for (int k = 0; k < nsteps; ++k) {
        const Step rs = load_release(table + (int64_t)k * tstride, r, win.nslots);
        if (!rs.active)
            continue;
        const double zi = p.zi, yi = p.yi, xi = p.xi;
        // a non-finite index fails the inclusive bounds too (comparisons with
        // NaN are false, infinities lie outside), so the finiteness test that
        // tells NONFINITE from OUT_OF_DOMAIN is needed only on that rare path;
        // non-short-circuit & keeps the six compares as lane masks
        const uint64_t in_mask =
            __builtin_amdgcn_ballot_w64(g.zidx_min <= zi) & __builtin_amdgcn_ballot_w64(zi <= g.zidx_max) &
            __builtin_amdgcn_ballot_w64(g.yidx_min <= yi) & __builtin_amdgcn_ballot_w64(yi <= g.yidx_max) &
            __builtin_amdgcn_ballot_w64(g.xidx_min <= xi) & __builtin_amdgcn_ballot_w64(xi <= g.xidx_max);
        const bool inb = (in_mask >> __lane_id()) & 1u;
        if (!inb) {
            const bool fin = __builtin_isfinite(zi) && __builtin_isfinite(yi) && __builtin_isfinite(xi);
            p.st = fin ? OPX_OUT_OF_DOMAIN : OPX_NONFINITE;
            break;
        }
        bool ok;
        if constexpr (SCHEME == kAB3)
            ok = ab3_step<T, LAYOUT>(p, c, rs, g, win, dt, Cw);
        else if constexpr (SCHEME == kAB3ISO)
            ok = ab3_step<T, LAYOUT, true>(p, c, rs, g, win, dt, Cw, iso, rho_t);
        else
            ok = rk2_step<T, LAYOUT, RING_OK>(p, c, rs, g, win, dt, alpha, Cw);
        if (!ok) {
            if (err)
                *err = OPX_FLAG_UNPACKED_PAIR;
            break;
        }
        moved = true;
    }
    // the state's addresses computed afresh from the index (through an empty
    // asm statement, so they are not kept from the loads above across the
    // steps, where they would hold VGPRs)
    asm volatile("" : "+v"(i));
    if (moved) {
        s.z[i] = p.z;
        const Axis az = encode_axis(p.zi, g.NZ), ay = encode_axis(p.yi, g.NY), ax = encode_axis(p.xi, g.NX);
        s.cell[i] = (int32_t)pack_cell(az, ay, ax, g);
        s.fz[i] = az.q;
        s.fy[i] = ay.q;
        s.fx[i] = ax.q;
        if constexpr (SCHEME != kRK2) {
            s.hz2[i] = p.hz2;
            s.hy2[i] = p.hy2;
            s.hx2[i] = p.hx2;
            s.hz1[i] = p.hz1;
            s.hy1[i] = p.hy1;
            s.hx1[i] = p.hx1;
        }
    }
    if (p.st != st0)
        s.status[i] = p.st;
}

template <typename T, int LAYOUT, int SCHEME, bool RING_OK>
__device__ __forceinline__ void fused_run(int64_t i, int64_t n, const opx_split_state &s, const opx_grid &g,
                                          const opx_window &win, const opx_release_step *__restrict__ table,
                                          int tstride, int nsteps, double dt, double alpha, const double *Cw,
                                          int32_t *err, IsoRate iso = IsoRate{})
{
    if (i >= n)
        return;
    const uint8_t st0 = s.status[i];
    if (st0 & OPX_INACTIVE)
        return;
    const int r = s.rel ? (int)s.rel[i] : 0;
    // a wave whose live particles all belong to one release steps it with the
    // release index, its table entries and window addresses in scalar registers
    const int r0 = __builtin_amdgcn_readfirstlane(r);
    if (__all(r == r0))
        fused_loop<T, LAYOUT, SCHEME, RING_OK>(i, s, g, win, table, tstride, nsteps, dt, alpha, Cw, err, iso, st0,
                                               r0);
    else
        fused_loop<T, LAYOUT, SCHEME, RING_OK>(i, s, g, win, table, tstride, nsteps, dt, alpha, Cw, err, iso, st0,
                                               r);
}

// Fused AB3 steps (one, or nsteps with MULTI), one particle per thread.  The
// grid covers s.n particles;
// the particles actually stepped are the first min(s.n, *n_active), so one
// captured graph serves any active count up to s.n (surplus blocks exit at
// once).  XCD k (blocks k, k+8, ...) takes the k-th contiguous eighth of the
// (sorted) particles in chunks of BLOCK, so neighbouring particles share one
// L2 and the resident blocks of an XCD sweep one compact region at a time.
// A loop over chunks would let a smaller grid cover s.n, but it lifts the
// kernel past 128 VGPRs (3 waves/SIMD instead of 4), which costs more than
// the surplus blocks.
template <typename T, int LAYOUT, int BLOCK, bool IND, bool MULTI>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MULTI ? 3 : 1)))
    k_ab3_fused(opx_split_state s, opx_grid g, opx_window win_arg, const opx_window *__restrict__ win_dev,
                const opx_release_step *__restrict__ table, int tstride, int nsteps, double dt,
                const int32_t *__restrict__ n_active, int32_t *err)
{
    // IND: the window descriptor lives in device memory, so a captured graph
    // stays valid when the host repacks windows between replays
    const opx_window win = IND ? *win_dev : win_arg;
    // the stretching curve C (NZ+1 <= kMaxC values) is read by every particle: keep it in LDS
    __shared__ double Cw[kLevels];
    stage_levels(Cw, win.C, g.NZ, threadIdx.x, BLOCK);
    __syncthreads();

    int64_t n = s.n;
    if (n_active)
        n = min(n, (int64_t)*n_active);
    const int64_t chunks = (n + BLOCK - 1) / BLOCK;
    // the host launches roundup8(ceil(s.n / BLOCK)) blocks, so gridDim.x / 8
    // >= ceil(chunks / 8) >= c1 - c0: every chunk has a block
    const uint32_t xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
    const int64_t c = chunks * xcd / kXcds + lb, c1 = chunks * (xcd + 1) / kXcds;
    if (c < c1) {
        if constexpr (MULTI)
            fused_run<T, LAYOUT, kAB3, false>(c * BLOCK + threadIdx.x, n, s, g, win, table, tstride, nsteps, dt, 0.0,
                                              Cw, err);
        else
            ab3_one<T, LAYOUT>(c * BLOCK + threadIdx.x, n, s, g, win, table, dt, Cw);
    }
}

// Fused isopycnal-following AB3 steps (FP32 storage; the window's third field
// holds the density at rho points), one particle per thread with the launch
// shape of k_ab3_fused; one step per launch runs the several-step path with
// nsteps = 1 (the same arithmetic).
template <int LAYOUT, int BLOCK, bool IND>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(3)))
    k_ab3_iso_fused(opx_split_state s, opx_grid g, opx_window win_arg, const opx_window *__restrict__ win_dev,
                    const opx_release_step *__restrict__ table, int tstride, int nsteps, double dt, IsoRate iso,
                    const int32_t *__restrict__ n_active, int32_t *err)
{
    const opx_window win = IND ? *win_dev : win_arg;
    __shared__ double Cw[kLevels];
    stage_levels(Cw, win.C, g.NZ, threadIdx.x, BLOCK);
    __syncthreads();

    int64_t n = s.n;
    if (n_active)
        n = min(n, (int64_t)*n_active);
    const int64_t chunks = (n + BLOCK - 1) / BLOCK;
    const uint32_t xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
    const int64_t c = chunks * xcd / kXcds + lb, c1 = chunks * (xcd + 1) / kXcds;
    if (c < c1)
        fused_run<float, LAYOUT, kAB3ISO, false>(c * BLOCK + threadIdx.x, n, s, g, win, table, tstride, nsteps, dt,
                                                 0.0, Cw, err, iso);
}

// ---------------------------------------------------------------- LDS-tiled fused AB3 (one wave per workgroup)

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void global_void;

constexpr int kTile3 = 4096; // LDS dwords for the u, v, w boxes of a wave (16 KB)
constexpr int kTile2 = 256;  // nodes of the 2-D box (5 dwords each: float4 + dy)

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
        v = min(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1)
        v = max(v, __shfl_xor(v, m, 64));
    return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Node range [lo, lo + d) holding both stencil nodes of every base the cells
// [cmin, cmax] produce on one axis; the base functions are those of ax_center
// (clamp(c, 0, N-2)), ax_inner (clamp(c - 1 + upper, 0, N-3)) and ax_outer
// (clamp(c + upper, 0, N-1)), all non-decreasing in c.
__device__ __forceinline__ void range_center(int cmin, int cmax, int N, int &lo, int &d)
{
    lo = max(0, min(cmin, N - 2));
    d = max(0, min(cmax, N - 2)) + 2 - lo;
}
__device__ __forceinline__ void range_inner(int cmin, int cmax, int N, int &lo, int &d)
{
    lo = max(0, min(cmin - 1, N - 3));
    d = max(0, min(cmax, N - 3)) + 2 - lo;
}
__device__ __forceinline__ void range_outer(int cmin, int cmax, int N, int &lo, int &d)
{
    lo = max(0, min(cmin, N - 1));
    d = max(0, min(cmax + 1, N - 1)) + 2 - lo;
}

// m = ceil(2^32 / d) for d >= 2 (from a 32-bit division), so that
// umulhi(e, m) = floor(e / d) for e * (d - 1) < 2^32; d = 1 is handled by the caller.
__device__ __forceinline__ uint32_t div_magic(uint32_t d) { return d > 1 ? 0xffffffffu / d + 1u : 0u; }
__device__ __forceinline__ uint32_t div_by(uint32_t e, uint32_t d, uint32_t m) { return d > 1 ? __umulhi(e, m) : e; }

// Copy box b (d0 planes, EPN dwords per node) of a field with dims (., n1, n2)
// into the LDS pool with direct-to-LDS loads: lane l of the instruction that
// covers pool dwords [base, base + 64) fetches dword base + l of the box.
template <int EPN>
__device__ __forceinline__ void stage_box(uint32_t *pool, const Box &b, int d0, const void *src, int n1, int n2,
                                          int lane)
{
    const uint32_t nd = (uint32_t)(d0 * b.d1 * b.d2 * EPN);
    const uint32_t d2 = (uint32_t)b.d2, d1 = (uint32_t)b.d1;
    const uint32_t m2 = div_magic(d2), m1 = div_magic(d1);
    const uint32_t *S = static_cast<const uint32_t *>(src);
    for (uint32_t base = 0; base < nd; base += 64) {
        const uint32_t dw = base + (uint32_t)lane;
        if (dw < nd) {
            const uint32_t e = dw / EPN;
            const uint32_t row = div_by(e, d2, m2);
            const uint32_t x = e - row * d2;
            const uint32_t zz = div_by(row, d1, m1);
            const uint32_t yy = row - zz * d1;
            const uint64_t node = ((uint64_t)(b.lo0 + (int)zz) * (uint64_t)n1 + (uint64_t)(b.lo1 + (int)yy)) *
                                      (uint64_t)n2 +
                                  (uint64_t)(b.lo2 + (int)x);
            const uint32_t *a = S + node * EPN + dw % EPN;
            __builtin_amdgcn_global_load_lds((global_void *)a, (lds_void *)(pool + b.off + base), 4, 0, 0);
        }
    }
}

// Fused AB3 step with the stencils of each wave staged in LDS (PAIR windows):
// the wave's particles are neighbours in the Morton order, so the boxes around
// their cells are small; the boxes are copied with coalesced direct-to-LDS
// loads and every stencil read then hits LDS.  A wave whose boxes exceed the
// pool, that holds particles of different packed windows or none reads from
// global memory as k_ab3_fused does (identical arithmetic either way).
template <typename T, bool IND>
__global__ void __launch_bounds__(64)
    k_ab3_tile(opx_split_state s, opx_grid g, opx_window win_arg, const opx_window *__restrict__ win_dev,
               const opx_release_step *__restrict__ table, double dt, const int32_t *__restrict__ n_active)
{
    constexpr int EPN = sizeof(typename Packed<T>::pair) / 4;
    const opx_window win = IND ? *win_dev : win_arg;
    __shared__ double Cw[kLevels];
    __shared__ __attribute__((aligned(16))) uint32_t pool[kTile3 + 5 * kTile2];
    const int lane = (int)threadIdx.x;
    stage_levels(Cw, win.C, g.NZ, lane, 64);

    int64_t n = s.n;
    if (n_active)
        n = min(n, (int64_t)*n_active);
    const int64_t chunks = (n + 63) / 64;
    const uint32_t xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
    const int64_t c = chunks * xcd / kXcds + lb, c1 = chunks * (xcd + 1) / kXcds;
    const int64_t i = c < c1 ? c * 64 + lane : n;
    const Pre p = ab3_pre(i, n, s, g, table, win.nslots);

    // ---- the wave's cell range and packed window
    constexpr int BIG = 1 << 30;
    const int j = p.sample ? packed_index<OPX_LAYOUT_PAIR>(win, p.rs.s0, p.rs.s1) : -1;
    const int zmin = uniform(wave_min(p.sample ? p.az.c : BIG)), zmax = uniform(wave_max(p.sample ? p.az.c : -BIG));
    const int ymin = uniform(wave_min(p.sample ? p.ay.c : BIG)), ymax = uniform(wave_max(p.sample ? p.ay.c : -BIG));
    const int xmin = uniform(wave_min(p.sample ? p.ax.c : BIG)), xmax = uniform(wave_max(p.sample ? p.ax.c : -BIG));
    const int jmin = uniform(wave_min(p.sample ? j : BIG)), jmax = uniform(wave_max(p.sample ? j : -BIG));
    Tile t;
    bool use = zmin <= zmax && jmin == jmax && jmin >= 0;
    if (use) {
        const int NZ = g.NZ, NY = g.NY, NX = g.NX;
        int du0, dv0, dw0;
        range_center(zmin, zmax, NZ, t.u.lo0, du0);
        range_center(ymin, ymax, NY, t.u.lo1, t.u.d1);
        range_inner(xmin, xmax, NX, t.u.lo2, t.u.d2);
        range_center(zmin, zmax, NZ, t.v.lo0, dv0);
        range_inner(ymin, ymax, NY, t.v.lo1, t.v.d1);
        range_center(xmin, xmax, NX, t.v.lo2, t.v.d2);
        range_outer(zmin, zmax, NZ, t.w.lo0, dw0);
        range_center(ymin, ymax, NY, t.w.lo1, t.w.d1);
        range_center(xmin, xmax, NX, t.w.lo2, t.w.d2);
        t.t2.lo0 = 0;
        range_center(ymin, ymax, NY, t.t2.lo1, t.t2.d1);
        range_center(xmin, xmax, NX, t.t2.lo2, t.t2.d2);
        // 64-bit volumes: a wave of scattered particles spans the whole grid
        const int64_t vu = (int64_t)du0 * t.u.d1 * t.u.d2, vv = (int64_t)dv0 * t.v.d1 * t.v.d2;
        const int64_t vw = (int64_t)dw0 * t.w.d1 * t.w.d2, v2 = (int64_t)t.t2.d1 * t.t2.d2;
        use = (vu + vv + vw) * EPN <= kTile3 && v2 <= kTile2;
        t.u.off = 0;
        t.v.off = (int)(vu * EPN);
        t.w.off = (int)((vu + vv) * EPN);
        t.t2.off = kTile3;
        t.j = jmin;
        if (use) {
            const int jj = t.j;
            stage_box<EPN>(pool, t.u, du0, window(win.pu, win.pu_stride, jj), NY, NX - 1, lane);
            stage_box<EPN>(pool, t.v, dv0, window(win.pv, win.pv_stride, jj), NY - 1, NX, lane);
            stage_box<EPN>(pool, t.w, dw0, window(win.pw, win.pw_stride, jj), NY, NX, lane);
            stage_box<4>(pool, t.t2, 1, window(win.p2, win.p2_stride, jj), 1, NX, lane);
            Box tdy = t.t2;
            tdy.off = kTile3 + 4 * (int)v2;
            stage_box<1>(pool, tdy, 1, win.dy, 1, NX, lane);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!p.sample)
        return;
    if (use)
        ab3_body<T, OPX_LAYOUT_PAIR, true>(i, p, s, g, win, dt, Cw, pool, t);
    else
        ab3_body<T, OPX_LAYOUT_PAIR, false>(i, p, s, g, win, dt, Cw, nullptr, t);
}

// ---------------------------------------------------------------- packed windows

// Packs node i of a field of n1 rows of n2 nodes per level from slots p, q.
template <typename T, int LAYOUT>
__device__ __forceinline__ void pack_node(const T *__restrict__ p, const T *__restrict__ q, void *__restrict__ out,
                                          int64_t i, int n1, int n2)
{
    if constexpr (LAYOUT == OPX_LAYOUT_OCT) {
        // neighbours beyond the last row / column are never stencil nodes of a base here
        const int64_t x1 = (i % n2) + 1 < n2 ? 1 : 0;
        const int64_t y1 = ((i / n2) % n1) + 1 < n1 ? n2 : 0;
        const int64_t c[4] = {i, i + x1, i + y1, i + y1 + x1};
        if constexpr (sizeof(T) == 4) {
            const float *P = (const float *)p, *Q = (const float *)q;
            float8 e;
            e.a = make_float4(P[c[0]], P[c[1]], P[c[2]], P[c[3]]);
            e.b = make_float4(Q[c[0]], Q[c[1]], Q[c[2]], Q[c[3]]);
            static_cast<float8 *>(out)[i] = e;
        } else {
            const uint16_t *P = (const uint16_t *)p, *Q = (const uint16_t *)q;
            static_cast<uint4 *>(out)[i] =
                make_uint4((uint32_t)P[c[0]] | ((uint32_t)P[c[1]] << 16), (uint32_t)P[c[2]] | ((uint32_t)P[c[3]] << 16),
                           (uint32_t)Q[c[0]] | ((uint32_t)Q[c[1]] << 16), (uint32_t)Q[c[2]] | ((uint32_t)Q[c[3]] << 16));
        }
        return;
    }
    if constexpr (LAYOUT == OPX_LAYOUT_YPAIR) {
        // the y+1 neighbour of the last row is never a stencil node of a base here
        const int64_t y1 = ((i / n2) % n1) + 1 < n1 ? i + n2 : i;
        if constexpr (sizeof(T) == 4) {
            const float *P = (const float *)p, *Q = (const float *)q;
            static_cast<float4 *>(out)[i] = make_float4(P[i], Q[i], P[y1], Q[y1]);
        } else {
            const uint16_t *P = (const uint16_t *)p, *Q = (const uint16_t *)q;
            static_cast<uint2 *>(out)[i] = make_uint2((uint32_t)P[i] | ((uint32_t)Q[i] << 16),
                                                      (uint32_t)P[y1] | ((uint32_t)Q[y1] << 16));
        }
        return;
    }
    if constexpr (LAYOUT == OPX_LAYOUT_PAIR) {
        if constexpr (sizeof(T) == 4) {
            static_cast<float2 *>(out)[i] = make_float2(((const float *)p)[i], ((const float *)q)[i]);
        } else {
            static_cast<uint32_t *>(out)[i] = (uint32_t)((const uint16_t *)p)[i] |
                                              ((uint32_t)((const uint16_t *)q)[i] << 16);
        }
    } else {
        // the x+1 neighbour of the last node of a row is never a stencil node
        const int64_t i1 = (i % n2) + 1 < n2 ? i + 1 : i;
        if constexpr (sizeof(T) == 4) {
            const float *P = (const float *)p, *Q = (const float *)q;
            static_cast<float4 *>(out)[i] = make_float4(P[i], P[i1], Q[i], Q[i1]);
        } else {
            const uint16_t *P = (const uint16_t *)p, *Q = (const uint16_t *)q;
            static_cast<uint2 *>(out)[i] = make_uint2((uint32_t)P[i] | ((uint32_t)P[i1] << 16),
                                                      (uint32_t)Q[i] | ((uint32_t)Q[i1] << 16));
        }
    }
}

// Packs the nodes of rows [row0, row0 + nrows) of every z-level (n rows of n2
// nodes per level, n1 rows per level): thread j of the launch packs node
// (z, row0 + r, x) with j = (z nrows + r) n2 + x, so a launch over all rows
// packs node j.  The window limit (4 GiB) keeps node indices below 2^32.
template <typename T, int LAYOUT>
__global__ void k_pack(const T *__restrict__ p, const T *__restrict__ q, void *__restrict__ out, uint32_t n, int n1,
                       int n2, int row0, int nrows)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n)
        return;
    const uint32_t band = (uint32_t)nrows * (uint32_t)n2;
    const uint32_t zl = j / band, r = j - zl * band;
    pack_node<T, LAYOUT>(p, q, out, ((int64_t)zl * n1 + row0) * n2 + r, n1, n2);
}

// Boxes of (y0, y1, x0, x1) of one field's nodes over every level, up to 128
// per launch, with the element prefix sums of the boxes (as for the tile
// gathers).
constexpr int kPackBoxes = 128;
struct PackBoxes {
    int32_t box[kPackBoxes][4];
    int64_t first[kPackBoxes + 1];
    int32_t n;
};

// Packs every node of the boxes (all levels): element e of the launch is node
// (z, y0 + yy, x0 + xx) of its box, found by binary search, x fastest.
template <typename T, int LAYOUT>
__global__ void k_pack_boxes(const T *__restrict__ p, const T *__restrict__ q, void *__restrict__ out, int n1, int n2,
                             PackBoxes B)
{
    const int64_t total = B.first[B.n];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = B.n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (B.first[mid] <= e)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int32_t *x = B.box[lo];
        const uint32_t w = (uint32_t)(x[3] - x[2]), h = (uint32_t)(x[1] - x[0]);
        uint32_t k = (uint32_t)(e - B.first[lo]); // a box of one window: fewer than 2^32 nodes
        const uint32_t xx = k % w;
        k /= w;
        const uint32_t yy = k % h, z = k / h;
        pack_node<T, LAYOUT>(p, q, out, ((int64_t)z * n1 + x[0] + yy) * n2 + x[2] + xx, n1, n2);
    }
}

// p2[i] = {h, zeta[s0], zeta[s1], dx} per rho point, for points i0 .. i0 + n - 1
__global__ void k_pack2(const float4 *__restrict__ hdd, const float *__restrict__ zp, const float *__restrict__ zq,
                        float4 *__restrict__ out, int64_t i0, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    i += i0;
    const float4 m = hdd[i];
    out[i] = make_float4(m.x, zp[i], zq[i], m.y);
}

// The QUAD copy of a static field for the drop-in interpolants: node i holds
// its corners (y, x), (y, x+1), (y+1, x), (y+1, x+1); neighbours beyond the
// last row / column repeat the node (never corners of a base, which the
// interpolants clamp to size - 2).
template <typename T>
__global__ void k_pack_quad(const T *__restrict__ p, void *__restrict__ out, uint32_t n, int n1, int n2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t x1 = (i % (uint32_t)n2) + 1 < (uint32_t)n2 ? 1u : 0u;
    const uint32_t y1 = ((i / (uint32_t)n2) % (uint32_t)n1) + 1 < (uint32_t)n1 ? (uint32_t)n2 : 0u;
    const uint32_t c[4] = {i, i + x1, i + y1, i + y1 + x1};
    if constexpr (sizeof(T) == 4) {
        const float *P = (const float *)p;
        static_cast<float4 *>(out)[i] = make_float4(P[c[0]], P[c[1]], P[c[2]], P[c[3]]);
    } else {
        const uint16_t *P = (const uint16_t *)p;
        static_cast<uint2 *>(out)[i] =
            make_uint2((uint32_t)P[c[0]] | ((uint32_t)P[c[1]] << 16), (uint32_t)P[c[2]] | ((uint32_t)P[c[3]] << 16));
    }
}

// The HEX copy of a snapshot pair for the drop-in interpolants: node i holds
// the corners (z, z+1) x (y, y+1) x (x, x+1) of p then of q, corner (dz, dy,
// dx) at (dz * 2 + dy) * 2 + dx; neighbours beyond the last level / row /
// column repeat the node (never corners of a base, which the interpolants
// clamp to size - 2).
template <typename T>
__global__ void k_pack_hex(const T *__restrict__ p, const T *__restrict__ q, void *__restrict__ out, uint32_t n, int n0,
                           int n1, int n2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t plane = (uint32_t)n1 * (uint32_t)n2;
    const uint32_t x1 = (i % (uint32_t)n2) + 1 < (uint32_t)n2 ? 1u : 0u;
    const uint32_t y1 = ((i / (uint32_t)n2) % (uint32_t)n1) + 1 < (uint32_t)n1 ? (uint32_t)n2 : 0u;
    const uint32_t z1 = i / plane + 1 < (uint32_t)n0 ? plane : 0u;
    const uint32_t c[8] = {i, i + x1, i + y1, i + y1 + x1, i + z1, i + z1 + x1, i + z1 + y1, i + z1 + y1 + x1};
    if constexpr (sizeof(T) == 4) {
        const float *P = (const float *)p, *Q = (const float *)q;
        float4 *o = static_cast<float4 *>(out) + 4 * (size_t)i;
        o[0] = make_float4(P[c[0]], P[c[1]], P[c[2]], P[c[3]]);
        o[1] = make_float4(P[c[4]], P[c[5]], P[c[6]], P[c[7]]);
        o[2] = make_float4(Q[c[0]], Q[c[1]], Q[c[2]], Q[c[3]]);
        o[3] = make_float4(Q[c[4]], Q[c[5]], Q[c[6]], Q[c[7]]);
    } else {
        const uint16_t *P = (const uint16_t *)p, *Q = (const uint16_t *)q;
        auto two = [](uint16_t a, uint16_t b) { return (uint32_t)a | ((uint32_t)b << 16); };
        uint4 *o = static_cast<uint4 *>(out) + 2 * (size_t)i;
        o[0] = make_uint4(two(P[c[0]], P[c[1]]), two(P[c[2]], P[c[3]]), two(P[c[4]], P[c[5]]), two(P[c[6]], P[c[7]]));
        o[1] = make_uint4(two(Q[c[0]], Q[c[1]]), two(Q[c[2]], Q[c[3]]), two(Q[c[4]], Q[c[5]]), two(Q[c[6]], Q[c[7]]));
    }
}

// ---------------------------------------------------------------- sparse snapshot streaming

__global__ void k_fill_u8(uint8_t *__restrict__ p, int64_t n, uint8_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// map[ty * ntx + tx] = 1 for every tile of tile x tile rho columns holding a
// particle whose status is not INACTIVE and whose horizontal cells are finite
// (cells outside the grid count in the nearest edge tile).  Threads that mark
// one tile store the same byte.
__global__ void k_tile_occupancy(opx_split_state s, opx_grid g, int tile, uint8_t *__restrict__ map, int ntx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n || (s.status[i] & OPX_INACTIVE))
        return;
    const uint32_t cell = (uint32_t)s.cell[i];
    const Axis y = decode_axis(cell, g.xbits, g.ybits), x = decode_axis(cell, 0, g.xbits);
    if (!(y.ok && x.ok))
        return;
    const int yy = min(max(y.c, 0), g.NY - 1), xx = min(max(x.c, 0), g.NX - 1);
    map[(yy / tile) * ntx + xx / tile] = 1;
}

// Boxes copied from mapped host memory: up to kGatherBoxes boxes per launch,
// passed by value with the prefix counts of their elements; thread e of the
// grid-stride loop copies element e of the launch (its box by binary search,
// then x fastest, so the lanes of a wave read consecutive addresses of a row).
constexpr int kGatherBoxes = 128;
struct GatherBoxes {
    int32_t box[kGatherBoxes][4]; // (y0, y1, x0, x1)
    int64_t first[kGatherBoxes + 1];
    int32_t n;
};

template <typename E>
__global__ void k_gather_boxes(E *__restrict__ dst, const E *__restrict__ src, int64_t ny, int64_t nx, GatherBoxes B)
{
    const int64_t total = B.first[B.n];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = B.n - 1; // the box b with first[b] <= e < first[b + 1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (B.first[mid] <= e)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int32_t *x = B.box[lo];
        const uint32_t w = (uint32_t)(x[3] - x[2]), h = (uint32_t)(x[1] - x[0]);
        uint32_t k = (uint32_t)(e - B.first[lo]); // below 2^31: one box of at most (nz, ny, nx) < 2^32 elements
        const uint32_t xx = k % w;
        k /= w;
        const uint32_t yy = k % h, z = k / h;
        const int64_t idx = ((int64_t)z * ny + x[0] + yy) * nx + x[2] + xx;
        dst[idx] = src[idx];
    }
}

// ---------------------------------------------------------------- conversions

__global__ void k_split_from_f64(opx_split_state s, opx_grid g, const double *zidx, const double *yidx,
                                 const double *xidx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n)
        return;
    Axis z = encode_axis(zidx[i], g.NZ), y = encode_axis(yidx[i], g.NY), x = encode_axis(xidx[i], g.NX);
    s.cell[i] = (int32_t)pack_cell(z, y, x, g);
    s.fz[i] = z.q;
    s.fy[i] = y.q;
    s.fx[i] = x.q;
    if (s.z)
        s.z[i] = __builtin_nan(""); // the first step computes z from zidx
}

__global__ void k_split_to_f64(opx_split_state s, opx_grid g, double *zidx, double *yidx, double *xidx)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n)
        return;
    const uint32_t cell = (uint32_t)s.cell[i];
    Axis z = decode_axis(cell, g.xbits + g.ybits, g.zbits);
    Axis y = decode_axis(cell, g.xbits, g.ybits);
    Axis x = decode_axis(cell, 0, g.xbits);
    z.q = s.fz[i];
    y.q = s.fy[i];
    x.q = s.fx[i];
    const double nan = __builtin_nan("");
    int64_t o = s.pid ? (int64_t)s.pid[i] : i;
    zidx[o] = z.ok ? z.pos() : nan;
    yidx[o] = y.ok ? y.pos() : nan;
    xidx[o] = x.ok ? x.pos() : nan;
}

// ---------------------------------------------------------------- output events at split positions

constexpr int kMaxSamples = 8;

// The reference's output kernels and compute_z (linear_interpolation.pyx:34-136,
// _roms_core.pyx:17-73) at the exact FP64 index of each split position, with
// the drop-in's FP64 arithmetic (opx_numerics.hpp); one thread per particle in
// the state's order, all samples of a particle in one pass, written as one
// whole record (NaN-padded) so that the scatter by pid writes full pieces.
// ONE_REL: a single release, whose field row every lane reads alike.
template <bool ONE_REL>
__global__ void __launch_bounds__(kBlock)
    k_sample(opx_split_state s, opx_grid g, const opx_field *__restrict__ fields, int nfields,
             const opx_sample *__restrict__ samples, int nsamples, double hc, double *__restrict__ out, int stride)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n)
        return;
    const double nan = __builtin_nan("");
    double v[kMaxSamples];
#pragma unroll
    for (int k = 0; k < kMaxSamples; ++k)
        v[k] = nan;
    if (!(s.status[i] & OPX_INACTIVE)) {
        const uint32_t cell = (uint32_t)s.cell[i];
        Axis z = decode_axis(cell, g.xbits + g.ybits, g.zbits);
        Axis y = decode_axis(cell, g.xbits, g.ybits);
        Axis x = decode_axis(cell, 0, g.xbits);
        z.q = s.fz[i];
        y.q = s.fy[i];
        x.q = s.fx[i];
        const double zi = z.ok ? z.pos() : nan, yi = y.ok ? y.pos() : nan, xi = x.ok ? x.pos() : nan;
        const opx_field *row = ONE_REL ? fields : fields + (int64_t)(s.rel ? (int)s.rel[i] : 0) * nfields;
#pragma unroll
        for (int k = 0; k < kMaxSamples; ++k) {
            if (k >= nsamples)
                break;
            const opx_sample sm = samples[k];
            if (sm.kind == OPX_SAMPLE_LINEAR) {
                v[k] = linear_f64(row[sm.field], zi);
            } else if (sm.kind == OPX_SAMPLE_BILINEAR) {
                v[k] = bilinear_f64(row[sm.field], yi, xi);
            } else if (sm.kind == OPX_SAMPLE_TRILINEAR) {
                v[k] = trilinear_f64(row[sm.field], zi, yi, xi);
            } else {
                const double h = bilinear_f64(row[sm.field], yi, xi);
                const double zeta = bilinear_f64(row[sm.field2], yi, xi);
                const double C = linear_f64(row[sm.field3], zi);
                v[k] = compute_z(zi, g.NZ, hc, h, C, zeta);
            }
        }
    }
    double *rec = out + (s.pid ? (int64_t)s.pid[i] : i) * stride;
    if (stride == 1) {
        rec[0] = v[0];
        return;
    }
    // stride 2, 4 or 8 with a 16-byte aligned `out`: 16-byte stores
    double2 *r2 = reinterpret_cast<double2 *>(rec);
#pragma unroll
    for (int k = 0; k < kMaxSamples / 2; ++k)
        if (2 * k < stride)
            r2[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// ---------------------------------------------------------------- fused RK2

// `(Py_ssize_t)p` clamped to [0, hi] (linear.pxd:3-9); NaN and values outside
// the int64 range convert to INT64_MIN on x86-64, i.e. clamp to 0.
__device__ __forceinline__ int trunc_clamp(double p, int hi)
{
    if (!(p > -9223372036854775808.0 && p < 9223372036854775808.0))
        return 0;
    const long long i = (long long)p;
    return i < 0 ? 0 : (i > hi ? hi : (int)i);
}

// Interpolation axis of an FP64 index: base truncate(idx + off) clamped to
// [0, size-2] and the unclamped FP64 weight of the upper node.
struct AxD {
    int b;
    double w;
};

__device__ __forceinline__ AxD axd(double idx, double off, int size)
{
    const double p = idx + off;
    const int b = trunc_clamp(p, size - 2);
    return {b, p - (double)b};
}


// u, v, w (FP32 trilinear) of pair (s0, s1) at weights
// (gt, ft) at an FP64 position (rk2_w_advection.pyx:124-150), through the
// stencil cache.
template <typename T, int LAYOUT, bool RING_OK>
__device__ __forceinline__ void rk2_uvw(const opx_window &win, const opx_grid &g, int j, int s0, int s1, float gt,
                                        float ft, double zi, double yi, double xi, StencilCache<T, LAYOUT> &c,
                                        float &u, float &v, double &w)
{
    const int NZ = g.NZ, NY = g.NY, NX = g.NX;
    const AxD cz = axd(zi, 0.0, NZ), cy = axd(yi, 0.0, NY), cx = axd(xi, 0.0, NX);
    const AxD ix = axd(xi, -0.5, NX - 1), iy = axd(yi, -0.5, NY - 1), oz = axd(zi, 0.5, NZ + 1);
    float vu[8], vv[8], vw[8];
    blend_uvw<T, LAYOUT, RING_OK>(win, g, j, s0, s1, gt, ft, cz.b, cy.b, cx.b, ix.b, iy.b, oz.b, c, vu, vv, vw);
    u = tri32(vu, (float)cz.w, (float)cy.w, (float)ix.w);
    v = tri32(vv, (float)cz.w, (float)iy.w, (float)cx.w);
    w = (double)tri32(vw, (float)oz.w, (float)cy.w, (float)cx.w);
}

// h and zeta (FP64 weights; zeta blended in FP32) where `hz` is set and,
// with DXDY, dx and dy (FP32) at an FP64 horizontal position, through the 2-D
// corner cache.
template <int LAYOUT, bool RING_OK, bool DXDY, typename Cache>
__device__ __forceinline__ void rk2_2d(const opx_window &win, int j, int s0, int s1, float gt, float ft, int NY,
                                       int NX, double yi, double xi, Cache &c, bool hz, double &h, double &zeta,
                                       float &dxv, float &dyv)
{
    const AxD cy = axd(yi, 0.0, NY), cx = axd(xi, 0.0, NX);
    sample2<LAYOUT, RING_OK>(win, j, s0, s1, NX, cy.b, cx.b, c.c2, c.s0_2d);
    if (hz)
        eval2(c.c2, gt, ft, cy.w, cx.w, h, zeta);
    if constexpr (DXDY) {
        dxv = bil32(c.c2.dx, (float)cy.w, (float)cx.w);
        dyv = bil32(c.c2.dy, (float)cy.w, (float)cx.w);
    }
}

// One fused RK2 step of a validated particle in registers: stage 1 at
// (x_n, t), stage 2 at (x*, t + alpha dt), update and zidx at t + dt.  Stage 2
// and the update usually sample the cells (and pair) of stage 1, so they blend
// its cached nodes at their own time weights.  Without RING_OK a packed LAYOUT
// reads only its packed windows and returns false, leaving the particle
// unchanged, if one of the step's pairs is not packed.
template <typename T, int LAYOUT, bool RING_OK>
__device__ __forceinline__ bool rk2_step(PState &p, StencilCache<T, LAYOUT> &c, const Step &rs, const opx_grid &g,
                                         const opx_window &win, double dt, double alpha, const double *Cw)
{
    const int NZ = g.NZ, NY = g.NY, NX = g.NX;
    const double zi = p.zi, yi = p.yi, xi = p.xi;
    const double hc = g.hc;
    const int j = packed_index<LAYOUT>(win, rs.s0, rs.s1), jm = packed_index<LAYOUT>(win, rs.s0m, rs.s1m);
    const int jp = packed_index<LAYOUT>(win, rs.s0p, rs.s1p);
    if (LAYOUT != OPX_LAYOUT_RING && !RING_OK && (j < 0 || jm < 0 || jp < 0))
        return false;

    // ---- stage 1 (rk2_w_advection.pyx:31-153): z at x_n (kept from the last
    // step, else compute_z) and k1 = (u/dx, v/dy, w)
    double h0, zeta0;
    float dxv, dyv;
    double z = p.z;
    const bool need_z = __builtin_isnan(z);
    rk2_2d<LAYOUT, RING_OK, true>(win, j, rs.s0, rs.s1, rs.gt, rs.ft, NY, NX, yi, xi, c, need_z, h0, zeta0, dxv,
                                  dyv);
    const AxD oz = axd(zi, 0.5, NZ + 1); // C = Cs_w at w-levels, offset +0.5
    if (need_z) {
        const double Cv = (1.0 - oz.w) * Cw[oz.b] + oz.w * Cw[oz.b + 1];
        const double sigma = (zi + 0.5) * inv_nz_here(Cw) - 1.0;
        const double S0 = div64(hc * sigma + h0 * Cv, hc + h0);
        z = zeta0 + (zeta0 + h0) * S0;
    }
    float u1, v1;
    double w1;
    rk2_uvw<T, LAYOUT, RING_OK>(win, g, j, rs.s0, rs.s1, rs.gt, rs.ft, zi, yi, xi, c, u1, v1, w1);
    const double kx1 = (double)div_nr(u1, dxv), ky1 = (double)div_nr(v1, dyv);

    // ---- stage 2 (:155-275) at x* = x_n + alpha dt k1, time t + alpha dt
    const double adt = alpha * dt;
    const double zs = z + adt * w1, ys = yi + adt * ky1, xs = xi + adt * kx1;
    double hs, zetas;
    float dxs, dys;
    rk2_2d<LAYOUT, RING_OK, true>(win, jm, rs.s0m, rs.s1m, rs.gtm, rs.ftm, NY, NX, ys, xs, c, true, hs, zetas, dxs,
                                  dys);
    const double zis = zidx_from_z(zs, hs, zetas, hc, NZ, Cw, oz.b);
    float u2, v2;
    double w2;
    rk2_uvw<T, LAYOUT, RING_OK>(win, g, jm, rs.s0m, rs.s1m, rs.gtm, rs.ftm, zis, ys, xs, c, u2, v2, w2);
    const double kx2 = (double)div_nr(u2, dxs), ky2 = (double)div_nr(v2, dys);

    // ---- update (:278-354): b1 = 1/(2 alpha) on stage 1, b2 = 1 - b1 on stage 2
    const double b1 = 1.0 / (2.0 * alpha), b2 = 1.0 - b1;
    const double zn = z + b1 * dt * w1 + b2 * dt * w2;
    const double yn = yi + b1 * dt * ky1 + b2 * dt * ky2;
    const double xn = xi + b1 * dt * kx1 + b2 * dt * kx2;
    const double yr = round32(yn), xr = round32(xn);
    double zr = __builtin_nan("");
    if (__builtin_isfinite(yr) && __builtin_isfinite(xr)) {
        double h1, zeta1;
        float unused0, unused1;
        rk2_2d<LAYOUT, RING_OK, false>(win, jp, rs.s0p, rs.s1p, rs.gt1, rs.ft1, NY, NX, yr, xr, c, true, h1, zeta1,
                                       unused0, unused1);
        const int k0 = (int)floor(zis + 0.5);
        zr = round32(zidx_from_z(zn, h1, zeta1, hc, NZ, Cw, k0));
    }
    p.zi = zr;
    p.yi = yr;
    p.xi = xr;
    p.z = zn;
    return true;
}

// Fused RK2 steps, one particle per thread (launch shape as k_ab3_fused); the
// one-step form (MULTI false) reads the rings for pairs that are not packed.
template <typename T, int LAYOUT, int BLOCK, bool IND, bool MULTI>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(MULTI ? 3 : 1)))
    k_rk2_fused(opx_split_state s, opx_grid g, opx_window win_arg, const opx_window *__restrict__ win_dev,
                const opx_release_step *__restrict__ table, int tstride, int nsteps, double dt, double alpha,
                const int32_t *__restrict__ n_active, int32_t *err)
{
    const opx_window win = IND ? *win_dev : win_arg;
    __shared__ double Cw[kLevels];
    stage_levels(Cw, win.C, g.NZ, threadIdx.x, BLOCK);
    __syncthreads();

    int64_t n = s.n;
    if (n_active)
        n = min(n, (int64_t)*n_active);
    const int64_t chunks = (n + BLOCK - 1) / BLOCK;
    const uint32_t xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
    const int64_t c = chunks * xcd / kXcds + lb, c1 = chunks * (xcd + 1) / kXcds;
    if (c < c1)
        fused_run<T, LAYOUT, kRK2, !MULTI>(c * BLOCK + threadIdx.x, n, s, g, win, table, tstride, MULTI ? nsteps : 1,
                                           dt, alpha, Cw, err);
}

// ---------------------------------------------------------------- re-binning

__device__ __forceinline__ uint32_t spatial_key(uint32_t x, uint32_t y, uint32_t z, int xbits, int ybits,
                                                int zbits, int xlow)
{
    const uint32_t xl = x & ((1u << xlow) - 1u);
    const uint32_t xh = x >> xlow;
    const int xhb = xbits - xlow;
    const int levels = max(xhb, max(ybits, zbits));
    uint32_t key = 0;
    int pos = 0;
    for (int b = 0; b < levels; ++b) {
        if (b < xhb)
            key |= ((xh >> b) & 1u) << pos++;
        if (b < ybits)
            key |= ((y >> b) & 1u) << pos++;
        if (b < zbits)
            key |= ((z >> b) & 1u) << pos++;
    }
    return (key << xlow) | xl;
}

// Sort key per particle: spatial key (active), inactive_key (parked) or
// inactive_key + 1 (free: retired release).
__global__ void k_keys(opx_split_state s, opx_grid g, int xlow, const opx_release_clock *__restrict__ clocks,
                       int nrel, uint32_t inactive_key, uint32_t *keys, uint32_t *vals)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n)
        return;
    const uint32_t cell = (uint32_t)s.cell[i];
    uint32_t key = inactive_key;
    bool live = true, parked = false;
    if (clocks) {
        const int r = s.rel ? (int)s.rel[i] : 0;
        if (r < nrel) {
            const int32_t left = clocks[r].steps_left;
            live = left >= 0;
            parked = left == 0;
        }
    }
    if (!live)
        key = inactive_key + 1;
    else if (!parked && !(s.status[i] & OPX_INACTIVE)) {
        uint32_t x = cell & ((1u << g.xbits) - 1);
        uint32_t y = (cell >> g.xbits) & ((1u << g.ybits) - 1);
        uint32_t z = cell >> (g.xbits + g.ybits);
        key = spatial_key(x, y, z, g.xbits, g.ybits, g.zbits, xlow);
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
}

// First index of the sorted keys holding a key >= k (binary search).
__device__ int64_t lower_bound_key(const uint32_t *keys, int64_t n, uint32_t k)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Active count = first parked key; live count = first free key (one thread).
__global__ void k_count_active(const uint32_t *keys, int64_t n, uint32_t inactive_key, int32_t *n_active,
                               int32_t *n_live)
{
    *n_active = (int32_t)lower_bound_key(keys, n, inactive_key);
    if (n_live)
        *n_live = (int32_t)lower_bound_key(keys, n, inactive_key + 1);
}

__global__ void k_extract_release(opx_split_state s, opx_grid g, uint32_t r, uint32_t pid_base, int64_t m,
                                  double *zidx, double *yidx, double *xidx, uint8_t *status)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.n || (uint32_t)s.rel[i] != r)
        return;
    const uint32_t j = (uint32_t)s.pid[i] - pid_base;
    if ((int64_t)j >= m)
        return;
    const uint32_t cell = (uint32_t)s.cell[i];
    Axis z = decode_axis(cell, g.xbits + g.ybits, g.zbits);
    Axis y = decode_axis(cell, g.xbits, g.ybits);
    Axis x = decode_axis(cell, 0, g.xbits);
    z.q = s.fz[i];
    y.q = s.fy[i];
    x.q = s.fx[i];
    const double nan = __builtin_nan("");
    zidx[j] = z.ok ? z.pos() : nan;
    yidx[j] = y.ok ? y.pos() : nan;
    xidx[j] = x.ok ? x.pos() : nan;
    if (status)
        status[j] = s.status[i];
}

__global__ void k_permute(opx_split_state src, opx_split_state dst, const uint32_t *__restrict__ perm)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= src.n)
        return;
    const uint32_t j = perm[i];
    dst.status[i] = src.status[j];
    dst.cell[i] = src.cell[j];
    dst.fz[i] = src.fz[j];
    dst.fy[i] = src.fy[j];
    dst.fx[i] = src.fx[j];
    dst.hz1[i] = src.hz1[j];
    dst.hy1[i] = src.hy1[j];
    dst.hx1[i] = src.hx1[j];
    dst.hz2[i] = src.hz2[j];
    dst.hy2[i] = src.hy2[j];
    dst.hx2[i] = src.hx2[j];
    if (src.pid)
        dst.pid[i] = src.pid[j];
    if (src.rel)
        dst.rel[i] = src.rel[j];
    if (src.z)
        dst.z[i] = src.z[j];
    if (src.rho_t)
        dst.rho_t[i] = src.rho_t[j];
}

__global__ void k_zero_int(int32_t *p) { *p = 0; }

// ---------------------------------------------------------------- host checks

int check_state(const opx_split_state *s, const char *what, bool need_hist = true)
{
    OPX_REQUIRE(s != nullptr, "%s: null state", what);
    OPX_REQUIRE(s->n >= 0 && s->n <= INT32_MAX, "%s: particle count %lld out of range", what, (long long)s->n);
    if (s->n == 0)
        return OPX_OK;
    OPX_REQUIRE(s->status && s->cell && s->fz && s->fy && s->fx, "%s: null status/cell/offset arrays", what);
    OPX_REQUIRE(!need_hist || (s->hz1 && s->hy1 && s->hx1 && s->hz2 && s->hy2 && s->hx2), "%s: null history arrays",
                what);
    return OPX_OK;
}

// The step entry points also keep each particle's z in the state.
int check_step_state(const opx_split_state *s, const char *what, bool need_hist = true)
{
    int rc;
    if ((rc = check_state(s, what, need_hist))) return rc;
    OPX_REQUIRE(s->n == 0 || s->z, "%s: null z array", what);
    return OPX_OK;
}

int check_grid(const opx_grid *g)
{
    OPX_REQUIRE(g != nullptr, "null grid");
    OPX_REQUIRE(g->NZ >= 2 && g->NY >= 3 && g->NX >= 3, "grid %dx%dx%d too small", g->NZ, g->NY, g->NX);
    OPX_REQUIRE(g->NZ + 1 <= kMaxC, "NZ = %d above the %d levels the fused step keeps in LDS", g->NZ, kMaxC - 1);
    OPX_REQUIRE((int64_t)(g->NZ + 1) * g->NY < (1 << 24) && g->NX < (1 << 24),
                "grid %dx%dx%d: the fused steps index nodes with 24-bit multiplies ((NZ + 1) NY and NX below 2^24)",
                g->NZ, g->NY, g->NX);
    OPX_REQUIRE(g->xbits > 0 && g->ybits > 0 && g->zbits > 0 && g->xbits + g->ybits + g->zbits <= 31 &&
                    (1 << g->xbits) >= g->NX + 5 && (1 << g->ybits) >= g->NY + 5 && (1 << g->zbits) >= g->NZ + 5,
                "cell bit widths (%d,%d,%d) must hold N+5 values per axis in 31 bits", g->zbits, g->ybits,
                g->xbits);
    // largest 3-D field (w: (NZ+1) NY NX) must be indexable with 32 bits
    OPX_REQUIRE((int64_t)(g->NZ + 1) * g->NY * g->NX < ((int64_t)1 << 32), "grid too large for 32-bit offsets");
    // Stored cells span [-kBias, N+1]; a position beyond them is clamped to that
    // range, which the domain check still flags only while every bound lies in
    // (-kBias, N+1] on its axis.
    const double lo[3] = {g->zidx_min, g->yidx_min, g->xidx_min};
    const double hi[3] = {g->zidx_max, g->yidx_max, g->xidx_max};
    const int n[3] = {g->NZ, g->NY, g->NX};
    for (int d = 0; d < 3; ++d)
        OPX_REQUIRE(lo[d] > -kBias && lo[d] <= n[d] + 1 && hi[d] > -kBias && hi[d] <= n[d] + 1,
                    "index bounds [%g, %g] of axis %d lie outside (%d, %d], the range the fused state holds",
                    lo[d], hi[d], d, -kBias, n[d] + 1);
    return OPX_OK;
}

// Levels of the window's third 3-D field (w: NZ + 1, density: NZ).
inline int w_levels(const opx_window *w, const opx_grid *g) { return w->wlevels ? w->wlevels : g->NZ + 1; }

// iso: the isopycnal-following step, whose third field is the density (NZ
// levels, FP32 storage); otherwise w (NZ + 1 levels).
int check_window(const opx_window *w, const opx_grid *g, bool iso = false)
{
    OPX_REQUIRE(w != nullptr, "null window");
    OPX_REQUIRE(w_levels(w, g) == (iso ? g->NZ : g->NZ + 1), "window's third field has %d levels; the %s step "
                "needs %d", w_levels(w, g), iso ? "isopycnal-following" : "w-advection", iso ? g->NZ : g->NZ + 1);
    OPX_REQUIRE(!iso || w->dtype3 == OPX_F32, "the isopycnal-following step needs FP32 storage (got %d)",
                w->dtype3);
    OPX_REQUIRE(w->u && w->v && w->w && w->zeta && w->hdd && w->C, "window has a null field pointer");
    OPX_REQUIRE(w->dtype3 == OPX_F32 || w->dtype3 == OPX_BF16, "fused step supports F32/BF16 u,v,w (got %d)",
                w->dtype3);
    OPX_REQUIRE(w->nslots >= 2 && w->nslots <= kMaxSlots, "window needs 2 to %d slots (got %d)", kMaxSlots,
                w->nslots);
    OPX_REQUIRE(w->npacked == 0 || w->npacked == w->nslots, "npacked %d must be 0 or nslots (%d)", w->npacked,
                w->nslots);
    if (w->npacked > 0) {
        OPX_REQUIRE(w->layout >= OPX_LAYOUT_PAIR && w->layout <= OPX_LAYOUT_YPAIR,
                    "packed windows need layout PAIR, QUAD, PAIR_LDS, OCT or YPAIR (got %d)", w->layout);
        OPX_REQUIRE(w->pu && w->pv && w->pw && w->p2 && w->dy, "pair ring has a null pointer");
        uint64_t bu, bv, bw, b2;
        int rc;
        if ((rc = opx_packed_bytes(g, w->dtype3, w->layout, &bu, &bv, &bw, &b2))) return rc;
        OPX_REQUIRE(w->pu_stride >= bu && w->pv_stride >= bv && w->pw_stride >= bw && w->p2_stride >= b2,
                    "pair ring strides below one packed window");
    }
    OPX_REQUIRE(w->u_stride >= (int64_t)g->NZ * g->NY * (g->NX - 1) &&
                    w->v_stride >= (int64_t)g->NZ * (g->NY - 1) * g->NX &&
                    w->w_stride >= (int64_t)w_levels(w, g) * g->NY * g->NX &&
                    w->zeta_stride >= (int64_t)g->NY * g->NX,
                "window slot strides smaller than the field sizes");
    return OPX_OK;
}

inline unsigned grid_xcd(int64_t n, int block)
{
    int64_t G = (n + block - 1) / block;
    G = (G + kXcds - 1) / kXcds * kXcds;
    return (unsigned)G;
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t sort_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 32);
    return bytes;
}

int launch_fused(const opx_split_state *s, const opx_grid *g, const opx_window &win_arg, const opx_window *win_dev,
                 int layout, int dtype3, const opx_release_step *table, int tstride, int nsteps, double dt,
                 const int32_t *n_active, int block, int32_t *err, void *stream)
{
    hipStream_t st = as_stream(stream);
    const bool ind = win_dev != nullptr;
    if (layout == OPX_LAYOUT_PAIR_LDS && nsteps > 1) // several steps per launch read PAIR windows directly
        layout = OPX_LAYOUT_PAIR;
    if (layout == OPX_LAYOUT_PAIR_LDS) {
        // one wave per workgroup, stencils staged in LDS (packed PAIR windows)
        const dim3 grid(grid_xcd(s->n, 64)), blk(64);
        if (dtype3 == OPX_F32) {
            if (ind)
                hipLaunchKernelGGL((k_ab3_tile<float, true>), grid, blk, 0, st, *s, *g, win_arg, win_dev, table, dt,
                                   n_active);
            else
                hipLaunchKernelGGL((k_ab3_tile<float, false>), grid, blk, 0, st, *s, *g, win_arg, win_dev, table, dt,
                                   n_active);
        } else {
            if (ind)
                hipLaunchKernelGGL((k_ab3_tile<uint16_t, true>), grid, blk, 0, st, *s, *g, win_arg, win_dev, table,
                                   dt, n_active);
            else
                hipLaunchKernelGGL((k_ab3_tile<uint16_t, false>), grid, blk, 0, st, *s, *g, win_arg, win_dev, table,
                                   dt, n_active);
        }
        return check_launch("opx_ab3_fused");
    }
    const dim3 grid(grid_xcd(s->n, block)), blk(block);
#define OPX_FUSED_LAUNCH(T, L, B, I)                                                                       \
    if (block == B && layout == L && ind == I) {                                                          \
        if (nsteps > 1)                                                                                   \
            hipLaunchKernelGGL((k_ab3_fused<T, L, B, I, true>), grid, blk, 0, st, *s, *g, win_arg, win_dev, \
                               table, tstride, nsteps, dt, n_active, err);                                \
        else                                                                                              \
            hipLaunchKernelGGL((k_ab3_fused<T, L, B, I, false>), grid, blk, 0, st, *s, *g, win_arg, win_dev, \
                               table, tstride, nsteps, dt, n_active, err);                                \
    }
#define OPX_FUSED_LAUNCH_L(T, L)                                                                           \
    OPX_FUSED_LAUNCH(T, L, 128, false) OPX_FUSED_LAUNCH(T, L, 256, false) OPX_FUSED_LAUNCH(T, L, 128, true) \
        OPX_FUSED_LAUNCH(T, L, 256, true)
    if (dtype3 == OPX_F32) {
        OPX_FUSED_LAUNCH_L(float, OPX_LAYOUT_RING)
        OPX_FUSED_LAUNCH_L(float, OPX_LAYOUT_PAIR)
        OPX_FUSED_LAUNCH_L(float, OPX_LAYOUT_QUAD)
        OPX_FUSED_LAUNCH_L(float, OPX_LAYOUT_OCT)
        OPX_FUSED_LAUNCH_L(float, OPX_LAYOUT_YPAIR)
    } else {
        OPX_FUSED_LAUNCH_L(uint16_t, OPX_LAYOUT_RING)
        OPX_FUSED_LAUNCH_L(uint16_t, OPX_LAYOUT_PAIR)
        OPX_FUSED_LAUNCH_L(uint16_t, OPX_LAYOUT_QUAD)
        OPX_FUSED_LAUNCH_L(uint16_t, OPX_LAYOUT_OCT)
        OPX_FUSED_LAUNCH_L(uint16_t, OPX_LAYOUT_YPAIR)
    }
#undef OPX_FUSED_LAUNCH_L
#undef OPX_FUSED_LAUNCH
    return check_launch("opx_ab3_fused");
}

int launch_iso(const opx_split_state *s, const opx_grid *g, const opx_window &win_arg, const opx_window *win_dev,
               int layout, const opx_release_step *table, int tstride, int nsteps, double dt, IsoRate iso,
               const int32_t *n_active, int block, int32_t *err, void *stream)
{
    if (layout == OPX_LAYOUT_PAIR_LDS) // the isopycnal step reads PAIR windows directly
        layout = OPX_LAYOUT_PAIR;
    const dim3 grid(grid_xcd(s->n, block)), blk(block);
    hipStream_t st = as_stream(stream);
    const bool ind = win_dev != nullptr;
#define OPX_ISO_LAUNCH(L, B, I)                                                                              \
    if (block == B && layout == L && ind == I)                                                              \
        hipLaunchKernelGGL((k_ab3_iso_fused<L, B, I>), grid, blk, 0, st, *s, *g, win_arg, win_dev, table, tstride, \
                           nsteps, dt, iso, n_active, err);
#define OPX_ISO_LAUNCH_L(L)                                                                                  \
    OPX_ISO_LAUNCH(L, 128, false) OPX_ISO_LAUNCH(L, 256, false) OPX_ISO_LAUNCH(L, 128, true)                 \
        OPX_ISO_LAUNCH(L, 256, true)
    OPX_ISO_LAUNCH_L(OPX_LAYOUT_RING)
    OPX_ISO_LAUNCH_L(OPX_LAYOUT_PAIR)
    OPX_ISO_LAUNCH_L(OPX_LAYOUT_QUAD)
    OPX_ISO_LAUNCH_L(OPX_LAYOUT_OCT)
    OPX_ISO_LAUNCH_L(OPX_LAYOUT_YPAIR)
#undef OPX_ISO_LAUNCH_L
#undef OPX_ISO_LAUNCH
    return check_launch("opx_ab3_iso_fused");
}

int launch_rk2(const opx_split_state *s, const opx_grid *g, const opx_window &win_arg, const opx_window *win_dev,
               int layout, int dtype3, const opx_release_step *table, int tstride, int nsteps, double dt, double alpha,
               const int32_t *n_active, int block, int32_t *err, void *stream)
{
    if (layout == OPX_LAYOUT_PAIR_LDS) // the RK2 step reads PAIR windows directly
        layout = OPX_LAYOUT_PAIR;
    const dim3 grid(grid_xcd(s->n, block)), blk(block);
    hipStream_t st = as_stream(stream);
    const bool ind = win_dev != nullptr;
#define OPX_RK2_LAUNCH(T, L, B, I)                                                                         \
    if (block == B && layout == L && ind == I) {                                                          \
        if (nsteps > 1)                                                                                   \
            hipLaunchKernelGGL((k_rk2_fused<T, L, B, I, true>), grid, blk, 0, st, *s, *g, win_arg, win_dev, \
                               table, tstride, nsteps, dt, alpha, n_active, err);                         \
        else                                                                                              \
            hipLaunchKernelGGL((k_rk2_fused<T, L, B, I, false>), grid, blk, 0, st, *s, *g, win_arg, win_dev, \
                               table, tstride, nsteps, dt, alpha, n_active, err);                         \
    }
#define OPX_RK2_LAUNCH_L(T, L)                                                                             \
    OPX_RK2_LAUNCH(T, L, 128, false) OPX_RK2_LAUNCH(T, L, 256, false) OPX_RK2_LAUNCH(T, L, 128, true)     \
        OPX_RK2_LAUNCH(T, L, 256, true)
    if (dtype3 == OPX_F32) {
        OPX_RK2_LAUNCH_L(float, OPX_LAYOUT_RING)
        OPX_RK2_LAUNCH_L(float, OPX_LAYOUT_PAIR)
        OPX_RK2_LAUNCH_L(float, OPX_LAYOUT_QUAD)
        OPX_RK2_LAUNCH_L(float, OPX_LAYOUT_OCT)
        OPX_RK2_LAUNCH_L(float, OPX_LAYOUT_YPAIR)
    } else {
        OPX_RK2_LAUNCH_L(uint16_t, OPX_LAYOUT_RING)
        OPX_RK2_LAUNCH_L(uint16_t, OPX_LAYOUT_PAIR)
        OPX_RK2_LAUNCH_L(uint16_t, OPX_LAYOUT_QUAD)
        OPX_RK2_LAUNCH_L(uint16_t, OPX_LAYOUT_OCT)
        OPX_RK2_LAUNCH_L(uint16_t, OPX_LAYOUT_YPAIR)
    }
#undef OPX_RK2_LAUNCH_L
#undef OPX_RK2_LAUNCH
    return check_launch("opx_rk2_fused");
}

} // namespace

namespace {

// ---- re-bin body shared by opx_rebin (caller's workspace) and
// opx_rebin_noalloc (the static scratch buffer below)

int check_rebin_args(const opx_split_state *src, const opx_split_state *dst, const opx_grid *g, int32_t xlow,
                     const opx_release_clock *clocks, int32_t nrel, const int32_t *n_active)
{
    int rc;
    if ((rc = check_state(src, "opx_rebin src"))) return rc;
    if ((rc = check_state(dst, "opx_rebin dst"))) return rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(dst->n == src->n, "src/dst sizes differ");
    OPX_REQUIRE(n_active != nullptr, "null n_active");
    OPX_REQUIRE(xlow >= 0 && xlow <= g->xbits, "xlow %d out of [0, %d]", xlow, g->xbits);
    OPX_REQUIRE(clocks == nullptr || nrel >= 0, "negative release count");
    OPX_REQUIRE((src->pid == nullptr) == (dst->pid == nullptr) && (src->rel == nullptr) == (dst->rel == nullptr) &&
                    (src->z == nullptr) == (dst->z == nullptr) && (src->rho_t == nullptr) == (dst->rho_t == nullptr),
                "src/dst optional arrays differ");
    return OPX_OK;
}

// `work` holds at least opx_rebin_workspace_bytes(src->n) bytes (checked by the callers).
int rebin(const opx_split_state *src, const opx_split_state *dst, const opx_grid *g, int32_t xlow,
          const opx_release_clock *clocks, int32_t nrel, void *work, uint64_t work_bytes, int32_t *n_active,
          int32_t *n_live, void *stream)
{
    int rc;
    const int64_t n = src->n;
    hipStream_t st = as_stream(stream);
    if (n == 0) {
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, n_active);
        if (n_live)
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, n_live);
        return check_launch("opx_rebin");
    }
    const size_t arr = align256((size_t)n * sizeof(uint32_t));
    char *base = static_cast<char *>(work);
    uint32_t *k_in = reinterpret_cast<uint32_t *>(base);
    uint32_t *k_out = reinterpret_cast<uint32_t *>(base + arr);
    uint32_t *v_in = reinterpret_cast<uint32_t *>(base + 2 * arr);
    uint32_t *v_out = reinterpret_cast<uint32_t *>(base + 3 * arr);
    void *temp = base + 4 * arr;
    size_t temp_bytes = (size_t)(work_bytes - 4 * arr);
    const int key_bits = g->xbits + g->ybits + g->zbits;
    const uint32_t inactive_key = 1u << key_bits;
    hipLaunchKernelGGL(k_keys, dim3(grid_for(n)), dim3(kBlock), 0, st, *src, *g, xlow, clocks, (int)nrel,
                       inactive_key, k_in, v_in);
    if ((rc = check_launch("opx_rebin keys"))) return rc;
    // keys span [0, inactive_key + 1] = key_bits + 1 bits
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k_in, k_out, v_in, v_out, (int)n, 0,
                                                      key_bits + 1, st);
    if (e != hipSuccess) {
        set_error("opx_rebin: radix sort failed: %s", hipGetErrorString(e));
        return OPX_ERR_HIP;
    }
    hipLaunchKernelGGL(k_count_active, dim3(1), dim3(1), 0, st, k_out, n, inactive_key, n_active, n_live);
    hipLaunchKernelGGL(k_permute, dim3(grid_for(n)), dim3(kBlock), 0, st, *src, *dst, v_out);
    return check_launch("opx_rebin permute");
}

// ---- the allocation-free re-bin's scratch: one file-scope device array, so
// opx_rebin_noalloc needs no allocation.  Each device holds its own copy at
// its own address, found once per device with hipGetSymbolAddress; host code
// uses that address, never the array's name (a host-side stand-in).  Every
// call on a device shares the device's copy: calls that overlap on one device
// race.  Nothing in it needs to start at zero (every key, value and the sort's
// temporary storage are written before they are read).
constexpr uint64_t kNoAllocScratchBytes = uint64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxScratchDevices] = {};

int noalloc_scratch(void **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("opx_rebin_noalloc: hipGetDevice failed: %s", hipGetErrorString(e));
        return OPX_ERR_HIP;
    }
    OPX_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "opx_rebin_noalloc: device %d beyond the %d supported", dev,
                kMaxScratchDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("opx_rebin_noalloc: hipGetSymbolAddress of the scratch buffer failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return OPX_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return OPX_OK;
}

} // namespace
} // namespace opx

using namespace opx;

extern "C" {

int opx_clock_advance_steps(opx_release_clock *clocks, opx_release_step *table, int32_t table_stride,
                            int32_t nsteps, int32_t nrel, const int32_t *nrel_dev, const double *time_array,
                            int32_t ntime, int32_t nslots, int32_t global_step, int32_t *step_dev, double dt,
                            double alpha, int32_t *err_flag, void *stream)
{
    OPX_REQUIRE(nrel >= 0, "negative release count");
    OPX_REQUIRE(alpha >= 0.0, "alpha must be >= 0");
    OPX_REQUIRE(nsteps >= 1 && nsteps <= kMaxLaunchSteps, "nsteps %d not in [1, %d]", nsteps, kMaxLaunchSteps);
    OPX_REQUIRE(nsteps == 1 || table_stride >= nrel, "table_stride %d below the release count %d", table_stride,
                nrel);
    if (nrel == 0 && step_dev == nullptr)
        return OPX_OK;
    OPX_REQUIRE(nrel == 0 || (clocks && table), "null clock/table");
    OPX_REQUIRE(time_array != nullptr, "null time array");
    OPX_REQUIRE(ntime >= 2, "time array needs at least 2 entries");
    OPX_REQUIRE(nslots >= 2 && nslots <= kMaxSlots, "nslots must be in [2, %d]", kMaxSlots);
    hipStream_t st = as_stream(stream);
    if (nrel > 0)
        hipLaunchKernelGGL(k_clock, dim3((nrel + 127) / 128), dim3(128), 0, st, clocks, table, table_stride, nsteps,
                           nrel, nrel_dev, time_array, ntime, nslots, global_step, step_dev, dt, alpha, err_flag);
    if (step_dev)
        hipLaunchKernelGGL(k_step_incr, dim3(1), dim3(1), 0, st, step_dev, nsteps);
    return check_launch("opx_clock_advance");
}

int opx_clock_advance(opx_release_clock *clocks, opx_release_step *table, int32_t nrel, const int32_t *nrel_dev,
                      const double *time_array, int32_t ntime, int32_t nslots, int32_t global_step,
                      int32_t *step_dev, double dt, double alpha, int32_t *err_flag, void *stream)
{
    return opx_clock_advance_steps(clocks, table, 0, 1, nrel, nrel_dev, time_array, ntime, nslots, global_step,
                                   step_dev, dt, alpha, err_flag, stream);
}

int opx_ab3_fused_steps(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                        const opx_release_step *table, int32_t table_stride, int32_t nsteps, double dt,
                        const int32_t *n_active, int32_t block, int32_t *err_flag, void *stream)
{
    int rc;
    if ((rc = check_step_state(s, "opx_ab3_fused"))) return rc;
    if ((rc = check_grid(g))) return rc;
    if ((rc = check_window(win, g))) return rc;
    OPX_REQUIRE(table != nullptr, "null release table");
    OPX_REQUIRE(nsteps >= 1 && nsteps <= kMaxLaunchSteps, "nsteps %d not in [1, %d]", nsteps, kMaxLaunchSteps);
    OPX_REQUIRE(nsteps == 1 || table_stride >= 1, "table_stride must be >= 1 for several steps");
    if (s->n == 0)
        return OPX_OK;
    if (block == 0)
        block = 128;
    OPX_REQUIRE(block == 128 || block == 256, "block size %d not in {128, 256}", block);
    const int layout = win->npacked > 0 ? win->layout : OPX_LAYOUT_RING;
    return launch_fused(s, g, *win, nullptr, layout, win->dtype3, table, table_stride, nsteps, dt, n_active, block,
                        err_flag, stream);
}

int opx_ab3_fused(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                  const opx_release_step *table, double dt, const int32_t *n_active, int32_t block, void *stream)
{
    return opx_ab3_fused_steps(s, g, win, table, 0, 1, dt, n_active, block, nullptr, stream);
}

int opx_ab3_fused_steps_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                            int32_t layout, const opx_release_step *table, int32_t table_stride, int32_t nsteps,
                            double dt, const int32_t *n_active, int32_t block, int32_t *err_flag, void *stream)
{
    int rc;
    if ((rc = check_step_state(s, "opx_ab3_fused_dev"))) return rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(win_dev != nullptr && table != nullptr, "null window or release table");
    OPX_REQUIRE(dtype3 == OPX_F32 || dtype3 == OPX_BF16, "dtype3 must be F32 or BF16");
    OPX_REQUIRE(layout >= OPX_LAYOUT_RING && layout <= OPX_LAYOUT_YPAIR, "unknown layout %d", layout);
    OPX_REQUIRE(nsteps >= 1 && nsteps <= kMaxLaunchSteps, "nsteps %d not in [1, %d]", nsteps, kMaxLaunchSteps);
    OPX_REQUIRE(nsteps == 1 || table_stride >= 1, "table_stride must be >= 1 for several steps");
    if (block == 0)
        block = 128;
    OPX_REQUIRE(block == 128 || block == 256, "block size %d not in {128, 256}", block);
    if (s->n == 0)
        return OPX_OK;
    opx_window unused{};
    return launch_fused(s, g, unused, win_dev, layout, dtype3, table, table_stride, nsteps, dt, n_active, block,
                        err_flag, stream);
}

int opx_ab3_fused_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                      int32_t layout, const opx_release_step *table, double dt, const int32_t *n_active, int32_t block,
                      void *stream)
{
    return opx_ab3_fused_steps_dev(s, g, win_dev, dtype3, layout, table, 0, 1, dt, n_active, block, nullptr,
                                   stream);
}

static int check_iso_args(const opx_split_state *s, const opx_release_step *table, int32_t table_stride,
                          int32_t nsteps, double w_max, double rho_scale, int32_t *block)
{
    OPX_REQUIRE(table != nullptr, "null release table");
    OPX_REQUIRE(nsteps >= 1 && nsteps <= kMaxLaunchSteps, "nsteps %d not in [1, %d]", nsteps, kMaxLaunchSteps);
    OPX_REQUIRE(nsteps == 1 || table_stride >= 1, "table_stride must be >= 1 for several steps");
    OPX_REQUIRE(s->n == 0 || s->rho_t != nullptr, "the isopycnal-following step needs rho_t");
    OPX_REQUIRE(std::isfinite(w_max) && std::isfinite(rho_scale) && rho_scale > 0.0,
                "w_max must be finite and rho_scale > 0 (got %g, %g)", w_max, rho_scale);
    if (*block == 0)
        *block = 128;
    OPX_REQUIRE(*block == 128 || *block == 256, "block size %d not in {128, 256}", *block);
    return OPX_OK;
}

int opx_ab3_iso_fused_steps(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                            const opx_release_step *table, int32_t table_stride, int32_t nsteps, double dt,
                            double w_max, double rho_scale, const int32_t *n_active, int32_t block,
                            int32_t *err_flag, void *stream)
{
    int rc;
    if ((rc = check_step_state(s, "opx_ab3_iso_fused"))) return rc;
    if ((rc = check_grid(g))) return rc;
    if ((rc = check_window(win, g, true))) return rc;
    if ((rc = check_iso_args(s, table, table_stride, nsteps, w_max, rho_scale, &block))) return rc;
    if (s->n == 0)
        return OPX_OK;
    const int layout = win->npacked > 0 ? win->layout : OPX_LAYOUT_RING;
    return launch_iso(s, g, *win, nullptr, layout, table, table_stride, nsteps, dt, IsoRate{w_max, rho_scale},
                      n_active, block, err_flag, stream);
}

int opx_ab3_iso_fused_steps_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev,
                                int32_t layout, const opx_release_step *table, int32_t table_stride, int32_t nsteps,
                                double dt, double w_max, double rho_scale, const int32_t *n_active, int32_t block,
                                int32_t *err_flag, void *stream)
{
    int rc;
    if ((rc = check_step_state(s, "opx_ab3_iso_fused_dev"))) return rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(win_dev != nullptr, "null window");
    OPX_REQUIRE(layout >= OPX_LAYOUT_RING && layout <= OPX_LAYOUT_YPAIR, "unknown layout %d", layout);
    if ((rc = check_iso_args(s, table, table_stride, nsteps, w_max, rho_scale, &block))) return rc;
    if (s->n == 0)
        return OPX_OK;
    opx_window unused{};
    return launch_iso(s, g, unused, win_dev, layout, table, table_stride, nsteps, dt, IsoRate{w_max, rho_scale},
                      n_active, block, err_flag, stream);
}

int opx_rk2_fused_steps(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                        const opx_release_step *table, int32_t table_stride, int32_t nsteps, double dt, double alpha,
                        const int32_t *n_active, int32_t block, int32_t *err_flag, void *stream)
{
    int rc;
    if ((rc = check_step_state(s, "opx_rk2_fused", false))) return rc;
    if ((rc = check_grid(g))) return rc;
    if ((rc = check_window(win, g))) return rc;
    OPX_REQUIRE(table != nullptr, "null release table");
    OPX_REQUIRE(alpha > 0.0, "alpha must be > 0");
    OPX_REQUIRE(nsteps >= 1 && nsteps <= kMaxLaunchSteps, "nsteps %d not in [1, %d]", nsteps, kMaxLaunchSteps);
    OPX_REQUIRE(nsteps == 1 || table_stride >= 1, "table_stride must be >= 1 for several steps");
    if (s->n == 0)
        return OPX_OK;
    if (block == 0)
        block = 128;
    OPX_REQUIRE(block == 128 || block == 256, "block size %d not in {128, 256}", block);
    const int layout = win->npacked > 0 ? win->layout : OPX_LAYOUT_RING;
    return launch_rk2(s, g, *win, nullptr, layout, win->dtype3, table, table_stride, nsteps, dt, alpha, n_active,
                      block, err_flag, stream);
}

int opx_rk2_fused(const opx_split_state *s, const opx_grid *g, const opx_window *win,
                  const opx_release_step *table, double dt, double alpha, const int32_t *n_active, int32_t block,
                  void *stream)
{
    return opx_rk2_fused_steps(s, g, win, table, 0, 1, dt, alpha, n_active, block, nullptr, stream);
}

int opx_rk2_fused_steps_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                            int32_t layout, const opx_release_step *table, int32_t table_stride, int32_t nsteps,
                            double dt, double alpha, const int32_t *n_active, int32_t block, int32_t *err_flag,
                            void *stream)
{
    int rc;
    if ((rc = check_step_state(s, "opx_rk2_fused_dev", false))) return rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(win_dev != nullptr && table != nullptr, "null window or release table");
    OPX_REQUIRE(dtype3 == OPX_F32 || dtype3 == OPX_BF16, "dtype3 must be F32 or BF16");
    OPX_REQUIRE(layout >= OPX_LAYOUT_RING && layout <= OPX_LAYOUT_YPAIR, "unknown layout %d", layout);
    OPX_REQUIRE(alpha > 0.0, "alpha must be > 0");
    OPX_REQUIRE(nsteps >= 1 && nsteps <= kMaxLaunchSteps, "nsteps %d not in [1, %d]", nsteps, kMaxLaunchSteps);
    OPX_REQUIRE(nsteps == 1 || table_stride >= 1, "table_stride must be >= 1 for several steps");
    if (block == 0)
        block = 128;
    OPX_REQUIRE(block == 128 || block == 256, "block size %d not in {128, 256}", block);
    if (s->n == 0)
        return OPX_OK;
    opx_window unused{};
    return launch_rk2(s, g, unused, win_dev, layout, dtype3, table, table_stride, nsteps, dt, alpha, n_active, block,
                      err_flag, stream);
}

int opx_rk2_fused_dev(const opx_split_state *s, const opx_grid *g, const opx_window *win_dev, int32_t dtype3,
                      int32_t layout, const opx_release_step *table, double dt, double alpha,
                      const int32_t *n_active, int32_t block, void *stream)
{
    return opx_rk2_fused_steps_dev(s, g, win_dev, dtype3, layout, table, 0, 1, dt, alpha, n_active, block, nullptr,
                                   stream);
}

static int elem_bytes(int dtype3, int layout)
{
    const int b = dtype3 == OPX_F32 ? 4 : 2;
    switch (layout) {
    case OPX_LAYOUT_PAIR: return 2 * b;
    case OPX_LAYOUT_QUAD: return 4 * b;
    case OPX_LAYOUT_OCT: return 8 * b;
    case OPX_LAYOUT_YPAIR: return 4 * b;
    default: return 0;
    }
}

int opx_packed_bytes(const opx_grid *g, int32_t dtype3, int32_t layout, uint64_t *bytes_u, uint64_t *bytes_v,
                     uint64_t *bytes_w, uint64_t *bytes_2d)
{
    int rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(dtype3 == OPX_F32 || dtype3 == OPX_BF16, "dtype3 must be F32 or BF16");
    if (layout == OPX_LAYOUT_PAIR_LDS) // same packed data as PAIR
        layout = OPX_LAYOUT_PAIR;
    OPX_REQUIRE(layout == OPX_LAYOUT_PAIR || layout == OPX_LAYOUT_QUAD || layout == OPX_LAYOUT_OCT ||
                    layout == OPX_LAYOUT_YPAIR,
                "layout must be PAIR, QUAD, OCT, YPAIR or PAIR_LDS");
    OPX_REQUIRE(bytes_u && bytes_v && bytes_w && bytes_2d, "null output");
    *bytes_2d = (uint64_t)16 * g->NY * g->NX;
    const uint64_t e = (uint64_t)elem_bytes(dtype3, layout);
    *bytes_u = e * (uint64_t)g->NZ * g->NY * (g->NX - 1);
    *bytes_v = e * (uint64_t)g->NZ * (g->NY - 1) * g->NX;
    *bytes_w = e * (uint64_t)(g->NZ + 1) * g->NY * g->NX;
    // the step kernels address a packed window with 32-bit byte offsets
    OPX_REQUIRE(*bytes_w <= kMaxWindowBytes && *bytes_u <= kMaxWindowBytes && *bytes_v <= kMaxWindowBytes,
                "a packed window of this grid in layout %d takes %llu bytes, above the 4 GiB of one packed window",
                layout, (unsigned long long)*bytes_w);
    return OPX_OK;
}

int opx_pack_window_rows(const opx_window *win, const opx_grid *g, int32_t layout, int32_t s0, int32_t s1,
                         int32_t y0, int32_t y1, void *pu, void *pv, void *pw, void *p2, void *stream)
{
    int rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(win && win->u && win->v && win->w, "null window");
    OPX_REQUIRE(win->dtype3 == OPX_F32 || win->dtype3 == OPX_BF16, "dtype3 must be F32 or BF16");
    if (layout == OPX_LAYOUT_PAIR_LDS) // same packed data as PAIR
        layout = OPX_LAYOUT_PAIR;
    OPX_REQUIRE(layout == OPX_LAYOUT_PAIR || layout == OPX_LAYOUT_QUAD || layout == OPX_LAYOUT_OCT ||
                    layout == OPX_LAYOUT_YPAIR,
                "layout must be PAIR, QUAD, OCT, YPAIR or PAIR_LDS");
    OPX_REQUIRE(s0 >= 0 && s0 < win->nslots && s1 >= 0 && s1 < win->nslots, "slots (%d, %d) outside [0, %d)", s0,
                s1, win->nslots);
    OPX_REQUIRE(0 <= y0 && y0 < y1 && y1 <= g->NY, "rows [%d, %d) outside [0, %d)", y0, y1, g->NY);
    OPX_REQUIRE(pu && pv && pw && p2, "null output window");
    OPX_REQUIRE(win->hdd && win->zeta, "window without hdd/zeta");
    uint64_t bu, bv, bw, b2;
    if ((rc = opx_packed_bytes(g, win->dtype3, layout, &bu, &bv, &bw, &b2))) return rc;
    hipStream_t st = as_stream(stream);
    const int64_t n2d = (int64_t)(y1 - y0) * g->NX;
    hipLaunchKernelGGL(k_pack2, dim3(grid_for(n2d)), dim3(kBlock), 0, st, reinterpret_cast<const float4 *>(win->hdd),
                       win->zeta + s0 * win->zeta_stride, win->zeta + s1 * win->zeta_stride,
                       static_cast<float4 *>(p2), (int64_t)y0 * g->NX, n2d);
    if ((rc = check_launch("opx_pack_window 2d"))) return rc;
    const int wl = w_levels(win, g);
    OPX_REQUIRE(wl == g->NZ || wl == g->NZ + 1, "third field levels %d not NZ or NZ + 1", wl);
    struct F {
        const void *base;
        int64_t stride;
        int nz, n1, n2;
        void *out;
    } f[3] = {{win->u, win->u_stride, g->NZ, g->NY, g->NX - 1, pu}, {win->v, win->v_stride, g->NZ, g->NY - 1, g->NX, pv},
              {win->w, win->w_stride, wl, g->NY, g->NX, pw}};
    for (auto &x : f) {
        // the field's rows in [y0, y1) (v has NY - 1 rows)
        const int r0 = y0, r1 = min(y1, x.n1);
        if (r0 >= r1)
            continue;
        const int64_t n = (int64_t)x.nz * (r1 - r0) * x.n2;
        // a packed window of at most 4 GiB holds fewer than 2^32 nodes
        OPX_REQUIRE(n < (int64_t(1) << 32), "window of %lld nodes", (long long)n);
        const unsigned grid = grid_for(n);
#define OPX_PACK(T, L)                                                                                      \
    hipLaunchKernelGGL((k_pack<T, L>), dim3(grid), dim3(kBlock), 0, st, b + s0 * x.stride, b + s1 * x.stride, \
                       x.out, (uint32_t)n, x.n1, x.n2, r0, r1 - r0)
        if (win->dtype3 == OPX_F32) {
            const float *b = static_cast<const float *>(x.base);
            if (layout == OPX_LAYOUT_PAIR)
                OPX_PACK(float, OPX_LAYOUT_PAIR);
            else if (layout == OPX_LAYOUT_QUAD)
                OPX_PACK(float, OPX_LAYOUT_QUAD);
            else if (layout == OPX_LAYOUT_YPAIR)
                OPX_PACK(float, OPX_LAYOUT_YPAIR);
            else
                OPX_PACK(float, OPX_LAYOUT_OCT);
        } else {
            const uint16_t *b = static_cast<const uint16_t *>(x.base);
            if (layout == OPX_LAYOUT_PAIR)
                OPX_PACK(uint16_t, OPX_LAYOUT_PAIR);
            else if (layout == OPX_LAYOUT_QUAD)
                OPX_PACK(uint16_t, OPX_LAYOUT_QUAD);
            else if (layout == OPX_LAYOUT_YPAIR)
                OPX_PACK(uint16_t, OPX_LAYOUT_YPAIR);
            else
                OPX_PACK(uint16_t, OPX_LAYOUT_OCT);
        }
#undef OPX_PACK
        if ((rc = check_launch("opx_pack_window"))) return rc;
    }
    return OPX_OK;
}

int opx_pack_window_boxes(const opx_window *win, const opx_grid *g, int32_t layout, int32_t s0, int32_t s1,
                          const int32_t *boxes, int32_t nboxes, void *pu, void *pv, void *pw, void *p2, void *stream)
{
    int rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(win && win->u && win->v && win->w, "null window");
    OPX_REQUIRE(win->dtype3 == OPX_F32 || win->dtype3 == OPX_BF16, "dtype3 must be F32 or BF16");
    if (layout == OPX_LAYOUT_PAIR_LDS) // same packed data as PAIR
        layout = OPX_LAYOUT_PAIR;
    OPX_REQUIRE(layout == OPX_LAYOUT_PAIR || layout == OPX_LAYOUT_QUAD || layout == OPX_LAYOUT_OCT ||
                    layout == OPX_LAYOUT_YPAIR,
                "layout must be PAIR, QUAD, OCT, YPAIR or PAIR_LDS");
    OPX_REQUIRE(s0 >= 0 && s0 < win->nslots && s1 >= 0 && s1 < win->nslots, "slots (%d, %d) outside [0, %d)", s0,
                s1, win->nslots);
    OPX_REQUIRE(nboxes >= 0 && (nboxes == 0 || boxes != nullptr), "%d boxes", nboxes);
    for (int32_t b = 0; b < nboxes; ++b) {
        const int32_t *x = boxes + 4 * b;
        OPX_REQUIRE(0 <= x[0] && x[0] < x[1] && x[1] <= g->NY && 0 <= x[2] && x[2] < x[3] && x[3] <= g->NX,
                    "box %d (rows [%d, %d), columns [%d, %d)) outside (%d, %d)", b, x[0], x[1], x[2], x[3], g->NY,
                    g->NX);
    }
    OPX_REQUIRE(pu && pv && pw && p2, "null output window");
    OPX_REQUIRE(win->hdd && win->zeta, "window without hdd/zeta");
    uint64_t bu, bv, bw, b2;
    if ((rc = opx_packed_bytes(g, win->dtype3, layout, &bu, &bv, &bw, &b2))) return rc;
    hipStream_t st = as_stream(stream);
    // the 2-D window whole (NY x NX points, small)
    const int64_t n2d = (int64_t)g->NY * g->NX;
    hipLaunchKernelGGL(k_pack2, dim3(grid_for(n2d)), dim3(kBlock), 0, st, reinterpret_cast<const float4 *>(win->hdd),
                       win->zeta + s0 * win->zeta_stride, win->zeta + s1 * win->zeta_stride,
                       static_cast<float4 *>(p2), (int64_t)0, n2d);
    if ((rc = check_launch("opx_pack_window_boxes 2d"))) return rc;
    if (nboxes == 0)
        return OPX_OK;
    const int wl = w_levels(win, g);
    OPX_REQUIRE(wl == g->NZ || wl == g->NZ + 1, "third field levels %d not NZ or NZ + 1", wl);
    struct F {
        const void *base;
        int64_t stride;
        int nz, n1, n2;
        void *out;
    } f[3] = {{win->u, win->u_stride, g->NZ, g->NY, g->NX - 1, pu}, {win->v, win->v_stride, g->NZ, g->NY - 1, g->NX, pv},
              {win->w, win->w_stride, wl, g->NY, g->NX, pw}};
    for (auto &x : f) {
        for (int32_t b0 = 0; b0 < nboxes; b0 += kPackBoxes) {
            // each box widened by the row above and the column to its left (whose
            // packed elements hold the box's nodes as y + 1 / x + 1 neighbours)
            // and clipped to the field's rows and columns
            PackBoxes B = {};
            int64_t total = 0;
            for (int32_t i = b0; i < min(nboxes, b0 + kPackBoxes); ++i) {
                const int32_t *bx = boxes + 4 * i;
                const int32_t y0 = max(0, bx[0] - 1), y1 = min(bx[1], x.n1);
                const int32_t c0 = max(0, bx[2] - 1), c1 = min(bx[3], x.n2);
                if (y0 >= y1 || c0 >= c1)
                    continue;
                B.box[B.n][0] = y0;
                B.box[B.n][1] = y1;
                B.box[B.n][2] = c0;
                B.box[B.n][3] = c1;
                B.first[B.n] = total;
                total += (int64_t)x.nz * (y1 - y0) * (c1 - c0);
                ++B.n;
            }
            if (B.n == 0)
                continue;
            B.first[B.n] = total;
            const unsigned grid = (unsigned)std::min<int64_t>((total + kBlock - 1) / kBlock, 8192);
#define OPX_PACKB(T, L)                                                                                            \
    hipLaunchKernelGGL((k_pack_boxes<T, L>), dim3(grid), dim3(kBlock), 0, st, b + s0 * x.stride, b + s1 * x.stride, \
                       x.out, x.n1, x.n2, B)
            if (win->dtype3 == OPX_F32) {
                const float *b = static_cast<const float *>(x.base);
                if (layout == OPX_LAYOUT_PAIR)
                    OPX_PACKB(float, OPX_LAYOUT_PAIR);
                else if (layout == OPX_LAYOUT_QUAD)
                    OPX_PACKB(float, OPX_LAYOUT_QUAD);
                else if (layout == OPX_LAYOUT_YPAIR)
                    OPX_PACKB(float, OPX_LAYOUT_YPAIR);
                else
                    OPX_PACKB(float, OPX_LAYOUT_OCT);
            } else {
                const uint16_t *b = static_cast<const uint16_t *>(x.base);
                if (layout == OPX_LAYOUT_PAIR)
                    OPX_PACKB(uint16_t, OPX_LAYOUT_PAIR);
                else if (layout == OPX_LAYOUT_QUAD)
                    OPX_PACKB(uint16_t, OPX_LAYOUT_QUAD);
                else if (layout == OPX_LAYOUT_YPAIR)
                    OPX_PACKB(uint16_t, OPX_LAYOUT_YPAIR);
                else
                    OPX_PACKB(uint16_t, OPX_LAYOUT_OCT);
            }
#undef OPX_PACKB
            if ((rc = check_launch("opx_pack_window_boxes"))) return rc;
        }
    }
    return OPX_OK;
}

int opx_pack_window(const opx_window *win, const opx_grid *g, int32_t layout, int32_t s0, int32_t s1, void *pu,
                    void *pv, void *pw, void *p2, void *stream)
{
    int rc;
    if ((rc = check_grid(g))) return rc;
    return opx_pack_window_rows(win, g, layout, s0, s1, 0, g->NY, pu, pv, pw, p2, stream);
}

int opx_packed_field_bytes(int32_t dtype, int32_t kind, int64_t nz, int64_t ny, int64_t nx, uint64_t *bytes)
{
    OPX_REQUIRE(bytes != nullptr, "null output");
    OPX_REQUIRE(dtype == OPX_F32 || dtype == OPX_BF16, "packed stencils need F32 or BF16, not dtype %d", dtype);
    OPX_REQUIRE(kind == OPX_PACK_QUAD || kind == OPX_PACK_OCT || kind == OPX_PACK_TPAIR || kind == OPX_PACK_HEX,
                "pack kind %d", kind);
    OPX_REQUIRE(nz >= 1 && ny >= 2 && nx >= 2, "array shape (%lld, %lld, %lld)", (long long)nz, (long long)ny,
                (long long)nx);
    OPX_REQUIRE(kind != OPX_PACK_HEX || nz >= 2, "a HEX copy needs a 3-D field (nz = %lld)", (long long)nz);
    const int per = kind == OPX_PACK_QUAD ? 4 : kind == OPX_PACK_OCT ? 8 : kind == OPX_PACK_HEX ? 16 : 2;
    *bytes = (uint64_t)nz * ny * nx * per * (dtype == OPX_F32 ? 4 : 2);
    return OPX_OK;
}

int opx_pack_field(const void *prev, const void *next, int32_t dtype, int64_t nz, int64_t ny, int64_t nx, void *out,
                   void *stream)
{
    return opx_pack_field_as(prev, next, dtype, next ? OPX_PACK_OCT : OPX_PACK_QUAD, nz, ny, nx, out, stream);
}

int opx_pack_field_as(const void *prev, const void *next, int32_t dtype, int32_t kind, int64_t nz, int64_t ny,
                      int64_t nx, void *out, void *stream)
{
    int rc;
    uint64_t bytes;
    if ((rc = opx_packed_field_bytes(dtype, kind, nz, ny, nx, &bytes))) return rc;
    OPX_REQUIRE(prev != nullptr && out != nullptr, "null array");
    OPX_REQUIRE((next != nullptr) == (kind != OPX_PACK_QUAD), "pack kind %d %s a next snapshot", kind,
                kind == OPX_PACK_QUAD ? "takes no" : "needs");
    const int64_t n = nz * ny * nx;
    OPX_REQUIRE(n < (int64_t(1) << 32), "field of %lld nodes", (long long)n);
    hipStream_t st = as_stream(stream);
    const unsigned grid = grid_for(n);
    if (kind == OPX_PACK_HEX) {
        if (dtype == OPX_F32)
            hipLaunchKernelGGL(k_pack_hex<float>, dim3(grid), dim3(kBlock), 0, st, static_cast<const float *>(prev),
                               static_cast<const float *>(next), out, (uint32_t)n, (int)nz, (int)ny, (int)nx);
        else
            hipLaunchKernelGGL(k_pack_hex<uint16_t>, dim3(grid), dim3(kBlock), 0, st,
                               static_cast<const uint16_t *>(prev), static_cast<const uint16_t *>(next), out,
                               (uint32_t)n, (int)nz, (int)ny, (int)nx);
    } else if (kind == OPX_PACK_TPAIR) {
        // the fused windows' PAIR layout: (prev[i], next[i]) per node
        if (dtype == OPX_F32)
            hipLaunchKernelGGL((k_pack<float, OPX_LAYOUT_PAIR>), dim3(grid), dim3(kBlock), 0, st,
                               static_cast<const float *>(prev), static_cast<const float *>(next), out, (uint32_t)n,
                               (int)ny, (int)nx, 0, (int)ny);
        else
            hipLaunchKernelGGL((k_pack<uint16_t, OPX_LAYOUT_PAIR>), dim3(grid), dim3(kBlock), 0, st,
                               static_cast<const uint16_t *>(prev), static_cast<const uint16_t *>(next), out,
                               (uint32_t)n, (int)ny, (int)nx, 0, (int)ny);
    } else if (next) {
        if (dtype == OPX_F32)
            hipLaunchKernelGGL((k_pack<float, OPX_LAYOUT_OCT>), dim3(grid), dim3(kBlock), 0, st,
                               static_cast<const float *>(prev), static_cast<const float *>(next), out, (uint32_t)n,
                               (int)ny, (int)nx, 0, (int)ny);
        else
            hipLaunchKernelGGL((k_pack<uint16_t, OPX_LAYOUT_OCT>), dim3(grid), dim3(kBlock), 0, st,
                               static_cast<const uint16_t *>(prev), static_cast<const uint16_t *>(next), out,
                               (uint32_t)n, (int)ny, (int)nx, 0, (int)ny);
    } else {
        if (dtype == OPX_F32)
            hipLaunchKernelGGL(k_pack_quad<float>, dim3(grid), dim3(kBlock), 0, st, static_cast<const float *>(prev),
                               out, (uint32_t)n, (int)ny, (int)nx);
        else
            hipLaunchKernelGGL(k_pack_quad<uint16_t>, dim3(grid), dim3(kBlock), 0, st,
                               static_cast<const uint16_t *>(prev), out, (uint32_t)n, (int)ny, (int)nx);
    }
    return check_launch("opx_pack_field");
}

int opx_tile_occupancy(const opx_split_state *s, const opx_grid *g, int32_t tile, uint8_t *map, int32_t nty,
                       int32_t ntx, void *stream)
{
    int rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(s != nullptr && map != nullptr, "null state or map");
    OPX_REQUIRE(tile >= 1, "tile %d < 1", tile);
    OPX_REQUIRE(nty == (g->NY + tile - 1) / tile && ntx == (g->NX + tile - 1) / tile,
                "map of %d x %d tiles, the grid needs %d x %d", nty, ntx, (g->NY + tile - 1) / tile,
                (g->NX + tile - 1) / tile);
    OPX_REQUIRE(s->n == 0 || (s->status && s->cell), "state without status or cells");
    hipStream_t st = as_stream(stream);
    const int64_t nm = (int64_t)nty * ntx;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(nm)), dim3(kBlock), 0, st, map, nm, (uint8_t)0);
    if ((rc = check_launch("opx_tile_occupancy fill"))) return rc;
    if (s->n > 0) {
        hipLaunchKernelGGL(k_tile_occupancy, dim3(grid_for(s->n)), dim3(kBlock), 0, st, *s, *g, (int)tile, map,
                           (int)ntx);
        if ((rc = check_launch("opx_tile_occupancy"))) return rc;
    }
    return OPX_OK;
}

int opx_gather_boxes_h2d(void *dst, const void *src, int64_t nz, int64_t ny, int64_t nx, int32_t elem_bytes,
                         const int32_t *boxes, int32_t nboxes, void *stream)
{
    int rc;
    OPX_REQUIRE(dst != nullptr && src != nullptr, "null array");
    OPX_REQUIRE(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8, "element size %d",
                elem_bytes);
    OPX_REQUIRE(nz >= 1 && ny >= 1 && nx >= 1, "array shape (%lld, %lld, %lld)", (long long)nz, (long long)ny,
                (long long)nx);
    OPX_REQUIRE(nboxes >= 0 && (nboxes == 0 || boxes != nullptr), "%d boxes", nboxes);
    for (int32_t b = 0; b < nboxes; ++b) {
        const int32_t *x = boxes + 4 * b;
        OPX_REQUIRE(0 <= x[0] && x[0] < x[1] && x[1] <= ny && 0 <= x[2] && x[2] < x[3] && x[3] <= nx,
                    "box %d (rows [%d, %d), columns [%d, %d)) outside (%lld, %lld)", b, x[0], x[1], x[2], x[3],
                    (long long)ny, (long long)nx);
    }
    if (nboxes == 0)
        return OPX_OK;
    // the kernel reads the host array through its device mapping: both ends of
    // the array must be mapped pinned memory of one contiguous mapping
    const size_t total = (size_t)nz * ny * nx * elem_bytes;
    void *d0 = nullptr, *d1 = nullptr;
    const char *h0 = static_cast<const char *>(src);
    if (hipHostGetDevicePointer(&d0, const_cast<char *>(h0), 0) != hipSuccess ||
        hipHostGetDevicePointer(&d1, const_cast<char *>(h0 + total - 1), 0) != hipSuccess ||
        static_cast<char *>(d1) - static_cast<char *>(d0) != (ptrdiff_t)(total - 1)) {
        (void)hipGetLastError();
        set_error("opx_gather_boxes_h2d: the source is not mapped pinned host memory");
        return OPX_ERR_INVALID;
    }
    OPX_REQUIRE(nz * ny * nx < (int64_t(1) << 31), "array of %lld elements", (long long)(nz * ny * nx));
    hipStream_t st = as_stream(stream);
    for (int32_t b0 = 0; b0 < nboxes; b0 += kGatherBoxes) {
        GatherBoxes B = {};
        B.n = min(kGatherBoxes, nboxes - b0);
        int64_t total = 0;
        for (int i = 0; i < B.n; ++i) {
            const int32_t *x = boxes + 4 * (b0 + i);
            for (int j = 0; j < 4; ++j)
                B.box[i][j] = x[j];
            B.first[i] = total;
            total += nz * (int64_t)(x[1] - x[0]) * (x[3] - x[2]);
        }
        B.first[B.n] = total;
        // 64 workgroups keep enough reads in flight for the PCIe link; a wider
        // grid only holds wave slots, which the stepping kernels of the other
        // streams then wait for (served FP32 1.41e10 -> 1.50e10 particle-steps/s
        // against 8192 workgroups; 32 or 16 starve the link)
        const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 64);
        switch (elem_bytes) {
        case 1: hipLaunchKernelGGL(k_gather_boxes<uint8_t>, dim3(grid), dim3(256), 0, st, static_cast<uint8_t *>(dst),
                                   static_cast<const uint8_t *>(d0), ny, nx, B); break;
        case 2: hipLaunchKernelGGL(k_gather_boxes<uint16_t>, dim3(grid), dim3(256), 0, st,
                                   static_cast<uint16_t *>(dst), static_cast<const uint16_t *>(d0), ny, nx, B); break;
        case 4: hipLaunchKernelGGL(k_gather_boxes<uint32_t>, dim3(grid), dim3(256), 0, st,
                                   static_cast<uint32_t *>(dst), static_cast<const uint32_t *>(d0), ny, nx, B); break;
        default: hipLaunchKernelGGL(k_gather_boxes<uint64_t>, dim3(grid), dim3(256), 0, st,
                                    static_cast<uint64_t *>(dst), static_cast<const uint64_t *>(d0), ny, nx, B); break;
        }
        if ((rc = check_launch("opx_gather_boxes_h2d"))) return rc;
    }
    return OPX_OK;
}

int opx_copy_boxes_h2d(void *dst, const void *src, int64_t nz, int64_t ny, int64_t nx, int32_t elem_bytes,
                       const int32_t *boxes, int32_t nboxes, void *stream)
{
    OPX_REQUIRE(dst != nullptr && src != nullptr, "null array");
    OPX_REQUIRE(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8, "element size %d",
                elem_bytes);
    OPX_REQUIRE(nz >= 1 && ny >= 1 && nx >= 1, "array shape (%lld, %lld, %lld)", (long long)nz, (long long)ny,
                (long long)nx);
    OPX_REQUIRE(nboxes >= 0 && (nboxes == 0 || boxes != nullptr), "%d boxes", nboxes);
    for (int32_t b = 0; b < nboxes; ++b) {
        const int32_t *x = boxes + 4 * b;
        OPX_REQUIRE(0 <= x[0] && x[0] < x[1] && x[1] <= ny && 0 <= x[2] && x[2] < x[3] && x[3] <= nx,
                    "box %d (rows [%d, %d), columns [%d, %d)) outside (%lld, %lld)", b, x[0], x[1], x[2], x[3],
                    (long long)ny, (long long)nx);
    }
    hipStream_t st = as_stream(stream);
    const size_t eb = (size_t)elem_bytes;
    for (int32_t b = 0; b < nboxes; ++b) {
        const int32_t *x = boxes + 4 * b;
        hipMemcpy3DParms p = {};
        p.srcPtr = make_hipPitchedPtr(const_cast<void *>(src), (size_t)nx * eb, (size_t)nx, (size_t)ny);
        p.dstPtr = make_hipPitchedPtr(dst, (size_t)nx * eb, (size_t)nx, (size_t)ny);
        p.srcPos = make_hipPos((size_t)x[2] * eb, (size_t)x[0], 0);
        p.dstPos = p.srcPos;
        p.extent = make_hipExtent((size_t)(x[3] - x[2]) * eb, (size_t)(x[1] - x[0]), (size_t)nz);
        p.kind = hipMemcpyHostToDevice;
        const hipError_t e = hipMemcpy3DAsync(&p, st);
        if (e != hipSuccess) {
            set_error("opx_copy_boxes_h2d: box %d: %s", b, hipGetErrorString(e));
            return OPX_ERR_HIP;
        }
    }
    return OPX_OK;
}

int opx_split_from_f64(const opx_split_state *s, const opx_grid *g, const double *zidx, const double *yidx,
                       const double *xidx, void *stream)
{
    OPX_REQUIRE(s != nullptr, "null state");
    int rc;
    if ((rc = check_grid(g))) return rc;
    if (s->n == 0)
        return OPX_OK;
    OPX_REQUIRE(s->cell && s->fz && s->fy && s->fx && zidx && yidx && xidx, "null array");
    hipLaunchKernelGGL(k_split_from_f64, dim3(grid_for(s->n)), dim3(kBlock), 0, as_stream(stream), *s, *g, zidx,
                       yidx, xidx);
    return check_launch("opx_split_from_f64");
}

int opx_sample_split(const opx_split_state *s, const opx_grid *g, const opx_field *fields_dev, int32_t nfields,
                     int32_t nrel, const opx_sample *samples_dev, int32_t nsamples, double hc, double *out,
                     int32_t out_stride, void *stream)
{
    int rc;
    if ((rc = check_state(s, "opx_sample_split", false))) return rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(nfields >= 1 && nrel >= 1, "need at least one field and release");
    OPX_REQUIRE(nsamples >= 1 && nsamples <= kMaxSamples, "nsamples %d not in [1, %d]", nsamples, kMaxSamples);
    OPX_REQUIRE(out_stride >= nsamples && (out_stride == 1 || out_stride == 2 || out_stride == 4 || out_stride == 8),
                "out_stride %d must be 1, 2, 4 or 8 and at least nsamples (%d)", out_stride, nsamples);
    OPX_REQUIRE(fields_dev != nullptr && samples_dev != nullptr && out != nullptr, "null table or output");
    OPX_REQUIRE(out_stride == 1 || reinterpret_cast<uintptr_t>(out) % 16 == 0, "out must be 16-byte aligned");
    if (s->n == 0)
        return OPX_OK;
    const dim3 grid(grid_for(s->n)), blk(kBlock);
    if (nrel == 1)
        hipLaunchKernelGGL(k_sample<true>, grid, blk, 0, as_stream(stream), *s, *g, fields_dev, (int)nfields,
                           samples_dev, (int)nsamples, hc, out, (int)out_stride);
    else
        hipLaunchKernelGGL(k_sample<false>, grid, blk, 0, as_stream(stream), *s, *g, fields_dev, (int)nfields,
                           samples_dev, (int)nsamples, hc, out, (int)out_stride);
    return check_launch("opx_sample_split");
}

int opx_split_to_f64(const opx_split_state *s, const opx_grid *g, double *zidx, double *yidx, double *xidx,
                     void *stream)
{
    OPX_REQUIRE(s != nullptr, "null state");
    int rc;
    if ((rc = check_grid(g))) return rc;
    if (s->n == 0)
        return OPX_OK;
    OPX_REQUIRE(s->cell && s->fz && s->fy && s->fx && zidx && yidx && xidx, "null array");
    hipLaunchKernelGGL(k_split_to_f64, dim3(grid_for(s->n)), dim3(kBlock), 0, as_stream(stream), *s, *g, zidx,
                       yidx, xidx);
    return check_launch("opx_split_to_f64");
}

int opx_extract_release(const opx_split_state *s, const opx_grid *g, int32_t r, uint32_t pid_base, int64_t m,
                        double *zidx, double *yidx, double *xidx, uint8_t *status, void *stream)
{
    OPX_REQUIRE(s != nullptr, "null state");
    int rc;
    if ((rc = check_grid(g))) return rc;
    OPX_REQUIRE(r >= 0 && r <= 0xffff, "release id %d out of [0, 65535]", r);
    OPX_REQUIRE(m >= 0, "negative release size");
    if (s->n == 0 || m == 0)
        return OPX_OK;
    OPX_REQUIRE(s->cell && s->fz && s->fy && s->fx && s->status && s->pid && s->rel, "null state array");
    OPX_REQUIRE(zidx && yidx && xidx, "null output");
    hipLaunchKernelGGL(k_extract_release, dim3(grid_for(s->n)), dim3(kBlock), 0, as_stream(stream), *s, *g,
                       (uint32_t)r, pid_base, m, zidx, yidx, xidx, status);
    return check_launch("opx_extract_release");
}

int opx_rebin_workspace_bytes(int64_t n, uint64_t *bytes)
{
    OPX_REQUIRE(bytes != nullptr, "null output");
    OPX_REQUIRE(n >= 0 && n <= INT32_MAX, "particle count out of range");
    size_t arr = align256((size_t)n * sizeof(uint32_t));
    *bytes = (uint64_t)(4 * arr + align256(sort_temp_bytes(n)));
    return OPX_OK;
}

int opx_rebin(const opx_split_state *src, const opx_split_state *dst, const opx_grid *g, int32_t xlow,
              const opx_release_clock *clocks, int32_t nrel, void *work, uint64_t work_bytes, int32_t *n_active,
              int32_t *n_live, void *stream)
{
    int rc;
    if ((rc = opx::check_rebin_args(src, dst, g, xlow, clocks, nrel, n_active))) return rc;
    uint64_t need = 0;
    opx_rebin_workspace_bytes(src->n, &need);
    OPX_REQUIRE(work != nullptr && work_bytes >= need, "rebin workspace too small (%llu < %llu bytes)",
                (unsigned long long)work_bytes, (unsigned long long)need);
    return opx::rebin(src, dst, g, xlow, clocks, nrel, work, work_bytes, n_active, n_live, stream);
}

int opx_rebin_noalloc(const opx_split_state *src, const opx_split_state *dst, const opx_grid *g, int32_t xlow,
                      const opx_release_clock *clocks, int32_t nrel, int32_t *n_active, int32_t *n_live,
                      void *stream)
{
    int rc;
    if ((rc = opx::check_rebin_args(src, dst, g, xlow, clocks, nrel, n_active))) return rc;
    // 16 bytes per particle alone exceed the buffer: refuse before querying the sort
    OPX_REQUIRE((uint64_t)src->n * 16u <= opx::kNoAllocScratchBytes,
                "opx_rebin_noalloc: %lld particles need more than the %llu-byte static scratch buffer",
                (long long)src->n, (unsigned long long)opx::kNoAllocScratchBytes);
    uint64_t need = 0;
    opx_rebin_workspace_bytes(src->n, &need);
    OPX_REQUIRE(need <= opx::kNoAllocScratchBytes,
                "opx_rebin_noalloc: %lld particles need %llu bytes of scratch, above the %llu-byte static buffer",
                (long long)src->n, (unsigned long long)need, (unsigned long long)opx::kNoAllocScratchBytes);
    void *work = nullptr;
    if ((rc = opx::noalloc_scratch(&work))) return rc;
    return opx::rebin(src, dst, g, xlow, clocks, nrel, work, opx::kNoAllocScratchBytes, n_active, n_live, stream);
}

} // extern "C"
