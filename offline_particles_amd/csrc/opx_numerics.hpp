// Device numerics shared by the libopx kernels: index truncation, stencil
// loads with the per-value time blend, linear/bilinear/trilinear interpolation
// and the ROMS Vtransform-2 vertical coordinate.
//
// The FP64 versions keep the reference's operation order (paths relative to
// src/offline_particles/ of the reference):
//   truncate_index / linear / bilinear / trilinear   kernels/_interpolation/linear.pxd:3-65
//   compute_z / compute_zidx                         kernels/roms/_vertical_coordinate.pxd:19-103
//   per-value blend prev*gt + next*ft                fields.py:423-428
// Translation units that need bit-exact FP64 results are compiled with
// -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "opx.h"

namespace opx {

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// `<Py_ssize_t>idx` then clamp to [0, max_idx] (linear.pxd:3-9).  NaN and
// values outside the int64 range convert to INT64_MIN on x86-64 and clamp to
// 0; that behaviour is reproduced here.
__device__ __forceinline__ int64_t truncate_index(double idx, int64_t max_idx)
{
    if (!(idx > -9223372036854775808.0 && idx < 9223372036854775808.0))
        return 0;
    int64_t i = (int64_t)idx;
    return i < 0 ? 0 : (i > max_idx ? max_idx : i);
}

// Stencil value i of a field, blended in its storage type and widened to FP64
// (fields.py:423-428 then np.ascontiguousarray(float64), _core.pyx:5-24).
__device__ __forceinline__ double load_blend(const opx_field &f, int64_t i)
{
    if (f.dtype == OPX_F64) {
        const double *p = static_cast<const double *>(f.prev);
        if (f.next == nullptr)
            return p[i];
        double a = p[i] * f.gt;
        double b = static_cast<const double *>(f.next)[i] * f.ft;
        return a + b;
    } else if (f.dtype == OPX_F32) {
        const float *p = static_cast<const float *>(f.prev);
        if (f.next == nullptr)
            return (double)p[i];
        float a = p[i] * (float)f.gt;
        float b = static_cast<const float *>(f.next)[i] * (float)f.ft;
        return (double)(a + b);
    } else {
        const uint16_t *p = static_cast<const uint16_t *>(f.prev);
        if (f.next == nullptr)
            return (double)bf16_bits_to_f32(p[i]);
        float a = bf16_bits_to_f32(p[i]) * (float)f.gt;
        float b = bf16_bits_to_f32(static_cast<const uint16_t *>(f.next)[i]) * (float)f.ft;
        return (double)(a + b);
    }
}

// Values i and i + 1 of a snapshot, read with one load: 8 bytes at 4-byte
// alignment (FP32) or 16 bytes at 8-byte alignment (FP64).
typedef uint32_t num_u32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef uint32_t num_u32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));

__device__ __forceinline__ void load_pair(const float *p, float &a, float &b)
{
    const num_u32x2_a4 t = *reinterpret_cast<const num_u32x2_a4 *>(p);
    const uint32_t x = t.x, y = t.y;
    a = __uint_as_float(x);
    b = __uint_as_float(y);
}

__device__ __forceinline__ void load_pair(const double *p, double &a, double &b)
{
    const num_u32x4_a8 t = *reinterpret_cast<const num_u32x4_a8 *>(p);
    const uint32_t x0 = t.x, x1 = t.y, y0 = t.z, y1 = t.w;
    a = __hiloint2double((int)x1, (int)x0);
    b = __hiloint2double((int)y1, (int)y0);
}

// load_blend of stencil values i and i + 1 from a TPAIR copy (prev[i],
// next[i], prev[i + 1], next[i + 1] adjacent): one 16-byte load at 8-byte
// alignment (FP32) or one 8-byte load at 4-byte alignment (BF16), blended in
// the same arithmetic.
__device__ __forceinline__ void tpair_blend2(const opx_field &f, int64_t i, double &va, double &vb)
{
    const float gt = (float)f.gt, ft = (float)f.ft;
    float p0, q0, p1, q1;
    if (f.dtype == OPX_F32) {
        const num_u32x4_a8 t = *reinterpret_cast<const num_u32x4_a8 *>(static_cast<const float *>(f.packed) + 2 * i);
        const uint32_t x0 = t.x, y0 = t.y, x1 = t.z, y1 = t.w;
        p0 = __uint_as_float(x0);
        q0 = __uint_as_float(y0);
        p1 = __uint_as_float(x1);
        q1 = __uint_as_float(y1);
    } else {
        const num_u32x2_a4 t = *reinterpret_cast<const num_u32x2_a4 *>(static_cast<const uint32_t *>(f.packed) + i);
        const uint32_t e0 = t.x, e1 = t.y;
        p0 = bf16_bits_to_f32((uint16_t)(e0 & 0xffffu));
        q0 = bf16_bits_to_f32((uint16_t)(e0 >> 16));
        p1 = bf16_bits_to_f32((uint16_t)(e1 & 0xffffu));
        q1 = bf16_bits_to_f32((uint16_t)(e1 >> 16));
    }
    float a0 = p0 * gt, b0 = q0 * ft;
    float a1 = p1 * gt, b1 = q1 * ft;
    va = (double)(a0 + b0);
    vb = (double)(a1 + b1);
}

// load_blend of stencil values i and i + 1 (the same values, in the same
// arithmetic), reading each snapshot's neighbours with one load for FP32 and
// FP64 fields, or both snapshots' with one load from a TPAIR copy.
__device__ __forceinline__ void load_blend2(const opx_field &f, int64_t i, double &va, double &vb)
{
    if (f.packed) {
        tpair_blend2(f, i, va, vb);
        return;
    }
    if (f.dtype == OPX_F64) {
        double p0, p1;
        load_pair(static_cast<const double *>(f.prev) + i, p0, p1);
        if (f.next == nullptr) {
            va = p0;
            vb = p1;
            return;
        }
        double n0, n1;
        load_pair(static_cast<const double *>(f.next) + i, n0, n1);
        double a0 = p0 * f.gt, b0 = n0 * f.ft;
        double a1 = p1 * f.gt, b1 = n1 * f.ft;
        va = a0 + b0;
        vb = a1 + b1;
    } else if (f.dtype == OPX_F32) {
        float p0, p1;
        load_pair(static_cast<const float *>(f.prev) + i, p0, p1);
        if (f.next == nullptr) {
            va = (double)p0;
            vb = (double)p1;
            return;
        }
        float n0, n1;
        load_pair(static_cast<const float *>(f.next) + i, n0, n1);
        const float gt = (float)f.gt, ft = (float)f.ft;
        float a0 = p0 * gt, b0 = n0 * ft;
        float a1 = p1 * gt, b1 = n1 * ft;
        va = (double)(a0 + b0);
        vb = (double)(a1 + b1);
    } else {
        va = load_blend(f, i);
        vb = load_blend(f, i + 1);
    }
}

// The stencil corners (y, x), (y, x+1), (y+1, x), (y+1, x+1) of `node` from a
// field's QUAD or OCT copy (opx_field.packed), blended and widened exactly as
// load_blend does: one 16-byte (QUAD F32, OCT BF16), 32-byte (OCT F32) or
// 8-byte (QUAD BF16) load instead of two rows per snapshot.  QUAD and OCT
// elements sit opx_field.packed_stride bytes apart (0: packed back to back).
__device__ __forceinline__ void packed_corners(const opx_field &f, int64_t node, double v[4])
{
    if (f.dtype == OPX_F32) {
        const float4 *e = static_cast<const float4 *>(f.packed);
        if (f.packed_kind == OPX_PACK_QUAD) {
            const int64_t stride = f.packed_stride ? f.packed_stride : 16;
            const float4 p = *reinterpret_cast<const float4 *>(static_cast<const char *>(f.packed) + node * stride);
            v[0] = (double)p.x;
            v[1] = (double)p.y;
            v[2] = (double)p.z;
            v[3] = (double)p.w;
        } else {
            const int64_t stride = f.packed_stride ? f.packed_stride : 32;
            const float4 *o = reinterpret_cast<const float4 *>(static_cast<const char *>(f.packed) + node * stride);
            const float4 p = o[0], q = o[1];
            const float gt = (float)f.gt, ft = (float)f.ft;
            float a0 = p.x * gt, b0 = q.x * ft;
            float a1 = p.y * gt, b1 = q.y * ft;
            float a2 = p.z * gt, b2 = q.z * ft;
            float a3 = p.w * gt, b3 = q.w * ft;
            v[0] = (double)(a0 + b0);
            v[1] = (double)(a1 + b1);
            v[2] = (double)(a2 + b2);
            v[3] = (double)(a3 + b3);
        }
    } else {
        if (f.packed_kind == OPX_PACK_QUAD) {
            const int64_t stride = f.packed_stride ? f.packed_stride : 8;
            const uint2 p = *reinterpret_cast<const uint2 *>(static_cast<const char *>(f.packed) + node * stride);
            v[0] = (double)bf16_bits_to_f32((uint16_t)(p.x & 0xffffu));
            v[1] = (double)bf16_bits_to_f32((uint16_t)(p.x >> 16));
            v[2] = (double)bf16_bits_to_f32((uint16_t)(p.y & 0xffffu));
            v[3] = (double)bf16_bits_to_f32((uint16_t)(p.y >> 16));
        } else {
            const int64_t stride = f.packed_stride ? f.packed_stride : 16;
            const uint4 e = *reinterpret_cast<const uint4 *>(static_cast<const char *>(f.packed) + node * stride);
            const uint32_t P[2] = {e.x, e.y}, Q[2] = {e.z, e.w};
            const float gt = (float)f.gt, ft = (float)f.ft;
            for (int c = 0; c < 4; ++c) {
                const uint16_t pb = (uint16_t)(c & 1 ? P[c >> 1] >> 16 : P[c >> 1] & 0xffffu);
                const uint16_t qb = (uint16_t)(c & 1 ? Q[c >> 1] >> 16 : Q[c >> 1] & 0xffffu);
                float a = bf16_bits_to_f32(pb) * gt;
                float b = bf16_bits_to_f32(qb) * ft;
                v[c] = (double)(a + b);
            }
        }
    }
}

// The 8 corners (dz, dy, dx) of `node` from a field's HEX copy (prev's then
// next's in one 64-byte FP32 / 32-byte BF16 element), each blended and
// widened exactly as load_blend does; corner (dz, dy, dx) at (dz*2+dy)*2+dx.
__device__ __forceinline__ void hex_corners(const opx_field &f, int64_t node, double v[8])
{
    const float gt = (float)f.gt, ft = (float)f.ft;
    if (f.dtype == OPX_F32) {
        const float4 *e = static_cast<const float4 *>(f.packed) + 4 * node;
        const float4 p0 = e[0], p1 = e[1], q0 = e[2], q1 = e[3];
        const float P[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
        const float Q[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
        for (int c = 0; c < 8; ++c) {
            float a = P[c] * gt, b = Q[c] * ft;
            v[c] = (double)(a + b);
        }
    } else {
        const uint4 *e = static_cast<const uint4 *>(f.packed) + 2 * node;
        const uint4 pe = e[0], qe = e[1];
        const uint32_t P[4] = {pe.x, pe.y, pe.z, pe.w}, Q[4] = {qe.x, qe.y, qe.z, qe.w};
        for (int c = 0; c < 8; ++c) {
            const uint16_t pb = (uint16_t)(c & 1 ? P[c >> 1] >> 16 : P[c >> 1] & 0xffffu);
            const uint16_t qb = (uint16_t)(c & 1 ? Q[c >> 1] >> 16 : Q[c >> 1] & 0xffffu);
            float a = bf16_bits_to_f32(pb) * gt;
            float b = bf16_bits_to_f32(qb) * ft;
            v[c] = (double)(a + b);
        }
    }
}

// linear_interpolation (linear.pxd:11-17) on a 1-D field at rho index idx.
__device__ __forceinline__ double linear_f64(const opx_field &f, double idx)
{
    idx += f.offsets[0];
    int64_t I0 = truncate_index(idx, f.shape[0] - 2);
    double f0 = idx - (double)I0;
    double g0 = 1.0 - f0;
    double v0, v1;
    load_blend2(f, I0, v0, v1);
    return g0 * v0 + f0 * v1;
}

// bilinear_interpolation (linear.pxd:19-33) on a 2-D field at rho (y, x).
__device__ __forceinline__ double bilinear_f64(const opx_field &f, double y, double x)
{
    double idx0 = y + f.offsets[0];
    double idx1 = x + f.offsets[1];
    const int64_t n1 = f.shape[1];
    int64_t I0 = truncate_index(idx0, f.shape[0] - 2);
    int64_t I1 = truncate_index(idx1, n1 - 2);
    double f0 = idx0 - (double)I0;
    double f1 = idx1 - (double)I1;
    double g0 = 1.0 - f0;
    double g1 = 1.0 - f1;
    int64_t b = I0 * n1 + I1;
    double v00, v01, v10, v11;
    if (f.packed && f.packed_kind != OPX_PACK_TPAIR) {
        double c[4];
        packed_corners(f, b, c);
        v00 = c[0];
        v01 = c[1];
        v10 = c[2];
        v11 = c[3];
    } else {
        load_blend2(f, b, v00, v01);
        load_blend2(f, b + n1, v10, v11);
    }
    return g0 * g1 * v00 + g0 * f1 * v01 + f0 * g1 * v10 + f0 * f1 * v11;
}

// trilinear_interpolation (linear.pxd:35-65), sum order of :58-65.
__device__ __forceinline__ double trilinear_f64(const opx_field &f, double z, double y, double x)
{
    double idx0 = z + f.offsets[0];
    double idx1 = y + f.offsets[1];
    double idx2 = x + f.offsets[2];
    const int64_t n1 = f.shape[1], n2 = f.shape[2];
    int64_t I0 = truncate_index(idx0, f.shape[0] - 2);
    int64_t I1 = truncate_index(idx1, n1 - 2);
    int64_t I2 = truncate_index(idx2, n2 - 2);
    double f0 = idx0 - (double)I0;
    double f1 = idx1 - (double)I1;
    double f2 = idx2 - (double)I2;
    double g0 = 1.0 - f0;
    double g1 = 1.0 - f1;
    double g2 = 1.0 - f2;
    int64_t b00 = (I0 * n1 + I1) * n2 + I2;
    int64_t b01 = b00 + n2;
    int64_t b10 = b00 + n1 * n2;
    int64_t b11 = b10 + n2;
    double v000, v001, v010, v011, v100, v101, v110, v111;
    if (f.packed && f.packed_kind == OPX_PACK_HEX) {
        double c[8];
        hex_corners(f, b00, c);
        v000 = c[0];
        v001 = c[1];
        v010 = c[2];
        v011 = c[3];
        v100 = c[4];
        v101 = c[5];
        v110 = c[6];
        v111 = c[7];
    } else if (f.packed && f.packed_kind != OPX_PACK_TPAIR) {
        double c0[4], c1[4];
        packed_corners(f, b00, c0);
        packed_corners(f, b10, c1);
        v000 = c0[0];
        v001 = c0[1];
        v010 = c0[2];
        v011 = c0[3];
        v100 = c1[0];
        v101 = c1[1];
        v110 = c1[2];
        v111 = c1[3];
    } else {
        load_blend2(f, b00, v000, v001);
        load_blend2(f, b01, v010, v011);
        load_blend2(f, b10, v100, v101);
        load_blend2(f, b11, v110, v111);
    }
    return (g0 * g1 * g2 * v000 +
            g0 * g1 * f2 * v001 +
            g0 * f1 * g2 * v010 +
            g0 * f1 * f2 * v011 +
            f0 * g1 * g2 * v100 +
            f0 * g1 * f2 * v101 +
            f0 * f1 * g2 * v110 +
            f0 * f1 * f2 * v111);
}

// ---- ROMS Vtransform = 2 vertical coordinate (_vertical_coordinate.pxd) ----

__device__ __forceinline__ double sigma_coordinate(double zidx, int NZ) { return (zidx + 0.5) / NZ - 1.0; }

__device__ __forceinline__ double S_coordinate(double hc, double sigma, double h, double C)
{
    return (hc * sigma + h * C) / (hc + h);
}

// compute_z (:31-35)
__device__ __forceinline__ double compute_z(double zidx, int NZ, double hc, double h, double C, double zeta)
{
    double sigma = sigma_coordinate(zidx, NZ);
    double S = S_coordinate(hc, sigma, h, C);
    return zeta + (zeta + h) * S;
}

// C[k] of the 1-D stretching field (always a static field).
__device__ __forceinline__ double C_at(const opx_field &C, int64_t k) { return load_blend(C, k); }

// compute_zidx (:98-103) with _compute_Cidx_from_S's edge returns and binary
// search (:50-84) and _zidx_from_S's linear inverse (:86-96).
__device__ __forceinline__ double compute_zidx(double z, double h, double zeta, double hc, int NZ,
                                               const opx_field &C)
{
    const double C_offset = C.offsets[0];
    const int64_t C_size = C.shape[0];
    double S = (z - zeta) / (zeta + h);
    int64_t Cidx;
    if (S <= S_coordinate(hc, sigma_coordinate(0.0 - C_offset, NZ), h, C_at(C, 0))) {
        Cidx = 0;
    } else if (S >= S_coordinate(hc, sigma_coordinate((double)(C_size - 2) - C_offset, NZ), h,
                                 C_at(C, C_size - 2))) {
        Cidx = C_size - 2;
    } else {
        int64_t lo = 0, hi = C_size - 2;
        while (lo < hi - 1) {
            int64_t mid = (lo + hi) / 2;
            double S_mid = S_coordinate(hc, sigma_coordinate((double)mid - C_offset, NZ), h, C_at(C, mid));
            if (S_mid <= S)
                lo = mid;
            else
                hi = mid;
        }
        Cidx = lo;
    }
    double S_low = S_coordinate(hc, sigma_coordinate((double)Cidx - C_offset, NZ), h, C_at(C, Cidx));
    double S_high = S_coordinate(hc, sigma_coordinate((double)(Cidx + 1) - C_offset, NZ), h, C_at(C, Cidx + 1));
    double f = (S - S_low) / (S_high - S_low);
    return (double)Cidx - C_offset + f;
}

} // namespace opx
