// Spatial ordering of the drop-in's particle arrays (opt-in, Simulation's
// spatial_order_every): a Morton permutation of the particles' cells and a
// gather that applies it to any particle array.  The reference processes
// particles in the caller's order (ab3_w_advection.pyx:80); each particle is
// independent, so the order changes no result, only how many cache lines the
// stencil gathers of one wave touch.
//
// Keys: the cell floor(index) per axis, clamped to [0, N-1], interleaved as
// Morton bits of (x >> xlow, y, z) with the low xlow bits of x kept
// contiguous (runs of 2^xlow cells along x); particles with the INACTIVE bit
// or a non-finite index get the key 2^key_bits and go last.  The radix sort is
// stable.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>

#include "opx.h"
#include "opx_runtime.hpp"

namespace opx {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int bits_for(int n)
{
    int b = 1;
    while ((1 << b) < n)
        ++b;
    return b;
}

size_t order_sort_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (int32_t *)nullptr, (int32_t *)nullptr, (int)n, 0, 32);
    return bytes;
}

__device__ __forceinline__ uint32_t cell_of(double v, int N)
{
    // floor of an index inside [0, N-1]; finite values only
    const double c = fmin(fmax(floor(v), 0.0), (double)(N - 1));
    return (uint32_t)c;
}

__global__ void __launch_bounds__(kBlock)
    k_order_keys(int64_t n, const uint8_t *__restrict__ status, const double *__restrict__ zidx,
                 const double *__restrict__ yidx, const double *__restrict__ xidx, int NZ, int NY, int NX, int zb,
                 int yb, int xb, int xlow, uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    vals[i] = (int32_t)i;
    const double z = zidx[i], y = yidx[i], x = xidx[i];
    const uint32_t last = 1u << (zb + yb + xb);
    if ((status[i] & OPX_INACTIVE) || !(isfinite(z) && isfinite(y) && isfinite(x))) {
        keys[i] = last;
        return;
    }
    const uint32_t cz = cell_of(z, NZ), cy = cell_of(y, NY), cx = cell_of(x, NX);
    // Morton interleave of (cx >> xlow, cy, cz), then the low xlow bits of cx
    uint32_t key = 0;
    int out = 0;
    const uint32_t hx = cx >> xlow;
    const int hb = xb - xlow;
    const int top = max(hb, max(yb, zb));
    for (int b = 0; b < top; ++b) {
        if (b < hb)
            key |= ((hx >> b) & 1u) << out++;
        if (b < yb)
            key |= ((cy >> b) & 1u) << out++;
        if (b < zb)
            key |= ((cz >> b) & 1u) << out++;
    }
    keys[i] = (key << xlow) | (cx & ((1u << xlow) - 1u));
}

template <typename E>
__global__ void __launch_bounds__(kBlock)
    k_gather(int64_t n, const int32_t *__restrict__ perm, const E *__restrict__ src, E *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n)
        dst[i] = src[perm[i]];
}

} // namespace
} // namespace opx

using namespace opx;

extern "C" {

int opx_spatial_order_workspace_bytes(int64_t n, uint64_t *bytes)
{
    OPX_REQUIRE(bytes != nullptr, "null output");
    OPX_REQUIRE(n >= 0 && n <= INT32_MAX, "particle count out of range");
    const size_t arr = align256((size_t)n * sizeof(uint32_t));
    *bytes = (uint64_t)(3 * arr + align256(order_sort_temp_bytes(n)));
    return OPX_OK;
}

int opx_spatial_order(int64_t n, const uint8_t *status, const double *zidx, const double *yidx, const double *xidx,
                      int32_t NZ, int32_t NY, int32_t NX, int32_t xlow, void *work, uint64_t work_bytes,
                      int32_t *perm, void *stream)
{
    OPX_REQUIRE(n >= 0 && n <= INT32_MAX, "particle count %lld out of range", (long long)n);
    OPX_REQUIRE(NZ >= 1 && NY >= 1 && NX >= 1, "grid %dx%dx%d invalid", NZ, NY, NX);
    const int zb = bits_for(NZ), yb = bits_for(NY), xb = bits_for(NX);
    OPX_REQUIRE(zb + yb + xb <= 31, "grid %dx%dx%d needs more than 31 key bits", NZ, NY, NX);
    OPX_REQUIRE(xlow >= 0 && xlow <= xb, "xlow %d out of [0, %d]", xlow, xb);
    if (n == 0)
        return OPX_OK;
    OPX_REQUIRE(status && zidx && yidx && xidx && perm, "null array");
    uint64_t need = 0;
    opx_spatial_order_workspace_bytes(n, &need);
    OPX_REQUIRE(work != nullptr && work_bytes >= need, "spatial-order workspace too small (%llu < %llu bytes)",
                (unsigned long long)work_bytes, (unsigned long long)need);
    const size_t arr = align256((size_t)n * sizeof(uint32_t));
    char *base = static_cast<char *>(work);
    uint32_t *k_in = reinterpret_cast<uint32_t *>(base);
    uint32_t *k_out = reinterpret_cast<uint32_t *>(base + arr);
    int32_t *v_in = reinterpret_cast<int32_t *>(base + 2 * arr);
    void *temp = base + 3 * arr;
    size_t temp_bytes = (size_t)(work_bytes - 3 * arr);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_order_keys, dim3(grid_for(n)), dim3(kBlock), 0, st, n, status, zidx, yidx, xidx, NZ, NY,
                       NX, zb, yb, xb, xlow, k_in, v_in);
    int rc;
    if ((rc = check_launch("opx_spatial_order keys")))
        return rc;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k_in, k_out, v_in, perm, (int)n, 0,
                                                      zb + yb + xb + 1, st);
    if (e != hipSuccess) {
        set_error("opx_spatial_order: radix sort failed: %s", hipGetErrorString(e));
        return OPX_ERR_HIP;
    }
    return check_launch("opx_spatial_order");
}

int opx_gather(int64_t n, const int32_t *perm, const void *src, void *dst, int32_t elem_bytes, void *stream)
{
    OPX_REQUIRE(n >= 0 && n <= INT32_MAX, "particle count %lld out of range", (long long)n);
    if (n == 0)
        return OPX_OK;
    OPX_REQUIRE(perm && src && dst, "null array");
    OPX_REQUIRE(src != dst, "opx_gather: src and dst must differ");
    hipStream_t st = as_stream(stream);
    const dim3 grid(grid_for(n)), block(kBlock);
    switch (elem_bytes) {
    case 1:
        hipLaunchKernelGGL(k_gather<uint8_t>, grid, block, 0, st, n, perm, (const uint8_t *)src, (uint8_t *)dst);
        break;
    case 2:
        hipLaunchKernelGGL(k_gather<uint16_t>, grid, block, 0, st, n, perm, (const uint16_t *)src, (uint16_t *)dst);
        break;
    case 4:
        hipLaunchKernelGGL(k_gather<uint32_t>, grid, block, 0, st, n, perm, (const uint32_t *)src, (uint32_t *)dst);
        break;
    case 8:
        hipLaunchKernelGGL(k_gather<uint64_t>, grid, block, 0, st, n, perm, (const uint64_t *)src, (uint64_t *)dst);
        break;
    default:
        OPX_REQUIRE(false, "opx_gather: element size %d not in {1, 2, 4, 8}", elem_bytes);
    }
    return check_launch("opx_gather");
}

} // extern "C"
