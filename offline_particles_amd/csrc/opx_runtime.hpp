// Host-side helpers shared by the libopx translation units: thread-local error
// message, argument checks and launch-error reporting.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "opx.h"

namespace opx {

// Defined in opx_plain.hip.
void set_error(const char *fmt, ...);

#define OPX_REQUIRE(cond, ...)                 \
    do {                                       \
        if (!(cond)) {                         \
            ::opx::set_error(__VA_ARGS__);     \
            return OPX_ERR_INVALID;            \
        }                                      \
    } while (0)

inline int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: HIP launch failed: %s", what, hipGetErrorString(e));
        return OPX_ERR_HIP;
    }
    return OPX_OK;
}

inline int dtype_size(int dtype)
{
    switch (dtype) {
    case OPX_F64: return 8;
    case OPX_F32: return 4;
    case OPX_BF16: return 2;
    default: return 0;
    }
}

// Validates one field descriptor: data pointer, dtype, ndim and sizes >= 2
// (the interpolation clamps base indices to [0, size-2]).
inline int check_field(const opx_field &f, int ndim, const char *name, bool allow_time)
{
    OPX_REQUIRE(f.prev != nullptr, "field '%s': null data pointer", name);
    OPX_REQUIRE(dtype_size(f.dtype) != 0, "field '%s': unknown dtype %d", name, f.dtype);
    OPX_REQUIRE(f.ndim == ndim, "field '%s': expected %d active dims, got %d", name, ndim, f.ndim);
    for (int d = 0; d < ndim; ++d)
        OPX_REQUIRE(f.shape[d] >= 2, "field '%s': dimension %d has size %lld < 2", name, d,
                    (long long)f.shape[d]);
    OPX_REQUIRE(allow_time || f.next == nullptr, "field '%s' must be static", name);
    if (f.packed != nullptr) {
        OPX_REQUIRE(f.dtype == OPX_F32 || f.dtype == OPX_BF16, "field '%s': packed stencils need F32 or BF16", name);
        OPX_REQUIRE(ndim >= 2, "field '%s': packed stencils need 2 or 3 dimensions", name);
        OPX_REQUIRE(f.next == nullptr ? f.packed_kind == OPX_PACK_QUAD
                                      : (f.packed_kind == OPX_PACK_OCT || f.packed_kind == OPX_PACK_TPAIR ||
                                         (f.packed_kind == OPX_PACK_HEX && ndim == 3)),
                    "field '%s': packed kind %d does not match a %s field", name, f.packed_kind,
                    f.next == nullptr ? "static" : "time-dependent");
        // QUAD elements are 16 (F32) or 8 (BF16) bytes, OCT 32 or 16, read in 16-byte loads
        const int quad = f.dtype == OPX_F32 ? 16 : 8;
        const int unit = f.packed_kind == OPX_PACK_QUAD ? quad : 16;
        const int size = f.packed_kind == OPX_PACK_QUAD ? quad : 2 * quad;
        const bool strided = f.packed_kind == OPX_PACK_QUAD || f.packed_kind == OPX_PACK_OCT;
        OPX_REQUIRE(f.packed_stride == 0 || (strided && f.packed_stride % unit == 0 && f.packed_stride >= size &&
                                             f.packed_stride <= 256),
                    "field '%s': packed_stride %d is not 0 or a QUAD / OCT stride (a multiple of %d from %d to 256)",
                    name, f.packed_stride, unit, size);
        OPX_REQUIRE(!strided || reinterpret_cast<uintptr_t>(f.packed) % unit == 0,
                    "field '%s': packed copy not %d-byte aligned", name, unit);
    } else {
        OPX_REQUIRE(f.packed_stride == 0, "field '%s': packed_stride without a packed copy", name);
    }
    return OPX_OK;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

} // namespace opx
