// Buffer-resource access for the stepper kernels (gfx950 raw buffer loads/stores).
//
// A raw buffer access whose byte offset lies past the resource's size reads 0 and drops a
// write.  The kernels use that as their bounds check: an element outside the grid gets the
// index Z_OOB instead of a branch around the access, so the loads of a phase stay in one basic
// block and are issued together.  Field sizes are bounded at context creation (nx, ny <= 512,
// nz <= 32: at most 2^23 elements), so every real byte offset and 8 * Z_OOB fit the 31-bit offset.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hqt {

typedef __amdgpu_buffer_rsrc_t Rs;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t Z_OOB = 0x0FFFFFFFu;      // element index past every field

__device__ __forceinline__ Rs rs_make(const void* p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
// element e of a field of type T (read as T)
template <typename T> __device__ __forceinline__ T bldr(Rs r, uint32_t e) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(e * 8u), 0, 0));
    else
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(e * 4u), 0, 0));
}
// element e of a field of type T, widened to double
template <typename T> __device__ __forceinline__ double bld(Rs r, uint32_t e) { return (double)bldr<T>(r, e); }
template <typename T> __device__ __forceinline__ void bst(Rs r, uint32_t e, double v) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)(e * 8u), 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)v), r, (int)(e * 4u), 0, 0);
}
template <typename T> __device__ __forceinline__ void bstr(Rs r, uint32_t e, T v) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)(e * 8u), 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)(e * 4u), 0, 0);
}

}  // namespace hqt
