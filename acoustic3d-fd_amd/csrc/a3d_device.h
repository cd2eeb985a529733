// a3d_device.h -- device-side plane addressing shared by the sweep kernels:
// buffer-resource loads and stores with the plane base in scalar registers and
// a 32-bit in-plane offset per lane, and the cache-policy hints.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace a3d {

// Buffer addressing of element o of the plane at offset zo: the plane base
// p + zo goes into a buffer resource (scalar registers, rebuilt per plane) and
// the lane passes its in-plane byte offset, so a load or store is one
// instruction with no per-lane 64-bit address arithmetic and no address
// registers live across the plane loop. In-plane byte offsets fit 31 bits (the
// host bounds a plane's size). Resource word 3 (0x00020000) is the untyped
// 32-bit data format of gfx9-family buffer resources.
template <class T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const T* p, int zo) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(p + (ptrdiff_t)zo), 0, 0x7fffffff, 0x00020000);
}
// Cache policy A: 0 default; NT (non-temporal) for streams no other tile re-reads
// (pointwise fields, the z-window leads, every store), so the L2 keeps the rows
// that neighbouring tiles read as halos.
constexpr int NT = 2;
template <int A = 0, class T>
__device__ __forceinline__ T bld(const T* p, int zo, int o) {
    const __amdgpu_buffer_rsrc_t r = plane_rsrc(p, zo);
    if constexpr (sizeof(T) == 2) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, o * 2, 0, A));
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, o * 4, 0, A));
    } else {
        static_assert(sizeof(T) == 8, "2, 4 or 8-byte elements");
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, o * 8, 0, A));
    }
}
template <int A = 0, class T>
__device__ __forceinline__ void bst(T* p, int zo, int o, T v) {
    const __amdgpu_buffer_rsrc_t r = plane_rsrc(p, zo);
    if constexpr (sizeof(T) == 2) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, o * 2, 0, A);
    } else if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, o * 4, 0, A);
    } else {
        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, o * 8, 0, A);
    }
}
// Two floats as one 8-byte element (bld with an 8-byte element type).
typedef float f32x2 __attribute__((ext_vector_type(2)));
// Four floats at element o of the plane at offset zo (16-byte elements).
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <int A = 0>
__device__ __forceinline__ f32x4 bld4(const f32x4* p, int zo, int o) {
    return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(plane_rsrc(p, zo), o * 16, 0, A));
}

}  // namespace a3d
