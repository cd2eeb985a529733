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

// sin t and cos t in FP32 for the TTI rotation terms evaluated in the sweeps (kernel op 4):
// t = q pi/2 + r with q = rint(2t/pi) and r from a two-part pi/2 (float(pi/2) and the
// rest), |r| <= pi/4; sin r and cos r from their Taylor series to r^9 and r^10 (the next
// terms are below 2e-9 there), Horner form with fused multiply-adds; the quadrant q mod 4
// swaps and negates them. Max. abs. error 6.2e-8 (sin) and 6.9e-8 (cos) against the exact
// values for |t| <= 7, the same on the GPU and the CPU (float rounding alone: 3.0e-8).
__device__ __forceinline__ void rot_sincos(float t, float& s, float& c) {
    const float q = rintf(t * 0.6366197723675814f);
    float r = fmaf(-q, 1.5707963705062866f, t);
    r = fmaf(-q, -4.371139000186243e-8f, r);
    const float r2 = r * r;
    float ps = fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);   // 1/9!, -1/7!
    ps = fmaf(r2, ps, 8.3333333e-3f);                      // 1/5!
    ps = fmaf(r2, ps, -1.6666667e-1f);                     // -1/3!
    ps = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);   // -1/10!, 1/8!
    pc = fmaf(r2, pc, -1.3888889e-3f);                     // -1/6!
    pc = fmaf(r2, pc, 4.1666667e-2f);                      // 1/4!
    pc = fmaf(r2, pc, -0.5f);
    pc = fmaf(r2, pc, 1.0f);
    const int iq = (int)q & 3;
    const float a = (iq & 1) ? pc : ps, b = (iq & 1) ? ps : pc;
    s = (iq & 2) ? -a : a;
    c = ((iq + 1) & 2) ? -b : b;
}

}  // namespace a3d
