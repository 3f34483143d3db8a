// Scalar conversions between FP32 and the 16-bit storage formats (bf16 with
// round-to-nearest-even, fp16 through the hardware conversion).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bem_lowp.hpp"

namespace bem {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_f16(float f)
{
    _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

template <int DT>
__device__ __forceinline__ uint16_t to_lp(float f)
{
    return DT == kDtBF16 ? f32_to_bf16_rne(f) : f32_to_f16(f);
}
template <int DT>
__device__ __forceinline__ float from_lp(uint16_t h)
{
    return DT == kDtBF16 ? bf16_to_f32(h) : f16_to_f32(h);
}

}  // namespace bem
