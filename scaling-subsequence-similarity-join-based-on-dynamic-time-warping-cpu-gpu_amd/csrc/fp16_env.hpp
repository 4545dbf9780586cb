// Packed FP16 envelope widths: one 32-bit word per sample holds wU = U - x (low half) and
// wL = x - L (high half), each rounded up to FP16, so that x + wU >= U and x - wL <= L after
// FP32 rounding.  The decode is x + wU and x - wL as single mixed-precision FMAs: the FP16
// operand is widened exactly and the sum rounded once, the same value as converting the half and
// adding (checked for every FP16 pattern by the fp16_env_check tool), in one instruction each.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dtwj {

__device__ __forceinline__ void fp16_env_decode(float x, uint32_t wb, float &up, float &lo) {
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(up) : "v"(wb), "v"(x));
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(lo) : "v"(wb), "v"(x));
}

// The same in plain arithmetic (reference for the check).
__device__ __forceinline__ void fp16_env_decode_ref(float x, uint32_t wb, float &up, float &lo) {
    up = x + (float)__builtin_bit_cast(_Float16, (uint16_t)(wb & 0xffffu));
    lo = x - (float)__builtin_bit_cast(_Float16, (uint16_t)(wb >> 16));
}

}  // namespace dtwj
