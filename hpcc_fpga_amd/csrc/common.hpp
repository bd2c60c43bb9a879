// Shared device helpers for the gfx950 LINPACK kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// v_mfma_f32_32x32x2_f32: D(32x32) += A(32x2) * B(2x32).
// Operand layout (wave64): lane l holds A[l%32][l/32] and B[l/32][l%32];
// accumulator register r of lane l holds D[row][col] with
// col = l%32, row = 8*(r/4) + 4*(l/32) + r%4.
__device__ __forceinline__ f32x16 mfma32(float a, float b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ int acc_row(int r, int lane) {
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// Ordering of LDS traffic between lanes of one wavefront (no hardware barrier
// needed when only one wave touches the region).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// splitmix64 finaliser and the counter-based uniform used by the generators.
__host__ __device__ __forceinline__ uint64_t hpl_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t hpl_stream_key(uint64_t seed, uint64_t stream) {
    return hpl_mix64(seed ^ (stream << 56));
}
__host__ __device__ __forceinline__ float hpl_u24(uint64_t key, uint64_t k) {
    uint64_t z = hpl_mix64(key + (k + 1) * 0x9E3779B97F4A7C15ULL);
    return (float)(uint32_t)(z >> 40) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// FP8 E4M3 (OCP "fn" variant: bias 7, no infinities, max 448, subnormals
// m * 2^-9), the operand format of gfx950's FP8 MFMA.
struct fp8e4m3 {
    uint8_t bits;
};

// Code of x rounded to nearest, ties to even; |x| <= 448 is the caller's job
// (the per-block power-of-two scale guarantees it), larger values saturate.
__device__ __forceinline__ uint32_t fp8_e4m3_code(float x) {
    const uint32_t u = __float_as_uint(x);
    const uint32_t s = (u >> 24) & 0x80u;
    const uint32_t ua = u & 0x7fffffffu;
    if (ua < 0x3c800000u)  // |x| < 2^-6: subnormal grid of step 2^-9 (8 = 2^-6 itself)
        return s | (uint32_t)__builtin_rintf(__uint_as_float(ua) * 512.0f);
    const uint32_t r = (ua + 0x7ffffu + ((ua >> 20) & 1u)) >> 20;  // 8-bit exponent : 3 bits
    uint32_t code = r - ((127u - 7u) << 3);
    if (code > 0x7eu) code = 0x7eu;
    return s | code;
}

// Exponent e of the power-of-two block scale 2^e: the smallest e with
// amax <= 448 * 2^e (0 for an all-zero block).
__device__ __forceinline__ int fp8_scale_exp(float amax) {
    if (!(amax > 0.0f)) return 0;
    const uint32_t u = __float_as_uint(amax);
    int k = (int)((u >> 23) & 0xffu) - 127;
    if (k < -118) k = -118;
    return k - 8 + ((u & 0x7fffffu) > 0x600000u ? 1 : 0);
}

__device__ __forceinline__ float exp2i(int e) { return __uint_as_float((uint32_t)(e + 127) << 23); }
