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
