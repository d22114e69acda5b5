// zero_pad.h -- drop-in for the reference's include/zero_pad.h
// (src/zero_pad.cu), launched through libbirdwing_hip.so.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <hip/hip_runtime_api.h>

#include "birdwing.h"
#include "config.h"

// d_dst[i] = i < L ? d_src[i] : 0 for i < N, widened to the NTT word, no reduction.
inline void zero_pad_gpu(const uint32_t* d_src, TestDataTypeUint* d_dst, size_t L, size_t N, hipStream_t stream) {
    if (bw_zero_pad(d_src, d_dst, (int)sizeof(TestDataTypeUint), L, N, (void*)stream) != BW_OK) {
        std::fprintf(stderr, "[zero_pad] %s\n", bw_last_error());
        std::exit(1);
    }
}

#if defined(NATIVE_HOST_LIMBS)
// 64-bit limbs reduced mod `modulus`.
inline void zero_pad_gpu_u64(const uint64_t* d_src, TestDataTypeUint* d_dst, size_t L, size_t N,
                             TestDataTypeUint modulus, hipStream_t stream) {
    if (bw_zero_pad_u64(d_src, (uint64_t*)d_dst, L, N, modulus, (void*)stream) != BW_OK) {
        std::fprintf(stderr, "[zero_pad] %s\n", bw_last_error());
        std::exit(1);
    }
}
#endif
