// misc_kernels.h -- elementwise kernels behind the stage-level entry points.
#pragma once
#include "bw_internal.h"

hipError_t bw_launch_zero_pad(const uint32_t* src, void* dst, int word_bytes, int64_t L, int64_t N, hipStream_t s);
hipError_t bw_launch_zero_pad_u64(const uint64_t* src, uint64_t* dst, int64_t L, int64_t N, const U64Prime& P,
                                  hipStream_t s);
hipError_t bw_launch_pointwise(int word_bytes, const void* a, const void* b, void* c, int64_t N, uint64_t p,
                               hipStream_t s);
// Multiply each of k planes by a per-prime constant: FP64 pairs (s, s/p) or
// u64 Montgomery constants (s 2^128 mod p).
hipError_t bw_launch_scale(int limb_bits, void* x, int64_t N, int64_t plane, int k, const void* consts,
                           const void* pairs, hipStream_t s);
