// carry_prop.h -- drop-in for the reference's include/carry_prop.h
// (src/carry_prop.cu).
//
// Host entry points (any C++ compiler): carry_propagate_gpu() and, on the
// 64-bit pipeline, carry_propagate_gpu_u160() turn coefficient arrays into
// limbs with libbirdwing_hip.so's single-pass carry scan.
//
// HIP translation units (__HIPCC__) also get the reference's three-pass
// __global__ kernels under their reference names and signatures, for callers
// that launch them directly.  Their semantics follow the reference except that
// carry_fixup_kernel* consume (clear) their incoming carry, so a second fixup
// launch adds only carries that escaped in the previous one (the reference
// re-adds every incoming carry on each fixup launch).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <cstdio>
#include <cstdlib>

#include "birdwing.h"
#include "config.h"

// 128-bit coefficients (hi, lo) -> n_out output limbs (OUTPUT_LIMB_BITS wide).
inline void carry_propagate_gpu(const uint64_t* d_C_hi, const uint64_t* d_C_lo, size_t N, OutputLimbType* d_out,
                                size_t n_out, void* stream = nullptr) {
    if (bw_carry128(d_C_hi, d_C_lo, N, d_out, n_out, OUTPUT_LIMB_BITS, stream) != BW_OK) {
        std::fprintf(stderr, "[carry] %s\n", bw_last_error());
        std::exit(1);
    }
}

#if defined(NATIVE_HOST_LIMBS)
inline void carry_propagate_gpu_u160(const uint64_t* d_C_lo, const uint64_t* d_C_mid, const uint32_t* d_C_hi,
                                     size_t N, OutputLimbType* d_out, size_t n_out, void* stream = nullptr) {
    if (bw_carry160(d_C_lo, d_C_mid, d_C_hi, N, d_out, n_out, stream) != BW_OK) {
        std::fprintf(stderr, "[carry] %s\n", bw_last_error());
        std::exit(1);
    }
}
#endif

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

// Pass 1: serial carry within each CARRY_SEG segment; segment carry-out stored.
static __global__ void carry_intra_segment_kernel(const uint64_t* __restrict__ C_hi, const uint64_t* __restrict__ C_lo,
                                                  OutputLimbType* __restrict__ out, int64_t* __restrict__ seg_carry,
                                                  size_t N) {
    if (threadIdx.x != 0) return;
    const size_t s0 = (size_t)blockIdx.x * CARRY_SEG, s1 = s0 + CARRY_SEG < N ? s0 + CARRY_SEG : N;
    unsigned __int128 c = 0;
    for (size_t i = s0; i < s1; i++) {
        const unsigned __int128 t = (((unsigned __int128)C_hi[i] << 64) | C_lo[i]) + c;
        out[i] = (OutputLimbType)(t & OUTPUT_LIMB_MASK);
        c = t >> OUTPUT_LIMB_BITS;
    }
    seg_carry[blockIdx.x] = (int64_t)c;
}

// Pass 2: exclusive shift of the segment carries.
static __global__ void carry_inter_segment_kernel(int64_t* __restrict__ seg_carry, size_t num_segs) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t run = 0;
    for (size_t s = 0; s < num_segs; s++) {
        const int64_t in = run;
        run = seg_carry[s];
        seg_carry[s] = in;
    }
}

// Pass 3: add the incoming carry (consumed); an overflow past the segment end
// is added to the next segment's carry and flags another launch.
static __global__ void carry_fixup_kernel(OutputLimbType* __restrict__ out, int64_t* __restrict__ seg_carry, size_t N,
                                          size_t num_segs, int* __restrict__ escape_flag) {
    if (threadIdx.x != 0) return;
    const size_t seg = blockIdx.x, s0 = seg * CARRY_SEG, s1 = s0 + CARRY_SEG < N ? s0 + CARRY_SEG : N;
    int64_t in = (int64_t)atomicExch(reinterpret_cast<unsigned long long*>(&seg_carry[seg]), 0ULL);
    if (in == 0) return;
    for (size_t i = s0; i < s1 && in != 0; i++) {
        const __int128 s = (__int128)out[i] + in;
        out[i] = (OutputLimbType)(s & OUTPUT_LIMB_MASK);
        in = (int64_t)(s >> OUTPUT_LIMB_BITS);
    }
    if (in != 0 && seg + 1 < num_segs) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&seg_carry[seg + 1]), (unsigned long long)in);
        *escape_flag = 1;
    }
}
#if defined(NATIVE_HOST_LIMBS)
// 160-bit variants: coefficient i = lo[i] | mid[i] << 64 | hi[i] << 128; the
// segment carries are held as (lo, mid, hi) triples.
static __device__ inline void bw_seg_u160_add(uint64_t* lo, uint64_t* mid, uint32_t* hi, size_t s,
                                              unsigned __int128 v) {
    const uint64_t vl = (uint64_t)v, vm = (uint64_t)(v >> 64);
    const uint64_t ol = atomicAdd(reinterpret_cast<unsigned long long*>(&lo[s]), (unsigned long long)vl);
    const uint64_t cm = vm + (ol + vl < ol ? 1 : 0);
    if (cm == 0) return;
    const uint64_t om = atomicAdd(reinterpret_cast<unsigned long long*>(&mid[s]), (unsigned long long)cm);
    if (om + cm < om) atomicAdd(&hi[s], 1u);
}

static __global__ void carry_intra_segment_kernel_u160(const uint64_t* __restrict__ C_lo,
                                                       const uint64_t* __restrict__ C_mid,
                                                       const uint32_t* __restrict__ C_hi,
                                                       OutputLimbType* __restrict__ out,
                                                       uint64_t* __restrict__ seg_carry_lo,
                                                       uint64_t* __restrict__ seg_carry_mid,
                                                       uint32_t* __restrict__ seg_carry_hi, size_t N) {
    if (threadIdx.x != 0) return;
    const size_t s0 = (size_t)blockIdx.x * CARRY_SEG, s1 = s0 + CARRY_SEG < N ? s0 + CARRY_SEG : N;
    // carry after emitting a 64-bit limb of (x + c), x < 2^160: c < 2^97, so
    // (c_hi_part, c_lo_part) fits 128 bits with the x's top word added apart.
    unsigned __int128 c = 0;
    for (size_t i = s0; i < s1; i++) {
        const unsigned __int128 low = (unsigned __int128)C_lo[i] + (uint64_t)c;
        out[i] = (OutputLimbType)(uint64_t)low;
        c = (c >> 64) + (low >> 64) + C_mid[i] + ((unsigned __int128)C_hi[i] << 64);
    }
    seg_carry_lo[blockIdx.x] = (uint64_t)c;
    seg_carry_mid[blockIdx.x] = (uint64_t)(c >> 64);
    seg_carry_hi[blockIdx.x] = 0;
}

static __global__ void carry_inter_segment_kernel_u160(uint64_t* __restrict__ seg_carry_lo,
                                                       uint64_t* __restrict__ seg_carry_mid,
                                                       uint32_t* __restrict__ seg_carry_hi, size_t num_segs) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t rl = 0, rm = 0;
    uint32_t rh = 0;
    for (size_t s = 0; s < num_segs; s++) {
        const uint64_t l = seg_carry_lo[s], m = seg_carry_mid[s];
        const uint32_t h = seg_carry_hi[s];
        seg_carry_lo[s] = rl;
        seg_carry_mid[s] = rm;
        seg_carry_hi[s] = rh;
        rl = l; rm = m; rh = h;
    }
}

static __global__ void carry_fixup_kernel_u160(OutputLimbType* __restrict__ out, uint64_t* __restrict__ seg_carry_lo,
                                               uint64_t* __restrict__ seg_carry_mid,
                                               uint32_t* __restrict__ seg_carry_hi, size_t N, size_t num_segs,
                                               int* __restrict__ escape_flag) {
    if (threadIdx.x != 0) return;
    const size_t seg = blockIdx.x, s0 = seg * CARRY_SEG, s1 = s0 + CARRY_SEG < N ? s0 + CARRY_SEG : N;
    const uint64_t l = atomicExch(reinterpret_cast<unsigned long long*>(&seg_carry_lo[seg]), 0ULL);
    const uint64_t m = atomicExch(reinterpret_cast<unsigned long long*>(&seg_carry_mid[seg]), 0ULL);
    const uint32_t h = atomicExch(&seg_carry_hi[seg], 0u);
    unsigned __int128 c = ((unsigned __int128)m << 64) | l;
    uint32_t top = h;  // carry bits 128..159
    for (size_t i = s0; i < s1 && (c != 0 || top != 0); i++) {
        const unsigned __int128 low = (unsigned __int128)(uint64_t)out[i] + (uint64_t)c;
        out[i] = (OutputLimbType)(uint64_t)low;
        c = (c >> 64) + (low >> 64) + ((unsigned __int128)top << 64);
        top = 0;
    }
    if ((c != 0 || top != 0) && seg + 1 < num_segs) {
        bw_seg_u160_add(seg_carry_lo, seg_carry_mid, seg_carry_hi, seg + 1, c);
        if (top) atomicAdd(&seg_carry_hi[seg + 1], top);
        *escape_flag = 1;
    }
}
#endif  // NATIVE_HOST_LIMBS
#endif  // __HIPCC__
