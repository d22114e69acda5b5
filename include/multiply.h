// multiply.h -- drop-in for the reference's include/multiply.h
// (src/host_multiply.cpp:27-134): one product of host limb vectors.
//
// C is resized to L_A + L_B limbs; `duration` is the wall time from plan
// creation (precompute + upload + allocate) to release, as in the reference.
// Unsupported sizes print to stderr and exit(1) (ensure_multiply_size_supported).
// The reference's timings_new.csv side effect is opt-in: define
// BW_WRITE_TIMINGS_CSV before including this header.
#pragma once
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "birdwing.h"
#include "config.h"
#include "ntt_limits.h"

using limb_t = OutputLimbType;

namespace bw_dropin {
template <class T>
inline void multiply_host(const std::vector<T>& A, const std::vector<T>& B, std::vector<T>& C, int limb_bits,
                          std::chrono::duration<double, std::milli>& duration) {
    ensure_multiply_size_supported(A.size(), B.size());
    C.assign(A.size() + B.size(), 0);
    double ms = 0.0;
    const auto t0 = std::chrono::high_resolution_clock::now();
    if (bw_multiply_host(limb_bits, A.data(), A.size(), B.data(), B.size(), C.data(), &ms) != BW_OK) {
        std::fprintf(stderr, "[NTT] multiply failed: %s\n", bw_last_error());
        std::exit(1);
    }
    duration = std::chrono::high_resolution_clock::now() - t0;
#ifdef BW_WRITE_TIMINGS_CSV
    if (FILE* f = std::fopen("timings_new.csv", "a")) {
        std::fseek(f, 0, SEEK_END);
        if (std::ftell(f) == 0) std::fprintf(f, "N,ntt_multiply_ms,carry_prop_ms\n");
        std::fprintf(f, "%zu,%f,%f\n", padded_ntt_size(A.size(), B.size()), duration.count(), 0.0);
        std::fclose(f);
    }
#endif
}
}  // namespace bw_dropin

#if !defined(NATIVE_HOST_LIMBS)
// 32-bit and hybrid pipelines (uint32_t host limbs)
inline void host_multiply_merge(const vector<uint32_t>& A, const vector<uint32_t>& B, vector<OutputLimbType>& C,
                                chrono::duration<double, milli>& duration) {
    bw_dropin::multiply_host(A, B, C, 32, duration);
}
#endif

#if defined(NATIVE_HOST_LIMBS)
// 64-bit pipeline (uint64_t host limbs)
inline void host_multiply_merge_native(const vector<uint64_t>& A, const vector<uint64_t>& B, vector<uint64_t>& C,
                                       chrono::duration<double, milli>& duration) {
    bw_dropin::multiply_host(A, B, C, 64, duration);
}
#endif
