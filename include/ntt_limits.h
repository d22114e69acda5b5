// ntt_limits.h -- drop-in for the reference's include/ntt_limits.h
// (src/ntt_limits.cpp:125-280), answered by libbirdwing_hip.so's size policy.
//
// By default every function answers exactly as the reference does for the
// configuration it is compiled in, from that configuration's own prime set
// (src/gpu_ntt.cu:28-47) and limb width:
//   -DLIMB_BITS=32                      3 primes near 2^30: max_supported_logN() == 23,
//                                       max_supported_limb_count() == 2^22
//   -DLIMB_BITS=64 (hybrid)             the first two 58-bit primes, 32-bit limbs:
//                                       33 and 2^32
//   -DLIMB_BITS=64 -DNATIVE_HOST_LIMBS  3 58-bit primes, 64-bit limbs: 32 and 2^31
// so a caller's size checks, and the reference's own tests/test_ntt_limits.cpp,
// see the reference's integers.
//
// Define BW_EXTENDED_SIZES before including this header to answer instead for
// the library's automatic prime choice, which reaches further: for 32-bit
// limbs N <= 2^26 under three primes and, as limb pairs under the 58-bit
// primes, 2^32 limbs per operand (max_supported_logN() == 33); for 64-bit
// limbs the reference 58-bit set.  host_multiply_merge* (multiply.h) check
// sizes with ensure_multiply_size_supported(), so they accept exactly the
// sizes the chosen policy accepts.
//
// The per-prime helpers are exact restatements.  ensure_*() print to stderr
// and exit(1), as in the reference.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "birdwing.h"
#include "config.h"

namespace bw_dropin {
#ifdef BW_EXTENDED_SIZES
// 32-bit limbs: lengths beyond the 32-bit-digit sets run as limb pairs
// (half as many 64-bit digits) under BW_SET_U64_PAIRS
inline int widest_set() { return BW_CFG_LIMB_BITS == 64 ? BW_SET_U64_REF : BW_SET_U64_PAIRS; }
inline bool paired(size_t N) { return BW_CFG_LIMB_BITS == 32 && N > ((size_t)1 << 26); }
inline int set_for(size_t N) { return paired(N) ? BW_SET_U64_PAIRS : bw_default_prime_set(BW_CFG_LIMB_BITS, N); }
#else
// the configuration's own prime set (src/gpu_ntt.cu:28-47, include/config.h:17-53)
inline int policy_set() {
#if LIMB_BITS == 64 && defined(NATIVE_HOST_LIMBS)
    return BW_SET_U64_REF;
#elif LIMB_BITS == 64
    return BW_SET_HYBRID;
#else
    return BW_SET_U32_REF;
#endif
}
#endif
}  // namespace bw_dropin

inline int max_logN_for_prime(uint64_t p) {
    int k = 0;
    uint64_t n = p - 1;
    while (n && !(n & 1)) { n >>= 1; k++; }
    return k;
}

inline int max_root_logN(TestDataTypeUint root, TestDataTypeUint p) {
    unsigned __int128 r = root % p;
    if (r == 0) return 0;
    for (int k = 0; k < 62; k++) {
        if (r == 1) return k;
        r = r * r % p;
    }
    return 62;
}

inline int max_supported_logN() {
#ifdef BW_EXTENDED_SIZES
    return bw_max_supported_logN(bw_dropin::widest_set()) + (BW_CFG_LIMB_BITS == 32 ? 1 : 0);   // limb positions
#else
    return bw_max_supported_logN(bw_dropin::policy_set());
#endif
}
inline size_t max_supported_N() { return size_t(1) << max_supported_logN(); }
inline size_t max_supported_limb_count() {
#ifdef BW_EXTENDED_SIZES
    return (size_t)bw_max_supported_limb_count(bw_dropin::widest_set(), INPUT_LIMB_BITS);
#else
    return (size_t)bw_max_supported_limb_count(bw_dropin::policy_set(), INPUT_LIMB_BITS);
#endif
}
inline size_t padded_ntt_size(size_t L_A, size_t L_B) { return (size_t)bw_padded_ntt_size(L_A, L_B); }

inline bool ntt_size_supported(size_t N, std::string* why = nullptr) {
    char buf[256] = {0};
#ifdef BW_EXTENDED_SIZES
    const int ok = (N && !(N & (N - 1)))
                       ? bw_ntt_size_supported(bw_dropin::set_for(N), bw_dropin::paired(N) ? N / 2 : N, buf, sizeof buf)
                       : bw_ntt_size_supported(BW_SET_U32_REF, N, buf, sizeof buf);
#else
    const int ok = bw_ntt_size_supported(bw_dropin::policy_set(), N, buf, sizeof buf);
#endif
    if (!ok && why) *why = buf;
    return ok != 0;
}

inline bool crt_coefficient_bound_satisfied(size_t L_A, size_t L_B, std::string* why = nullptr) {
    char buf[256] = {0};
#ifdef BW_EXTENDED_SIZES
    const int ok = bw_crt_bound_satisfied(bw_dropin::set_for(padded_ntt_size(L_A ? L_A : 1, L_B ? L_B : 1)),
                                          INPUT_LIMB_BITS, L_A, L_B, buf, sizeof buf);
#else
    const int ok = bw_crt_bound_satisfied(bw_dropin::policy_set(), INPUT_LIMB_BITS, L_A, L_B, buf, sizeof buf);
#endif
    if (!ok && why) *why = buf;
    return ok != 0;
}

inline bool multiply_size_supported(size_t L_A, size_t L_B, std::string* why = nullptr) {
    char buf[256] = {0};
#ifdef BW_EXTENDED_SIZES
    const int ok = bw_multiply_size_supported(-1, INPUT_LIMB_BITS, L_A, L_B, buf, sizeof buf);
#else
    const int ok = bw_multiply_size_supported(bw_dropin::policy_set(), INPUT_LIMB_BITS, L_A, L_B, buf, sizeof buf);
#endif
    if (!ok && why) *why = buf;
    return ok != 0;
}

inline void ensure_ntt_size_supported(size_t N) {
    std::string why;
    if (!ntt_size_supported(N, &why)) {
        std::fprintf(stderr, "[NTT] Unsupported size: %s\n[NTT] Max supported N=2^%d (%zu), max L=2^%d (%zu)\n",
                     why.c_str(), max_supported_logN(), max_supported_N(), max_supported_logN() - 1,
                     max_supported_limb_count());
        std::exit(1);
    }
}

inline void ensure_multiply_size_supported(size_t L_A, size_t L_B) {
    std::string why;
    if (!multiply_size_supported(L_A, L_B, &why)) {
        std::fprintf(stderr, "[NTT] Unsupported multiply: %s\n[NTT] Max supported N=2^%d (%zu), max L (square)=2^%d (%zu)\n",
                     why.c_str(), max_supported_logN(), max_supported_N(), max_supported_logN() - 1,
                     max_supported_limb_count());
        std::exit(1);
    }
}
