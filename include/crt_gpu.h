// crt_gpu.h -- drop-in for the reference's include/crt_gpu.h (src/crt_gpu.cu).
//
// compute_garner_params() fills the same constants as the reference (and
// prefix_M, which the reference leaves unwritten).  upload_garner_params() and
// upload_residue_ptrs() record the parameters and residue arrays for the next
// crt_combine_gpu*() call (the reference keeps them in process-global
// __constant__ memory; here they are a thread-local record, so threads with
// their own parameters do not interfere).  The Garner kernel itself is
// libbirdwing_hip.so's; the primes must be one of its prime sets.
#pragma once
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <hip/hip_runtime_api.h>

#include "birdwing.h"
#include "config.h"
#include "crt_utils.h"

struct CRTGarnerParams {
    uint64_t primes[NUM_MODULI];
    uint64_t inv[NUM_MODULI];       // inv[i] = (p0 ... p_{i-1})^-1 mod p_i
    uint64_t prefix_M[NUM_MODULI];  // (p0 ... p_{i-1}) mod p_i
    uint64_t M_mod_table[NUM_MODULI][NUM_MODULI];
    uint64_t barrett_m[NUM_MODULI];
};

inline CRTGarnerParams compute_garner_params(const std::vector<TestDataTypeUint>& primes) {
    assert(primes.size() == NUM_MODULI);
    CRTGarnerParams g{};
    for (int i = 0; i < NUM_MODULI; i++) {
        const uint64_t pi = primes[i];
        uint64_t pre = 1 % pi;
        for (int j = 0; j < i; j++) pre = (uint64_t)(((unsigned __int128)pre * (primes[j] % pi)) % pi);
        g.primes[i] = pi;
        g.prefix_M[i] = pre;
        g.inv[i] = i == 0 ? 1 : modinv_u64(pre, pi);
        for (int k = 0; k < NUM_MODULI; k++) {
            uint64_t mk = 1 % pi;
            for (int j = 0; j < k; j++) mk = (uint64_t)(((unsigned __int128)mk * (primes[j] % pi)) % pi);
            g.M_mod_table[k][i] = mk;
        }
        g.barrett_m[i] = pi < (1ULL << 32) ? (uint64_t)(((unsigned __int128)1 << 64) / pi) : 0;
    }
    return g;
}

namespace bw_dropin {
struct CrtState {
    CRTGarnerParams params{};
    bool have_params = false;
    const void* res[3] = {nullptr, nullptr, nullptr};
};
inline CrtState& crt_state() {
    static thread_local CrtState s;
    return s;
}
// prime set of libbirdwing_hip.so whose primes are exactly params.primes
inline int crt_set_of(const CRTGarnerParams& g) {
    for (int s = 0; s <= 3; s++) {
        int k = 0;
        uint64_t P[3] = {0, 0, 0};
        if (bw_prime_set_info(s, &k, P, nullptr) != BW_OK || k != NUM_MODULI) continue;
        bool same = true;
        for (int i = 0; i < k; i++) same = same && P[i] == g.primes[i];
        if (same) return s;
    }
    return -1;
}
}  // namespace bw_dropin

inline void upload_garner_params(const CRTGarnerParams& params) {
    bw_dropin::crt_state().params = params;
    bw_dropin::crt_state().have_params = true;
}

inline void upload_residue_ptrs(const std::vector<TestDataTypeUint*>& c_dev) {
    for (int j = 0; j < NUM_MODULI && j < (int)c_dev.size(); j++) bw_dropin::crt_state().res[j] = c_dev[j];
}

// Residues layout [NUM_MODULI][N]; coefficient i = (C_hi[i] << 64) | C_lo[i].
inline void crt_combine_gpu(uint64_t* d_C_hi, uint64_t* d_C_lo, int N) {
    auto& st = bw_dropin::crt_state();
    const int set = bw_dropin::crt_set_of(st.params);
    if (!st.have_params || set < 0 || (NUM_MODULI != 3 && set != BW_SET_HYBRID)) {
        std::fprintf(stderr, "[CRT] Garner parameters must be one of the library's 3-prime sets or the hybrid pair\n");
        std::exit(1);
    }
    if (bw_crt_combine(set, st.res[0], st.res[1], st.res[2], (uint64_t)N, d_C_lo, d_C_hi, nullptr, nullptr) != BW_OK) {
        std::fprintf(stderr, "[CRT] %s\n", bw_last_error());
        std::exit(1);
    }
    (void)hipDeviceSynchronize();
}

#if defined(NATIVE_HOST_LIMBS)
// U160 output: x = lo | (mid << 64) | ((uint64_t)hi << 128)
inline void crt_combine_gpu_u160(uint64_t* d_C_lo, uint64_t* d_C_mid, uint32_t* d_C_hi, int N) {
    auto& st = bw_dropin::crt_state();
    const int set = bw_dropin::crt_set_of(st.params);
    if (!st.have_params || set < 0) {
        std::fprintf(stderr, "[CRT] Garner parameters must be one of the library's 3-prime sets\n");
        std::exit(1);
    }
    if (bw_crt_combine(set, st.res[0], st.res[1], st.res[2], (uint64_t)N, d_C_lo, d_C_mid, d_C_hi, nullptr) != BW_OK) {
        std::fprintf(stderr, "[CRT] %s\n", bw_last_error());
        std::exit(1);
    }
    (void)hipDeviceSynchronize();
}
#endif
