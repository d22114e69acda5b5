// ref_glue.cpp -- C-ABI shim over the reference's own CPU CRT sources
// (src/crt.cpp, src/crt_utils.cpp), compiled where they lie in the reference
// tree by oracle/Makefile.  TEST INFRASTRUCTURE ONLY: used by tests/ to pin
// the oracle's Garner restatement against the reference's crt_combine_many
// (valid only while the prime product is below 2^128, src/crt.cpp:42-87).
#include "crt.h"
#include <cstdint>
#include <vector>

extern "C" {
int ref_limb_bits() { return LIMB_BITS; }

void ref_crt_combine_many(const uint64_t* primes, const uint64_t* residues, int k,
                          uint64_t* lo, uint64_t* hi) {
    std::vector<TestDataTypeUint> P(k), R(k);
    for (int i = 0; i < k; i++) { P[i] = (TestDataTypeUint)primes[i]; R[i] = (TestDataTypeUint)residues[i]; }
    unsigned __int128 x = crt_combine_many(P, R);
    *lo = (uint64_t)x;
    *hi = (uint64_t)(x >> 64);
}

void ref_crt_combine_2(uint64_t p1, uint64_t p2, uint64_t a, uint64_t b, uint64_t* lo, uint64_t* hi) {
    CRT2Params params((TestDataTypeUint)p1, (TestDataTypeUint)p2);
    unsigned __int128 x = crt_combine_2(params, (TestDataTypeUint)a, (TestDataTypeUint)b);
    *lo = (uint64_t)x;
    *hi = (uint64_t)(x >> 64);
}

uint64_t ref_modinv_u64(uint64_t a, uint64_t m) { return modinv_u64(a, m); }
}
