// crt.h -- drop-in for the reference's include/crt.h (src/crt.cpp): host CRT.
#pragma once
#include <cstdint>
#include <vector>

#include "crt_utils.h"

struct CRT2Params {
    TestDataTypeUint p1;
    TestDataTypeUint p2;
    TestDataTypeUint p1_inv_mod_p2;
    unsigned __int128 modulus;
    CRT2Params(TestDataTypeUint a, TestDataTypeUint b)
        : p1(a), p2(b), p1_inv_mod_p2((TestDataTypeUint)modinv_u64(a % b, b)),
          modulus((unsigned __int128)a * b) {}
};

// x in [0, p1 p2) with x = a mod p1, x = b mod p2.
inline unsigned __int128 crt_combine_2(const CRT2Params& c, TestDataTypeUint a, TestDataTypeUint b) {
    const TestDataTypeUint t0 = b >= (TestDataTypeUint)(a % c.p2) ? b - (TestDataTypeUint)(a % c.p2)
                                                                  : b + c.p2 - (TestDataTypeUint)(a % c.p2);
    const TestDataTypeUint t = (TestDataTypeUint)(((unsigned __int128)t0 * c.p1_inv_mod_p2) % c.p2);
    return (unsigned __int128)a + (unsigned __int128)c.p1 * t;
}

// Garner over k primes; valid while prod(p) < 2^128 (as in the reference).
inline unsigned __int128 crt_combine_many(const std::vector<TestDataTypeUint>& primes,
                                          const std::vector<TestDataTypeUint>& residues) {
    unsigned __int128 M = 1, x = 0;
    for (size_t i = 0; i < primes.size(); i++) {
        const TestDataTypeUint p = primes[i], r = residues[i];
        const TestDataTypeUint xm = (TestDataTypeUint)(x % p);
        const TestDataTypeUint t = r >= xm ? r - xm : r + (p - xm);
        const TestDataTypeUint inv = (TestDataTypeUint)modinv_u64((uint64_t)(M % p), p);
        const TestDataTypeUint k = (TestDataTypeUint)(((unsigned __int128)t * inv) % p);
        x += M * k;
        M *= p;
    }
    return x;
}
