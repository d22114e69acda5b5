// crt_utils.h -- drop-in for the reference's include/crt_utils.h
// (src/crt_utils.cpp): 128-bit printing and the modular inverse.
#pragma once
#include <cstdint>
#include <cstdio>

#include "config.h"

// Decimal print of an unsigned 128-bit value.
inline void print_u128(unsigned __int128 x) {
    char buf[48];
    int i = 47;
    buf[i--] = '\0';
    if (x == 0) { std::printf("0"); return; }
    while (x > 0) {
        buf[i--] = (char)('0' + (unsigned)(x % 10));
        x /= 10;
    }
    std::printf("%s", &buf[i + 1]);
}

// a^-1 mod m by the extended Euclidean algorithm (a, m coprime); 128-bit
// intermediates so every 64-bit modulus is safe.
inline uint64_t modinv_u64(uint64_t a, uint64_t m) {
    __int128 t0 = 0, t1 = 1, r0 = m, r1 = a % m;
    while (r1 != 0) {
        const __int128 q = r0 / r1;
        __int128 tmp = r0 - q * r1; r0 = r1; r1 = tmp;
        tmp = t0 - q * t1; t0 = t1; t1 = tmp;
    }
    if (t0 < 0) t0 += m;
    return (uint64_t)t0;
}
