// wide_int.h -- drop-in for the reference's include/wide_int.h: the U160 value
// (bits [0,64), [64,128), [128,160)) and its host helpers.
#pragma once
#include <cstdint>

struct U160 {
    uint64_t lo;
    uint64_t mid;
    uint32_t hi;
};

inline uint64_t u160_to_u64_lo(const U160& x) { return x.lo; }
inline uint64_t u160_to_u64_mid(const U160& x) { return x.mid; }
inline uint32_t u160_to_u32_hi(const U160& x) { return x.hi; }
inline bool u160_eq(const U160& a, const U160& b) { return a.lo == b.lo && a.mid == b.mid && a.hi == b.hi; }
inline void u160_from_u128(uint64_t lo, uint64_t hi, U160& out) { out.lo = lo; out.mid = hi; out.hi = 0; }

// a += b (mod 2^160)
inline void add160_host(U160& a, uint64_t b_lo, uint64_t b_mid, uint32_t b_hi) {
    const unsigned __int128 s0 = (unsigned __int128)a.lo + b_lo;
    const unsigned __int128 s1 = (unsigned __int128)a.mid + b_mid + (uint64_t)(s0 >> 64);
    a.lo = (uint64_t)s0;
    a.mid = (uint64_t)s1;
    a.hi = (uint32_t)(a.hi + b_hi + (uint32_t)(s1 >> 64));
}

// (M_hi:M_lo) * k -> U160 (mod 2^160)
inline void mul128x64_host(uint64_t M_hi, uint64_t M_lo, uint64_t k, uint64_t& out_lo, uint64_t& out_mid,
                           uint32_t& out_hi) {
    const unsigned __int128 p0 = (unsigned __int128)M_lo * k;
    const unsigned __int128 p1 = (unsigned __int128)M_hi * k + (uint64_t)(p0 >> 64);
    out_lo = (uint64_t)p0;
    out_mid = (uint64_t)p1;
    out_hi = (uint32_t)(p1 >> 64);
}
