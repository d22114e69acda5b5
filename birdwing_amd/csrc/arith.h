// arith.h -- modular arithmetic used by the gfx950 kernels.
//
// u32 digit path (primes p = c*2^m + 1 < 2^31): residues are held in FP64
// registers as signed integers.  FP64 FMA is full rate on CDNA while 32-bit
// integer high multiplies are not, and the FP64 form needs no lazy-reduction
// corrections: a product y*w with |y| < 2^45 is reduced to (-p, p) by
//
//     q = rint(y * (w/p))           (one FMA against 1.5*2^52, one subtract)
//     r = fma(y, w, -q*(p-1)) - q   (= y*w - q*p exactly)
//
// fma(y, w, -q*(p-1)) is exact because q*(p-1) = (q*c)*2^m has a significand
// of at most 53 bits (q < 2^46, c < 2^7) and the exact result y*w - q*c*2^m =
// r + q is an integer below 2^53.  Sums of such values grow by at most p per
// radix-2 level, so a 2^26-point transform stays far below 2^45.
//
// u64 digit path (58-bit primes): 64-bit integers with Shoup's precomputed-
// quotient product (Shoup; Harvey 2014) for known twiddles and Montgomery's
// product (Montgomery 1985) for the pointwise step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BW_MAGIC_RND 6755399441055744.0  // 1.5 * 2^52: adding it rounds |x| < 2^51 to an integer

struct F64Prime {
    double p;      // prime
    double pinv;   // 1/p (rounded)
    double cm;     // p - 1 = c * 2^m
};

__device__ __forceinline__ double f64_round(double x) {
    return __builtin_fma(x, 1.0, BW_MAGIC_RND) - BW_MAGIC_RND;
}

// y * w mod p, result in (-p, p) (centered to within a unit).  wp = w/p.
__device__ __forceinline__ double f64_mulmod(double y, double w, double wp, const F64Prime& P) {
    double q = __builtin_fma(y, wp, BW_MAGIC_RND) - BW_MAGIC_RND;
    double t = q * P.cm;
    return __builtin_fma(y, w, -t) - q;
}

// Reduce |x| < 2^51 to a centered representative in [-p/2-1, p/2+1].
__device__ __forceinline__ double f64_reduce(double x, const F64Prime& P) {
    double q = __builtin_fma(x, P.pinv, BW_MAGIC_RND) - BW_MAGIC_RND;
    return __builtin_fma(-q, P.p, x);
}

// Canonical representative in [0, p).
__device__ __forceinline__ uint32_t f64_canon(double x, const F64Prime& P) {
    double r = f64_reduce(x, P);
    r = r < 0.0 ? r + P.p : r;
    r = r >= P.p ? r - P.p : r;
    return (uint32_t)r;
}

// ---------------------------------------------------------------------------
// 64-bit path.
struct U64Prime {
    uint64_t p;
    uint64_t pinv_neg;   // -p^-1 mod 2^64 (Montgomery)
    uint64_t r2;         // 2^128 mod p (to enter Montgomery form)
    uint64_t floor_r;    // floor(2^64 / p) (ingress reduction)
};

__device__ __forceinline__ uint64_t u64_mulhi(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

// a*w mod p with wq = floor(w * 2^64 / p); a < 2^64, w < p < 2^62 -> [0, p)
__device__ __forceinline__ uint64_t u64_mulmod_shoup(uint64_t a, uint64_t w, uint64_t wq, uint64_t p) {
    uint64_t q = u64_mulhi(a, wq);
    uint64_t r = a * w - q * p;
    return r >= p ? r - p : r;
}

// Montgomery product a*b*2^-64 mod p for a, b < p < 2^62 -> [0, p)
__device__ __forceinline__ uint64_t u64_montmul(uint64_t a, uint64_t b, const U64Prime& P) {
    uint64_t lo = a * b, hi = u64_mulhi(a, b);
    uint64_t m = lo * P.pinv_neg;
    uint64_t mh = u64_mulhi(m, P.p);
    // (lo + m*p) is divisible by 2^64; its low word is 0 with a carry out iff lo != 0
    uint64_t r = hi + mh + (lo != 0 ? 1ULL : 0ULL);
    return r >= P.p ? r - P.p : r;
}

// x mod p for any 64-bit x (p < 2^62): q = mulhi(x, floor(2^64/p)) undershoots by < 4.
__device__ __forceinline__ uint64_t u64_reduce(uint64_t x, const U64Prime& P) {
    uint64_t q = u64_mulhi(x, P.floor_r);
    uint64_t r = x - q * P.p;
    while (r >= P.p) r -= P.p;
    return r;
}

__device__ __forceinline__ uint64_t u64_add(uint64_t a, uint64_t b, uint64_t p) {
    uint64_t s = a + b;
    return s >= p ? s - p : s;
}
__device__ __forceinline__ uint64_t u64_sub(uint64_t a, uint64_t b, uint64_t p) {
    return a >= b ? a - b : a + p - b;
}
