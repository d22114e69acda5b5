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

// y * w mod p straight to a 32-bit word in (-p, p): four FP64 operations and
// no conversion.  With M = 1.5*2^52:
//   qr = fma(y, -w/p, M)         = M + q, q = rint(-y*w/p)
//   t  = fma(qr, p-1, -M*(p-1))  = q*(p-1) exactly (q*c < 2^53, M*(p-1) exact)
//   u  = fma(y, -w, -t)          = r + q with r = -y*w - q*p, exactly as above
//   v  = u - qr                  = r - M, exactly (an integer in (-2^53, -2^52))
// v's low 32 significand bits are (2^51 - r) mod 2^32 = -r mod 2^32, so the
// low word read as int32 is -r = y*w + q*p, congruent to y*w (|r| < p < 2^31).
// The negations are FMA source modifiers.  Same input range as f64_mulmod.
__device__ __forceinline__ int32_t f64_mulmod_i32(double y, double w, double wp, const F64Prime& P) {
    const double qr = __builtin_fma(y, -wp, BW_MAGIC_RND);
    const double t = __builtin_fma(qr, P.cm, -(BW_MAGIC_RND * P.cm));
    const double u = __builtin_fma(y, -w, -t);
    const double v = u - qr;
    return (int32_t)(uint32_t)__builtin_bit_cast(uint64_t, v);
}

// Reduce |x| < 2^51 to a centered representative in [-p/2-1, p/2+1].
__device__ __forceinline__ double f64_reduce(double x, const F64Prime& P) {
    double q = __builtin_fma(x, P.pinv, BW_MAGIC_RND) - BW_MAGIC_RND;
    return __builtin_fma(-q, P.p, x);
}

// Canonical representative in [0, p), as a double: the reduction leaves
// |r| <= p/2 + 1 (q is the rounded quotient up to an error far below one
// unit for |x| < 2^51), so adding p to a negative r is the only correction.
__device__ __forceinline__ double f64_canon_d(double x, const F64Prime& P) {
    const double r = f64_reduce(x, P);
    return r < 0.0 ? r + P.p : r;
}
__device__ __forceinline__ uint32_t f64_canon(double x, const F64Prime& P) { return (uint32_t)f64_canon_d(x, P); }

// Element i of an array no kernel writes, read through the constant address
// space (8-byte words): with a wave-uniform index the loads are scalar loads.
template <class T>
__device__ __forceinline__ T bw_ldconst(const T* p, int64_t i) {
    static_assert(sizeof(T) % 8 == 0, "8-byte words");
    const __attribute__((address_space(4))) uint64_t* q = (const __attribute__((address_space(4))) uint64_t*)(p + i);
    T r;
    uint64_t* d = reinterpret_cast<uint64_t*>(&r);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 8); k++) d[k] = q[k];
    return r;
}

// 32-bit words of a tile through a raw buffer resource on the tile's first
// element: each access adds a per-lane byte offset (VGPR) and a wave-uniform
// byte offset (SGPR) in the memory instruction itself, so a strided sequence
// of accesses costs no 64-bit vector address arithmetic.  Both offsets and
// their sum stay below 2^32 (the tile spans less than 4 GiB).  No range
// check: num_records is the maximum.
struct TileBuf {
    __amdgpu_buffer_rsrc_t r;
    __device__ explicit TileBuf(const void* base)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000)) {}
    __device__ uint32_t ld(uint32_t lane_off, uint32_t uni_off) const {
        return __builtin_amdgcn_raw_buffer_load_b32(r, lane_off, uni_off, 0);
    }
    __device__ uint64_t ld64(uint32_t lane_off, uint32_t uni_off) const {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, uni_off, 0);
        return (uint64_t)v[0] | ((uint64_t)v[1] << 32);
    }
    __device__ void st(uint32_t v, uint32_t lane_off, uint32_t uni_off) const {
        __builtin_amdgcn_raw_buffer_store_b32(v, r, lane_off, uni_off, 0);
    }
};

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
