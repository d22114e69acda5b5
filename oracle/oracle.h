/*
 * oracle.h -- CPU restatement of the reference big-integer multiply path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: it is the checker used by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg.  Every function cites the reference
 * file:line whose behaviour it restates (paths relative to the reference repo
 * nikitaborisov/birdwing).
 *
 * Pinning (what checks each part, see tests/test_oracle.py):
 *   mt64_*            -- C++ standard [rand.predef] known answer (10000th
 *                        output of mt19937_64 seeded 5489) and std::mt19937_64
 *                        compiled with g++.
 *   orc_* size policy -- the reference's own tests/test_ntt_limits.cpp
 *                        expectations (restated as pytest assertions), plus
 *                        the reference test binary built from its sources.
 *   orc_ntt_*         -- brute-force O(N^2) DFT for N <= 64, INTT(NTT(x)) = x,
 *                        and cyclic convolution == schoolbook mod p.
 *   orc_garner_*      -- reference KATs from tests/test_crt_gpu.cu and
 *                        tests/test_crt.cpp, uniqueness of x < prod(p), and
 *                        the reference's own crt_combine_many built from
 *                        src/crt.cpp (oracle/_ref) where the product < 2^128.
 *   orc_carry_*       -- reference KATs from tests/test_carry_prop.cu and
 *                        Python big-integer sums.
 *   orc_multiply      -- GMP mpz_mul and schoolbook, bit-exact.
 */
#ifndef BW_ORACLE_H
#define BW_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MT19937-64 (Matsumoto & Nishimura 2004; == std::mt19937_64) ---- */
typedef struct { uint64_t mt[312]; int mti; } orc_mt64;
void     orc_mt64_seed(orc_mt64* s, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64* s);
/* tests/test_full_multiply.cpp:61-73: i%8==0 -> 0, i%8==1 -> 1, else rng()%2^30 */
void orc_ref_random_limbs_u32(uint32_t* out, size_t n, uint64_t seed);
/* tests/test_full_multiply.cpp:431-441: i%8==0 -> 0, ==1 -> 1, ==2 -> 2^64-1, else rng() */
void orc_ref_random_limbs_u64(uint64_t* out, size_t n, uint64_t seed);
/* tests/test_pipeline_crt.cpp:41-47: rng() % 2^30 for every limb */
void orc_ref_pipeline_limbs_u32(uint32_t* out, size_t n, uint64_t seed);
/* Uniform synthetic limbs: low 32 bits (or all 64) of rng(); top limb |= 1 if 0. */
void orc_uniform_limbs(void* out, size_t n, int limb_bits, uint64_t seed);

/* ---- prime sets (src/gpu_ntt.cu:28-47; wide u32 set from tests/test_crt.cpp:19-21) ---- */
enum { ORC_SET_U32_REF = 0, ORC_SET_U32_WIDE = 1, ORC_SET_U64_REF = 2, ORC_SET_HYBRID = 3 };
int orc_prime_set(int set, uint64_t* primes, uint64_t* roots); /* returns k */

/* ---- size policy (src/ntt_limits.cpp) ---- */
int      orc_valuation2(uint64_t n);                                  /* :14-22 */
int      orc_max_logN_for_prime(uint64_t p);                          /* :24-27 */
int      orc_max_root_logN(uint64_t root, uint64_t p);                /* :37-56 */
int      orc_max_supported_logN(int set);                             /* :125-135 */
uint64_t orc_padded_ntt_size(uint64_t L_A, uint64_t L_B);             /* :145-152 */
/* min(L_A,L_B)*(2^in_bits-1)^2 < prod(p): :186-214 (exact, no long double) */
int      orc_crt_bound_ok(int set, int in_bits, uint64_t L_A, uint64_t L_B);
int      orc_ntt_size_supported(int set, uint64_t N);                 /* :216-239 */
int      orc_multiply_size_supported(int set, int in_bits, uint64_t L_A, uint64_t L_B); /* :241-254 */
uint64_t orc_max_limb_count_by_crt(int set, int in_bits);             /* :154-174 */
uint64_t orc_max_supported_limb_count(int set, int in_bits);          /* :176-184 */

/* ---- modular helpers ---- */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p);
uint64_t orc_modinv(uint64_t a, uint64_t m);                          /* src/crt_utils.cpp:29-51 */
/* omega_N = root_max^(2^(maxlog-logN)), maxlog = min(v2(p-1), ord(root)); src/gpu_ntt.cu:171-192 */
uint64_t orc_root_for_N(uint64_t p, uint64_t root_max, int logN);

/* ---- transforms: X[k] = sum_n x[n] w^(nk), natural order in and out ---- */
void orc_ntt_forward(uint64_t* x, uint64_t N, uint64_t p, uint64_t w);
void orc_ntt_inverse(uint64_t* x, uint64_t N, uint64_t p, uint64_t w); /* includes N^-1 */
void orc_dft_bruteforce(const uint64_t* x, uint64_t* X, uint64_t N, uint64_t p, uint64_t w);
/* src/gpu_ntt.cu:216-238 */
void orc_pointwise(const uint64_t* a, const uint64_t* b, uint64_t* c, uint64_t N, uint64_t p);
/* linear convolution mod p, truncated to N terms (tests/test_pipeline_crt.cpp:106-123) */
void orc_conv_schoolbook_mod(const uint64_t* a, uint64_t La, const uint64_t* b, uint64_t Lb,
                             uint64_t* c, uint64_t N, uint64_t p);

/* ---- Garner CRT (src/crt_gpu.cu:70-105 constants, :157-203 and :231-276 kernels) ---- */
typedef struct {
    int k;
    uint64_t primes[3], inv[3], prefix_M[3], M_mod_table[3][3], barrett_m[3];
} orc_garner;
void orc_garner_params(orc_garner* g, const uint64_t* primes, int k);
/* residues: k rows of N (row = prime).  128-bit result (src/crt_gpu.cu:157-203). */
void orc_garner128(const orc_garner* g, const uint64_t* residues, uint64_t N,
                   uint64_t* hi, uint64_t* lo);
/* U160 result (tests/test_crt_gpu.cu:288-322, src/crt_gpu.cu:231-276). */
void orc_garner160(const orc_garner* g, const uint64_t* residues, uint64_t N,
                   uint64_t* lo, uint64_t* mid, uint32_t* hi);

/* ---- carry (tests/test_carry_prop.cu:20-35 and :123-151; NOT the 3-pass GPU scheme) ---- */
/* out has n_out limbs of out_bits (32 or 64); coefficients beyond N are zero. */
void orc_carry128(const uint64_t* hi, const uint64_t* lo, uint64_t N, void* out,
                  uint64_t n_out, int out_bits);
void orc_carry160(const uint64_t* lo, const uint64_t* mid, const uint32_t* hi, uint64_t N,
                  void* out, uint64_t n_out);

/* ---- whole product ---- */
/* Prime set chosen by orc_choose_set (reference u32 set while N <= 2^23, wide set above). */
int orc_choose_set(int limb_bits, uint64_t N);
/* 1 (default): the primes of one product on their own OpenMP threads; 0: one
 * thread per product (thread pools of products). */
void orc_set_prime_threads(int on);
/* out has L_A+L_B limbs.  Returns 0, or -1 if the size is unsupported. */
int orc_multiply(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                 int limb_bits, int set);
/* Same, but stops after Garner and returns the coefficients (N of them). */
int orc_multiply_crt(const void* A, uint64_t L_A, const void* B, uint64_t L_B, int limb_bits,
                     int set, uint64_t* c0, uint64_t* c1, uint32_t* c2);
/* O(L_A*L_B) schoolbook, tests/test_full_multiply.cpp:27-58; out has L_A+L_B limbs. */
void orc_schoolbook(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                    int limb_bits);

/* ---- GMP mpz_mul (tests/test_full_multiply.cpp:173-237, :443-455) ---- */
int orc_gmp_available(void);
/* out has L_A+L_B limbs; returns the mpz_mul time in seconds (import/export excluded). */
double orc_gmp_mul(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                   int limb_bits);

#ifdef __cplusplus
}
#endif
#endif
