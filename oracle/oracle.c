/*
 * oracle.c -- CPU restatement of the reference big-integer multiply path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Straight-line, obviously-correct
 * code: 128-bit products reduced with '%', textbook radix-2 transforms, the
 * serial carry.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* MT19937-64.  Published algorithm of Matsumoto & Nishimura (2004), the
 * engine the reference seeds with std::mt19937_64 (tests/test_full_multiply.cpp:62). */
#define MT_NN 312
#define MT_MM 156
#define MT_A  0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

void orc_mt64_seed(orc_mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < MT_NN; i++)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = MT_NN;
}

uint64_t orc_mt64_next(orc_mt64* s) {
    static const uint64_t mag01[2] = {0ULL, MT_A};
    uint64_t x;
    if (s->mti >= MT_NN) {
        int i;
        for (i = 0; i < MT_NN - MT_MM; i++) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + MT_MM] ^ (x >> 1) ^ mag01[x & 1];
        }
        for (; i < MT_NN - 1; i++) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[x & 1];
        }
        x = (s->mt[MT_NN - 1] & MT_UM) | (s->mt[0] & MT_LM);
        s->mt[MT_NN - 1] = s->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[x & 1];
        s->mti = 0;
    }
    x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

void orc_ref_random_limbs_u32(uint32_t* out, size_t n, uint64_t seed) {
    orc_mt64 s; orc_mt64_seed(&s, seed);
    for (size_t i = 0; i < n; i++) {
        if (i % 8 == 0) out[i] = 0;
        else if (i % 8 == 1) out[i] = 1;
        else out[i] = (uint32_t)orc_mt64_next(&s) % (1u << 30);   /* (uint32_t)rng() % 2^30 */
    }
}

void orc_ref_random_limbs_u64(uint64_t* out, size_t n, uint64_t seed) {
    orc_mt64 s; orc_mt64_seed(&s, seed);
    for (size_t i = 0; i < n; i++) {
        if (i % 8 == 0) out[i] = 0;
        else if (i % 8 == 1) out[i] = 1;
        else if (i % 8 == 2) out[i] = UINT64_MAX;
        else out[i] = orc_mt64_next(&s);
    }
}

void orc_ref_pipeline_limbs_u32(uint32_t* out, size_t n, uint64_t seed) {
    orc_mt64 s; orc_mt64_seed(&s, seed);
    for (size_t i = 0; i < n; i++) out[i] = (uint32_t)(orc_mt64_next(&s) % (1ULL << 30));
}

void orc_uniform_limbs(void* out, size_t n, int limb_bits, uint64_t seed) {
    orc_mt64 s; orc_mt64_seed(&s, seed);
    if (limb_bits == 64) {
        uint64_t* o = (uint64_t*)out;
        for (size_t i = 0; i < n; i++) o[i] = orc_mt64_next(&s);
        if (n && o[n - 1] == 0) o[n - 1] |= 1;
    } else {
        uint32_t* o = (uint32_t*)out;
        for (size_t i = 0; i < n; i++) o[i] = (uint32_t)orc_mt64_next(&s);
        if (n && o[n - 1] == 0) o[n - 1] |= 1;
    }
}

/* ------------------------------------------------------------------------ */
/* Prime sets.  u32 reference and u64 reference sets: src/gpu_ntt.cu:28-47.
 * Wide u32 set {2013265921, 1811939329, 469762049}: tests/test_crt.cpp:19-21;
 * its roots are g^((p-1)/2^v2) for the primitive roots g = 31, 13, 3. */
int orc_prime_set(int set, uint64_t* primes, uint64_t* roots) {
    static const uint64_t u32ref_p[3] = {0x2d000001ULL, 0x23800001ULL, 0x26800001ULL};
    static const uint64_t u32ref_r[3] = {663, 721, 19};
    static const uint64_t u32w_p[3] = {2013265921ULL, 1811939329ULL, 469762049ULL};
    static const uint64_t u32w_r[3] = {440564289ULL, 72705542ULL, 2187ULL};
    static const uint64_t u64_p[3] = {0x400002600000001ULL, 0x400004200000001ULL, 0x400001100000001ULL};
    static const uint64_t u64_r[3] = {273765203699653965ULL, 26231613454922890ULL, 144261359744151243ULL};
    const uint64_t *P, *R; int k;
    switch (set) {
    case ORC_SET_U32_REF:  P = u32ref_p; R = u32ref_r; k = 3; break;
    case ORC_SET_U32_WIDE: P = u32w_p;   R = u32w_r;   k = 3; break;
    case ORC_SET_U64_REF:  P = u64_p;    R = u64_r;    k = 3; break;
    case ORC_SET_HYBRID:   P = u64_p;    R = u64_r;    k = 2; break;
    default: return 0;
    }
    for (int i = 0; i < k; i++) { if (primes) primes[i] = P[i]; if (roots) roots[i] = R[i]; }
    return k;
}

/* ------------------------------------------------------------------------ */
/* Size policy, restating src/ntt_limits.cpp with the prime set as a parameter. */
int orc_valuation2(uint64_t n) {
    int k = 0;
    if (n == 0) return 64;
    while ((n & 1) == 0) { k++; n >>= 1; }
    return k;
}
int orc_max_logN_for_prime(uint64_t p) { return orc_valuation2(p - 1); }

uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }

uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p; a %= p;
    while (e) { if (e & 1) r = orc_mulmod(r, a, p); a = orc_mulmod(a, a, p); e >>= 1; }
    return r;
}

int orc_max_root_logN(uint64_t root, uint64_t p) {
    uint64_t r = root % p;
    if (r == 0) return 0;
    int k = 0;
    while (k < 62) {
        if (r == 1) return k;
        r = orc_mulmod(r, r, p);
        k++;
    }
    return k;
}

static int set_maxlog(const uint64_t* P, const uint64_t* R, int k) {
    int lim = 1 << 30;
    for (int i = 0; i < k; i++) {
        int a = orc_max_logN_for_prime(P[i]), b = orc_max_root_logN(R[i], P[i]);
        int m = a < b ? a : b;
        if (m < lim) lim = m;
    }
    return lim;
}

int orc_max_supported_logN(int set) {
    uint64_t P[3], R[3]; int k = orc_prime_set(set, P, R);
    return k ? set_maxlog(P, R, k) : -1;
}

uint64_t orc_padded_ntt_size(uint64_t L_A, uint64_t L_B) {
    uint64_t L_C = L_A + L_B - 1, N = 1;
    while (N < L_C) N <<= 1;
    return N;
}

/* Exact comparison  min(L)*(2^b-1)^2 < prod(p), with 256-bit arithmetic
 * (the reference uses long double for 64-bit inputs, :156-162,194-198; exact
 * integers give the same answer wherever long double is decisive). */
typedef struct { uint64_t w[4]; } u256;
static void u256_mul64(u256* x, uint64_t m) {
    u128 c = 0;
    for (int i = 0; i < 4; i++) { c += (u128)x->w[i] * m; x->w[i] = (uint64_t)c; c >>= 64; }
}
static int u256_lt(const u256* a, const u256* b) {
    for (int i = 3; i >= 0; i--) { if (a->w[i] != b->w[i]) return a->w[i] < b->w[i]; }
    return 0;
}
static u256 limb_sq_times(int in_bits, uint64_t terms) {
    u256 x = {{terms, 0, 0, 0}};
    uint64_t lm = in_bits == 64 ? UINT64_MAX : 0xFFFFFFFFULL;
    u256_mul64(&x, lm); u256_mul64(&x, lm);
    return x;
}
static u256 prime_prod(int set) {
    uint64_t P[3]; int k = orc_prime_set(set, P, NULL);
    u256 M = {{1, 0, 0, 0}};
    for (int i = 0; i < k; i++) u256_mul64(&M, P[i]);
    return M;
}

int orc_crt_bound_ok(int set, int in_bits, uint64_t L_A, uint64_t L_B) {
    if (L_A == 0 || L_B == 0) return 0;
    uint64_t terms = L_A < L_B ? L_A : L_B;
    u256 c = limb_sq_times(in_bits, terms), M = prime_prod(set);
    return u256_lt(&c, &M);
}

int orc_ntt_size_supported(int set, uint64_t N) {
    if (N == 0 || (N & (N - 1))) return 0;
    int logN = orc_valuation2(N);
    return logN <= orc_max_supported_logN(set);
}

int orc_multiply_size_supported(int set, int in_bits, uint64_t L_A, uint64_t L_B) {
    if (L_A == 0 || L_B == 0) return 0;
    if (!orc_ntt_size_supported(set, orc_padded_ntt_size(L_A, L_B))) return 0;
    return orc_crt_bound_ok(set, in_bits, L_A, L_B);
}

uint64_t orc_max_limb_count_by_crt(int set, int in_bits) {
    /* largest L with L*(2^b-1)^2 < M, i.e. floor((M-1)/(2^b-1)^2): binary search
     * (every prime set here gives a bound below 2^63) */
    uint64_t lo = 0, hi = 1ULL << 63;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo + 1) / 2;
        u256 c = limb_sq_times(in_bits, mid), M = prime_prod(set);
        if (u256_lt(&c, &M)) lo = mid; else hi = mid - 1;
    }
    return lo;
}

uint64_t orc_max_supported_limb_count(int set, int in_bits) {
    int ml = orc_max_supported_logN(set);
    if (ml <= 0) return 0;
    uint64_t by_ntt = 1ULL << (ml - 1), by_crt = orc_max_limb_count_by_crt(set, in_bits);
    return by_ntt < by_crt ? by_ntt : by_crt;
}

/* ------------------------------------------------------------------------ */
/* Extended Euclid, src/crt_utils.cpp:29-51 (unsigned-safe restatement). */
uint64_t orc_modinv(uint64_t a, uint64_t m) {
    __int128 t0 = 0, t1 = 1, r0 = m, r1 = a % m;
    while (r1 != 0) {
        __int128 q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
        r0 = r1; r1 = r2; t0 = t1; t1 = t2;
    }
    if (r0 != 1) return 0;
    if (t0 < 0) t0 += m;
    return (uint64_t)t0;
}

uint64_t orc_root_for_N(uint64_t p, uint64_t root_max, int logN) {
    int a = orc_max_logN_for_prime(p), b = orc_max_root_logN(root_max, p);
    int maxlog = a < b ? a : b;
    if (logN > maxlog) return 0;
    uint64_t w = root_max % p;
    for (int j = 0; j < maxlog - logN; j++) w = orc_mulmod(w, w, p);
    return w;
}

/* ------------------------------------------------------------------------ */
/* Transforms.  Iterative radix-2 Cooley-Tukey: bit-reverse, then butterflies
 * with twiddles w^(j*N/len).  X[k] = sum_n x[n] w^(nk). */
static void bitrev_permute(uint64_t* x, uint64_t N) {
    for (uint64_t i = 1, j = 0; i < N; i++) {
        uint64_t bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { uint64_t t = x[i]; x[i] = x[j]; x[j] = t; }
    }
}

/* Shoup's precomputed-quotient product (Shoup; Harvey 2014): for p < 2^63,
 * a < 2^64, b < p and bq = floor(b*2^64/p), a*b - mulhi(a,bq)*p lies in [0, 2p). */
static inline uint64_t shoup_mul(uint64_t a, uint64_t b, uint64_t bq, uint64_t p) {
    uint64_t q = (uint64_t)(((u128)a * bq) >> 64);
    uint64_t r = a * b - q * p;
    return r >= p ? r - p : r;
}

static void ntt_core(uint64_t* x, uint64_t N, uint64_t p, uint64_t w) {
    if (N <= 1) return;
    for (uint64_t i = 0; i < N; i++) x[i] %= p;
    bitrev_permute(x, N);
    /* twiddle table w^j, j < N/2, with Shoup quotients */
    uint64_t half = N / 2;
    uint64_t* tw = (uint64_t*)malloc(sizeof(uint64_t) * half * 2);
    uint64_t* twq = tw + half;
    tw[0] = 1;
    for (uint64_t j = 1; j < half; j++) tw[j] = orc_mulmod(tw[j - 1], w, p);
    for (uint64_t j = 0; j < half; j++) twq[j] = (uint64_t)(((u128)tw[j] << 64) / p);
    for (uint64_t len = 2; len <= N; len <<= 1) {
        uint64_t h = len / 2, step = N / len;
        for (uint64_t blk = 0; blk < N; blk += len) {
            for (uint64_t j = 0; j < h; j++) {
                uint64_t u = x[blk + j];
                uint64_t v = shoup_mul(x[blk + j + h], tw[j * step], twq[j * step], p);
                uint64_t s = u + v; if (s >= p) s -= p;
                uint64_t d = u >= v ? u - v : u + p - v;
                x[blk + j] = s; x[blk + j + h] = d;
            }
        }
    }
    free(tw);
}

void orc_ntt_forward(uint64_t* x, uint64_t N, uint64_t p, uint64_t w) { ntt_core(x, N, p, w); }

void orc_ntt_inverse(uint64_t* x, uint64_t N, uint64_t p, uint64_t w) {
    if (N == 0) return;
    ntt_core(x, N, p, orc_modinv(w % p, p));
    uint64_t ninv = orc_modinv(N % p, p);
    for (uint64_t i = 0; i < N; i++) x[i] = orc_mulmod(x[i], ninv, p);
}

void orc_dft_bruteforce(const uint64_t* x, uint64_t* X, uint64_t N, uint64_t p, uint64_t w) {
    for (uint64_t k = 0; k < N; k++) {
        uint64_t acc = 0, wk = orc_powmod(w, k, p), wnk = 1;
        for (uint64_t n = 0; n < N; n++) {
            acc = (uint64_t)(((u128)acc + orc_mulmod(x[n] % p, wnk, p)) % p);
            wnk = orc_mulmod(wnk, wk, p);
        }
        X[k] = acc;
    }
}

/* Pipeline transforms without permutations: decimation in frequency forward
 * (natural order in, bit-reversed order out) and decimation in time inverse
 * (bit-reversed in, natural out, times N^-1).  The pointwise product between
 * them does not depend on the order.  Twiddles are stored per level,
 * tw[m + j] = w_{2m}^j for j < m, so every level reads them contiguously. */
typedef struct { uint64_t* w; uint64_t* q; } orc_tw;

static orc_tw tw_levels(uint64_t N, uint64_t p, uint64_t root) {
    orc_tw t;
    t.w = (uint64_t*)malloc(sizeof(uint64_t) * (N ? N : 1));
    t.q = (uint64_t*)malloc(sizeof(uint64_t) * (N ? N : 1));
    for (uint64_t m = 1; m < N; m <<= 1) {
        uint64_t wm = orc_powmod(root, N / (2 * m), p), x = 1;
        for (uint64_t j = 0; j < m; j++) {
            t.w[m + j] = x;
            t.q[m + j] = (uint64_t)(((u128)x << 64) / p);
            x = orc_mulmod(x, wm, p);
        }
    }
    return t;
}

static void ntt_dif_noperm(uint64_t* x, uint64_t N, uint64_t p, const orc_tw* t) {
    for (uint64_t m = N / 2; m >= 1; m >>= 1) {
        for (uint64_t blk = 0; blk < N; blk += 2 * m) {
            for (uint64_t j = 0; j < m; j++) {
                uint64_t u = x[blk + j], v = x[blk + j + m];
                uint64_t s = u + v; if (s >= p) s -= p;
                uint64_t d = u >= v ? u - v : u + p - v;
                x[blk + j] = s;
                x[blk + j + m] = shoup_mul(d, t->w[m + j], t->q[m + j], p);
            }
        }
        if (m == 1) break;
    }
}

static void ntt_dit_noperm(uint64_t* x, uint64_t N, uint64_t p, const orc_tw* it) {
    for (uint64_t m = 1; m < N; m <<= 1) {
        for (uint64_t blk = 0; blk < N; blk += 2 * m) {
            for (uint64_t j = 0; j < m; j++) {
                uint64_t u = x[blk + j];
                uint64_t v = shoup_mul(x[blk + j + m], it->w[m + j], it->q[m + j], p);
                uint64_t s = u + v; if (s >= p) s -= p;
                uint64_t d = u >= v ? u - v : u + p - v;
                x[blk + j] = s;
                x[blk + j + m] = d;
            }
        }
    }
}

void orc_pointwise(const uint64_t* a, const uint64_t* b, uint64_t* c, uint64_t N, uint64_t p) {
    for (uint64_t i = 0; i < N; i++) c[i] = orc_mulmod(a[i], b[i], p);
}

void orc_conv_schoolbook_mod(const uint64_t* a, uint64_t La, const uint64_t* b, uint64_t Lb,
                             uint64_t* c, uint64_t N, uint64_t p) {
    memset(c, 0, sizeof(uint64_t) * N);
    for (uint64_t i = 0; i < La; i++)
        for (uint64_t j = 0; j < Lb; j++)
            if (i + j < N) c[i + j] = (uint64_t)(((u128)c[i + j] + orc_mulmod(a[i], b[j], p)) % p);
}

/* ------------------------------------------------------------------------ */
/* Garner.  Constants as src/crt_gpu.cu:70-105 (prefix_M filled, which the
 * reference leaves unwritten); the per-coefficient recurrence follows the
 * kernel at src/crt_gpu.cu:157-203: t = (r_j - x mod p_j) mod p_j,
 * k_j = t*inv_j, x_mod[i>j] += M_mod_table[j][i]*k_j, x += M_j*k_j. */
void orc_garner_params(orc_garner* g, const uint64_t* primes, int k) {
    memset(g, 0, sizeof(*g));
    g->k = k;
    for (int i = 0; i < k; i++) {
        uint64_t pi = primes[i];
        g->primes[i] = pi;
        /* prefix product mod p_i, computed incrementally (it can exceed 128 bits) */
        uint64_t Mmod = 1 % pi;
        for (int j = 0; j < i; j++) Mmod = orc_mulmod(Mmod, primes[j] % pi, pi);
        g->prefix_M[i] = Mmod;
        g->inv[i] = (i == 0) ? 1 : orc_modinv(Mmod, pi);
        for (int kk = 0; kk < k; kk++) {
            uint64_t Mk = 1 % pi;
            for (int j = 0; j < kk; j++) Mk = orc_mulmod(Mk, primes[j] % pi, pi);
            g->M_mod_table[kk][i] = Mk;           /* M_kk mod p_i */
        }
        g->barrett_m[i] = pi < (1ULL << 32) ? (uint64_t)(((u128)1 << 64) / pi) : 0;
    }
}

/* 192-bit accumulator x += M * kj where M = prod_{i<j} p_i (<= 2 primes, < 2^128) */
static void garner_one(const orc_garner* g, const uint64_t* r, uint64_t out[3]) {
    uint64_t x[3] = {0, 0, 0};
    uint64_t M[3] = {1, 0, 0};
    uint64_t x_mod[3] = {0, 0, 0};
    for (int j = 0; j < g->k; j++) {
        uint64_t p = g->primes[j];
        uint64_t rj = r[j] % p;
        uint64_t t = rj >= x_mod[j] ? rj - x_mod[j] : rj + p - x_mod[j];
        uint64_t kj = orc_mulmod(t, g->inv[j], p);
        for (int i = j + 1; i < g->k; i++) {
            uint64_t pk = g->primes[i];
            uint64_t c = orc_mulmod(g->M_mod_table[j][i], kj, pk);
            x_mod[i] += c;
            if (x_mod[i] >= pk) x_mod[i] -= pk;
        }
        /* x += M*kj (192-bit) */
        u128 carry = 0;
        for (int w = 0; w < 3; w++) {
            u128 t2 = (u128)M[w] * kj + x[w] + carry;
            x[w] = (uint64_t)t2; carry = t2 >> 64;
        }
        /* M *= p */
        carry = 0;
        for (int w = 0; w < 3; w++) {
            u128 t2 = (u128)M[w] * p + carry;
            M[w] = (uint64_t)t2; carry = t2 >> 64;
        }
    }
    out[0] = x[0]; out[1] = x[1]; out[2] = x[2];
}

void orc_garner128(const orc_garner* g, const uint64_t* residues, uint64_t N,
                   uint64_t* hi, uint64_t* lo) {
    uint64_t r[3], x[3];
    for (uint64_t n = 0; n < N; n++) {
        for (int j = 0; j < g->k; j++) r[j] = residues[(uint64_t)j * N + n];
        garner_one(g, r, x);
        lo[n] = x[0]; hi[n] = x[1];
    }
}

void orc_garner160(const orc_garner* g, const uint64_t* residues, uint64_t N,
                   uint64_t* lo, uint64_t* mid, uint32_t* hi) {
    uint64_t r[3], x[3];
    for (uint64_t n = 0; n < N; n++) {
        for (int j = 0; j < g->k; j++) r[j] = residues[(uint64_t)j * N + n];
        garner_one(g, r, x);
        lo[n] = x[0]; mid[n] = x[1]; hi[n] = (uint32_t)x[2];
    }
}

/* ------------------------------------------------------------------------ */
/* Serial carry, tests/test_carry_prop.cu:20-35 (128-bit) and :123-151 (U160):
 * running (coefficient + carry), emit the low limb, shift the rest down.  The
 * carry is kept in 192 bits so no coefficient size can overflow it. */
static void carry_generic(const uint64_t* w0, const uint64_t* w1, const uint32_t* w2h,
                          const uint64_t* w2, uint64_t N, void* out, uint64_t n_out, int out_bits) {
    uint64_t c[3] = {0, 0, 0};     /* 192-bit running carry */
    for (uint64_t i = 0; i < n_out; i++) {
        uint64_t v0 = 0, v1 = 0, v2 = 0;
        if (i < N) { v0 = w0[i]; v1 = w1[i]; v2 = w2h ? w2h[i] : (w2 ? w2[i] : 0); }
        u128 s = (u128)v0 + c[0];
        uint64_t t0 = (uint64_t)s; s >>= 64;
        s += (u128)v1 + c[1];
        uint64_t t1 = (uint64_t)s; s >>= 64;
        uint64_t t2 = (uint64_t)(s + v2 + c[2]);
        if (out_bits == 64) {
            ((uint64_t*)out)[i] = t0;
            c[0] = t1; c[1] = t2; c[2] = 0;
        } else {
            ((uint32_t*)out)[i] = (uint32_t)t0;
            c[0] = (t0 >> 32) | (t1 << 32);
            c[1] = (t1 >> 32) | (t2 << 32);
            c[2] = t2 >> 32;
        }
    }
}

void orc_carry128(const uint64_t* hi, const uint64_t* lo, uint64_t N, void* out,
                  uint64_t n_out, int out_bits) {
    carry_generic(lo, hi, NULL, NULL, N, out, n_out, out_bits);
}

void orc_carry160(const uint64_t* lo, const uint64_t* mid, const uint32_t* hi, uint64_t N,
                  void* out, uint64_t n_out) {
    carry_generic(lo, mid, hi, NULL, N, out, n_out, 64);
}

/* ------------------------------------------------------------------------ */
int orc_choose_set(int limb_bits, uint64_t N) {
    if (limb_bits == 64) return ORC_SET_U64_REF;
    return N <= (1ULL << 23) ? ORC_SET_U32_REF : ORC_SET_U32_WIDE;
}

static uint64_t limb_at(const void* X, uint64_t i, int limb_bits) {
    return limb_bits == 64 ? ((const uint64_t*)X)[i] : ((const uint32_t*)X)[i];
}

/* Per-product parallelism: the primes of one product run on their own threads
 * unless orc_set_prime_threads(0) asks for one thread per product (the
 * thread-pool CPU baselines run one product per pool thread). */
static int g_prime_threads = 1;
void orc_set_prime_threads(int on) { g_prime_threads = on != 0; }

int orc_multiply_crt(const void* A, uint64_t L_A, const void* B, uint64_t L_B, int limb_bits,
                     int set, uint64_t* c0, uint64_t* c1, uint32_t* c2) {
    if (!orc_multiply_size_supported(set, limb_bits, L_A, L_B)) return -1;
    uint64_t N = orc_padded_ntt_size(L_A, L_B);
    int logN = orc_valuation2(N);
    uint64_t P[3], R[3];
    int k = orc_prime_set(set, P, R);
    uint64_t* res = (uint64_t*)malloc(sizeof(uint64_t) * N * (uint64_t)k);
    /* one prime per thread (the primes are independent until Garner) */
    #pragma omp parallel for schedule(static, 1) if (g_prime_threads)
    for (int j = 0; j < k; j++) {
        uint64_t p = P[j], w = orc_root_for_N(p, R[j], logN);
        uint64_t* fa = res + (uint64_t)j * N;
        uint64_t* fb = (uint64_t*)malloc(sizeof(uint64_t) * N);
        /* ingress: zero-pad and reduce (src/zero_pad.cu:3-42) */
        for (uint64_t i = 0; i < N; i++) fa[i] = i < L_A ? limb_at(A, i, limb_bits) % p : 0;
        for (uint64_t i = 0; i < N; i++) fb[i] = i < L_B ? limb_at(B, i, limb_bits) % p : 0;
        orc_tw tw = tw_levels(N, p, w);
        orc_tw itw = tw_levels(N, p, orc_modinv(w, p));
        ntt_dif_noperm(fa, N, p, &tw);
        ntt_dif_noperm(fb, N, p, &tw);
        uint64_t ninv = orc_modinv(N % p, p);
        for (uint64_t i = 0; i < N; i++) fa[i] = orc_mulmod(orc_mulmod(fa[i], fb[i], p), ninv, p);
        ntt_dit_noperm(fa, N, p, &itw);
        free(tw.w); free(tw.q); free(itw.w); free(itw.q);
        free(fb);
    }
    orc_garner g; orc_garner_params(&g, P, k);
    uint64_t r[3], x[3];
    for (uint64_t n = 0; n < N; n++) {
        for (int j = 0; j < k; j++) r[j] = res[(uint64_t)j * N + n];
        garner_one(&g, r, x);
        c0[n] = x[0]; c1[n] = x[1]; if (c2) c2[n] = (uint32_t)x[2];
    }
    free(res);
    return 0;
}

int orc_multiply(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                 int limb_bits, int set) {
    if (!orc_multiply_size_supported(set, limb_bits, L_A, L_B)) return -1;
    uint64_t N = orc_padded_ntt_size(L_A, L_B);
    uint64_t* c0 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint64_t* c1 = (uint64_t*)malloc(sizeof(uint64_t) * N);
    uint32_t* c2 = (uint32_t*)malloc(sizeof(uint32_t) * N);
    int rc = orc_multiply_crt(A, L_A, B, L_B, limb_bits, set, c0, c1, c2);
    if (rc == 0) carry_generic(c0, c1, c2, NULL, N, out, L_A + L_B, limb_bits);
    free(c0); free(c1); free(c2);
    return rc;
}

void orc_schoolbook(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                    int limb_bits) {
    uint64_t n = L_A + L_B;
    /* 192-bit column sums, then the serial carry */
    uint64_t* w0 = (uint64_t*)calloc(n, sizeof(uint64_t));
    uint64_t* w1 = (uint64_t*)calloc(n, sizeof(uint64_t));
    uint64_t* w2 = (uint64_t*)calloc(n, sizeof(uint64_t));
    for (uint64_t i = 0; i < L_A; i++) {
        uint64_t a = limb_at(A, i, limb_bits);
        for (uint64_t j = 0; j < L_B; j++) {
            u128 pr = (u128)a * limb_at(B, j, limb_bits);
            u128 s = (u128)w0[i + j] + (uint64_t)pr;
            w0[i + j] = (uint64_t)s;
            s = (s >> 64) + w1[i + j] + (uint64_t)(pr >> 64);
            w1[i + j] = (uint64_t)s;
            w2[i + j] += (uint64_t)(s >> 64);
        }
    }
    carry_generic(w0, w1, NULL, w2, n, out, n, limb_bits);
    free(w0); free(w1); free(w2);
}

/* ------------------------------------------------------------------------ */
#ifdef ORC_HAVE_GMP
#include <gmp.h>
int orc_gmp_available(void) { return 1; }
double orc_gmp_mul(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                   int limb_bits) {
    size_t sz = limb_bits == 64 ? 8 : 4;
    mpz_t a, b, c;
    mpz_init(a); mpz_init(b); mpz_init(c);
    mpz_import(a, L_A, -1, sz, 0, 0, A);
    mpz_import(b, L_B, -1, sz, 0, 0, B);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    mpz_mul(c, a, b);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    memset(out, 0, sz * (L_A + L_B));
    size_t count = 0;
    mpz_export(out, &count, -1, sz, 0, 0, c);
    mpz_clear(a); mpz_clear(b); mpz_clear(c);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
#else
int orc_gmp_available(void) { return 0; }
double orc_gmp_mul(const void* A, uint64_t L_A, const void* B, uint64_t L_B, void* out,
                   int limb_bits) {
    (void)A; (void)L_A; (void)B; (void)L_B; (void)out; (void)limb_bits;
    return -1.0;
}
#endif
