// host_math.cpp -- host-side number theory for plans (see host_math.h).
#include "host_math.h"
#include <cstdio>
#include <cstring>

namespace bw {

// u32 reference set and u64 reference set: src/gpu_ntt.cu:28-47 (the hybrid
// set is the first two u64 primes, :40-43).  The wide u32 set {15*2^27+1,
// 27*2^26+1, 7*2^26+1} (tests/test_crt.cpp:19-21) reaches N = 2^26; its roots
// are g^((p-1)/2^v2) for the primitive roots g = 31, 13, 3.
//
// The 64-bit-digit set: five primes c*2^m + 1 < 2^31 with c < 128 and m >= 25
// (every such prime there is except two smaller ones), product 2^153.36.  A
// 64-bit digit is a pair of 32-bit limbs or one 64-bit limb; with N = 2^25 the
// coefficients stay below 2^24 * 2^128 = 2^152.  Against three primes on
// 32-bit digits this halves N for five transforms instead of three (about 0.8
// of the butterflies and of the residue bytes) and still runs in FP64.  Roots
// are g^((p-1)/2^v2) for the primitive roots g = 5, 31, 13, 29, 10.
//
// The limb-pair set runs the reference's 58-bit primes on 64-bit digits (a
// pair of 32-bit limbs or one 64-bit limb, coefficients < N * 2^128 < 2^174):
// the reach of the reference's hybrid pipeline (32-bit limbs under 58-bit
// primes, include/config.h:26-30, N <= 2^33 limb positions) at three
// transforms of N/2 64-bit words.
static const PrimeSet kSets[SET_COUNT] = {
    {SET_U32_REF, 32, 3, {0x2d000001ULL, 0x23800001ULL, 0x26800001ULL}, {663, 721, 19}, 0},
    {SET_U32_WIDE, 32, 3, {2013265921ULL, 1811939329ULL, 469762049ULL}, {440564289ULL, 72705542ULL, 2187ULL}, 0},
    {SET_U64_REF, 64, 3, {0x400002600000001ULL, 0x400004200000001ULL, 0x400001100000001ULL},
     {273765203699653965ULL, 26231613454922890ULL, 144261359744151243ULL}, 0},
    {SET_HYBRID, 64, 2, {0x400002600000001ULL, 0x400004200000001ULL},
     {273765203699653965ULL, 26231613454922890ULL}, 0},
    {SET_D64, 32, 5, {2113929217ULL, 2013265921ULL, 1811939329ULL, 1711276033ULL, 1107296257ULL},
     {1971140334ULL, 440564289ULL, 72705542ULL, 969788637ULL, 1087287097ULL}, 64},
    {SET_U64_PAIRS, 64, 3, {0x400002600000001ULL, 0x400004200000001ULL, 0x400001100000001ULL},
     {273765203699653965ULL, 26231613454922890ULL, 144261359744151243ULL}, 64},
};

const PrimeSet* prime_set(int id) { return (id >= 0 && id < SET_COUNT) ? &kSets[id] : nullptr; }

int default_set(int limb_bits, uint64_t N) {
    if (limb_bits == 64) return SET_U64_REF;
    return N <= (1ULL << 23) ? SET_U32_REF : SET_U32_WIDE;
}

uint64_t digits_of(const PrimeSet& s, int in_bits, uint64_t L) {
    return (s.digit_bits == 64 && in_bits == 32) ? (L + 1) / 2 : L;
}

uint64_t padded_size_for(const PrimeSet& s, int in_bits, uint64_t L_A, uint64_t L_B) {
    return padded_ntt_size(digits_of(s, in_bits, L_A), digits_of(s, in_bits, L_B));
}

int auto_set(int limb_bits, uint64_t L_A, uint64_t L_B) {
    if (multiply_size_supported(kSets[SET_D64], limb_bits, L_A, L_B, nullptr)) return SET_D64;
    const int d = default_set(limb_bits, padded_ntt_size(L_A, L_B));
    if (limb_bits == 32 && !multiply_size_supported(kSets[d], limb_bits, L_A, L_B, nullptr) &&
        multiply_size_supported(kSets[SET_U64_PAIRS], limb_bits, L_A, L_B, nullptr))
        return SET_U64_PAIRS;
    return d;
}

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)(((u128)a * b) % p); }

uint64_t powmod(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = mulmod(r, a, p);
        a = mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

uint64_t invmod(uint64_t a, uint64_t p) { return powmod(a, p - 2, p); }

int valuation2(uint64_t n) {
    if (n == 0) return 64;
    int k = 0;
    while (!(n & 1)) { n >>= 1; k++; }
    return k;
}

int max_root_log(uint64_t root, uint64_t p) {   // src/ntt_limits.cpp:37-51
    uint64_t r = root % p;
    if (r == 0) return 0;
    for (int k = 0; k < 62; k++) {
        if (r == 1) return k;
        r = mulmod(r, r, p);
    }
    return 62;
}

int set_max_log(const PrimeSet& s) {
    int lim = 1 << 30;
    for (int i = 0; i < s.k; i++) {
        int a = valuation2(s.p[i] - 1), b = max_root_log(s.root_max[i], s.p[i]);
        int m = a < b ? a : b;
        if (m < lim) lim = m;
    }
    return lim;
}

uint64_t root_of_order(const PrimeSet& s, int i, int logN) {
    const uint64_t p = s.p[i];
    int a = valuation2(p - 1), b = max_root_log(s.root_max[i], p);
    int maxlog = a < b ? a : b;
    uint64_t w = s.root_max[i] % p;
    for (int j = 0; j < maxlog - logN; j++) w = mulmod(w, w, p);
    return w;
}

uint64_t padded_ntt_size(uint64_t L_A, uint64_t L_B) {
    uint64_t L_C = L_A + L_B - 1, N = 1;
    while (N < L_C) N <<= 1;
    return N;
}

bool ntt_size_supported(const PrimeSet& s, uint64_t N, std::string* why) {
    if (N == 0) { if (why) *why = "NTT size N must be positive"; return false; }
    if (N & (N - 1)) { if (why) *why = "NTT size N=" + std::to_string(N) + " is not a power of two"; return false; }
    const int logN = valuation2(N);
    for (int i = 0; i < s.k; i++) {
        int a = valuation2(s.p[i] - 1), b = max_root_log(s.root_max[i], s.p[i]);
        int m = a < b ? a : b;
        if (logN > m) {
            if (why)
                *why = "NTT size N=2^" + std::to_string(logN) + " exceeds modulus " + std::to_string(i) +
                       " (p=" + std::to_string(s.p[i]) + ", max 2^" + std::to_string(m) + ")";
            return false;
        }
    }
    return true;
}

// prod(p) and min(L)*(2^b-1)^2 compared exactly in 256-bit arithmetic.
struct U256 { uint64_t w[4]; };
static void mul64(U256& x, uint64_t m) {
    u128 c = 0;
    for (int i = 0; i < 4; i++) { c += (u128)x.w[i] * m; x.w[i] = (uint64_t)c; c >>= 64; }
}
static bool lt(const U256& a, const U256& b) {
    for (int i = 3; i >= 0; i--) if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
    return false;
}
static U256 coeff_bound(int in_bits, uint64_t terms) {
    U256 x = {{terms, 0, 0, 0}};
    const uint64_t lm = in_bits == 64 ? ~0ULL : 0xFFFFFFFFULL;
    mul64(x, lm);
    mul64(x, lm);
    return x;
}
static U256 modulus_product(const PrimeSet& s) {
    U256 M = {{1, 0, 0, 0}};
    for (int i = 0; i < s.k; i++) mul64(M, s.p[i]);
    return M;
}

bool crt_bound_ok(const PrimeSet& s, int in_bits, uint64_t L_A, uint64_t L_B, std::string* why) {
    if (L_A == 0 || L_B == 0) { if (why) *why = "operand limb count must be positive"; return false; }
    const uint64_t dA = digits_of(s, in_bits, L_A), dB = digits_of(s, in_bits, L_B);
    const uint64_t terms = dA < dB ? dA : dB;
    const int dbits = s.digit_bits ? s.digit_bits : in_bits;
    if (lt(coeff_bound(dbits, terms), modulus_product(s))) return true;
    if (why)
        *why = "max convolution coefficient for L_A=" + std::to_string(L_A) + ", L_B=" + std::to_string(L_B) +
               " (terms=" + std::to_string(terms) + ", limb=" + std::to_string(in_bits) +
               "-bit) exceeds CRT modulus product";
    return false;
}

bool multiply_size_supported(const PrimeSet& s, int in_bits, uint64_t L_A, uint64_t L_B, std::string* why) {
    if (L_A == 0 || L_B == 0) { if (why) *why = "operand limb count must be positive"; return false; }
    if (s.digit_bits == 64 && in_bits != 32 && in_bits != 64) { if (why) *why = "limbs must be 32 or 64 bits"; return false; }
    if (!ntt_size_supported(s, padded_size_for(s, in_bits, L_A, L_B), why)) return false;
    return crt_bound_ok(s, in_bits, L_A, L_B, why);
}

uint64_t max_limb_count_by_crt(const PrimeSet& s, int in_bits) {
    uint64_t lo = 0, hi = 1ULL << 63;     // every set's bound is below 2^63
    const U256 M = modulus_product(s);
    const int dbits = s.digit_bits ? s.digit_bits : in_bits;
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo + 1) / 2;
        if (lt(coeff_bound(dbits, mid), M)) lo = mid; else hi = mid - 1;
    }
    // in limbs: a 64-bit digit holds two 32-bit limbs
    return (s.digit_bits == 64 && in_bits == 32) ? (lo > (1ULL << 62) ? (1ULL << 63) : 2 * lo) : lo;
}

uint64_t max_supported_limb_count(const PrimeSet& s, int in_bits) {
    const int ml = set_max_log(s);
    if (ml <= 0) return 0;
    const uint64_t by_ntt = ((s.digit_bits == 64 && in_bits == 32) ? 2ULL : 1ULL) << (ml - 1);
    const uint64_t by_crt = max_limb_count_by_crt(s, in_bits);
    return by_ntt < by_crt ? by_ntt : by_crt;
}

// ---------------------------------------------------------------------------
void f64_prime(uint64_t p, F64Prime& out) {
    out.p = (double)p;
    out.pinv = 1.0 / (double)p;
    out.cm = (double)(p - 1);
}

static void put_f64(std::vector<unsigned char>& v, size_t idx, uint64_t w, uint64_t p) {
    double pair[2] = {(double)w, (double)w / (double)p};
    std::memcpy(v.data() + idx * 16, pair, 16);
}
static void put_u64(std::vector<unsigned char>& v, size_t idx, uint64_t w, uint64_t p) {
    uint64_t pair[2] = {w, (uint64_t)(((u128)w << 64) / p)};
    std::memcpy(v.data() + idx * 16, pair, 16);
}

void build_tables(const PrimeSet& s, HostTables& out) {
    const int L = set_max_log(s);
    out.table_log = L;
    const size_t nstage = 1u << BW_STAGE_LOG, nlo = 1u << BW_TW_LO_LOG, nhi = (size_t)1 << (L - BW_TW_LO_LOG);
    out.tw_stage.assign((size_t)s.k * nstage * 16, 0);
    out.tw_lo.assign((size_t)s.k * nlo * 16, 0);
    out.tw_hi.assign((size_t)s.k * nhi * 16, 0);
    const bool f64 = s.limb_bits == 32;
    out.consts.assign((size_t)s.k * (f64 ? sizeof(F64PrimeDev) : sizeof(U64PrimeDev)), 0);
    auto put = f64 ? put_f64 : put_u64;
    for (int i = 0; i < s.k; i++) {
        const uint64_t p = s.p[i];
        const uint64_t wL = root_of_order(s, i, L);
        // stage table: omega_{2^13}^e
        const uint64_t w13 = powmod(wL, 1ULL << (L - BW_STAGE_LOG), p);
        uint64_t x = 1;
        for (size_t e = 0; e < nstage; e++) { put(out.tw_stage, i * nstage + e, x, p); x = mulmod(x, w13, p); }
        x = 1;
        for (size_t e = 0; e < nlo; e++) { put(out.tw_lo, i * nlo + e, x, p); x = mulmod(x, wL, p); }
        const uint64_t wstep = powmod(wL, nlo, p);
        x = 1;
        for (size_t e = 0; e < nhi; e++) { put(out.tw_hi, i * nhi + e, x, p); x = mulmod(x, wstep, p); }
        const uint64_t w32 = powmod(wL, 1ULL << (L - 5), p), iw32 = invmod(w32, p);
        if (f64) {
            F64PrimeDev c;
            std::memset(&c, 0, sizeof(c));
            f64_prime(p, c.P);
            uint64_t a = 1, b = 1;
            for (int e = 0; e < 16; e++) {
                c.w32[e][0] = (double)a; c.w32[e][1] = (double)a / (double)p;
                c.iw32[e][0] = (double)b; c.iw32[e][1] = (double)b / (double)p;
                a = mulmod(a, w32, p); b = mulmod(b, iw32, p);
            }
            const uint64_t c32 = (uint64_t)((((u128)1) << 32) % p);
            c.c32[0] = (double)c32;
            c.c32[1] = (double)c32 / (double)p;
            std::memcpy(out.consts.data() + i * sizeof(F64PrimeDev), &c, sizeof(c));
        } else {
            U64PrimeDev c;
            std::memset(&c, 0, sizeof(c));
            c.P.p = p;
            uint64_t inv = 1;                       // Newton: p^-1 mod 2^64
            for (int it = 0; it < 7; it++) inv *= 2 - p * inv;
            c.P.pinv_neg = (uint64_t)0 - inv;
            c.P.r2 = (uint64_t)((((u128)1 << 64) % p) * (u128)(((u128)1 << 64) % p) % p);
            c.P.floor_r = (uint64_t)(~0ULL / p);
            uint64_t a = 1, b = 1;
            for (int e = 0; e < 16; e++) {
                c.w32[e][0] = a; c.w32[e][1] = (uint64_t)(((u128)a << 64) / p);
                c.iw32[e][0] = b; c.iw32[e][1] = (uint64_t)(((u128)b << 64) / p);
                a = mulmod(a, w32, p); b = mulmod(b, iw32, p);
            }
            std::memcpy(out.consts.data() + i * sizeof(U64PrimeDev), &c, sizeof(c));
        }
    }
}

GarnerDev garner_constants(const PrimeSet& s) {
    GarnerDev g;
    std::memset(&g, 0, sizeof(g));
    g.k = s.k;
    for (int i = 0; i < s.k; i++) g.p[i] = s.p[i];
    g.p0 = s.p[0];
    const u128 p01 = (u128)s.p[0] * s.p[1];
    g.p01_lo = (uint64_t)p01;
    g.p01_hi = (uint64_t)(p01 >> 64);
    const uint64_t p1 = s.p[1];
    const uint64_t inv1 = invmod(s.p[0] % p1, p1);
    g.inv1[0] = (double)inv1; g.inv1[1] = (double)inv1 / (double)p1;
    g.inv1_s[0] = inv1; g.inv1_s[1] = (uint64_t)(((u128)inv1 << 64) / p1);
    if (s.k > 2) {
        const uint64_t p2 = s.p[2];
        const uint64_t p0m = s.p[0] % p2;
        const uint64_t inv2 = invmod(mulmod(p0m, s.p[1] % p2, p2), p2);
        g.inv2[0] = (double)inv2; g.inv2[1] = (double)inv2 / (double)p2;
        g.inv2_s[0] = inv2; g.inv2_s[1] = (uint64_t)(((u128)inv2 << 64) / p2);
        g.p0mod2[0] = (double)p0m; g.p0mod2[1] = (double)p0m / (double)p2;
        g.p0mod2_s[0] = p0m; g.p0mod2_s[1] = (uint64_t)(((u128)p0m << 64) / p2);
    }
    return g;
}

GarnerD64Dev garner_d64_constants(const PrimeSet& s) {
    GarnerD64Dev g;
    std::memset(&g, 0, sizeof(g));
    for (int j = 0; j < s.k && j < 5; j++) {
        f64_prime(s.p[j], g.fp[j]);
        g.p[j] = (uint32_t)s.p[j];
        for (int i = 0; i < j; i++) {
            const uint64_t v = invmod(s.p[i] % s.p[j], s.p[j]);
            g.inv[j][i][0] = (double)v;
            g.inv[j][i][1] = (double)v / (double)s.p[j];
        }
    }
    return g;
}

void ninv_pairs(const PrimeSet& s, uint64_t N, std::vector<uint64_t>& out64, std::vector<double>& outf) {
    out64.assign(2 * s.k, 0);
    outf.assign(2 * s.k, 0.0);
    for (int i = 0; i < s.k; i++) {
        const uint64_t p = s.p[i];
        const uint64_t ni = invmod(N % p, p);
        outf[2 * i] = (double)ni;
        outf[2 * i + 1] = (double)ni / (double)p;
        // Montgomery: montmul(montmul(a, b), K) = a b N^-1 needs K = N^-1 2^128 mod p
        const uint64_t r = (uint64_t)((((u128)1 << 64)) % p);
        out64[2 * i] = mulmod(mulmod(ni, r, p), r, p);
        out64[2 * i + 1] = 0;
    }
}

// ---------------------------------------------------------------------------
// table images (precompute cache)
static const char kImageMagic[8] = {'B', 'W', 'T', 'A', 'B', 'L', 'E', 'S'};
static const uint32_t kImageVersion = 1;

static uint64_t fnv1a(uint64_t h, const unsigned char* p, size_t n) {
    for (size_t i = 0; i < n; i++) h = (h ^ p[i]) * 0x100000001b3ULL;
    return h;
}

namespace {
struct ImageHeader {
    char magic[8];
    uint32_t version;
    int32_t set_id, k, table_log;
    uint64_t primes[BW_MAX_PRIMES];
    uint64_t sizes[4];
};
}  // namespace

bool write_table_image(const char* path, int set_id, const HostTables& h, std::string* why) {
    const PrimeSet* s = prime_set(set_id);
    if (!s) { if (why) *why = "unknown prime set"; return false; }
    ImageHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    std::memcpy(hd.magic, kImageMagic, 8);
    hd.version = kImageVersion;
    hd.set_id = set_id;
    hd.k = s->k;
    hd.table_log = h.table_log;
    for (int i = 0; i < s->k; i++) hd.primes[i] = s->p[i];
    const std::vector<unsigned char>* blobs[4] = {&h.consts, &h.tw_stage, &h.tw_lo, &h.tw_hi};
    for (int i = 0; i < 4; i++) hd.sizes[i] = blobs[i]->size();
    FILE* f = std::fopen(path, "wb");
    if (!f) { if (why) *why = std::string("cannot open ") + path + " for writing"; return false; }
    uint64_t sum = fnv1a(0xcbf29ce484222325ULL, reinterpret_cast<const unsigned char*>(&hd), sizeof(hd));
    bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1;
    for (int i = 0; i < 4 && ok; i++) {
        sum = fnv1a(sum, blobs[i]->data(), blobs[i]->size());
        ok = blobs[i]->empty() || std::fwrite(blobs[i]->data(), blobs[i]->size(), 1, f) == 1;
    }
    ok = ok && std::fwrite(&sum, sizeof(sum), 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok && why) *why = std::string("write to ") + path + " failed";
    return ok;
}

bool read_table_image(const char* path, int* set_id, HostTables& h, std::string* why) {
    auto bad = [&](const std::string& m) { if (why) *why = m; return false; };
    FILE* f = std::fopen(path, "rb");
    if (!f) return bad(std::string("cannot open ") + path);
    ImageHeader hd;
    if (std::fread(&hd, sizeof(hd), 1, f) != 1) { std::fclose(f); return bad("truncated header"); }
    if (std::memcmp(hd.magic, kImageMagic, 8) != 0 || hd.version != kImageVersion) {
        std::fclose(f);
        return bad("not a table image of this format");
    }
    const PrimeSet* s = prime_set(hd.set_id);
    bool same = s && hd.k == s->k && hd.table_log == set_max_log(*s);
    for (int i = 0; same && i < BW_MAX_PRIMES; i++) same = hd.primes[i] == (i < s->k ? s->p[i] : 0);
    if (same) {   // the sizes build_tables produces for this set
        const uint64_t k = (uint64_t)s->k;
        const uint64_t want[4] = {k * (s->limb_bits == 32 ? sizeof(F64PrimeDev) : sizeof(U64PrimeDev)),
                                  k * ((uint64_t)1 << BW_STAGE_LOG) * 16, k * ((uint64_t)1 << BW_TW_LO_LOG) * 16,
                                  k * ((uint64_t)1 << (hd.table_log - BW_TW_LO_LOG)) * 16};
        for (int i = 0; i < 4; i++) same = same && hd.sizes[i] == want[i];
    }
    if (!same) { std::fclose(f); return bad("image does not match this library's prime set"); }
    std::vector<unsigned char>* blobs[4] = {&h.consts, &h.tw_stage, &h.tw_lo, &h.tw_hi};
    uint64_t sum = fnv1a(0xcbf29ce484222325ULL, reinterpret_cast<const unsigned char*>(&hd), sizeof(hd));
    for (int i = 0; i < 4; i++) {
        blobs[i]->resize(hd.sizes[i]);
        if (hd.sizes[i] && std::fread(blobs[i]->data(), hd.sizes[i], 1, f) != 1) {
            std::fclose(f);
            return bad("truncated tables");
        }
        sum = fnv1a(sum, blobs[i]->data(), blobs[i]->size());
    }
    uint64_t stored = 0;
    const bool have = std::fread(&stored, sizeof(stored), 1, f) == 1;
    const bool extra = std::fgetc(f) != EOF;
    std::fclose(f);
    if (!have || extra || stored != sum) return bad("checksum mismatch (damaged image)");
    h.table_log = hd.table_log;
    *set_id = hd.set_id;
    return true;
}

}  // namespace bw
