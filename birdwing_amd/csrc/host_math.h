// host_math.h -- host-side number theory for plans: prime sets, roots,
// twiddle tables, Garner constants, size policy (src/ntt_limits.cpp).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>
#include "bw_internal.h"

namespace bw {

enum { SET_U32_REF = 0, SET_U32_WIDE = 1, SET_U64_REF = 2, SET_HYBRID = 3, SET_D64 = 4, SET_U64_PAIRS = 5,
       SET_COUNT = 6 };

struct PrimeSet {
    int id;
    int limb_bits;       // arithmetic: 32 -> FP64 kernels, 64 -> integer kernels
    int k;
    uint64_t p[BW_MAX_PRIMES];
    uint64_t root_max[BW_MAX_PRIMES];
    int digit_bits;      // 0: a digit is one input limb; 64: 64-bit digits (pairs of 32-bit limbs)
};

const PrimeSet* prime_set(int id);
int default_set(int limb_bits, uint64_t N);
// Prime set a product without a requested set runs under: the five-prime
// 64-bit-digit set when it supports the sizes, else default_set() when that
// does, else (32-bit limbs beyond both) the 58-bit primes on limb pairs.
int auto_set(int limb_bits, uint64_t L_A, uint64_t L_B);
// Digits of an operand of L limbs of in_bits under set s.
uint64_t digits_of(const PrimeSet& s, int in_bits, uint64_t L);
// Transform length of a product under set s (from its digit counts).
uint64_t padded_size_for(const PrimeSet& s, int in_bits, uint64_t L_A, uint64_t L_B);

typedef unsigned __int128 u128;
uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t powmod(uint64_t a, uint64_t e, uint64_t p);
uint64_t invmod(uint64_t a, uint64_t p);   // p prime

int valuation2(uint64_t n);
int max_root_log(uint64_t root, uint64_t p);
int set_max_log(const PrimeSet& s);
// root of order 2^logN (src/gpu_ntt.cu:171-192: square root_max down from the max order)
uint64_t root_of_order(const PrimeSet& s, int i, int logN);

uint64_t padded_ntt_size(uint64_t L_A, uint64_t L_B);
bool ntt_size_supported(const PrimeSet& s, uint64_t N, std::string* why);
bool crt_bound_ok(const PrimeSet& s, int in_bits, uint64_t L_A, uint64_t L_B, std::string* why);
bool multiply_size_supported(const PrimeSet& s, int in_bits, uint64_t L_A, uint64_t L_B, std::string* why);
uint64_t max_limb_count_by_crt(const PrimeSet& s, int in_bits);
uint64_t max_supported_limb_count(const PrimeSet& s, int in_bits);

// Host images of the device tables.
struct HostTables {
    int table_log;
    std::vector<unsigned char> consts;    // F64PrimeDev[k] / U64PrimeDev[k]
    std::vector<unsigned char> tw_stage;  // k * 2^13 pairs (16 B)
    std::vector<unsigned char> tw_lo;     // k * 2^12 pairs
    std::vector<unsigned char> tw_hi;     // k * 2^(L-12) pairs
};
void build_tables(const PrimeSet& s, HostTables& out);
// Serialised host tables of one prime set (a precompute cache on disk):
// magic, format version, set id, k, table_log, the primes, the four blob
// sizes, the blobs, and an FNV-1a 64 checksum of everything before it.
// read_table_image rejects files of another format, of a prime set whose
// primes or sizes differ from this library's, and damaged files.
bool write_table_image(const char* path, int set_id, const HostTables& h, std::string* why);
bool read_table_image(const char* path, int* set_id, HostTables& h, std::string* why);
GarnerDev garner_constants(const PrimeSet& s);
GarnerD64Dev garner_d64_constants(const PrimeSet& s);
void f64_prime(uint64_t p, F64Prime& out);
// Per-plan (N^-1, aux) pairs, layout [k][2].
void ninv_pairs(const PrimeSet& s, uint64_t N, std::vector<uint64_t>& out64, std::vector<double>& outf);

}  // namespace bw
