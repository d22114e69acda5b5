// bw_internal.h -- structures shared by the host orchestration and kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "arith.h"

#define BW_MAX_PRIMES 5
#define BW_STAGE_LOG 13          // stage twiddle table: omega_{2^13}^e, e < 2^13
#define BW_TW_LO_LOG 12          // pass twiddles: 2-level table split at 2^12
#define BW_MAX_PASS_LOG 13       // largest sub-transform held in one workgroup

// Per-prime constants of the FP64 (u32 digit) path: roots of order up to 32.
struct F64PrimeDev {
    F64Prime P;
    double w32[16][2];    // omega_32^e, e < 16: (w, w/p)
    double iw32[16][2];   // omega_32^-e
    double c32[2];        // 2^32 mod p as (w, w/p): 64-bit digits enter as hi * 2^32 + lo
};

// Per-prime constants of the u64 digit path (Shoup pairs (w, floor(w 2^64/p))).
struct U64PrimeDev {
    U64Prime P;
    uint64_t w32[16][2];
    uint64_t iw32[16][2];
};

// Device tables for one prime set, shared by every plan that uses the set.
struct SetTables {
    int limb_bits;         // 32 (FP64 arithmetic) or 64 (integer arithmetic)
    int k;                 // number of primes
    int table_log;         // omega_{2^table_log} is the table root
    void* consts;          // F64PrimeDev[k] or U64PrimeDev[k]
    void* tw_stage;        // [k][2^BW_STAGE_LOG] pairs: omega_{2^13}^e
    void* tw_lo;           // [k][2^BW_TW_LO_LOG] pairs: omega_{2^L}^e
    void* tw_hi;           // [k][2^(L-12)] pairs: omega_{2^L}^(e*2^12)
};

// Garner constants for the fused CRT+carry kernel.
struct GarnerDev {
    int k;
    uint64_t p[3];
    // u32 path (FP64 arithmetic): inverses as (w, w/p) pairs
    double inv1[2], inv2[2], p0mod2[2];
    // u64 path (Shoup pairs)
    uint64_t inv1_s[2], inv2_s[2], p0mod2_s[2];
    uint64_t p0;           // p0
    uint64_t p01_lo, p01_hi;   // p0*p1 as 128-bit
};

// Garner constants of the five-prime 64-bit-digit set (FP64 arithmetic):
// inv[j][i] = p_i^-1 mod p_j (i < j) as (w, w/p_j) pairs.
struct GarnerD64Dev {
    F64Prime fp[5];
    double inv[5][4][2];
    uint32_t p[5];
};

// One pass of the multi-pass transform.  Element (o, t, s) of the N-array lives
// at (o*T + t)*S + s; the pass transforms over t for every (o, s).
struct PassGeom {
    int logT;
    int64_t S;        // stride
    int64_t O;        // outer count
    int logM;         // log2(T*S): sub-problem whose twiddles follow this pass
    int G;            // lines per workgroup
};

enum { BW_MODE_FWD = 0, BW_MODE_INV = 1, BW_MODE_FUSED = 2 };

// Batched ingress: products stacked along the outer dimension of the first
// pass (product o owns elements [o N, (o+1) N) of every prime plane) read their
// raw limbs at in[0] + a_off (a_len limbs) and in[1] + b_off (b_len limbs).
struct BwIngress {
    int64_t a_off, b_off;
    int64_t a_len, b_len;
};

struct PassParams {
    // operands: FWD handles nops operands (A then B); FUSED reads both, writes out0;
    // INV reads/writes operand 0.
    const void* in[2];
    void* out[2];
    int64_t in_len[2];     // raw limb counts when ingress != 0
    int ingress;           // 1: inputs are raw limbs (zero-padded to N, pass 1 only)
    int canon_out;         // 1: store canonical residues in [0, p); 0: centered (32-bit path).
                           // Forward column passes of the register-staged kernels store
                           // centered residues whatever this says (never a transform's last pass).
    int fold_ninv;         // FUSED: skip N^-1 (the next INV pass applies it); INV: apply N^-1
    int nops;              // FWD: operands in this launch
    int64_t N;             // transform length
    int64_t plane;         // elements per prime plane (N, or batch stride)
    int64_t S, O;
    int logM;
    int table_log;
    int prime0;            // first prime index handled (grid.y adds to it)
    const void* consts;    // F64PrimeDev* / U64PrimeDev*
    const void* tw_stage;
    const void* tw_lo;
    const void* tw_hi;
    const void* ninv;      // per-prime (N^-1, aux) pairs for FUSED mode
    const BwIngress* bdesc;  // ingress of stacked products (null: one product, in_len)
    int rev;               // 1: the launch walks its workgroups' tiles, primes and operands in
                           // reverse order (reads first what the launch before it wrote last,
                           // while the Infinity Cache still holds it)
};

// Status codes (BW_OK, BW_ERR_*) come from the public header.
#include "../../include/birdwing.h"
