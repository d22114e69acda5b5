// carry_kernels.h -- Garner CRT fused with a single-pass carry scan.
//
// A product coefficient c_n (< 2^(NW*LB)) is split into limb-width words
// w_0..w_{NW-1}.  The output value is sum_n s_n 2^(LB n) with
//     s_n = sum_j w_j(c_{n-j}),   s_n < NW * 2^LB,
// so s_n = lo_n + hi_n 2^LB with hi_n < NW.  Folding hi_{n-1} into lo_n gives
// t_n = lo_n + hi_{n-1} <= 2^LB + NW - 2, whose carry into the next limb is
// binary: generate g_n = (t_n >= 2^LB), propagate p_n = (t_n == 2^LB - 1).
// The carries then follow c_{n+1} = g_n | (p_n & c_n), which a wavefront
// resolves for 64 limbs at once with two ballots and one 64-bit add
// (carries = (X + Y + cin) ^ X ^ Y with X = G|P, Y = G), and workgroups chain
// with a decoupled look-back over (generate, propagate) tile summaries
// (Merrill & Garland 2016, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back"), tile ids taken from an atomic counter so every predecessor is
// already running.
#pragma once
#include "bw_internal.h"

#ifndef BW_CARRY_ROWS
#define BW_CARRY_ROWS 8                                   // rows of 64 limbs per wavefront
#endif
// Wavefronts per workgroup: every workgroup starts with one device-scope
// atomic (its tile id) and ends with a look-back; eight waves (4096 limbs a
// tile) pay them half as often per limb as four and measured 116 against
// 126 us on the 2^24-limb product, batches unchanged.
#ifndef BW_CARRY_WAVES
#define BW_CARRY_WAVES 8
#endif
#define BW_CARRY_TILE (BW_CARRY_WAVES * 64 * BW_CARRY_ROWS)   // limbs per workgroup

enum {
    BW_SRC_GARNER_U32 = 0,   // 3 canonical u32 residue planes -> 32-bit limbs
    BW_SRC_GARNER_U64 = 1,   // 3 canonical u64 residue planes -> 64-bit limbs
    BW_SRC_HILO128 = 2,      // (hi, lo) u64 arrays -> 32-bit limbs
    BW_SRC_U160 = 3,         // (lo, mid, hi32) arrays -> 64-bit limbs
    BW_SRC_HILO128_L64 = 4,  // (hi, lo) u64 arrays -> 64-bit limbs
    BW_SRC_GARNER_D64 = 5,   // 5 canonical u32 residue planes of 64-bit digits -> 64-bit limbs
    BW_SRC_GARNER_U64X2 = 6, // 2 canonical u64 residue planes of 32-bit digits -> 32-bit limbs
};

// Batched carry: one launch covers many products; product j owns tiles
// [tile0, tile0 + ceil(n_out / tile)) of the launch and reads its 3 residue
// planes at res + i * plane (i = prime), N coefficients each.
struct BwCarryJob {
    const void* res;
    int64_t plane;
    int64_t N;
    int64_t n_out;
    int64_t out_off;      // limbs from CarryParams::out
    int64_t tile0;
    int64_t half_last;    // 1: the last limb is written as its low 32 bits only
};

struct CarryParams {
    int src;
    const void* a0;       // residues base / hi / lo arrays (see src)
    const void* a1;
    const void* a2;
    int64_t plane;        // residue plane stride (elements)
    int64_t N;            // coefficients (c_n = 0 for n >= N)
    int64_t n_out;        // output limbs
    void* out;
    GarnerDev g;
    F64Prime fp[3];       // FP64 constants of the three primes (u32 path)
    GarnerD64Dev gd;      // five-prime constants (64-bit digits)
    int half_last;        // 1: the last output limb is written as its low 32 bits only
                          // (a 32-bit-limb product of odd length, computed in 64-bit limbs)
    unsigned long long* flags;   // one per tile
    int* counter;         // tile-id allocator, returns to 0 after the launch
    uint32_t epoch;       // tags this launch's flags
    int64_t ntiles;
    int* err;             // set to 1 if a look-back wait exceeded its bound
    const BwCarryJob* jobs;        // batched launch (null: the single product above)
    const int* tile_job;           // batched: product of each tile
};

hipError_t bw_launch_garner_carry(const CarryParams& p, hipStream_t stream);
int64_t bw_carry_tiles(int64_t n_out);
// Clears the tile flags and the tile-id counter of a carry launch, and the
// look-back error word when `err` is given (a kernel, so that the sequence can
// live in a hipGraph).
hipError_t bw_launch_carry_reset(unsigned long long* flags, int64_t ntiles, int* counter, hipStream_t stream,
                                 int* err = nullptr);
// Standalone Garner (coefficients out): 32-bit-prime sets -> (hi, lo); the
// three 58-bit primes -> U160 (lo, mid, hi32); the two 58-bit primes of the
// hybrid set -> (hi, lo).  arith_bits: 32 or 64; g.k gives the prime count.
hipError_t bw_launch_garner_only(int arith_bits, const void* r0, const void* r1, const void* r2, int64_t N,
                                 const GarnerDev& g, const F64Prime* fp,
                                 uint64_t* o0, uint64_t* o1, uint32_t* o2, hipStream_t stream);
