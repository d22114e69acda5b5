// garner_dev.h -- Garner reconstruction of the five-prime 64-bit-digit set,
// shared by the carry kernels and the whole-product kernel.
#pragma once
#include "bw_internal.h"

// Five-prime Garner for 64-bit digits (residues canonical or centered, |r| < p):
// mixed-radix digits
//   k_j = (...((r_j - k_0) p_0^-1 - k_1) p_1^-1 ... - k_{j-1}) p_{j-1}^-1 mod p_j
// in FP64 products, then x = k_0 + p_0 (k_1 + p_1 (k_2 + p_2 (k_3 + p_3 k_4)))
// by Horner steps on 32-bit words; x < 2^154 -> three 64-bit words.
// r: the five residue words of one digit (canonical or centered residues).
__device__ __forceinline__ void garner_d64(const uint32_t* r, const GarnerD64Dev& g, uint64_t* w) {
    uint32_t k[5];
    double kd[5];      // the same digits as doubles (no conversion back)
#pragma unroll
    for (int j = 0; j < 5; j++) {
        double t = (double)(int32_t)r[j];
#pragma unroll
        for (int i = 0; i < j; i++) t = f64_mulmod(t - kd[i], g.inv[j][i][0], g.inv[j][i][1], g.fp[j]);
        // t is in (-p_j, p_j): a residue word (canonical, or centered with
        // |r| <= p/2 + 1) for j = 0, the output of a modular product for
        // j > 0; one conditional add makes it canonical
        kd[j] = t < 0.0 ? t + g.fp[j].p : t;
        k[j] = (uint32_t)kd[j];
    }
    // Before step j, x = k_{j+1} + p_{j+1} (k_{j+2} + ...) < p_{j+1} ... p_4
    // < 2^(31 (4 - j)) fills at most 4 - j words, and k_j + p_j x at most one
    // more: the step multiplies only those words (ten multiply-adds, not 20).
    uint32_t x[5] = {k[4], 0u, 0u, 0u, 0u};
#pragma unroll
    for (int j = 3; j >= 0; j--) {
        // the prime as a wave-uniform 32-bit scalar: each word is one
        // 32 x 32 + 64-bit multiply-add (v_mad_u64_u32)
        const uint32_t pj = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.p[j]);
        const int nw = 4 - j;
        uint64_t c = k[j];
#pragma unroll
        for (int i = 0; i < nw; i++) {
            c += (uint64_t)x[i] * pj;
            x[i] = (uint32_t)c;
            c >>= 32;
        }
        x[nw] = (uint32_t)c;
    }
    w[0] = x[0] | ((uint64_t)x[1] << 32);
    w[1] = x[2] | ((uint64_t)x[3] << 32);
    w[2] = x[4];
}
