// ntt_kernels.h -- multi-pass number-theoretic transform kernels for gfx950.
//
// A length-N transform is split into passes of T_j = 2^logT points (four-step
// decomposition applied recursively).  Pass j views the N-array as
// (O_j, T_j, S_j) and, for each of the O_j*S_j "lines" (o, s), computes
//
//   forward:  y[k] = sum_t x[(o,t,s)] w_T^{t k};   out[(o,q(k),s)] = y[k] * w_M^{k s}
//   inverse:  z[k] = in[(o,q(k),s)] * w_M^{-k s};  out[(o,t,s)] = sum_k z[k] w_T^{-t k}
//
// with M = T*S.  Inside a workgroup the line is staged through LDS and the
// T-point transform is done as radix-16 register codelets (decimation in
// frequency forward, decimation in time inverse), so the forward output lands
// at the digit-reversed position q(k) and the inverse consumes that order: no
// permutation pass ever runs.  The final layout of the spectrum is therefore a
// fixed permutation of natural order (pointwise products do not care).
#pragma once
#include "bw_internal.h"

// Radix schedule of a 2^LOGT-point sub-transform: stage 0 has radix
// 2^(LOGT mod 4) (16 when LOGT is a multiple of 4), every later stage 16.
__host__ __device__ constexpr int sched_r0log(int logT) { return logT % 4 == 0 ? 4 : logT % 4; }
__host__ __device__ constexpr int sched_nst(int logT) { return logT == 0 ? 0 : (logT - sched_r0log(logT)) / 4 + 1; }
__host__ __device__ constexpr int sched_rlog(int logT, int i) { return i == 0 ? sched_r0log(logT) : 4; }
// log2 of the span (block size) that stage i works on
__host__ __device__ constexpr int sched_spanlog(int logT, int i) { return i == 0 ? logT : logT - sched_r0log(logT) - 4 * (i - 1); }

// Frequency index of the value a forward pass leaves at position t of a line.
__host__ __device__ inline int64_t sched_freq_of_pos(int logT, int64_t t) {
    int64_t f = 0;
    int shift = 0;
    const int nst = sched_nst(logT);
    for (int i = 0; i < nst; i++) {
        const int rl = sched_rlog(logT, i);
        const int stridelog = sched_spanlog(logT, i) - rl;
        const int64_t k = (t >> stridelog) & ((1 << rl) - 1);
        f |= k << shift;
        shift += rl;
    }
    return f;
}

// LDS padding: one slot per 16 elements breaks the stride-16 bank pattern of
// the last stage.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_line(int logT) { return (1 << logT) + ((1 << logT) >> 4); }

// Launch one pass.  Returns hipSuccess or the launch error.
struct PassLaunch {
    int mode;          // BW_MODE_*
    int logT;
    int G;
    int limb_bits;     // 32 -> FP64 arithmetic, 64 -> integer arithmetic
    int nprimes;       // grid.y
    int nops;          // grid.z (FWD)
    int64_t tiles;     // grid.x
};
hipError_t bw_launch_pass(const PassLaunch& L, const PassParams& prm, hipStream_t stream);
// Largest lines-per-workgroup the kernel set provides for a pass of 2^logT points.
int bw_pass_max_G(int logT, int limb_bits, bool cols);
// Passes of up to 2^BW_REG2_MAX_LOG points use the register-staged kernels
// (ntt2_impl.h), larger ones the LDS radix-16 kernels (ntt_pass_impl.h).
#define BW_REG2_MAX_LOG 10
// Radix split T = R0 * R1 of the register-staged kernels: R1 (stage 1,
// stride 1) takes up to 16 points and 32 above, so the stage-0 radix R0 stays
// <= 32.  Column passes keep R1 = 16 through T = 2^9, so a 2^9-point column
// pass still runs 32 adjacent columns (128-byte rows) in 512 threads; row
// passes switch to R1 = 32 at 2^9, which keeps the fused pass's stage-0
// registers at radix 16.
__host__ __device__ constexpr int t2_r1log(int logT, bool cols) {
    return logT <= 5 ? logT : (logT <= (cols ? 9 : 8) ? 4 : 5);
}
__host__ __device__ constexpr int t2_r0log(int logT, bool cols) { return logT - t2_r1log(logT, cols); }
// LDS line stride (values) of a tile line: odd for column tiles (adjacent
// threads own adjacent lines); for row tiles with 16-wide stage-1 groups,
// 16 mod 32 so the two 16-thread groups of a half-wavefront hit disjoint banks.
__host__ __device__ constexpr int t2_lds_line(int logT, bool cols) {
    return (!cols && t2_r1log(logT, cols) == 4)
               ? ((1 << logT) + ((1 << logT) >> 4) + 1 + 15) / 32 * 32 + 16
               : (1 << logT) + ((1 << logT) >> t2_r1log(logT, cols)) + 1;
}
int bw_reg2_lds_line(int logT, bool cols);
// Row-resident fused pass (row_fused.hip): one row of 2^logT points per
// workgroup (logT = 14, 15; FP64 arithmetic, 32-bit words).  LDS: the padded
// line of int32 residues (one pad slot per 16), then 256 twiddle pairs.
#define BW_ROW_MIN_LOG 14
#define BW_ROW_MAX_LOG 15
__host__ __device__ constexpr int row_fused_line_bytes(int logT) {
    return (((1 << logT) + ((1 << logT) >> 4)) * 4 + 15) / 16 * 16;
}
__host__ __device__ constexpr int row_fused_lds_bytes(int logT) { return row_fused_line_bytes(logT) + 256 * 16; }
hipError_t bw_launch_row_fused(int logT, int64_t rows, int nprimes, const PassParams& prm, hipStream_t stream);
// LDS of a register-staged workgroup: the tile (G lines of LT values of vbytes,
// rounded up to 16 bytes), then the T stage twiddles (16-byte pairs), then
// for the forward ingress pass of R1 * R1-point column tiles on the FP64
// path (vbytes 4) the column twiddle tables: R0*G powers w^e and R0 + 1
// powers w^(b*s0) (16-byte pairs).
__host__ __device__ constexpr int t2_tile_bytes(int G, int lt, int vbytes) { return (G * lt * vbytes + 15) / 16 * 16; }
__host__ __device__ constexpr int t2_coltab_entries(int logT, bool cols, int G, int vbytes, int mode, int ingress) {
    return (cols && vbytes == 4 && mode == BW_MODE_FWD && ingress != 0 && t2_r0log(logT, cols) == t2_r1log(logT, cols))
               ? (G << t2_r0log(logT, cols)) + (1 << t2_r0log(logT, cols)) + 1
               : 0;
}
__host__ __device__ constexpr int t2_lds_bytes(int logT, bool cols, int G, int vbytes, int mode, int ingress) {
    return t2_tile_bytes(G, t2_lds_line(logT, cols), vbytes) + (16 << logT) +
           16 * t2_coltab_entries(logT, cols, G, vbytes, mode, ingress);
}
// Inverse column passes of at most 256 points (R1 = 16) on the FP64 path run
// 64-column tiles -- 256-byte row segments, 1024-thread workgroups -- when
// their row stride is at least 2^BW_INV_WIDE_MIN_LOGS elements: with 128-byte
// segments at a 256 KiB stride (the last inverse pass of 2^24 points) every
// segment lands on its own DRAM page, and the pass ran 13% slower; at short
// strides the wider workgroup measured 3% slower.
#define BW_INV_WIDE_MIN_LOGS 12
__host__ __device__ constexpr bool t2_inv_wide_ok(int logT, int vbytes) {
    return vbytes == 4 && logT <= 8 && t2_r1log(logT, true) == 4;
}
// Frequency held at position q of a pass line (either kernel family).
__host__ __device__ inline int64_t bw_pass_freq_of_pos(int logT, int64_t q, bool cols) {
    if (logT <= BW_REG2_MAX_LOG) {
        const int r1l = t2_r1log(logT, cols);
        const int r0l = logT - r1l;
        return (q >> r1l) + ((q & ((1 << r1l) - 1)) << r0l);
    }
    return sched_freq_of_pos(logT, q);
}
