// ntt2_impl.h -- register-staged pass kernels for sub-transforms of T <= 1024.
//
// Same pass contract as ntt_pass_impl.h (element (o, t, s) at (o*T + t)*S + s,
// forward output at digit-reversed positions, inverse consuming them), built
// for throughput:
//
//   T = R0 * R1 (R0, R1 <= 32).  Stage 0 (radix R0, stride R1) reads its group
//   straight from HBM into registers; one padded LDS transpose hands stage 1
//   (radix R1, stride 1) its groups; stage 1 writes straight to HBM.  A tile
//   is G lines (G adjacent columns when S > 1, so every row segment is G*4
//   contiguous bytes; G adjacent rows when S == 1).
//
//   Position q = k0*R1 + k1 of a line holds frequency f = k0 + R0*k1.  The pass
//   twiddle w_M^{f s} for the R1 outputs of one stage-1 group (fixed k0, s) is
//   base * step^k1 with base = w_M^{k0 s} and step = w_M^{R0 s}: two table
//   look-ups and one modular product per element, no gathers.
#pragma once
#include <type_traits>
#include "ntt_pass_impl.h"

// Arithmetic policies keyed on the prime constants themselves.
struct ArF64x {
    using V = double;          // register value: signed integer held in FP64
    using W = uint32_t;        // storage word (canonical residue)
    using Raw = uint32_t;      // raw input limb (< 2^32: enters unreduced)
    using TW = double2;        // twiddle (w, w/p)
    using NI = double;
    using PC32 = F64PrimeDev;
    using PR = F64Prime;
    using LV = int32_t;        // LDS element: centered residue (halves LDS per wavefront)
    __device__ static LV to_lds(V x, const PR& P) { return (int32_t)f64_reduce(x, P); }
    __device__ static V from_lds(LV x) { return (double)x; }
    // words hold canonical [0, p) or centered residues; both read as int32
    __device__ static V from_word(W x, const PR&) { return (double)(int32_t)x; }
    __device__ static V from_raw(Raw x, const PR&) { return (double)x; }
    // a 64-bit digit hi * 2^32 + lo: hi * (2^32 mod p) in (-p, p) plus lo (< 2^33 in all)
    __device__ static V from_raw64(uint64_t d, const PC32& c) {
        return f64_mulmod((double)(uint32_t)(d >> 32), c.c32[0], c.c32[1], c.P) + (double)(uint32_t)d;
    }
    __device__ static W to_word(V x, const PR& P) { return f64_canon(x, P); }
    __device__ static W to_word_c(V x, const PR& P, bool canon) {
        return canon ? f64_canon(x, P) : (uint32_t)(int32_t)f64_reduce(x, P);
    }
    // values already in (-p, p) -- outputs of a modular product -- need no
    // reduction before the 32-bit conversion
    __device__ static LV to_lds_r(V x) { return (int32_t)x; }
    __device__ static W to_word_r(V x) { return (uint32_t)(int32_t)x; }
    // a twiddle product reduced straight to the 32-bit LDS / storage word
    // (four FP64 operations and no conversion, f64_mulmod_i32)
    __device__ static LV lds_mul(V x, TW w, const PR& P) { return f64_mulmod_i32(x, w.x, w.y, P); }
    __device__ static W word_mul(V x, double w, const PR& P) { return (uint32_t)f64_mulmod_i32(x, w, w * P.pinv, P); }
    // a*b (no N^-1) for a, b < 2^46 in magnitude
    __device__ static V mul_plain(V a, V b, const PR& P) {
        const double br = f64_reduce(b, P);
        return f64_mulmod(a, br, br * P.pinv, P);
    }
    __device__ static TW tw_mul_ninv(TW t, const NI* ninv, const PR& P) {
        const double w = f64_mulmod(t.x, ninv[0], ninv[1], P);
        return make_double2(w, w * P.pinv);
    }
    __device__ static V add(V a, V b, const PR&) { return a + b; }
    __device__ static V sub(V a, V b, const PR&) { return a - b; }
    __device__ static V mul(V y, TW w, const PR& P) { return f64_mulmod(y, w.x, w.y, P); }
    __device__ static V mulc(V y, const double* w, const PR& P) { return f64_mulmod(y, w[0], w[1], P); }
    __device__ static V mul_scaled(V a, V b, const NI* ninv, const PR& P) {
        const double br = f64_reduce(b, P);
        const double r = f64_mulmod(a, br, br * P.pinv, P);
        return f64_mulmod(r, ninv[0], ninv[1], P);
    }
    __device__ static V zero() { return 0.0; }
    // Fused pass: A's spectrum waits in registers as centered int32 residues
    // (half the registers of FP64) and then serves as the reduced multiplier
    // of the pointwise product, so B's unreduced value needs no reduction.
    using KV = int32_t;
    __device__ static KV to_keep(V x, const PR& P) { return (int32_t)f64_reduce(x, P); }
    __device__ static V from_keep(KV k) { return (double)k; }
    __device__ static KV keep_mul(KV k, V x, const PR& P) {
        const double w = (double)k;
        return f64_mulmod_i32(x, w, w * P.pinv, P);
    }
    __device__ static KV keep_mul_scaled(KV k, V x, const NI* ninv, const PR& P) {
        const double w = (double)k;
        return f64_mulmod_i32(f64_mulmod(x, w, w * P.pinv, P), ninv[0], ninv[1], P);
    }
    struct Gen {          // w_k = base * step^k, values in (-p, p)
        double w, step, stepp;
        __device__ void init(double2 base, double2 st, const PR&) { w = base.x; step = st.x; stepp = st.y; }
        __device__ double apply(double x, const PR& P) const { return f64_mulmod(x, w, w * P.pinv, P); }
        __device__ void next(const PR& P) { w = f64_mulmod(w, step, stepp, P); }
    };
    __device__ static TW twprod(TW a, TW b, const PR& P) {
        const double w = f64_mulmod(a.x, b.x, b.y, P);
        return make_double2(w, w * P.pinv);
    }
    // twiddles base * step^k, k < n (n a multiple of 4), as four interleaved
    // chains of stride step^4 (short dependency chains)
    template <int NTW>
    __device__ static void gen_powers(TW base, TW step, double* w, const PR& P) {
        const double s = step.x, sp = step.y;
        const double s2 = f64_mulmod(s, s, sp, P), s2p = s2 * P.pinv;
        const double s4 = f64_mulmod(s2, s2, s2p, P), s4p = s4 * P.pinv;
        // w holds NTW entries: the first four are written only where they exist
        w[0] = base.x;
        if constexpr (NTW > 1) w[1] = f64_mulmod(base.x, s, sp, P);
        if constexpr (NTW > 2) w[2] = f64_mulmod(base.x, s2, s2p, P);
        if constexpr (NTW > 3) w[3] = f64_mulmod(w[1], s2, s2p, P);
#pragma unroll
        for (int k = 4; k < NTW; k++) w[k] = f64_mulmod(w[k - 4], s4, s4p, P);
    }
    // the same twiddles handed to f(k, w_k) as the chains produce them: only
    // the four chain heads stay live (fewer registers than an array of NTW)
    template <int NTW, class F>
    __device__ static void for_powers(TW base, TW step, const PR& P, F&& f) {
        const double s = step.x, sp = step.y;
        const double s2 = f64_mulmod(s, s, sp, P), s2p = s2 * P.pinv;
        const double s4 = f64_mulmod(s2, s2, s2p, P), s4p = s4 * P.pinv;
        double c[4];
        c[0] = base.x;
        if constexpr (NTW > 1) c[1] = f64_mulmod(base.x, s, sp, P);
        if constexpr (NTW > 2) c[2] = f64_mulmod(base.x, s2, s2p, P);
        if constexpr (NTW > 3) c[3] = f64_mulmod(c[1], s2, s2p, P);
#pragma unroll
        for (int k = 0; k < NTW; k++) {
            f(k, c[k & 3]);
            if (k + 4 < NTW) c[k & 3] = f64_mulmod(c[k & 3], s4, s4p, P);
        }
    }
    __device__ static V apply_tw(V x, double w, const PR& P) { return f64_mulmod(x, w, w * P.pinv, P); }
    __device__ static const PR& prime(const PC32& c) { return c.P; }
};

struct ArU64x {
    using V = uint64_t;
    using W = uint64_t;
    using Raw = uint64_t;
    using TW = ulonglong2;     // Shoup pair (w, floor(w 2^64/p))
    using NI = uint64_t;
    using PC32 = U64PrimeDev;
    using PR = U64Prime;
    using LV = uint64_t;
    __device__ static LV to_lds(V x, const PR&) { return x; }
    __device__ static V from_lds(LV x) { return x; }
    __device__ static V from_word(W x, const PR&) { return x; }
    __device__ static V from_raw(Raw x, const PR& P) { return u64_reduce(x, P); }
    // a 64-bit digit (a pair of 32-bit limbs or one 64-bit limb)
    __device__ static V from_raw64(uint64_t d, const PC32& c) { return u64_reduce(d, c.P); }
    __device__ static W to_word(V x, const PR&) { return x; }
    __device__ static W to_word_c(V x, const PR&, bool) { return x; }
    __device__ static LV to_lds_r(V x) { return x; }
    __device__ static W to_word_r(V x) { return x; }
    __device__ static LV lds_mul(V x, TW w, const PR& P) { return mul(x, w, P); }
    __device__ static W word_mul(V x, uint64_t w, const PR& P) { return apply_tw(x, w, P); }
    // montmul(montmul(a, b), 2^128) = a*b
    __device__ static V mul_plain(V a, V b, const PR& P) { return u64_montmul(u64_montmul(a, b, P), P.r2, P); }
    // ninv[0] = N^-1 2^128 mod p: montmul(t, ninv) = t N^-1 2^64 -> montmul by 1 -> t N^-1
    __device__ static TW tw_mul_ninv(TW t, const NI* ninv, const PR& P) {
        ulonglong2 r;
        r.x = u64_montmul(u64_montmul(t.x, ninv[0], P), 1, P);
        r.y = 0;
        return r;
    }
    __device__ static V add(V a, V b, const PR& P) { return u64_add(a, b, P.p); }
    __device__ static V sub(V a, V b, const PR& P) { return u64_sub(a, b, P.p); }
    __device__ static V mul(V y, TW w, const PR& P) { return u64_mulmod_shoup(y, w.x, w.y, P.p); }
    __device__ static V mulc(V y, const uint64_t* w, const PR& P) { return u64_mulmod_shoup(y, w[0], w[1], P.p); }
    __device__ static V mul_scaled(V a, V b, const NI* ninv, const PR& P) {
        return u64_montmul(u64_montmul(a, b, P), ninv[0], P);
    }
    __device__ static V zero() { return 0; }
    using KV = uint64_t;
    __device__ static KV to_keep(V x, const PR&) { return x; }
    __device__ static V from_keep(KV k) { return k; }
    __device__ static KV keep_mul(KV k, V x, const PR& P) { return mul_plain(k, x, P); }
    __device__ static KV keep_mul_scaled(KV k, V x, const NI* ninv, const PR& P) { return mul_scaled(k, x, ninv, P); }
    struct Gen {          // Montgomery form: wm = w 2^64 mod p
        uint64_t wm, stepm;
        __device__ void init(ulonglong2 base, ulonglong2 st, const PR& P) {
            wm = u64_montmul(base.x, P.r2, P);
            stepm = u64_montmul(st.x, P.r2, P);
        }
        __device__ uint64_t apply(uint64_t x, const PR& P) const { return u64_montmul(x, wm, P); }
        __device__ void next(const PR& P) { wm = u64_montmul(wm, stepm, P); }
    };
    __device__ static TW twprod(TW a, TW b, const PR& P) {
        ulonglong2 r;
        r.x = u64_mulmod_shoup(a.x, b.x, b.y, P.p);
        r.y = 0;   // only .x is consumed
        return r;
    }
    // Montgomery-form twiddles base * step^k (four interleaved chains)
    template <int NTW>
    __device__ static void gen_powers(TW base, TW step, uint64_t* w, const PR& P) {
        const uint64_t s = u64_montmul(step.x, P.r2, P);
        const uint64_t s2 = u64_montmul(s, s, P), s4 = u64_montmul(s2, s2, P);
        w[0] = u64_montmul(base.x, P.r2, P);
        if constexpr (NTW > 1) w[1] = u64_montmul(w[0], s, P);
        if constexpr (NTW > 2) w[2] = u64_montmul(w[0], s2, P);
        if constexpr (NTW > 3) w[3] = u64_montmul(w[1], s2, P);
#pragma unroll
        for (int k = 4; k < NTW; k++) w[k] = u64_montmul(w[k - 4], s4, P);
    }
    template <int NTW, class F>
    __device__ static void for_powers(TW base, TW step, const PR& P, F&& f) {
        const uint64_t s = u64_montmul(step.x, P.r2, P);
        const uint64_t s2 = u64_montmul(s, s, P), s4 = u64_montmul(s2, s2, P);
        uint64_t c[4];
        c[0] = u64_montmul(base.x, P.r2, P);
        if constexpr (NTW > 1) c[1] = u64_montmul(c[0], s, P);
        if constexpr (NTW > 2) c[2] = u64_montmul(c[0], s2, P);
        if constexpr (NTW > 3) c[3] = u64_montmul(c[1], s2, P);
#pragma unroll
        for (int k = 0; k < NTW; k++) {
            f(k, c[k & 3]);
            if (k + 4 < NTW) c[k & 3] = u64_montmul(c[k & 3], s4, P);
        }
    }
    __device__ static V apply_tw(V x, uint64_t w, const PR& P) { return u64_montmul(x, w, P); }
    __device__ static const PR& prime(const PC32& c) { return c.P; }
};

// Codelets with roots up to order 32 (R <= 32).  HZ: the upper half of the
// input is zero (a zero-padded operand), so the first level reduces to copies
// and twiddle products and x[R/2 ..] is never read.
template <class AR, int RLOG, bool HZ = false>
__device__ __forceinline__ void dft32_dif(typename AR::V* x, const typename AR::PC32& pc) {
    constexpr int R = 1 << RLOG;
#pragma unroll
    for (int q = 0; q < RLOG; q++) {
        const int h = R >> (q + 1);
        if (HZ && q == 0) {
#pragma unroll
            for (int i = 0; i < h; i++) x[i + h] = (i == 0) ? x[i] : AR::mulc(x[i], pc.w32[i * (16 / h)], pc.P);
            continue;
        }
#pragma unroll
        for (int blk = 0; blk < R; blk += 2 * h) {
#pragma unroll
            for (int i = 0; i < h; i++) {
                auto a = x[blk + i], b = x[blk + i + h];
                x[blk + i] = AR::add(a, b, pc.P);
                auto d = AR::sub(a, b, pc.P);
                x[blk + i + h] = (i == 0) ? d : AR::mulc(d, pc.w32[i * (16 / h)], pc.P);
            }
        }
    }
}

template <class AR, int RLOG>
__device__ __forceinline__ void dft32_dit_inv(typename AR::V* x, const typename AR::PC32& pc) {
    constexpr int R = 1 << RLOG;
#pragma unroll
    for (int q = RLOG - 1; q >= 0; q--) {
        const int h = R >> (q + 1);
#pragma unroll
        for (int blk = 0; blk < R; blk += 2 * h) {
#pragma unroll
            for (int i = 0; i < h; i++) {
                auto a = x[blk + i], b = x[blk + i + h];
                if (i != 0) b = AR::mulc(b, pc.iw32[i * (16 / h)], pc.P);
                x[blk + i] = AR::add(a, b, pc.P);
                x[blk + i + h] = AR::sub(a, b, pc.P);
            }
        }
    }
}


// LDS line: pad one slot per R1 elements, odd line stride.
template <int LOGT, bool COLS>
struct T2Lds {
    static constexpr int R1 = 1 << t2_r1log(LOGT, COLS);
    static constexpr int LT = (1 << LOGT) + (1 << LOGT) / R1 + 1;
    __device__ static int pos(int q) { return q + q / R1; }
    // pos(k * R1 + j) for j < R1 (stage-0 element k of group j), with no division
    __device__ static int pos_kj(int k, int j) { return k * (R1 + 1) + j; }
};


// w_M^e from the two-level table, e < 2^table_log.
// Fused pass: load both operands' stage-0 words before transforming A.
#ifndef BW_FUSED_PREFETCH
#define BW_FUSED_PREFETCH 1
#endif

template <class AR>
__device__ __forceinline__ typename AR::TW tw_lookup(int64_t e, const typename AR::TW* lo, const typename AR::TW* hi,
                                                     const typename AR::PC32& pc) {
    const typename AR::TW a = lo[e & ((1 << BW_TW_LO_LOG) - 1)];
    const int64_t eh = e >> BW_TW_LO_LOG;
    return eh ? AR::twprod(a, hi[eh], AR::prime(pc)) : a;
}

// One workgroup = one tile of G lines (G a power of two, runtime) with
// NT = G * R1 threads: thread i owns stage-0 group i and stage-1 groups
// i, i + NT, ...  COLS (S > 1) and INGRESS (0: residue words in, 1: one raw
// limb per digit, 2: 64-bit digits from pairs of 32-bit limbs, 3: 32-bit limbs
// into 64-bit residue words) are compile-time so the inner loops carry no
// branches.
template <class AR, int LOGT, int MODE, bool COLS, int INGRESS>
__device__ __forceinline__ void ntt2_body(const PassParams& prm, int bx, int prime) {
    using V = typename AR::V;
    using W = typename AR::W;
    using TW = typename AR::TW;
    constexpr int T = 1 << LOGT;
    constexpr int R1LOG = t2_r1log(LOGT, COLS), R0LOG = t2_r0log(LOGT, COLS);
    constexpr int R1 = 1 << R1LOG, R0 = 1 << R0LOG;
    constexpr int LT = t2_lds_line(LOGT, COLS);   // the host sizes LDS with the same function
    constexpr int IT1 = R0 > R1 ? R0 / R1 : 1;     // stage-1 groups per thread (upper bound)
    constexpr int TWSH = BW_STAGE_LOG - LOGT;      // w_T = w_8192^(2^TWSH)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typename AR::LV* lds = reinterpret_cast<typename AR::LV*>(smem_raw);

    const int NT = blockDim.x;
    const int logG = __builtin_ctz(NT) - R1LOG;
    const int G = 1 << logG;
    // stage twiddles w_T^e (e < T) staged once per workgroup after the tile
    TW* twl = reinterpret_cast<TW*>(smem_raw + t2_tile_bytes(G, LT, (int)sizeof(typename AR::LV)));
    const int op = (MODE == BW_MODE_FWD) ? (prm.rev ? (int)gridDim.z - 1 - (int)blockIdx.z : (int)blockIdx.z) : 0;
    // read-only inputs of the launch through the constant address space, so
    // they load as scalars however the kernel's stores might alias them
    const typename AR::PC32 pc = bw_ldconst(reinterpret_cast<const typename AR::PC32*>(prm.consts), prime);
    const auto& P = AR::prime(pc);
    const TW* tws = reinterpret_cast<const TW*>(prm.tw_stage) + ((int64_t)prime << BW_STAGE_LOG);
    const TW* twlo = reinterpret_cast<const TW*>(prm.tw_lo) + ((int64_t)prime << BW_TW_LO_LOG);
    const TW* twhi = reinterpret_cast<const TW*>(prm.tw_hi) + ((int64_t)prime << (prm.table_log - BW_TW_LO_LOG));
    const int64_t S = prm.S;
    const int64_t l0 = (int64_t)bx << logG;
    const int64_t o0 = COLS ? l0 / S : l0;
    const int64_t s0 = COLS ? l0 - o0 * S : 0;
    const int twshift = prm.table_log - prm.logM;
    // The forward ingress pass of 256-point column tiles on the FP64 path
    // takes the pass-twiddle bases of column s = s0 + g from two LDS tables
    // filled once per workgroup: w^(b*g) and w^(R0*g) from ztab (e < R0*G, the
    // same for every tile) times w^(b*s0) and w^(R0*s0) from utab.  Looked up
    // per lane, they scatter every lane over the twiddle tables once M is
    // large (the ingress pass has M = N).  The other passes keep the per-lane
    // look-ups: the tables' LDS costs them waves per SIMD (measured: the
    // second forward pass of 2^24 points 5% slower with them, the larger
    // tiles lose a workgroup per CU, the small batched tiles 1-2%).
    constexpr bool CTAB = COLS && sizeof(W) == 4 && MODE == BW_MODE_FWD && R0 == R1 && INGRESS != 0;
    TW* ztab = twl + T;
    TW* utab = ztab + (R0 << logG);
    const int64_t plane_off = (int64_t)prime * prm.plane;
    const int tid = threadIdx.x;
    const bool canon = prm.canon_out != 0;
    const typename AR::NI* ninv = reinterpret_cast<const typename AR::NI*>(prm.ninv) + 2 * prime;

    // element t of line g lives at line_base(g) + t * tstride
    const int64_t tstride = COLS ? S : 1;
    auto line_base = [&](int g) -> int64_t { return COLS ? o0 * T * S + s0 + g : (o0 + g) * (int64_t)T; };
    // stage-0 group of this thread: line g0, offset j
    const int g0 = COLS ? (tid & (G - 1)) : (tid >> R1LOG);
    const int j0 = COLS ? (tid >> logG) : (tid & (R1 - 1));
    // Column tiles of 32-bit words: element (g, t) of the tile sits at the
    // wave-uniform tile0 plus g + t*S, below T*S <= 2^29 (the launcher
    // checks; 8 bytes per raw digit of the ingress pass still stay below
    // 2^32), so accesses go through a TileBuf with 32-bit byte offsets.
    constexpr bool LANE32 = COLS && sizeof(W) == 4;
    const int64_t tile0 = o0 * T * S + s0;
    // the column stride in bytes, pinned to a scalar register: every
    // wave-uniform buffer offset is a multiple of it
    const uint32_t sb = LANE32 ? (uint32_t)__builtin_amdgcn_readfirstlane((int)((uint32_t)S * (uint32_t)sizeof(W))) : 0u;
    auto lane_bytes = [&](int g, int t) -> uint32_t { return (uint32_t)g * (uint32_t)sizeof(W) + (uint32_t)t * sb; };

    // raw limbs of operand `which` for this thread's line, their count, and
    // the shift that turns a line position into a position inside the
    // product (stacked products: line o reads product o, shift o * M); the
    // raw limb type: a 64-bit digit is two 32-bit limbs.  Column tiles hold
    // one product per tile, so raw, len and the shift are wave-uniform.
    using RL = typename std::conditional<INGRESS == 2 || INGRESS == 3, uint32_t, typename AR::Raw>::type;
    auto ingress_src = [&](int which, const RL*& raw, int64_t& len, int64_t& shift) {
        raw = reinterpret_cast<const RL*>(prm.in[which]);
        len = prm.in_len[which];
        shift = 0;
        if (prm.bdesc) {
            const int64_t o = COLS ? o0 : o0 + g0;
            const BwIngress d = bw_ldconst(prm.bdesc, o);
            raw += which ? d.b_off : d.a_off;
            len = which ? d.b_len : d.a_len;
            shift = o << prm.logM;
        }
    };
    // stage-0 input words of one operand (raw digits or residue words)
    using IW = typename std::conditional<
        INGRESS == 2, uint64_t,
        typename std::conditional<INGRESS == 1, typename AR::Raw,
                                  typename std::conditional<INGRESS == 3, uint32_t, W>::type>::type>::type;
    auto digit_value = [&](IW w) -> V {
        if constexpr (INGRESS == 2) return AR::from_raw64(w, pc);
        else return AR::from_raw((typename AR::Raw)w, P);
    };
    // This thread's R0 stage-0 input words of one operand: digits base + m*step.
    // Raw operands are zero past their end.  Cases: every digit inside (loads
    // without per-element tests, 64-bit when the limb pair is 8-byte aligned),
    // every digit past the end (no loads), the lower half inside and the upper
    // half past the end (a zero-padded operand), and digits straddling the end
    // elsewhere (per-element tests).
    auto load0 = [&](int which, IW* w) -> bool {
        int64_t base = line_base(g0) + j0 * tstride;
        const int64_t step = (int64_t)R1 * tstride;
        if constexpr (INGRESS != 0) {
            const RL* raw;
            int64_t len, shift;
            ingress_src(which, raw, len, shift);
            base -= shift;
            constexpr int LW = INGRESS == 2 ? 2 : 1;          // limbs per digit
            const int64_t last = base + (int64_t)(R0 - 1) * step;
            const bool upper_zero = R0 > 1 && LW * (base + (int64_t)(R0 / 2) * step) >= len;
            // LANE32: digit base + m*step = (tile0 - shift) + (g0 + j0*S) + m*step,
            // the first term wave-uniform; byte offsets below 2^32 (T*S*8 <= 2^32)
            const TileBuf rb(raw + LW * (tile0 - shift));
            const uint32_t lb = LANE32 ? lane_bytes(g0, j0) * LW : 0u;
            auto lane_load = [&](int mlo, int mhi) {
                if constexpr (INGRESS == 2) {
                    if ((reinterpret_cast<uintptr_t>(raw) & 7) == 0) {
#pragma unroll
                        for (int m = mlo; m < mhi; m++) w[m] = rb.ld64(lb, (uint32_t)(m * R1 * LW) * sb);
                    } else {
#pragma unroll
                        for (int m = mlo; m < mhi; m++) {
                            const uint32_t uo = (uint32_t)(m * R1 * LW) * sb;
                            w[m] = (uint64_t)rb.ld(lb, uo) | ((uint64_t)rb.ld(lb + 4, uo) << 32);
                        }
                    }
                } else {
#pragma unroll
                    for (int m = mlo; m < mhi; m++) w[m] = (IW)rb.ld(lb, (uint32_t)(m * R1 * LW) * sb);
                }
            };
            if (LANE32 && LW * last + (LW - 1) < len) {
                lane_load(0, R0);
            } else if (LANE32 && upper_zero && LW * (base + (int64_t)(R0 / 2 - 1) * step) + (LW - 1) < len) {
                lane_load(0, R0 / 2);
#pragma unroll
                for (int m = R0 / 2; m < R0; m++) w[m] = (IW)0;
            } else if (LW * last + (LW - 1) < len) {
                if constexpr (INGRESS == 2) {
                    if ((reinterpret_cast<uintptr_t>(raw) & 7) == 0) {
                        const uint64_t* r64 = reinterpret_cast<const uint64_t*>(raw);
#pragma unroll
                        for (int m = 0; m < R0; m++) w[m] = r64[base + m * step];
                    } else {
#pragma unroll
                        for (int m = 0; m < R0; m++) {
                            const int64_t a = 2 * (base + m * step);
                            w[m] = (uint64_t)raw[a] | ((uint64_t)raw[a + 1] << 32);
                        }
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < R0; m++) w[m] = (IW)raw[base + m * step];
                }
            } else if (LW * base >= len) {
#pragma unroll
                for (int m = 0; m < R0; m++) w[m] = (IW)0;
            } else if (upper_zero && LW * (base + (int64_t)(R0 / 2 - 1) * step) + (LW - 1) < len) {
                // the common zero-padded case: lower half inside, upper half past the end
                if constexpr (INGRESS == 2) {
                    if ((reinterpret_cast<uintptr_t>(raw) & 7) == 0) {
                        const uint64_t* r64 = reinterpret_cast<const uint64_t*>(raw);
#pragma unroll
                        for (int m = 0; m < R0 / 2; m++) w[m] = r64[base + m * step];
                    } else {
#pragma unroll
                        for (int m = 0; m < R0 / 2; m++) {
                            const int64_t a = 2 * (base + m * step);
                            w[m] = (uint64_t)raw[a] | ((uint64_t)raw[a + 1] << 32);
                        }
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < R0 / 2; m++) w[m] = (IW)raw[base + m * step];
                }
#pragma unroll
                for (int m = R0 / 2; m < R0; m++) w[m] = (IW)0;
            } else {
                // straddles the end: digits below dfull are whole, digit dfull
                // may hold only its low limb (odd limb count)
                const int64_t dfull = len / LW;
#pragma unroll
                for (int m = 0; m < R0; m++) {
                    const int64_t a = base + m * step;
                    if constexpr (INGRESS == 2) {
                        w[m] = a < dfull ? ((uint64_t)raw[2 * a] | ((uint64_t)raw[2 * a + 1] << 32))
                                         : (a == dfull && (len & 1)) ? (uint64_t)raw[2 * a] : (uint64_t)0;
                    } else {
                        w[m] = a < dfull ? (IW)raw[a] : (IW)0;
                    }
                }
            }
            return upper_zero;
        } else if constexpr (LANE32) {
            const TileBuf in(reinterpret_cast<const W*>(prm.in[which]) + plane_off + tile0);
            const uint32_t lb = lane_bytes(g0, j0);
#pragma unroll
            for (int m = 0; m < R0; m++) w[m] = (W)in.ld(lb, (uint32_t)(m * R1) * sb);
            return false;
        } else {
            const W* in = reinterpret_cast<const W*>(prm.in[which]) + plane_off;
#pragma unroll
            for (int m = 0; m < R0; m++) w[m] = in[base + m * step];
            return false;
        }
    };
    // hz: the upper half of the group is zero padding (raw operands only)
    auto fwd_stage0_from = [&](const IW* w, bool hz) {
        V x[R0];
        if constexpr (INGRESS != 0 && R0 > 1) {
            if (hz) {
#pragma unroll
                for (int m = 0; m < R0 / 2; m++) x[m] = digit_value(w[m]);
                dft32_dif<AR, R0LOG, true>(x, pc);
            } else {
#pragma unroll
                for (int m = 0; m < R0; m++) x[m] = digit_value(w[m]);
                dft32_dif<AR, R0LOG>(x, pc);
            }
        } else {
#pragma unroll
            for (int m = 0; m < R0; m++) {
                if constexpr (INGRESS != 0) x[m] = digit_value(w[m]);
                else x[m] = AR::from_word(w[m], P);
            }
            dft32_dif<AR, R0LOG>(x, pc);
        }
        // k > 0: the stage twiddle product goes straight to the LDS word.  All
        // products first, then all writes: the twiddle reads share the LDS
        // buffer with the tile, so a write between two reads would order them.
        typename AR::LV v[R0];
        v[0] = AR::to_lds(x[0], P);
#pragma unroll
        for (int k = 1; k < R0; k++) v[k] = AR::lds_mul(x[brev(k, R0LOG)], twl[j0 * k], P);
#pragma unroll
        for (int k = 0; k < R0; k++) lds[g0 * LT + T2Lds<LOGT, COLS>::pos_kj(k, j0)] = v[k];
    };
    auto fwd_stage0 = [&](int which) {
        IW w[R0];
        const bool hz = load0(which, w);
        fwd_stage0_from(w, hz);
    };
    // fused pass: A's stage-1 spectrum, then the pointwise product, held in registers
    constexpr int KEEPN = IT1 * R1;
    typename AR::KV keep[KEEPN];
#pragma unroll
    for (int i = 0; i < KEEPN; i++) keep[i] = 0;
    auto dec1 = [&](int gi, int& g, int& b) {
        if (COLS) { g = gi & (G - 1); b = gi >> logG; } else { g = gi >> R0LOG; b = gi & (R0 - 1); }
    };
    // stage 1 forward: LDS -> DFT_R1 -> keep (fused) or pass twiddle -> HBM
    // KM (compile-time): 0 store to HBM with the pass twiddle, 1 keep[i] = x[i],
    // 2 keep[i] *= x[i], 3 keep[i] *= x[i] * N^-1 (the pointwise product)
    auto fwd_stage1 = [&](int which, auto km) {
        constexpr int KM = decltype(km)::value;
#pragma unroll
        for (int it = 0; it < IT1; it++) {
            const int gi = tid + it * NT;
            if (IT1 * NT > G * R0 && gi >= G * R0) continue;   // idle stage-1 slot
            int g, b;
            dec1(gi, g, b);
            V x[R1];
#pragma unroll
            for (int m = 0; m < R1; m++) x[m] = AR::from_lds(lds[g * LT + T2Lds<LOGT, COLS>::pos(b * R1 + m)]);
            dft32_dif<AR, R1LOG>(x, pc);
            if constexpr (KM == 1) {
#pragma unroll
                for (int k = 0; k < R1; k++) keep[it * R1 + k] = AR::to_keep(x[brev(k, R1LOG)], P);
                continue;
            } else if constexpr (KM == 2) {
#pragma unroll
                for (int k = 0; k < R1; k++) keep[it * R1 + k] = AR::keep_mul(keep[it * R1 + k], x[brev(k, R1LOG)], P);
                continue;
            } else if constexpr (KM == 3) {
#pragma unroll
                for (int k = 0; k < R1; k++)
                    keep[it * R1 + k] = AR::keep_mul_scaled(keep[it * R1 + k], x[brev(k, R1LOG)], ninv, P);
                continue;
            }
            W* out = reinterpret_cast<W*>(prm.out[which]) + plane_off + line_base(g) + (int64_t)b * R1 * tstride;
            const TileBuf out_u(reinterpret_cast<W*>(prm.out[which]) + plane_off + tile0);   // LANE32
            const uint32_t lb = LANE32 ? lane_bytes(g, b * R1) : 0u;
            auto put = [&](int k, W v) {
                if constexpr (LANE32) out_u.st((uint32_t)v, lb, (uint32_t)k * sb);
                else out[k * tstride] = v;
            };
            if constexpr (COLS) {
                const int64_t s = s0 + g;
                const TW base = CTAB ? AR::twprod(utab[b], ztab[b * g], P)
                                     : tw_lookup<AR>(((int64_t)b * s) << twshift, twlo, twhi, pc);
                const TW step = CTAB ? AR::twprod(utab[R0], ztab[R0 * g], P)
                                     : tw_lookup<AR>(((int64_t)R0 * s) << twshift, twlo, twhi, pc);
                // each pass-twiddle product goes straight to its storage word
                // (centered: a forward column pass is never a transform's last)
                AR::template for_powers<R1>(base, step, P,
                                            [&](int k, auto w) { put(k, AR::word_mul(x[brev(k, R1LOG)], w, P)); });
            } else if (canon) {
#pragma unroll
                for (int k = 0; k < R1; k++) put(k, AR::to_word_c(x[brev(k, R1LOG)], P, true));
            } else {
#pragma unroll
                for (int k = 0; k < R1; k++) put(k, AR::to_word_c(x[brev(k, R1LOG)], P, false));
            }
        }
    };
    // stage 1 inverse: keep (fused) or HBM with the inverse pass twiddle -> inverse DFT_R1
    // -> inverse stage twiddle -> LDS
    // The inverse pass (IT1 == 1) waits for the stage-twiddle fill only here,
    // after its loads and its DFT: every thread, idle slots included, reaches
    // that barrier.
    auto inv_stage1 = [&](auto km) {
        constexpr bool KEEP = decltype(km)::value != 0;
        constexpr bool LATE_SYNC = !KEEP && MODE == BW_MODE_INV && IT1 == 1;
#pragma unroll
        for (int it = 0; it < IT1; it++) {
            const int gi = tid + it * NT;
            const bool idle = IT1 * NT > G * R0 && gi >= G * R0;   // idle stage-1 slot
            if (idle && !LATE_SYNC) continue;
            int g, b;
            dec1(gi, g, b);
            V x[R1];
            if (idle) {
                // LATE_SYNC only: no work, but the barrier below
            } else if constexpr (KEEP) {
#pragma unroll
                for (int k = 0; k < R1; k++) x[brev(k, R1LOG)] = AR::from_keep(keep[it * R1 + k]);
            } else {
                const W* in = reinterpret_cast<const W*>(prm.in[0]) + plane_off + line_base(g) + (int64_t)b * R1 * tstride;
                const TileBuf in_u(reinterpret_cast<const W*>(prm.in[0]) + plane_off + tile0);   // LANE32
                const uint32_t lb = LANE32 ? lane_bytes(g, b * R1) : 0u;
                auto get = [&](int k) -> W {
                    if constexpr (LANE32) return (W)in_u.ld(lb, (uint32_t)k * sb);
                    else return in[k * tstride];
                };
                if constexpr (COLS) {
                    const int64_t s = s0 + g;
                    const int64_t mask = ((int64_t)1 << prm.table_log) - 1;
                    TW base = CTAB ? AR::twprod(utab[b], ztab[b * g], P)
                                   : tw_lookup<AR>((-(((int64_t)b * s) << twshift)) & mask, twlo, twhi, pc);
                    if (prm.fold_ninv) base = AR::tw_mul_ninv(base, ninv, P);
                    const TW step = CTAB ? AR::twprod(utab[R0], ztab[R0 * g], P)
                                         : tw_lookup<AR>((-(((int64_t)R0 * s) << twshift)) & mask, twlo, twhi, pc);
                    AR::template for_powers<R1>(base, step, P, [&](int k, auto w) {
                        x[brev(k, R1LOG)] = AR::apply_tw(AR::from_word(get(k), P), w, P);
                    });
                } else {
#pragma unroll
                    for (int k = 0; k < R1; k++) x[brev(k, R1LOG)] = AR::from_word(get(k), P);
                }
            }
            if (!idle) dft32_dit_inv<AR, R1LOG>(x, pc);
            if constexpr (LATE_SYNC) {
                __syncthreads();
                if (idle) continue;
            }
            // the inverse stage twiddle of position b*R1 + m (stage-0 group m,
            // index b) is applied here: its product leaves the value reduced.
            // m = 0 is untwiddled (a reduction); b = 0 multiplies by twl[0] = 1,
            // which keeps the loop free of branches.
            // all products (their twiddle reads) before the LDS writes
            typename AR::LV v[R1];
            v[0] = AR::to_lds(x[0], P);
#pragma unroll
            for (int m = 1; m < R1; m++) v[m] = AR::lds_mul(x[m], twl[(T - m * b) & (T - 1)], P);
#pragma unroll
            for (int m = 0; m < R1; m++) lds[g * LT + T2Lds<LOGT, COLS>::pos(b * R1 + m)] = v[m];
        }
    };
    // stage 0 inverse: LDS (stage twiddles already applied) -> inverse DFT_R0 -> HBM
    auto inv_stage0 = [&]() {
        V x[R0];
#pragma unroll
        for (int k = 0; k < R0; k++) x[brev(k, R0LOG)] = AR::from_lds(lds[g0 * LT + T2Lds<LOGT, COLS>::pos_kj(k, j0)]);
        dft32_dit_inv<AR, R0LOG>(x, pc);     // stage twiddles were applied before the transpose
        W* out = reinterpret_cast<W*>(prm.out[0]) + plane_off + line_base(g0) + j0 * tstride;
        const TileBuf out_u(reinterpret_cast<W*>(prm.out[0]) + plane_off + tile0);    // LANE32
        const uint32_t lb = LANE32 ? lane_bytes(g0, j0) : 0u;
        const int64_t step = (int64_t)R1 * tstride;
        auto put = [&](int m, W v) {
            if constexpr (LANE32) out_u.st((uint32_t)v, lb, (uint32_t)(m * R1) * sb);
            else out[m * step] = v;
        };
        if (canon) {
#pragma unroll
            for (int m = 0; m < R0; m++) put(m, AR::to_word_c(x[m], P, true));
        } else {
#pragma unroll
            for (int m = 0; m < R0; m++) put(m, AR::to_word_c(x[m], P, false));
        }
    };

    auto fill_twiddles = [&]() {
        for (int i = tid; i < T; i += NT) twl[i] = tws[i << TWSH];
        if constexpr (CTAB) {
            // conjugate powers for the inverse
            const int64_t mask = ((int64_t)1 << prm.table_log) - 1;
            const int64_t sg = MODE == BW_MODE_INV ? -1 : 1;
            for (int i = tid; i < (R0 << logG); i += NT)
                ztab[i] = tw_lookup<AR>((sg * ((int64_t)i << twshift)) & mask, twlo, twhi, pc);
            for (int i = tid; i <= R0; i += NT)
                utab[i] = tw_lookup<AR>((sg * (((int64_t)i * s0) << twshift)) & mask, twlo, twhi, pc);
        }
    };
    if constexpr (MODE == BW_MODE_FWD) {
        IW w[R0];
        const bool hz = load0(op, w);
        fill_twiddles();
        __syncthreads();
        fwd_stage0_from(w, hz);
        __syncthreads();
        fwd_stage1(op, std::integral_constant<int, 0>());
    } else if constexpr (MODE == BW_MODE_INV) {
        fill_twiddles();
        if constexpr (IT1 > 1) __syncthreads();   // else inside inv_stage1, after its loads
        inv_stage1(std::integral_constant<int, 0>());
        __syncthreads();
        inv_stage0();
    } else {
        // prefetch B's stage-0 words while A is transformed, unless the two
        // register sets would cost occupancy (radix-32 stage 0)
        constexpr bool PREFETCH = BW_FUSED_PREFETCH && R0 <= 16;
        IW wb[PREFETCH ? R0 : 1];
        bool hzb = false;
        if constexpr (PREFETCH) {
            IW wa[R0];
            const bool hza = load0(0, wa);
            hzb = load0(1, wb);
            fill_twiddles();
            __syncthreads();
            fwd_stage0_from(wa, hza);
        } else {
            IW wa[R0];
            const bool hza = load0(0, wa);
            fill_twiddles();
            __syncthreads();
            fwd_stage0_from(wa, hza);
        }
        __syncthreads();
        fwd_stage1(0, std::integral_constant<int, 1>());
        __syncthreads();
        if constexpr (PREFETCH) fwd_stage0_from(wb, hzb);
        else fwd_stage0(1);
        __syncthreads();
        // B's spectrum multiplies into A's
        if (prm.fold_ninv) fwd_stage1(1, std::integral_constant<int, 2>());
        else fwd_stage1(1, std::integral_constant<int, 3>());
        __syncthreads();
        inv_stage1(std::integral_constant<int, 1>());
        __syncthreads();
        inv_stage0();
    }
}

// The kernel: which tile and prime a workgroup takes.  The ingress pass
// walks the tiles of one prime after another (blockIdx.x fastest: the raw
// operands, read by every prime, stream through in order once per prime);
// the other passes take the primes of a tile back to back, which measured
// 3-7% faster for them.  prm.rev reverses the whole walk (last tile, prime
// and operand first), so that a pass starts on the data its predecessor
// wrote last.
template <class AR, int LOGT, int MODE, bool COLS, int INGRESS>
constexpr int ntt2_max_threads() {
    return (MODE == BW_MODE_INV && COLS && t2_inv_wide_ok(LOGT, (int)sizeof(typename AR::LV))) ? 1024 : 512;
}

template <class AR, int LOGT, int MODE, bool COLS, int INGRESS>
__global__ __launch_bounds__((ntt2_max_threads<AR, LOGT, MODE, COLS, INGRESS>())) void ntt2_kernel(PassParams prm) {
    const uint32_t bx = prm.rev ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
    const uint32_t by = prm.rev ? gridDim.y - 1 - blockIdx.y : blockIdx.y;
    if constexpr (INGRESS == 0) {
        const uint32_t id = bx + by * gridDim.x;
        ntt2_body<AR, LOGT, MODE, COLS, INGRESS>(prm, (int)(id / gridDim.y), prm.prime0 + (int)(id % gridDim.y));
    } else {
        ntt2_body<AR, LOGT, MODE, COLS, INGRESS>(prm, (int)bx, prm.prime0 + (int)by);
    }
}

typedef void (*PassKernel2)(PassParams);

template <class AR, int MODE, int LOGT>
static PassKernel2 pick2_v(bool cols, int ingress) {
    if constexpr (MODE == BW_MODE_INV) {
        if (ingress) return nullptr;
        return cols ? ntt2_kernel<AR, LOGT, MODE, true, 0> : ntt2_kernel<AR, LOGT, MODE, false, 0>;
    } else if constexpr (MODE == BW_MODE_FUSED) {
        if (cols) return nullptr;
        if (ingress == 3) {
            if constexpr (std::is_same<AR, ArU64x>::value) return ntt2_kernel<AR, LOGT, MODE, false, 3>;
            return nullptr;
        }
        if (ingress == 2) return ntt2_kernel<AR, LOGT, MODE, false, 2>;
        return ingress ? ntt2_kernel<AR, LOGT, MODE, false, 1> : ntt2_kernel<AR, LOGT, MODE, false, 0>;
    } else {
        if (ingress == 3) {
            if constexpr (std::is_same<AR, ArU64x>::value)
                return cols ? ntt2_kernel<AR, LOGT, MODE, true, 3> : ntt2_kernel<AR, LOGT, MODE, false, 3>;
            return nullptr;
        }
        if (ingress == 2) return cols ? ntt2_kernel<AR, LOGT, MODE, true, 2> : ntt2_kernel<AR, LOGT, MODE, false, 2>;
        if (cols) return ingress ? ntt2_kernel<AR, LOGT, MODE, true, 1> : ntt2_kernel<AR, LOGT, MODE, true, 0>;
        return ingress ? ntt2_kernel<AR, LOGT, MODE, false, 1> : ntt2_kernel<AR, LOGT, MODE, false, 0>;
    }
}

template <class AR, int MODE>
static PassKernel2 pick2(int logT, bool cols, int ingress) {
    switch (logT) {
    case 1: return pick2_v<AR, MODE, 1>(cols, ingress);
    case 2: return pick2_v<AR, MODE, 2>(cols, ingress);
    case 3: return pick2_v<AR, MODE, 3>(cols, ingress);
    case 4: return pick2_v<AR, MODE, 4>(cols, ingress);
    case 5: return pick2_v<AR, MODE, 5>(cols, ingress);
    case 6: return pick2_v<AR, MODE, 6>(cols, ingress);
    case 7: return pick2_v<AR, MODE, 7>(cols, ingress);
    case 8: return pick2_v<AR, MODE, 8>(cols, ingress);
    case 9: return pick2_v<AR, MODE, 9>(cols, ingress);
    case 10: return pick2_v<AR, MODE, 10>(cols, ingress);
    default: return nullptr;
    }
}
