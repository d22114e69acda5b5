// ntt_pass_impl.h -- templates of the NTT pass kernels (see ntt_kernels.h).
// Included by the per-(arithmetic, mode) translation units ntt_<ar>_<mode>.hip.
#pragma once
#include "ntt_kernels.h"

// ---------------------------------------------------------------------------
// Arithmetic policies.
struct ArF64 {
    using V = double;          // register value: signed integer held in FP64
    using W = uint32_t;        // storage word (canonical residue)
    using Raw = uint32_t;      // raw input limb
    using TW = double2;        // twiddle (w, w/p)
    using PC = F64PrimeDev;
    using NI = double;         // element type of (N^-1, aux) pairs
    __device__ static V from_word(W x, const PC&) { return (double)(int32_t)x; }   // canonical or centered
    __device__ static V from_raw(Raw x, const PC&) { return (double)x; }   // < 2^32: no reduction needed
    __device__ static W to_word(V x, const PC& c) { return f64_canon(x, c.P); }
    __device__ static V add(V a, V b, const PC&) { return a + b; }
    __device__ static V sub(V a, V b, const PC&) { return a - b; }
    __device__ static V mul(V y, TW w, const PC& c) { return f64_mulmod(y, w.x, w.y, c.P); }
    __device__ static V mulc(V y, const double* w, const PC& c) { return f64_mulmod(y, w[0], w[1], c.P); }
    // a*b*N^-1 for a, b < 2^46 in magnitude
    __device__ static V mul_scaled(V a, V b, const double* ninv, const PC& c) {
        double br = f64_reduce(b, c.P);
        double r = f64_mulmod(a, br, br * c.P.pinv, c.P);
        return f64_mulmod(r, ninv[0], ninv[1], c.P);
    }
    __device__ static V zero() { return 0.0; }
};

struct ArU64 {
    using V = uint64_t;
    using W = uint64_t;
    using Raw = uint64_t;
    using TW = ulonglong2;     // Shoup pair (w, floor(w 2^64/p))
    using PC = U64PrimeDev;
    using NI = uint64_t;
    __device__ static V from_word(W x, const PC&) { return x; }
    __device__ static V from_raw(Raw x, const PC& c) { return u64_reduce(x, c.P); }
    __device__ static W to_word(V x, const PC&) { return x; }
    __device__ static V add(V a, V b, const PC& c) { return u64_add(a, b, c.P.p); }
    __device__ static V sub(V a, V b, const PC& c) { return u64_sub(a, b, c.P.p); }
    __device__ static V mul(V y, TW w, const PC& c) { return u64_mulmod_shoup(y, w.x, w.y, c.P.p); }
    __device__ static V mulc(V y, const uint64_t* w, const PC& c) { return u64_mulmod_shoup(y, w[0], w[1], c.P.p); }
    // montmul(montmul(a, b), ninv[0]) with ninv[0] = N^-1 * 2^128 mod p
    __device__ static V mul_scaled(V a, V b, const uint64_t* ninv, const PC& c) {
        return u64_montmul(u64_montmul(a, b, c.P), ninv[0], c.P);
    }
    __device__ static V zero() { return 0; }
};

// ---------------------------------------------------------------------------
// Register codelets: 2^RLOG-point DFT (RLOG <= 4).
// Forward: decimation in frequency, natural input, output X[k] in x[bitrev(k)].
template <class AR, int RLOG>
__device__ __forceinline__ void dft_dif(typename AR::V* x, const typename AR::PC& pc) {
    constexpr int R = 1 << RLOG;
#pragma unroll
    for (int q = 0; q < RLOG; q++) {
        const int h = R >> (q + 1);
#pragma unroll
        for (int blk = 0; blk < R; blk += 2 * h) {
#pragma unroll
            for (int i = 0; i < h; i++) {
                auto a = x[blk + i], b = x[blk + i + h];
                x[blk + i] = AR::add(a, b, pc);
                auto d = AR::sub(a, b, pc);
                x[blk + i + h] = (i == 0) ? d : AR::mulc(d, pc.w32[i * (16 / h)], pc);
            }
        }
    }
}

// Inverse: decimation in time with inverse roots, input Z[k] in x[bitrev(k)],
// output R * z[m] in natural order.
template <class AR, int RLOG>
__device__ __forceinline__ void dft_dit_inv(typename AR::V* x, const typename AR::PC& pc) {
    constexpr int R = 1 << RLOG;
#pragma unroll
    for (int q = RLOG - 1; q >= 0; q--) {
        const int h = R >> (q + 1);
#pragma unroll
        for (int blk = 0; blk < R; blk += 2 * h) {
#pragma unroll
            for (int i = 0; i < h; i++) {
                auto a = x[blk + i], b = x[blk + i + h];
                if (i != 0) b = AR::mulc(b, pc.iw32[i * (16 / h)], pc);
                x[blk + i] = AR::add(a, b, pc);
                x[blk + i + h] = AR::sub(a, b, pc);
            }
        }
    }
}

__device__ __forceinline__ constexpr int brev(int k, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((k >> i) & 1) << (bits - 1 - i);
    return r;
}

// ---------------------------------------------------------------------------
// One radix stage of a line held in LDS.  Groups of R elements spaced by
// stride = span/R; after the codelet, output k is twiddled by w_span^{j k}.
template <class AR, int LOGT, int G, int NT, int I, bool KEEP>
__device__ __forceinline__ void stage_fwd(typename AR::V* lds, const typename AR::PC& pc,
                                          const typename AR::TW* tws, typename AR::V* keep) {
    constexpr int T = 1 << LOGT;
    constexpr int LT = lds_line(LOGT);
    constexpr int RLOG = sched_rlog(LOGT, I);
    constexpr int R = 1 << RLOG;
    constexpr int SPANLOG = sched_spanlog(LOGT, I);
    constexpr int STRIDELOG = SPANLOG - RLOG;
    constexpr int STRIDE = 1 << STRIDELOG;
    constexpr int GROUPS = G * (T / R);
    constexpr int TWSHIFT = BW_STAGE_LOG - SPANLOG;   // w_span = w_8192^(2^TWSHIFT)
    const int tid = threadIdx.x;
#pragma unroll
    for (int it = 0; it < (GROUPS + NT - 1) / NT; it++) {
        const int gi = tid + it * NT;
        if (GROUPS % NT != 0 && gi >= GROUPS) break;
        const int g = gi / (T / R);
        const int gl = gi % (T / R);
        const int b = gl >> STRIDELOG;
        const int j = gl & (STRIDE - 1);
        const int base = (b << SPANLOG) + j;
        typename AR::V x[R];
#pragma unroll
        for (int m = 0; m < R; m++) x[m] = lds[g * LT + lds_pad(base + m * STRIDE)];
        dft_dif<AR, RLOG>(x, pc);
        if (STRIDE > 1) {
#pragma unroll
            for (int k = 1; k < R; k++) {
                const int e = (j * k) << TWSHIFT;
                if (e != 0) x[brev(k, RLOG)] = AR::mul(x[brev(k, RLOG)], tws[e], pc);
            }
        }
        if (KEEP) {
#pragma unroll
            for (int k = 0; k < R; k++) keep[it * R + k] = x[brev(k, RLOG)];
        } else {
#pragma unroll
            for (int k = 0; k < R; k++) lds[g * LT + lds_pad(base + k * STRIDE)] = x[brev(k, RLOG)];
        }
    }
}

// Inverse of stage_fwd.  FROMKEEP: the stage input comes from registers
// (fused pointwise product) instead of LDS.
template <class AR, int LOGT, int G, int NT, int I, bool FROMKEEP>
__device__ __forceinline__ void stage_inv(typename AR::V* lds, const typename AR::PC& pc,
                                          const typename AR::TW* tws, const typename AR::V* keep) {
    constexpr int T = 1 << LOGT;
    constexpr int LT = lds_line(LOGT);
    constexpr int RLOG = sched_rlog(LOGT, I);
    constexpr int R = 1 << RLOG;
    constexpr int SPANLOG = sched_spanlog(LOGT, I);
    constexpr int STRIDELOG = SPANLOG - RLOG;
    constexpr int STRIDE = 1 << STRIDELOG;
    constexpr int GROUPS = G * (T / R);
    constexpr int TWSHIFT = BW_STAGE_LOG - SPANLOG;
    constexpr int TWMASK = (1 << BW_STAGE_LOG) - 1;
    const int tid = threadIdx.x;
#pragma unroll
    for (int it = 0; it < (GROUPS + NT - 1) / NT; it++) {
        const int gi = tid + it * NT;
        if (GROUPS % NT != 0 && gi >= GROUPS) break;
        const int g = gi / (T / R);
        const int gl = gi % (T / R);
        const int b = gl >> STRIDELOG;
        const int j = gl & (STRIDE - 1);
        const int base = (b << SPANLOG) + j;
        typename AR::V x[R];
#pragma unroll
        for (int k = 0; k < R; k++)
            x[brev(k, RLOG)] = FROMKEEP ? keep[it * R + k] : lds[g * LT + lds_pad(base + k * STRIDE)];
        if (STRIDE > 1) {
#pragma unroll
            for (int k = 1; k < R; k++) {
                const int e = (j * k) << TWSHIFT;
                if (e != 0) x[brev(k, RLOG)] = AR::mul(x[brev(k, RLOG)], tws[(-e) & TWMASK], pc);
            }
        }
        dft_dit_inv<AR, RLOG>(x, pc);
#pragma unroll
        for (int m = 0; m < R; m++) lds[g * LT + lds_pad(base + m * STRIDE)] = x[m];
    }
}

template <class AR, int LOGT, int G, int NT, int I>
struct FwdStages {
    __device__ static void run(typename AR::V* lds, const typename AR::PC& pc, const typename AR::TW* tws, int upto) {
        if constexpr (I < sched_nst(LOGT)) {
            if (I < upto) {
                stage_fwd<AR, LOGT, G, NT, I, false>(lds, pc, tws, nullptr);
                __syncthreads();
                FwdStages<AR, LOGT, G, NT, I + 1>::run(lds, pc, tws, upto);
            }
        }
    }
};

template <class AR, int LOGT, int G, int NT, int I>
struct InvStages {   // runs stages I, I-1, ..., 0
    __device__ static void run(typename AR::V* lds, const typename AR::PC& pc, const typename AR::TW* tws) {
        if constexpr (I >= 0) {
            stage_inv<AR, LOGT, G, NT, I, false>(lds, pc, tws, nullptr);
            __syncthreads();
            InvStages<AR, LOGT, G, NT, I - 1>::run(lds, pc, tws);
        }
    }
};

// Pass twiddle w_M^{f s}; e = f*s*2^(L - logM) < 2^L, negated for the inverse.
template <class AR>
__device__ __forceinline__ typename AR::V apply_pass_tw(typename AR::V v, int64_t e, int table_log, bool inverse,
                                                       const typename AR::TW* lo, const typename AR::TW* hi,
                                                       const typename AR::PC& pc) {
    const int64_t mask = (int64_t(1) << table_log) - 1;
    if (inverse) e = (-e) & mask;
    if (e == 0) return v;
    const int64_t elo = e & ((1 << BW_TW_LO_LOG) - 1);
    const int64_t ehi = e >> BW_TW_LO_LOG;
    if (elo) v = AR::mul(v, lo[elo], pc);
    if (ehi) v = AR::mul(v, hi[ehi], pc);
    return v;
}

// ---------------------------------------------------------------------------
template <class AR, int LOGT, int G, int NT, int MODE>
__global__ __launch_bounds__(NT) void ntt_pass_kernel(PassParams prm) {
    using V = typename AR::V;
    using W = typename AR::W;
    using TW = typename AR::TW;
    constexpr int T = 1 << LOGT;
    constexpr int LT = lds_line(LOGT);
    constexpr int NST = sched_nst(LOGT);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);

    const int prime = prm.prime0 + blockIdx.y;
    const int op = (MODE == BW_MODE_FWD) ? blockIdx.z : 0;
    const typename AR::PC pc = reinterpret_cast<const typename AR::PC*>(prm.consts)[prime];
    const TW* tws = reinterpret_cast<const TW*>(prm.tw_stage) + ((int64_t)prime << BW_STAGE_LOG);
    const TW* twlo = reinterpret_cast<const TW*>(prm.tw_lo) + ((int64_t)prime << BW_TW_LO_LOG);
    const TW* twhi = reinterpret_cast<const TW*>(prm.tw_hi) + ((int64_t)prime << (prm.table_log - BW_TW_LO_LOG));
    const int64_t S = prm.S;
    const bool cols = S > 1;
    const int64_t l0 = (int64_t)blockIdx.x * G;
    const int64_t o0 = cols ? l0 / S : l0;
    const int64_t s0 = cols ? l0 % S : 0;
    const int twshift = prm.table_log - prm.logM;
    const int tid = threadIdx.x;

    // address (within a prime plane) of element (g, t) of this tile
    auto addr_of = [&](int g, int t) -> int64_t {
        return cols ? (o0 * T + t) * S + s0 + g : (o0 + g) * (int64_t)T + t;
    };
    auto load_tile = [&](int which, bool inv_tw) {
        const int64_t plane_off = (int64_t)prime * prm.plane;
#pragma unroll 4
        for (int e = tid; e < G * T; e += NT) {
            const int g = cols ? e % G : e / T;
            const int t = cols ? e / G : e % T;
            const int64_t a = addr_of(g, t);
            V v;
            if (prm.ingress) {
                const typename AR::Raw* raw = reinterpret_cast<const typename AR::Raw*>(prm.in[which]);
                v = a < prm.in_len[which] ? AR::from_raw(raw[a], pc) : AR::zero();
            } else {
                v = AR::from_word(reinterpret_cast<const W*>(prm.in[which])[plane_off + a], pc);
            }
            if (inv_tw && cols) {
                const int64_t f = sched_freq_of_pos(LOGT, t);
                v = apply_pass_tw<AR>(v, (f * (s0 + g)) << twshift, prm.table_log, true, twlo, twhi, pc);
            }
            lds[g * LT + lds_pad(t)] = v;
        }
    };
    auto store_tile = [&](int which, bool fwd_tw) {
        const int64_t plane_off = (int64_t)prime * prm.plane;
        W* out = reinterpret_cast<W*>(prm.out[which]);
#pragma unroll 4
        for (int e = tid; e < G * T; e += NT) {
            const int g = cols ? e % G : e / T;
            const int t = cols ? e / G : e % T;
            V v = lds[g * LT + lds_pad(t)];
            if (fwd_tw && cols) {
                const int64_t f = sched_freq_of_pos(LOGT, t);
                v = apply_pass_tw<AR>(v, (f * (s0 + g)) << twshift, prm.table_log, false, twlo, twhi, pc);
            }
            out[plane_off + addr_of(g, t)] = AR::to_word(v, pc);
        }
    };

    if constexpr (MODE == BW_MODE_FWD) {
        load_tile(op, false);
        __syncthreads();
        FwdStages<AR, LOGT, G, NT, 0>::run(lds, pc, tws, NST);
        store_tile(op, true);
    } else if constexpr (MODE == BW_MODE_INV) {
        load_tile(0, true);
        __syncthreads();
        InvStages<AR, LOGT, G, NT, NST - 1>::run(lds, pc, tws);
        store_tile(0, false);
    } else {
        // FUSED (last pass, S == 1): forward A and B, product scaled by N^-1,
        // then the inverse of this pass.  The last forward stage's outputs stay
        // in registers; the same thread owns the same positions for A and B.
        constexpr int LR = 1 << sched_rlog(LOGT, NST - 1);
        constexpr int LGROUPS = G * (T / LR);
        constexpr int KEEPN = ((LGROUPS + NT - 1) / NT) * LR;
        V keepA[KEEPN], keepB[KEEPN];
#pragma unroll
        for (int i = 0; i < KEEPN; i++) { keepA[i] = AR::zero(); keepB[i] = AR::zero(); }
        load_tile(0, false);
        __syncthreads();
        FwdStages<AR, LOGT, G, NT, 0>::run(lds, pc, tws, NST - 1);
        stage_fwd<AR, LOGT, G, NT, NST - 1, true>(lds, pc, tws, keepA);
        __syncthreads();
        load_tile(1, false);
        __syncthreads();
        FwdStages<AR, LOGT, G, NT, 0>::run(lds, pc, tws, NST - 1);
        stage_fwd<AR, LOGT, G, NT, NST - 1, true>(lds, pc, tws, keepB);
        const typename AR::NI* ninv = reinterpret_cast<const typename AR::NI*>(prm.ninv) + 2 * prime;
#pragma unroll
        for (int i = 0; i < KEEPN; i++) keepA[i] = AR::mul_scaled(keepA[i], keepB[i], ninv, pc);
        __syncthreads();
        stage_inv<AR, LOGT, G, NT, NST - 1, true>(lds, pc, tws, keepA);
        __syncthreads();
        InvStages<AR, LOGT, G, NT, NST - 2>::run(lds, pc, tws);
        store_tile(0, false);
    }
}

// ---------------------------------------------------------------------------
// Dispatch table.
typedef void (*PassKernel)(PassParams);

template <class AR, int MODE, int LOGT, int G>
static PassKernel pick_kernel() { return ntt_pass_kernel<AR, LOGT, G, 256, MODE>; }

template <class AR, int MODE, int LOGT>
static PassKernel pick_G(int G) {
    // lines per workgroup: G * 2^LOGT <= 8192 keeps LDS at <= 68 KiB (two workgroups per CU)
    if constexpr ((1 << LOGT) * 16 <= 8192) { if (G == 16) return pick_kernel<AR, MODE, LOGT, 16>(); }
    if constexpr ((1 << LOGT) * 8 <= 8192) { if (G == 8) return pick_kernel<AR, MODE, LOGT, 8>(); }
    if constexpr ((1 << LOGT) * 4 <= 8192) { if (G == 4) return pick_kernel<AR, MODE, LOGT, 4>(); }
    if constexpr ((1 << LOGT) * 2 <= 8192) { if (G == 2) return pick_kernel<AR, MODE, LOGT, 2>(); }
    if (G == 1) return pick_kernel<AR, MODE, LOGT, 1>();
    return nullptr;
}

template <class AR, int MODE>
static PassKernel pick_logT(int logT, int G) {
    switch (logT) {
    case 1: return pick_G<AR, MODE, 1>(G);
    case 2: return pick_G<AR, MODE, 2>(G);
    case 3: return pick_G<AR, MODE, 3>(G);
    case 4: return pick_G<AR, MODE, 4>(G);
    case 5: return pick_G<AR, MODE, 5>(G);
    case 6: return pick_G<AR, MODE, 6>(G);
    case 7: return pick_G<AR, MODE, 7>(G);
    case 8: return pick_G<AR, MODE, 8>(G);
    case 9: return pick_G<AR, MODE, 9>(G);
    case 10: return pick_G<AR, MODE, 10>(G);
    case 11: return pick_G<AR, MODE, 11>(G);
    case 12: return pick_G<AR, MODE, 12>(G);
    case 13: return pick_G<AR, MODE, 13>(G);
    default: return nullptr;
    }
}

