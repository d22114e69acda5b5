// product_fused.hip -- whole products in one workgroup (product_fused.h).
//
// A bundle of K products of transform length n = 2^L (64-bit digits, five
// primes) runs in one workgroup:
//   for the five primes at once (one chunk):
//     their twiddle tables omega_n^e (e < n) into LDS;
//     forward transforms of every product's A (into the prime's residue
//     plane) and B (into a chunk buffer): decimation-in-frequency stages of
//     radix 8 and 16 (register-resident codelets; two stages up to 2^8),
//     in place in LDS, the first stage reading the operand digits from HBM;
//     pointwise product times n^-1 into the residue plane;
//     inverse transform of the plane: the same stages mirrored (inverse
//     twiddles before each radix-R inverse DFT, last stage first);
//   Garner: the five residues of each coefficient -> three 64-bit words (LDS);
//   carry: t_k = lo(s_k) + hi(s_{k-1}) with s_k = sum_j w_j(c_{k-j}), binary
//   generate / propagate per limb resolved by wave ballots per row of 64
//   limbs and one sequential scan over the bundle's rows (at most 80).
// Values between stages are centered int32 residues; inside a stage they are
// FP64 integers (exact modular products, arith.h).  LDS positions of
// transform elements and twiddles are XOR-swizzled (sw()).
//
// LDS (32-bit words): twiddle tables [5][n] at 0, then product k's region
// at 5n + 10kn: five residue planes, then five B buffers.
#include <algorithm>
#include <mutex>

#include "garner_dev.h"
#include "ntt2_impl.h"
#include "product_fused.h"

namespace {

using AR = ArF64x;
constexpr int NT = BW_FUSED_THREADS;
constexpr int NWAVES = NT / 64;
constexpr int TW_N = 1 << BW_STAGE_LOG;
// rows of 64 output limbs per wave: K * ((n >> 6) + 1) <= 80 rows per bundle
constexpr int CARRY_ROWS = 5;
constexpr int MAX_ROWS = CARRY_ROWS * NWAVES;
// coefficients per thread in the Garner step: K * n <= 3072
constexpr int CPT = 3;
// primes transformed together: all five (five B buffers beside the five
// residue planes of every product of the bundle)
constexpr int NB = 5;

__host__ __device__ constexpr int rows_per(int L) { return ((1 << L) >> 6) + 1; }
static_assert(bw_fused_bundle_size(11) * rows_per(11) <= MAX_ROWS, "carry rows");
static_assert(bw_fused_bundle_size(10) * rows_per(10) <= MAX_ROWS, "carry rows");
static_assert(bw_fused_bundle_size(9) * rows_per(9) <= MAX_ROWS, "carry rows");
static_assert(bw_fused_bundle_size(8) * rows_per(8) <= MAX_ROWS, "carry rows");
static_assert(bw_fused_bundle_size(7) * rows_per(7) <= MAX_ROWS, "carry rows");
static_assert(bw_fused_bundle_size(6) * rows_per(6) <= MAX_ROWS, "carry rows");
static_assert(BW_FUSED_MAX_BUNDLE * rows_per(5) <= MAX_ROWS, "carry rows");
static_assert(bw_fused_bundle_size(11) * (1 << 11) <= CPT * NT, "Garner coefficients");
static_assert(bw_fused_bundle_size(10) * (1 << 10) <= CPT * NT, "Garner coefficients");
static_assert(bw_fused_bundle_size(9) * (1 << 9) <= CPT * NT, "Garner coefficients");
static_assert(bw_fused_bundle_size(8) * (1 << 8) <= CPT * NT, "Garner coefficients");
static_assert(BW_FUSED_MAX_BUNDLE * (1 << 7) <= CPT * NT, "Garner coefficients");
static_assert((5 + 10) << BW_FUSED_MAX_LOG <= BW_FUSED_LDS_WORDS, "LDS of the longest transform");

// LDS position of element p of a transform or twiddle table: the column
// within each 32-word row is XORed with the row index, so the power-of-two
// strides of the stages (and of the twiddle look-ups) spread over the banks;
// a bijection within every row.
__device__ __forceinline__ int sw(int p) { return p ^ ((p >> 5) & 31); }

// Stages of a 2^L transform: ceil(L / 4) of radix up to 16, as even as
// possible (the larger first): [4, 4] for L = 8, [4, 3, 3] for L = 10,
// [3, 2] for L = 5.
// (Radix 32 holds 64 FP64 registers per value set and spilled.)
__device__ __forceinline__ int stage_count(int L) { return (L + 3) / 4; }
__device__ __forceinline__ int stage_rlog(int L, int s) {
    const int ns = stage_count(L);
    return L / ns + (s < L % ns ? 1 : 0);
}
// log2 span of stage s: L minus the radices before it
__device__ __forceinline__ int stage_logm(int L, int s) {
    const int ns = stage_count(L);
    return L - (s * (L / ns) + (s < L % ns ? s : L % ns));
}

// 64-bit digit e of an operand (zero past its end): one 64-bit limb, or a
// pair of 32-bit limbs
__device__ __forceinline__ uint64_t digit(const unsigned char* base, int64_t nl, int bits, int e) {
    if (bits == 64) return e < nl ? reinterpret_cast<const uint64_t*>(base)[e] : 0ull;
    const uint32_t* p = reinterpret_cast<const uint32_t*>(base);
    const int64_t i = 2 * (int64_t)e;
    const uint64_t lo = i < nl ? p[i] : 0u;
    const uint64_t hi = i + 1 < nl ? p[i + 1] : 0u;
    return lo | (hi << 32);
}

struct Bundle {
    int K;          // products
    int L;          // log2 transform length
    int q0;         // first prime of the current chunk
};

// Forward stage over span 2^logm: group (block, j) of a transform holds the
// elements base + u * 2^(logm - RLOG), u < R; a radix-R DFT over u, then
// frequency u times omega_m^(j u), stored back at position u.  A bundle has
// K * 2 * NB transforms: product k, then A of the chunk's primes (their
// residue planes), then B (the chunk buffers).
template <int NB, int RLOG, bool FIRST, bool LAST>
__device__ void fwd_stage(const FusedParams& prm, const F64PrimeDev* spc, const BwProductJob* jobs, int32_t* lds,
                          const Bundle& bd, int logm) {
    constexpr int R = 1 << RLOG;
    const int L = bd.L;
    const int n = 1 << L;
    const int logst = logm - RLOG;
    const int lper = L - RLOG;                 // log2 groups per transform
    const int total = (bd.K * 2 * NB) << lper;
    const int tsh = L - logm;                  // omega_m^e = omega_n^(e n/m)
    for (int g = threadIdx.x; g < total; g += NT) {
        const int ta = g >> lper;
        const int k = ta / (2 * NB);
        const int t = ta - k * (2 * NB);
        const int gi = g & ((1 << lper) - 1);
        const int j = gi & ((1 << logst) - 1);
        const int base = ((gi >> logst) << logm) + j;
        const bool opb = t >= NB;
        const int qi = opb ? t - NB : t;
        const int q = bd.q0 + qi;
        int32_t* buf = lds + NB * n + k * (5 + NB) * n + (opb ? 5 + qi : q) * n;
        const F64PrimeDev& pc = spc[q];
        double x[R];
        if constexpr (FIRST) {
            const BwProductJob& J = jobs[k];
            const unsigned char* src = opb ? prm.B + J.b_off : prm.A + J.a_off;
            const int64_t nl = opb ? J.lb : J.la;
#pragma unroll
            for (int u = 0; u < R; u++) x[u] = AR::from_raw64(digit(src, nl, J.limb_bits, base + (u << logst)), pc);
        } else {
#pragma unroll
            for (int u = 0; u < R; u++) x[u] = (double)buf[sw(base + (u << logst))];
        }
        dft32_dif<AR, RLOG>(x, pc);          // register i holds frequency brev(i)
        const int32_t* tw = lds + qi * n;
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int u = brev(i, RLOG);
            if (i == 0 || LAST) {
                // omega^0, and the last stage (j = 0): no twiddle, a reduction
                buf[sw(base + (u << logst))] = AR::to_lds(x[i], pc.P);
            } else {
                // a modular product leaves the value in (-p, p): no reduction
                const double w = (double)tw[sw((j * u) << tsh)];
                buf[sw(base + (u << logst))] = AR::to_lds_r(f64_mulmod(x[i], w, w * pc.P.pinv, pc.P));
            }
        }
    }
}

// Inverse of fwd_stage on the residue planes of the chunk's primes:
// frequency u times omega_m^-(j u), an inverse radix-R DFT (R times the
// inverse), element t stored at position t.
template <int NB, int RLOG, bool LAST>
__device__ void inv_stage(const F64PrimeDev* spc, int32_t* lds, const Bundle& bd, int logm) {
    constexpr int R = 1 << RLOG;
    const int L = bd.L;
    const int n = 1 << L;
    const int logst = logm - RLOG;
    const int lper = L - RLOG;
    const int total = (bd.K * NB) << lper;
    const int tsh = L - logm;
    for (int g = threadIdx.x; g < total; g += NT) {
        const int ta = g >> lper;
        const int k = ta / NB;
        const int qi = ta - k * NB;
        const int gi = g & ((1 << lper) - 1);
        const int j = gi & ((1 << logst) - 1);
        const int base = ((gi >> logst) << logm) + j;
        const int q = bd.q0 + qi;
        int32_t* buf = lds + NB * n + k * (5 + NB) * n + q * n;
        const F64PrimeDev& pc = spc[q];
        const int32_t* tw = lds + qi * n;
        double x[R];
#pragma unroll
        for (int i = 0; i < R; i++) {
            const int u = brev(i, RLOG);
            const double v = (double)buf[sw(base + (u << logst))];
            if (i == 0 || LAST) {
                x[i] = v;
            } else {
                const double w = (double)tw[sw((n - ((j * u) << tsh)) & (n - 1))];     // omega_m^-(j u)
                x[i] = f64_mulmod(v, w, w * pc.P.pinv, pc.P);
            }
        }
        dft32_dit_inv<AR, RLOG>(x, pc);
#pragma unroll
        for (int u = 0; u < R; u++) buf[sw(base + (u << logst))] = AR::to_lds(x[u], pc.P);
    }
}

template <int NB>
__device__ void fwd_dispatch(const FusedParams& prm, const F64PrimeDev* spc, const BwProductJob* jobs, int32_t* lds,
                             const Bundle& bd, int s) {
    const int r = stage_rlog(bd.L, s);
    const int logm = stage_logm(bd.L, s);
    const bool first = s == 0;
    const bool last = logm == r;        // span R: j = 0, no twiddles
#define BW_FWD(RL, F, LA) fwd_stage<NB, RL, F, LA>(prm, spc, jobs, lds, bd, logm)
#define BW_FWD_R(F, LA)                                   \
    switch (r) {                                          \
    case 1: BW_FWD(1, F, LA); break;                      \
    case 2: BW_FWD(2, F, LA); break;                      \
    case 3: BW_FWD(3, F, LA); break;                      \
    default: BW_FWD(4, F, LA); break;                     \
    }
    // radices 2 .. 16 in every position (L = 5 runs [8, 4])
    if (first && last) {
        BW_FWD_R(true, true)
    } else if (first) {
        BW_FWD_R(true, false)
    } else if (last) {
        BW_FWD_R(false, true)
    } else {
        BW_FWD_R(false, false)
    }
#undef BW_FWD_R
#undef BW_FWD
}

template <int NB>
__device__ void inv_dispatch(const F64PrimeDev* spc, int32_t* lds, const Bundle& bd, int s) {
    const int r = stage_rlog(bd.L, s);
    const int logm = stage_logm(bd.L, s);
#define BW_INV(RL, LA) inv_stage<NB, RL, LA>(spc, lds, bd, logm)
    const bool last = logm == r;
    switch (r) {
    case 1: if (last) BW_INV(1, true); else BW_INV(1, false); break;
    case 2: if (last) BW_INV(2, true); else BW_INV(2, false); break;
    case 3: if (last) BW_INV(3, true); else BW_INV(3, false); break;
    default: if (last) BW_INV(4, true); else BW_INV(4, false); break;
    }
#undef BW_INV
}

// Transforms, pointwise products and inverse transforms of the bundle's
// primes [q0, q0 + NB): the centered residues of every product's
// coefficients under those primes end in their residue planes.
template <int NB>
__device__ void chunk_transforms(const FusedParams& prm, const F64PrimeDev* spc, const BwProductJob* jobs,
                                 int32_t* lds, const Bundle& bd) {
    const int L = bd.L, n = 1 << L;
    const int ns = stage_count(L);
    const int tid = threadIdx.x;
    for (int idx = tid; idx < NB << L; idx += NT) {
        const int qi = idx >> L, e = idx & (n - 1);
        lds[(qi << L) + sw(e)] =
            (int32_t)prm.tw_stage[(size_t)(bd.q0 + qi) * TW_N + ((size_t)e << (BW_STAGE_LOG - L))].x;
    }
    __syncthreads();
    for (int s = 0; s < ns; s++) {
        fwd_dispatch<NB>(prm, spc, jobs, lds, bd, s);
        __syncthreads();
    }
    // pointwise product with n^-1 = p - (p - 1)/n (exact in FP64); A and
    // B are swizzled alike, so positions pair up directly
    for (int idx = tid; idx < (bd.K * NB) << L; idx += NT) {
        const int ta = idx >> L, e = idx & (n - 1);
        const int k = ta / NB, qi = ta - k * NB;
        const F64Prime P = spc[bd.q0 + qi].P;
        const double ni = P.p - P.cm * (1.0 / (double)n);
        const double nip[2] = {ni, ni * P.pinv};
        int32_t* a = lds + NB * n + k * (5 + NB) * n + (bd.q0 + qi) * n;
        const int32_t* b = lds + NB * n + k * (5 + NB) * n + (5 + qi) * n;
        a[e] = (int32_t)AR::mul_scaled((double)a[e], (double)b[e], nip, P);
    }
    __syncthreads();
    for (int s = ns - 1; s >= 0; s--) {
        inv_dispatch<NB>(spc, lds, bd, s);
        __syncthreads();
    }
}

// All five primes of a bundle, NB at a time.
template <int NB>
__device__ void bundle_transforms(const FusedParams& prm, const F64PrimeDev* spc, const BwProductJob* jobs,
                                  int32_t* lds, Bundle& bd) {
    for (bd.q0 = 0; bd.q0 < 5; bd.q0 += NB) chunk_transforms<NB>(prm, spc, jobs, lds, bd);
}

// Garner and carry of a bundle whose five residue planes per product are in
// LDS (product k's at lds + PB n + k (5 + PB) n, position sw(c)): Garner's
// three 64-bit words of every coefficient, then the carry by rows of 64
// limbs (product k owns rows [k rp, (k + 1) rp)), result limbs to HBM.
// Overwrites the planes.  s_row, s_cin: MAX_ROWS words of LDS each.
template <int PB>
__device__ void bundle_garner_carry(const FusedParams& prm, const BwProductJob* s_jobs, int32_t* lds,
                                    const Bundle& bd, uint32_t* s_row, uint32_t* s_cin) {
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int L = bd.L, n = 1 << L;
    // Garner: coefficient words into registers, then over the planes
    const int ncoef = bd.K << L;
    uint64_t cw[CPT][3];
#pragma unroll
    for (int r = 0; r < CPT; r++) {
        const int i = tid + r * NT;
        if (i < ncoef) {
            const int k = i >> L, c = i & (n - 1);
            const int32_t* planes = lds + PB * n + k * (5 + PB) * n;
            uint32_t rr[5];
#pragma unroll
            for (int q = 0; q < 5; q++) rr[q] = (uint32_t)planes[q * n + sw(c)];
            garner_d64(rr, prm.gd, cw[r]);
        }
    }
    __syncthreads();
    // product k's words at W + 3 k n: word j of coefficient c at j n + c
    uint64_t* W = reinterpret_cast<uint64_t*>(lds);
#pragma unroll
    for (int r = 0; r < CPT; r++) {
        const int i = tid + r * NT;
        if (i < ncoef) {
            const int k = i >> L, c = i & (n - 1);
            uint64_t* Wk = W + (size_t)3 * k * n;
            Wk[c] = cw[r][0];
            Wk[n + c] = cw[r][1];
            Wk[2 * n + c] = cw[r][2];
        }
    }
    __syncthreads();
    // carry: limb i of product k = t_i + carry_i
    const int rp = rows_per(L);
    const int nrows = bd.K * rp;
    uint64_t tv[CARRY_ROWS], Gm[CARRY_ROWS], Pm[CARRY_ROWS];
#pragma unroll
    for (int r = 0; r < CARRY_ROWS; r++) {
        const int row = wave + r * NWAVES;
        const int k = row / rp, lr = row - k * rp;
        const bool live = row < nrows;
        const uint64_t* Wk = W + (size_t)3 * (live ? k : 0) * n;
        auto word = [&](int i, int c) -> uint64_t { return (live && i >= 0 && i < n) ? Wk[c * n + i] : 0ull; };
        const int i = lr * 64 + lane;
        const uint64_t a0 = word(i, 0), a1 = word(i - 1, 1), a2 = word(i - 2, 2);
        const uint64_t l1 = a0 + a1, lo = l1 + a2;
        const uint64_t b0 = word(i - 1, 0), b1 = word(i - 2, 1), b2 = word(i - 3, 2);
        const uint64_t m1 = b0 + b1, m2 = m1 + b2;
        const uint64_t hprev = (uint64_t)(m1 < b0) + (uint64_t)(m2 < m1);   // hi(s_{i-1}) <= 2
        const uint64_t t = lo + hprev;
        const bool g = t < lo;
        const bool p = !g && t == ~0ull;
        tv[r] = t;
        Gm[r] = __ballot(g);
        Pm[r] = __ballot(p);
        if (lane == 0 && live) {
            const uint64_t X = Gm[r] | Pm[r];
            const uint64_t sum = X + Gm[r];
            s_row[row] = (sum < X ? 1u : 0u) | (Pm[r] == ~0ull ? 2u : 0u);
        }
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t c = 0;
        for (int row = 0; row < nrows; row++) {
            if (row % rp == 0) c = 0;           // a product's first row
            s_cin[row] = c;
            const uint32_t a = s_row[row];
            c = (a & 1u) | ((a >> 1) & c);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < CARRY_ROWS; r++) {
        const int row = wave + r * NWAVES;
        if (row < nrows) {
            const int k = row / rp, lr = row - k * rp;
            const BwProductJob& J = s_jobs[k];
            const int64_t nl = J.la + J.lb;
            const int64_t n_out = J.limb_bits == 32 ? (nl + 1) >> 1 : nl;
            const bool half_last = J.limb_bits == 32 && (nl & 1);
            const uint64_t X = Gm[r] | Pm[r];
            const uint64_t s2 = X + Gm[r] + (uint64_t)s_cin[row];
            const uint64_t carries = s2 ^ Pm[r];
            const int i = lr * 64 + lane;
            const uint64_t limb = tv[r] + ((carries >> lane) & 1ull);
            uint64_t* out = reinterpret_cast<uint64_t*>(prm.C + J.c_off);
            if (i < n_out) {
                if (half_last && i == n_out - 1) reinterpret_cast<uint32_t*>(out + i)[0] = (uint32_t)limb;
                else out[i] = limb;
            }
        }
    }
}

// The five primes' constants into LDS (every thread takes part; a barrier
// must follow before they are read).
__device__ void load_consts(const FusedParams& prm, F64PrimeDev* spc) {
    constexpr int WORDS = (int)(sizeof(F64PrimeDev) / 8);
    static_assert(sizeof(F64PrimeDev) % 8 == 0, "8-byte words");
    const uint64_t* src = reinterpret_cast<const uint64_t*>(prm.consts);
    uint64_t* dst = reinterpret_cast<uint64_t*>(spc);
    for (int i = threadIdx.x; i < 5 * WORDS; i += NT) dst[i] = src[i];
}

}  // namespace

__global__ __launch_bounds__(NT) void product_fused_kernel(FusedParams prm) {
    extern __shared__ __align__(16) int32_t lds[];
    __shared__ int s_bundle;
    __shared__ uint32_t s_row[MAX_ROWS];
    __shared__ uint32_t s_cin[MAX_ROWS];
    __shared__ BwProductJob s_jobs[BW_FUSED_MAX_BUNDLE];
    // the primes' constants in LDS: lanes of one wave may work on different
    // primes (transforms shorter than 64 groups), which global loads would
    // serve one lane at a time
    __shared__ F64PrimeDev spc[5];
    const int tid = threadIdx.x;
    load_consts(prm, spc);
    for (;;) {
        if (tid == 0) {
            // every workgroup takes exactly one id past the last bundle; the
            // last id of all resets the counter for the next launch
            const int t = atomicAdd(prm.counter, 1);
            if (t == prm.nbundles + (int)gridDim.x - 1) atomicExch(prm.counter, 0);
            s_bundle = t;
        }
        __syncthreads();
        const int bi = s_bundle;
        if (bi >= prm.nbundles) break;
        const BwProductBundle B = prm.bundles[bi];
        if (tid < B.count && tid < BW_FUSED_MAX_BUNDLE) s_jobs[tid] = prm.jobs[B.first + tid];
        __syncthreads();
        Bundle bd;
        bd.K = B.count;
        bd.L = s_jobs[0].logN;
        bd.q0 = 0;
        const int L = bd.L, n = 1 << L;
        // a bundle the launch was not sized for is skipped (its LDS would not fit)
        bool ok = L >= 1 && L <= BW_FUSED_MAX_LOG && bd.K >= 1 && bd.K <= bw_fused_bundle_size(L);
        for (int k = 0; ok && k < bd.K; k++) {
            const int64_t nl = s_jobs[k].la + s_jobs[k].lb;
            const int64_t n_out = s_jobs[k].limb_bits == 32 ? (nl + 1) >> 1 : nl;
            ok = s_jobs[k].logN == L && n_out <= (int64_t)n + 1;
        }
        if (!ok) {
            __syncthreads();
            continue;
        }
        bundle_transforms<NB>(prm, spc, s_jobs, lds, bd);
        bundle_garner_carry<NB>(prm, s_jobs, lds, bd, s_row, s_cin);
        __syncthreads();
    }
}

hipError_t bw_launch_product_fused(const FusedParams& p, int grid, hipStream_t stream) {
    if (grid < 1 || p.nbundles < 1) return hipErrorInvalidValue;
    const size_t lds = (size_t)BW_FUSED_LDS_WORDS * 4;
    static std::once_flag once;
    hipError_t e = hipSuccess;
    std::call_once(once, [&] {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(product_fused_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(product_fused_kernel, dim3((unsigned)grid), dim3(NT), lds, stream, p);
    return hipGetLastError();
}
