// carry_kernels.hip -- Garner CRT + single-pass carry scan (see carry_kernels.h).
#include "carry_kernels.h"
#include "garner_dev.h"


// ---------------------------------------------------------------------------
// Garner's mixed-radix reconstruction x = k0 + p0 k1 + p0 p1 k2 with
// k1 = (r1 - k0) p0^-1 mod p1 and k2 = (r2 - (k0 + p0 k1)) (p0 p1)^-1 mod p2:
// the recurrence of the reference kernel (src/crt_gpu.cu:157-203), with the
// x mod p_j bookkeeping done by the same M_mod_table products.

// u32 path: residues < 2^31, FP64 modular products; x < 2^96 -> 3 words.
__device__ __forceinline__ void garner_u32(uint32_t r0, uint32_t r1, uint32_t r2, const GarnerDev& g,
                                           const F64Prime* fp, uint64_t& lo, uint32_t& hi32) {
    const double d1 = (double)r1 - (double)r0;
    const uint32_t k1 = f64_canon(f64_mulmod(d1, g.inv1[0], g.inv1[1], fp[1]), fp[1]);
    const double xm = (double)r0 + f64_mulmod((double)k1, g.p0mod2[0], g.p0mod2[1], fp[2]);
    const uint32_t k2 = f64_canon(f64_mulmod((double)r2 - xm, g.inv2[0], g.inv2[1], fp[2]), fp[2]);
    const uint64_t v = (uint64_t)r0 + g.p0 * (uint64_t)k1;     // < 2^63
    const uint64_t plo = g.p01_lo * (uint64_t)k2;
    const uint64_t phi = __umul64hi(g.p01_lo, (uint64_t)k2);
    lo = v + plo;
    hi32 = (uint32_t)(phi + (lo < v ? 1ULL : 0ULL));
}

// u64 path: residues < p_j < 2^62, Shoup products; x < 2^174 in general but
// < 2^160 for every supported size -> 3 words (lo, mid, hi).
__device__ __forceinline__ uint64_t cond_sub(uint64_t x, uint64_t p) { return x >= p ? x - p : x; }

__device__ __forceinline__ void garner_u64(uint64_t r0, uint64_t r1, uint64_t r2, const GarnerDev& g,
                                           uint64_t& w0, uint64_t& w1, uint64_t& w2) {
    const uint64_t p0 = g.p[0], p1 = g.p[1], p2 = g.p[2];
    const uint64_t k1 = u64_mulmod_shoup(u64_sub(r1, cond_sub(r0, p1), p1), g.inv1_s[0], g.inv1_s[1], p1);
    const uint64_t xm = u64_add(cond_sub(r0, p2),
                                u64_mulmod_shoup(cond_sub(k1, p2), g.p0mod2_s[0], g.p0mod2_s[1], p2), p2);
    const uint64_t k2 = u64_mulmod_shoup(u64_sub(r2, xm, p2), g.inv2_s[0], g.inv2_s[1], p2);
    typedef unsigned __int128 u128;
    u128 a = (u128)p0 * k1 + r0;                     // < 2^117
    u128 b = (u128)g.p01_lo * k2;                    // low product
    u128 c = (u128)g.p01_hi * k2;                    // shifted by 64
    u128 s0 = (u128)(uint64_t)a + (uint64_t)b;
    w0 = (uint64_t)s0;
    u128 s1 = (s0 >> 64) + (a >> 64) + (b >> 64) + (uint64_t)c;
    w1 = (uint64_t)s1;
    w2 = (uint64_t)((s1 >> 64) + (c >> 64));
}

// Two-prime Garner of the hybrid set (58-bit primes, p0 < p1, 32-bit digits):
// x = r0 + p0 k1 with k1 = (r1 - r0) p0^-1 mod p1; x < p0 p1 < 2^117.
__device__ __forceinline__ void garner_u64x2(uint64_t r0, uint64_t r1, const GarnerDev& g, uint64_t& lo,
                                             uint64_t& hi) {
    const uint64_t p1 = g.p[1];
    const uint64_t k1 = u64_mulmod_shoup(u64_sub(r1, r0, p1), g.inv1_s[0], g.inv1_s[1], p1);
    typedef unsigned __int128 u128;
    const u128 x = (u128)g.p0 * k1 + r0;
    lo = (uint64_t)x;
    hi = (uint64_t)(x >> 64);
}

// ---------------------------------------------------------------------------
// Coefficient sources: words w_j(c_n) in limb width, zero outside [0, N).
// The view names the coefficient arrays of the product a tile belongs to.
struct CarryView {
    const void* a0;
    const void* a1;
    const void* a2;
    int64_t plane;
    int64_t N;
};
// WAVES: the least waves per SIMD the kernel is compiled for (a register
// cap; each source at the most that compiles without spills).
//
// fetch() issues the loads of coefficient n into a Fetch record and words()
// turns the record into limb words, so the kernel can put the loads of all
// its rows in flight before the first conversion.  LateFetch: the record is
// the index, and words() loads and converts.
template <class S>
struct LateFetch {
    struct Fetch { int64_t n; };
    __device__ static void fetch(const CarryView&, int64_t n, Fetch& f) { f.n = n; }
    template <class Wd>
    __device__ static void words(const CarryView& p, const CarryParams& c, int64_t, const Fetch& f, Wd* w) {
        S::load(p, c, f.n, w);
    }
};
struct SrcGarnerU32 : LateFetch<SrcGarnerU32> {
    using Word = uint32_t;
    static constexpr int NW = 3;
    static constexpr int WAVES = 5;
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        const bool in = n >= 0 && n < p.N;      // no branch: loads of later rows can issue early
        const int64_t m = in ? n : 0;
        const uint32_t* r = reinterpret_cast<const uint32_t*>(p.a0);
        uint64_t lo; uint32_t hi;
        garner_u32(r[m], r[p.plane + m], r[2 * p.plane + m], c.g, c.fp, lo, hi);
        w[0] = in ? (uint32_t)lo : 0u; w[1] = in ? (uint32_t)(lo >> 32) : 0u; w[2] = in ? hi : 0u;
    }
};
struct SrcGarnerU64 : LateFetch<SrcGarnerU64> {
    using Word = uint64_t;
    static constexpr int NW = 3;
    static constexpr int WAVES = 4;
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        const bool in = n >= 0 && n < p.N;
        const int64_t m = in ? n : 0;
        const uint64_t* r = reinterpret_cast<const uint64_t*>(p.a0);
        garner_u64(r[m], r[p.plane + m], r[2 * p.plane + m], c.g, w[0], w[1], w[2]);
        if (!in) w[0] = w[1] = w[2] = 0;
    }
};
struct SrcGarnerU64x2 : LateFetch<SrcGarnerU64x2> {
    using Word = uint32_t;
    static constexpr int NW = 4;
    static constexpr int WAVES = 5;
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        const bool in = n >= 0 && n < p.N;
        const int64_t m = in ? n : 0;
        const uint64_t* r = reinterpret_cast<const uint64_t*>(p.a0);
        uint64_t lo, hi;
        garner_u64x2(r[m], r[p.plane + m], c.g, lo, hi);
        w[0] = in ? (uint32_t)lo : 0u; w[1] = in ? (uint32_t)(lo >> 32) : 0u;
        w[2] = in ? (uint32_t)hi : 0u; w[3] = in ? (uint32_t)(hi >> 32) : 0u;
    }
};
struct SrcGarnerD64 {
    using Word = uint64_t;
    static constexpr int NW = 3;
    static constexpr int WAVES = 4;   // room for the residues of all rows in flight
    // the five residue words of one coefficient (index clamped into range:
    // the loads carry no branch)
    struct Fetch { uint32_t r[5]; };
    __device__ static void fetch(const CarryView& p, int64_t n, Fetch& f) {
        typedef const __attribute__((address_space(1))) uint32_t GU32;
        const bool in = n >= 0 && n < p.N;
        const GU32* a = (const GU32*)p.a0 + (in ? n : 0);
#pragma unroll
        for (int j = 0; j < 5; j++) f.r[j] = a[j * p.plane];
    }
    __device__ static void words(const CarryView& p, const CarryParams& c, int64_t n, const Fetch& f, Word* w) {
        garner_d64(f.r, c.gd, w);
        if (!(n >= 0 && n < p.N)) w[0] = w[1] = w[2] = 0;
    }
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        Fetch f;
        fetch(p, n, f);
        words(p, c, n, f, w);
    }
};
struct SrcHiLo128 : LateFetch<SrcHiLo128> {     // a0 = hi, a1 = lo
    using Word = uint32_t;
    static constexpr int NW = 4;
    static constexpr int WAVES = 5;
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        if (n < 0 || n >= p.N) { w[0] = w[1] = w[2] = w[3] = 0; return; }
        const uint64_t hi = reinterpret_cast<const uint64_t*>(p.a0)[n];
        const uint64_t lo = reinterpret_cast<const uint64_t*>(p.a1)[n];
        w[0] = (uint32_t)lo; w[1] = (uint32_t)(lo >> 32); w[2] = (uint32_t)hi; w[3] = (uint32_t)(hi >> 32);
    }
};
struct SrcHiLo128L64 : LateFetch<SrcHiLo128L64> {  // a0 = hi, a1 = lo, 64-bit limbs
    using Word = uint64_t;
    static constexpr int NW = 2;
    static constexpr int WAVES = 5;
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        if (n < 0 || n >= p.N) { w[0] = w[1] = 0; return; }
        w[1] = reinterpret_cast<const uint64_t*>(p.a0)[n];
        w[0] = reinterpret_cast<const uint64_t*>(p.a1)[n];
    }
};
struct SrcU160 : LateFetch<SrcU160> {        // a0 = lo, a1 = mid, a2 = hi32
    using Word = uint64_t;
    static constexpr int NW = 3;
    static constexpr int WAVES = 5;
    __device__ static void load(const CarryView& p, const CarryParams& c, int64_t n, Word* w) {
        if (n < 0 || n >= p.N) { w[0] = w[1] = w[2] = 0; return; }
        w[0] = reinterpret_cast<const uint64_t*>(p.a0)[n];
        w[1] = reinterpret_cast<const uint64_t*>(p.a1)[n];
        w[2] = reinterpret_cast<const uint32_t*>(p.a2)[n];
    }
};

// Lane l receives x from lane l-j of this row; lanes < j take prev from lane l+64-j.
template <class Wd>
__device__ __forceinline__ Wd shift_up(Wd cur, Wd prev, int j, int lane) {
    const Wd a = __shfl(cur, (lane - j) & 63);
    const Wd b = __shfl(prev, (lane + 64 - j) & 63);
    return lane >= j ? a : b;
}

// Flag word: [63:32] epoch, [3] propagate, [2] generate/carry, [1:0] status.
#define FLAG_AGG 1ULL
#define FLAG_INC 2ULL

// K rows of 64 limbs per wave, NWV waves: a tile of 64 K NWV limbs.
template <class SRC, int K = BW_CARRY_ROWS, int NWV = BW_CARRY_WAVES>
__global__ __launch_bounds__(64 * NWV) __attribute__((amdgpu_waves_per_eu(SRC::WAVES))) void garner_carry_kernel(CarryParams prm) {
    using Wd = typename SRC::Word;
    constexpr int NW = SRC::NW;
    constexpr int64_t TILE = (int64_t)64 * K * NWV;
    constexpr Wd ONES = (Wd)~(Wd)0;
    __shared__ int64_t s_tile;
    __shared__ uint32_t s_wagg[NWV];
    __shared__ uint32_t s_wcin[NWV];

    if (threadIdx.x == 0) {
        if (prm.ntiles == 1) {
            s_tile = 0;                        // one tile: no id to take, nothing to chain
        } else {
            const int t = atomicAdd(prm.counter, 1);
            if ((int64_t)t == prm.ntiles - 1) atomicExch(prm.counter, 0);   // every id is taken: reset for the next launch
            s_tile = t;
        }
    }
    __syncthreads();
    const int64_t tile = s_tile;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    // the product this tile belongs to, and the tile's index inside it
    CarryView view = {prm.a0, prm.a1, prm.a2, prm.plane, prm.N};
    int64_t n_out = prm.n_out;
    Wd* out = reinterpret_cast<Wd*>(prm.out);
    int64_t ltile = tile;
    bool half_last = prm.half_last != 0;
    if (prm.jobs) {
        const BwCarryJob J = prm.jobs[prm.tile_job[tile]];
        view.a0 = J.res;
        view.plane = J.plane;
        view.N = J.N;
        n_out = J.n_out;
        out += J.out_off;
        ltile = tile - J.tile0;
        half_last = J.half_last != 0;
    }
    const int64_t wbase = ltile * TILE + (int64_t)wave * 64 * K;

    Wd tval[K];
    bool tg[K], tp[K];
    // Limb n sums word j of coefficient n - j (j < NW), so row 0 of a wave
    // needs the last NW - 1 coefficients of the wave before it (words j of
    // its lanes 64 - j .. 63) and the high part of that row's last sum.
    // Every wave hands those to its successor through LDS (slot wave + 1);
    // wave 0 computes them for the tile before (slot 0).  Row 0 is therefore
    // finished last, after a barrier, from its coefficient words kept in cw0.
    __shared__ Wd s_tw[NWV + 1][NW][NW];   // [slot][j][lane - (64 - NW)]: word j of the last lanes
    __shared__ Wd s_thi[NWV + 1];          // [slot]: high part of the last row's sum at lane 63
    auto row_sums = [&](const Wd* cw, const Wd* pw, Wd& lo, Wd& hi) {
        lo = cw[0];
        hi = 0;
#pragma unroll
        for (int j = 1; j < NW; j++) {
            const Wd v = shift_up(cw[j], pw[j], j, lane);
            const Wd s = lo + v;
            hi += (s < lo) ? 1 : 0;
            lo = s;
        }
    };
    auto publish_tail = [&](int slot, const Wd* cw, Wd hi) {
        if (lane >= 64 - NW) {
#pragma unroll
            for (int j = 0; j < NW; j++) s_tw[slot][j][lane - (64 - NW)] = cw[j];
        }
        if (lane == 63) s_thi[slot] = hi;
    };
    Wd zero[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) zero[j] = 0;
    // the loads of every row (and, in wave 0, of the previous tile's last
    // row) are issued before any coefficient is converted
    typename SRC::Fetch fprev;
    if (wave == 0) SRC::fetch(view, wbase - 64 + lane, fprev);
    typename SRC::Fetch fr[K];
#pragma unroll
    for (int i = 0; i < K; i++) SRC::fetch(view, wbase + i * 64 + lane, fr[i]);
    if (wave == 0) {
        // the previous tile's last 64 limbs (zero before the product's start)
        Wd cw[NW];
        SRC::words(view, prm, wbase - 64 + lane, fprev, cw);
        Wd lo, hi;
        row_sums(cw, zero, lo, hi);
        publish_tail(0, cw, hi);
    }
    Wd cw0[NW];
    SRC::words(view, prm, wbase + lane, fr[0], cw0);
    Wd pw[NW];
    Wd phi;
    {
        // row 0's high part at lane 63 does not depend on the row before it
        Wd lo;
        row_sums(cw0, zero, lo, phi);
#pragma unroll
        for (int j = 0; j < NW; j++) pw[j] = cw0[j];
    }
#pragma unroll
    for (int i = 1; i < K; i++) {
        Wd cw[NW];
        SRC::words(view, prm, wbase + i * 64 + lane, fr[i], cw);
        Wd lo, hi;
        row_sums(cw, pw, lo, hi);
        const Wd h1 = shift_up(hi, phi, 1, lane);
        const Wd t = lo + h1;
        tg[i] = t < lo;                        // carry out of the limb
        tp[i] = !tg[i] && (t == ONES);
        tval[i] = t;
#pragma unroll
        for (int j = 0; j < NW; j++) pw[j] = cw[j];
        phi = hi;
    }
    if (wave < NWV - 1) publish_tail(wave + 1, pw, phi);
    __syncthreads();
    {
        // row 0 with the tail of the row before it
        Wd prev[NW];
        const int ti = lane >= 64 - NW ? lane - (64 - NW) : 0;     // in range for every lane
#pragma unroll
        for (int j = 0; j < NW; j++) prev[j] = s_tw[wave][j][ti];
        Wd lo, hi;
        row_sums(cw0, prev, lo, hi);           // lanes < j take word j of lane 64 - j + lane of prev
        const Wd hs = shift_up(hi, hi, 1, lane);
        const Wd h1 = lane == 0 ? s_thi[wave] : hs;
        const Wd t = lo + h1;
        tg[0] = t < lo;
        tp[0] = !tg[0] && (t == ONES);
        tval[0] = t;
    }

    // Wave summary over its K rows (cin = 0).
    uint32_t wgen = 0, wprop = 1;
#pragma unroll
    for (int i = 0; i < K; i++) {
        const uint64_t G = __ballot(tg[i]);
        const uint64_t P = __ballot(tp[i]);
        const uint64_t X = G | P;
        const uint64_t sum = X + G;
        const uint32_t rgen = sum < X ? 1u : 0u;
        const uint32_t rprop = (P == ~0ULL) ? 1u : 0u;
        wgen = rgen | (rprop & wgen);
        wprop &= rprop;
    }
    if (lane == 0) s_wagg[wave] = wgen | (wprop << 1);
    __syncthreads();

    if (threadIdx.x == 0) {
        uint32_t gen = 0, prop = 1;
        for (int w = 0; w < NWV; w++) {
            const uint32_t a = s_wagg[w];
            gen = (a & 1) | ((a >> 1) & gen);
            prop &= (a >> 1);
        }
        const uint32_t epoch = prm.epoch;
        const unsigned long long ep = (unsigned long long)epoch << 32;
        uint32_t cin = 0;
        if (ltile == 0) {
            __hip_atomic_store(&prm.flags[tile], ep | FLAG_INC | ((unsigned long long)gen << 2),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&prm.flags[tile], ep | FLAG_AGG | ((unsigned long long)gen << 2) |
                                   ((unsigned long long)prop << 3),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t agen = 0, aprop = 1;
            int64_t pred = tile - 1;
            uint32_t spins = 0;
            for (;;) {
                const unsigned long long f =
                    __hip_atomic_load(&prm.flags[pred], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((uint32_t)(f >> 32) != epoch || (f & 3ULL) == 0) {
                    // Predecessors publish their summary without waiting, so this
                    // wait is short; the bound only turns a logic error into a
                    // reported failure instead of a hang.
                    if (++spins > (1u << 24)) {
                        __hip_atomic_store(prm.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        cin = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                const uint32_t fg = (uint32_t)(f >> 2) & 1u;
                if ((f & 3ULL) == FLAG_INC) {
                    cin = agen | (aprop & fg);
                    break;
                }
                const uint32_t fp = (uint32_t)(f >> 3) & 1u;
                agen = agen | (aprop & fg);
                aprop = aprop & fp;
                if (!aprop) { cin = agen; break; }   // the chain cannot reach further back
                pred--;
            }
            const uint32_t cout = gen | (prop & cin);
            __hip_atomic_store(&prm.flags[tile], ep | FLAG_INC | ((unsigned long long)cout << 2),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        uint32_t c = cin;
        for (int w = 0; w < NWV; w++) {
            s_wcin[w] = c;
            const uint32_t a = s_wagg[w];
            c = (a & 1) | ((a >> 1) & c);
        }
    }
    __syncthreads();

    uint64_t cin = s_wcin[wave];
#pragma unroll
    for (int i = 0; i < K; i++) {
        const uint64_t G = __ballot(tg[i]);
        const uint64_t P = __ballot(tp[i]);
        const uint64_t X = G | P;
        const uint64_t s1 = X + G;
        const uint64_t s2 = s1 + cin;
        const uint64_t cout = (s1 < X || s2 < s1) ? 1ULL : 0ULL;
        const uint64_t carries = s2 ^ P;           // carry into each lane
        const Wd limb = tval[i] + (Wd)((carries >> lane) & 1ULL);
        const int64_t n = wbase + i * 64 + lane;
        if (n < n_out) {
            if (half_last && n == n_out - 1) reinterpret_cast<uint32_t*>(out + n)[0] = (uint32_t)limb;
            else out[n] = limb;
        }
        cin = cout;
    }
}

int64_t bw_carry_tiles(int64_t n_out) { return (n_out + BW_CARRY_TILE - 1) / BW_CARRY_TILE; }

__global__ void carry_reset_kernel(unsigned long long* flags, int64_t ntiles, int* counter, int* err) {
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = i0; i < ntiles; i += (int64_t)gridDim.x * blockDim.x) flags[i] = 0ULL;
    if (i0 == 0) {
        *counter = 0;
        if (err) *err = 0;
    }
}

hipError_t bw_launch_carry_reset(unsigned long long* flags, int64_t ntiles, int* counter, hipStream_t stream,
                                 int* err) {
    int64_t blocks = (ntiles + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(carry_reset_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, flags, ntiles, counter, err);
    return hipGetLastError();
}

// A single product of at most 4096 output limbs is one tile whatever its
// shape; there the latency of the one workgroup decides, so it runs sixteen
// waves of two or four rows each (the Garner work of the tile spread over
// twice the waves of the default shape).
template <class SRC>
static void launch_carry(const CarryParams& p, hipStream_t stream) {
    if (!p.jobs && p.ntiles == 1 && p.n_out <= 64 * 2 * 16)
        hipLaunchKernelGGL((garner_carry_kernel<SRC, 2, 16>), dim3(1), dim3(1024), 0, stream, p);
    else if (!p.jobs && p.ntiles == 1 && p.n_out <= 64 * 4 * 16)
        hipLaunchKernelGGL((garner_carry_kernel<SRC, 4, 16>), dim3(1), dim3(1024), 0, stream, p);
    else
        hipLaunchKernelGGL(garner_carry_kernel<SRC>, dim3((unsigned)p.ntiles), dim3(64 * BW_CARRY_WAVES), 0, stream,
                           p);
}

hipError_t bw_launch_garner_carry(const CarryParams& p, hipStream_t stream) {
    dim3 grid((unsigned)p.ntiles), block(64 * BW_CARRY_WAVES);
    switch (p.src) {
    case BW_SRC_GARNER_U32: launch_carry<SrcGarnerU32>(p, stream); break;
    case BW_SRC_GARNER_U64: launch_carry<SrcGarnerU64>(p, stream); break;
    case BW_SRC_HILO128: hipLaunchKernelGGL(garner_carry_kernel<SrcHiLo128>, grid, block, 0, stream, p); break;
    case BW_SRC_HILO128_L64: hipLaunchKernelGGL(garner_carry_kernel<SrcHiLo128L64>, grid, block, 0, stream, p); break;
    case BW_SRC_U160: hipLaunchKernelGGL(garner_carry_kernel<SrcU160>, grid, block, 0, stream, p); break;
    case BW_SRC_GARNER_D64: launch_carry<SrcGarnerD64>(p, stream); break;
    case BW_SRC_GARNER_U64X2: hipLaunchKernelGGL(garner_carry_kernel<SrcGarnerU64x2>, grid, block, 0, stream, p); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Standalone Garner: coefficients out (the reference's crt_combine_gpu layouts).
__global__ void garner_u32_kernel(const uint32_t* r0, const uint32_t* r1, const uint32_t* r2, int64_t N,
                                  GarnerDev g, F64Prime f0, F64Prime f1, F64Prime f2, uint64_t* hi, uint64_t* lo) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const F64Prime fp[3] = {f0, f1, f2};
    uint64_t l; uint32_t h;
    garner_u32(r0[n], r1[n], r2[n], g, fp, l, h);
    lo[n] = l;
    hi[n] = h;
}

__global__ void garner_u64_kernel(const uint64_t* r0, const uint64_t* r1, const uint64_t* r2, int64_t N,
                                  GarnerDev g, uint64_t* lo, uint64_t* mid, uint32_t* hi) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    uint64_t w0, w1, w2;
    garner_u64(r0[n], r1[n], r2[n], g, w0, w1, w2);
    lo[n] = w0;
    mid[n] = w1;
    hi[n] = (uint32_t)w2;
}

__global__ void garner_u64x2_kernel(const uint64_t* r0, const uint64_t* r1, int64_t N, GarnerDev g, uint64_t* hi,
                                    uint64_t* lo) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    uint64_t l, h;
    garner_u64x2(r0[n], r1[n], g, l, h);
    lo[n] = l;
    hi[n] = h;
}

hipError_t bw_launch_garner_only(int arith_bits, const void* r0, const void* r1, const void* r2, int64_t N,
                                 const GarnerDev& g, const F64Prime* fp,
                                 uint64_t* o0, uint64_t* o1, uint32_t* o2, hipStream_t stream) {
    if (N <= 0) return hipSuccess;
    dim3 grid((unsigned)((N + 255) / 256)), block(256);
    if (arith_bits == 64 && g.k == 2)
        hipLaunchKernelGGL(garner_u64x2_kernel, grid, block, 0, stream, (const uint64_t*)r0, (const uint64_t*)r1, N, g,
                           o1, o0);
    else if (arith_bits == 64)
        hipLaunchKernelGGL(garner_u64_kernel, grid, block, 0, stream, (const uint64_t*)r0, (const uint64_t*)r1,
                           (const uint64_t*)r2, N, g, o0, o1, o2);
    else
        hipLaunchKernelGGL(garner_u32_kernel, grid, block, 0, stream, (const uint32_t*)r0, (const uint32_t*)r1,
                           (const uint32_t*)r2, N, g, fp[0], fp[1], fp[2], o1, o0);
    return hipGetLastError();
}
