// row_fused.hip -- the last forward pass, the pointwise product and the first
// inverse pass of a product on whole rows of T = 2^14 or 2^15 points, one row
// per workgroup, the row resident in LDS (FP64 arithmetic, 32-bit words).
//
// The row's T-point DFT runs as four decimation-in-frequency stages of radix
// R0 = T/4096, 16, 16, 16 (strides 4096, 256, 16, 1) through one padded LDS
// line of centered int32 residues; the stage twiddles w_L^{j k} (L = the
// stage's block length, j the offset in the block's first stride, k the
// output index) multiply each stage's outputs before they are written back.
// The last stage leaves each thread 32 spectrum values in registers: A's wait
// there (int32), B's multiply into them with N^-1, and the inverse runs the
// same stages backwards (decimation in time, conjugate twiddles applied after
// each inverse stage, before the write), so the spectrum order inside the row
// never leaves the workgroup.  Rows arrive from the previous forward pass
// (pass twiddles applied there) and leave for the next inverse pass.
//
// Against the register-staged passes this removes one pass over HBM in each
// direction (a 2^9 x 2^15 split of a 2^24-point transform runs three launches
// instead of five) and one pass-twiddle boundary per transform.
#include "ntt2_impl.h"

#include <mutex>

#define BW_ROW_THREADS 1024

template <class AR, int LOGT>
__global__ __launch_bounds__(BW_ROW_THREADS) void row_fused_kernel(PassParams prm) {
    using V = typename AR::V;
    using W = typename AR::W;
    using TW = typename AR::TW;
    constexpr int T = 1 << LOGT;
    constexpr int R0LOG = LOGT - 12;               // stage radices R0, 16, 16, 16
    constexpr int R0 = 1 << R0LOG;
    constexpr int NT = BW_ROW_THREADS;
    constexpr int G0 = 4096 / NT;                  // stage-0 groups per thread
    constexpr int G16 = T / 16 / NT;               // radix-16 groups per thread
    static_assert(LOGT >= 14 && LOGT <= 15 && G0 >= 1 && G16 >= 1, "row sizes 2^14, 2^15");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    int32_t* lds = reinterpret_cast<int32_t*>(smem_raw);
    TW* tw256 = reinterpret_cast<TW*>(smem_raw + row_fused_line_bytes(LOGT));   // w_256^e, e < 256
    auto pos = [](int i) { return i + (i >> 4); };                              // one pad slot per 16

    const int tid = threadIdx.x;
    const int64_t row = blockIdx.x;
    const int prime = prm.prime0 + blockIdx.y;
    const typename AR::PC32 pc = reinterpret_cast<const typename AR::PC32*>(prm.consts)[prime];
    const auto& P = AR::prime(pc);
    const TW* tws = reinterpret_cast<const TW*>(prm.tw_stage) + ((int64_t)prime << BW_STAGE_LOG);
    const TW* twlo = reinterpret_cast<const TW*>(prm.tw_lo) + ((int64_t)prime << BW_TW_LO_LOG);
    const TW* twhi = reinterpret_cast<const TW*>(prm.tw_hi) + ((int64_t)prime << (prm.table_log - BW_TW_LO_LOG));
    const typename AR::NI* ninv = reinterpret_cast<const typename AR::NI*>(prm.ninv) + 2 * prime;
    const int64_t base = (int64_t)prime * prm.plane + row * T;
    const int tshift = prm.table_log - LOGT;                  // w_T = w_{2^table_log}^(2^tshift)
    const int64_t tmask = ((int64_t)1 << prm.table_log) - 1;

    for (int e = tid; e < 256; e += NT) tw256[e] = tws[e << (BW_STAGE_LOG - 8)];
    // w_4096^e from the w_8192 table; w_T^e from the two-level table
    auto w4096 = [&](int e) -> TW { return tws[(2 * e) & ((1 << BW_STAGE_LOG) - 1)]; };
    auto wT = [&](int64_t e) -> TW { return tw_lookup<AR>((e << tshift) & tmask, twlo, twhi, pc); };

    // forward of one operand; after the last stage the 32 values go to keep
    typename AR::KV keep[G16 * 16];
    auto forward = [&](const W* in, auto km) {
        constexpr int KM = decltype(km)::value;   // 1: keep = X, 2: keep *= X, 3: keep *= X N^-1
        // stage 0: radix R0, stride 4096, straight from HBM
#pragma unroll 1
        for (int q = 0; q < G0; q++) {
            const int j = tid + q * NT;
            V x[R0];
#pragma unroll
            for (int m = 0; m < R0; m++) x[m] = AR::from_word(in[j + m * 4096], P);
            dft32_dif<AR, R0LOG>(x, pc);
            const TW st = wT(j);                                          // w_T^{j k} = st^k
            typename AR::Gen gen;
            gen.init(st, st, P);
            lds[pos(j)] = AR::to_lds(x[0], P);
#pragma unroll
            for (int k = 1; k < R0; k++) {
                lds[pos(j + k * 4096)] = AR::to_lds_r(gen.apply(x[brev(k, R0LOG)], P));
                gen.next(P);
            }
        }
        __syncthreads();
        // stage 1: radix 16, stride 256, blocks of 4096
#pragma unroll 1
        for (int q = 0; q < G16; q++) {
            const int g = tid + q * NT, b = g >> 8, j = g & 255;
            const int o = b * 4096 + j;
            V x[16];
#pragma unroll
            for (int m = 0; m < 16; m++) x[m] = AR::from_lds(lds[pos(o + m * 256)]);
            dft32_dif<AR, 4>(x, pc);
            const TW st = w4096(j);                                       // w_4096^{j k} = st^k
            typename AR::Gen gen;
            gen.init(st, st, P);
            lds[pos(o)] = AR::to_lds(x[0], P);
#pragma unroll
            for (int k = 1; k < 16; k++) {
                lds[pos(o + k * 256)] = AR::to_lds_r(gen.apply(x[brev(k, 4)], P));
                gen.next(P);
            }
        }
        __syncthreads();
        // stage 2: radix 16, stride 16, blocks of 256 (twiddles from the LDS table)
#pragma unroll 1
        for (int q = 0; q < G16; q++) {
            const int g = tid + q * NT, b = g >> 4, j = g & 15;
            const int o = b * 256 + j;
            V x[16];
#pragma unroll
            for (int m = 0; m < 16; m++) x[m] = AR::from_lds(lds[pos(o + m * 16)]);
            dft32_dif<AR, 4>(x, pc);
            lds[pos(o)] = AR::to_lds(x[0], P);
#pragma unroll
            for (int k = 1; k < 16; k++) lds[pos(o + k * 16)] = AR::to_lds_r(AR::mul(x[brev(k, 4)], tw256[j * k], P));
        }
        __syncthreads();
        // stage 3: radix 16, stride 1 -> registers
#pragma unroll
        for (int q = 0; q < G16; q++) {
            const int o = (tid + q * NT) * 16;
            V x[16];
#pragma unroll
            for (int m = 0; m < 16; m++) x[m] = AR::from_lds(lds[pos(o + m)]);
            dft32_dif<AR, 4>(x, pc);
#pragma unroll
            for (int k = 0; k < 16; k++) {
                auto& kv = keep[q * 16 + k];
                if constexpr (KM == 1) kv = AR::to_keep(x[brev(k, 4)], P);
                else if constexpr (KM == 2) kv = AR::keep_mul(kv, x[brev(k, 4)], P);
                else kv = AR::keep_mul_scaled(kv, x[brev(k, 4)], ninv, P);
            }
        }
        __syncthreads();        // the line is free for the next operand
    };

    const W* inA = reinterpret_cast<const W*>(prm.in[0]) + base;
    const W* inB = reinterpret_cast<const W*>(prm.in[1]) + base;
    forward(inA, std::integral_constant<int, 1>());
    if (prm.fold_ninv) forward(inB, std::integral_constant<int, 2>());
    else forward(inB, std::integral_constant<int, 3>());

    // inverse stage 3 (from registers), then the conjugate stage-2 twiddle of
    // position o + m: w_256^{-m (b mod 16)}
#pragma unroll
    for (int q = 0; q < G16; q++) {
        const int b = tid + q * NT, o = b * 16;
        V x[16];
#pragma unroll
        for (int k = 0; k < 16; k++) x[brev(k, 4)] = AR::from_keep(keep[q * 16 + k]);
        dft32_dit_inv<AR, 4>(x, pc);
        const int c = b & 15;
        lds[pos(o)] = AR::to_lds(x[0], P);
#pragma unroll
        for (int m = 1; m < 16; m++) lds[pos(o + m)] = AR::to_lds_r(AR::mul(x[m], tw256[(256 - m * c) & 255], P));
    }
    __syncthreads();
    // inverse stage 2, then the conjugate stage-1 twiddle of position
    // o + 16 m: w_4096^{-(j + 16 m) c}, c = b mod 16
#pragma unroll 1
    for (int q = 0; q < G16; q++) {
        const int g = tid + q * NT, b = g >> 4, j = g & 15;
        const int o = b * 256 + j;
        V x[16];
#pragma unroll
        for (int k = 0; k < 16; k++) x[brev(k, 4)] = AR::from_lds(lds[pos(o + k * 16)]);
        dft32_dit_inv<AR, 4>(x, pc);
        const int c = b & 15;
        typename AR::Gen gen;
        gen.init(w4096(4096 - j * c), w4096(4096 - 16 * c), P);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            lds[pos(o + m * 16)] = AR::to_lds_r(gen.apply(x[m], P));
            gen.next(P);
        }
    }
    __syncthreads();
    // inverse stage 1, then the conjugate stage-0 twiddle of position
    // o + 256 m: w_T^{-(j + 256 m) b}
#pragma unroll 1
    for (int q = 0; q < G16; q++) {
        const int g = tid + q * NT, b = g >> 8, j = g & 255;
        const int o = b * 4096 + j;
        V x[16];
#pragma unroll
        for (int k = 0; k < 16; k++) x[brev(k, 4)] = AR::from_lds(lds[pos(o + k * 256)]);
        dft32_dit_inv<AR, 4>(x, pc);
        typename AR::Gen gen;
        gen.init(wT(-(int64_t)j * b), wT(-(int64_t)256 * b), P);
#pragma unroll
        for (int m = 0; m < 16; m++) {
            lds[pos(o + m * 256)] = AR::to_lds_r(gen.apply(x[m], P));
            gen.next(P);
        }
    }
    __syncthreads();
    // inverse stage 0 -> HBM (centered residues for the next inverse pass)
    W* out = reinterpret_cast<W*>(prm.out[0]) + base;
#pragma unroll 1
    for (int q = 0; q < G0; q++) {
        const int j = tid + q * NT;
        V x[R0];
#pragma unroll
        for (int k = 0; k < R0; k++) x[brev(k, R0LOG)] = AR::from_lds(lds[pos(j + k * 4096)]);
        dft32_dit_inv<AR, R0LOG>(x, pc);
#pragma unroll
        for (int m = 0; m < R0; m++) out[j + m * 4096] = AR::to_word_c(x[m], P, prm.canon_out != 0);
    }
}

hipError_t bw_launch_row_fused(int logT, int64_t rows, int nprimes, const PassParams& prm, hipStream_t stream) {
    void (*k)(PassParams) = nullptr;
    switch (logT) {
    case 14: k = row_fused_kernel<ArF64x, 14>; break;
    case 15: k = row_fused_kernel<ArF64x, 15>; break;
    default: return hipErrorInvalidValue;
    }
    if (prm.S != 1 || rows < 1 || nprimes < 1) return hipErrorInvalidValue;
    const size_t lds = (size_t)row_fused_lds_bytes(logT);
    static std::once_flag once[2];
    hipError_t e = hipSuccess;
    std::call_once(once[logT - 14], [&] {
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((unsigned)rows, (unsigned)nprimes), dim3(BW_ROW_THREADS), lds, stream, prm);
    return hipGetLastError();
}
