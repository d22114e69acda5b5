// mfma_probe.hip -- one 16-point DFT stage mod p on int8 MFMA against the
// VALU codelet of the product kernels (measurement only; not part of
// libbirdwing_hip.so).
//
// Both kernels transform the same lines of 16 centered residues (|x| < p/2,
// p = 2113929217 = 63 * 2^25 + 1, the first prime of the product's five-prime
// set) and store canonical residues X_k = sum_j x_j w^(jk) mod p in natural
// order.
//
// VALU: the product's radix-16 codelet (dft32_dif, FP64 modular arithmetic),
// one line per thread.
//
// MFMA (v_mfma_i32_16x16x64_i8): a residue is four balanced base-256 digits
// d_a in [-128, 128) (the bytes of x + 0x80808080, each XOR 0x80: two VALU
// operations), and so is every twiddle w^(jk).  One MFMA multiplies a block of
// 16 lines (A: 16 x 64, the digits of their 16 residues) by B_s (64 x 16:
// digit s - a of w^(jk) at row (j, a), column k) and accumulates
// D_s = sum_{j, a} d_a(x_j) d_{s-a}(w^(jk)), |D_s| < 2^20, for s = 0..6: seven
// MFMAs give X_k = sum_s D_s 256^s exactly, reduced mod p by Horner steps in
// FP64 (seven conversions, six FMAs, three reductions per output).  The
// pairing of A's and B's K index is the instruction's own (lane group g, byte
// j of the lane), so the (j, a) <-> K assignment only has to agree between A
// and B; D element r of lane l is row 4 (l / 16) + r, column l % 16.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstring>
#include <vector>

#include "../csrc/ntt2_impl.h"

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t canon_u32(double x, const F64Prime& P) { return f64_canon(x, P); }

// `rounds` stages per line: each a 16-point DFT followed by a stage boundary
// like the product's (reduce, 32-bit words through LDS, back to FP64); the
// stored output is the last round's, canonical.
// VALU: one line per thread, the product's codelet.
__global__ __launch_bounds__(256) void dft16_valu(const int32_t* in, uint32_t* out, int64_t lines, F64PrimeDev pc,
                                                  int rounds) {
    __shared__ int32_t lds[256 * 17];
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= lines) return;                  // lines is a multiple of 256 (the host checks)
    double x[16];
    const int4* src = reinterpret_cast<const int4*>(in + line * 16);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int4 v = src[q];
        x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
    }
    int32_t* my = lds + threadIdx.x * 17;
#pragma unroll 1
    for (int r = 0; r < rounds; r++) {
        dft32_dif<ArF64x, 4>(x, pc);
        if (r + 1 == rounds) break;
        // natural order X_k -> words -> LDS -> FP64 (the next stage's input)
#pragma unroll
        for (int k = 0; k < 16; k++) my[k] = (int32_t)f64_reduce(x[brev(k, 4)], pc.P);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k++) x[k] = (double)my[k];
    }
    uint4* dst = reinterpret_cast<uint4*>(out + line * 16);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint4 v;
        v.x = canon_u32(x[brev(4 * q, 4)], pc.P);
        v.y = canon_u32(x[brev(4 * q + 1, 4)], pc.P);
        v.z = canon_u32(x[brev(4 * q + 2, 4)], pc.P);
        v.w = canon_u32(x[brev(4 * q + 3, 4)], pc.P);
        dst[q] = v;
    }
}

// MFMA: blocks of 16 lines per wave; lane l = 16 g + m holds residues
// 4g .. 4g+3 of line m as 16 digit bytes; btab: the seven B_s operands, one
// v4i per lane each.  Between rounds the outputs (lane l: lines 4g + r,
// output l % 16) go through LDS back to the A layout.
__global__ __launch_bounds__(256) void dft16_mfma(const int32_t* in, uint32_t* out, int64_t lines, const v4i* btab,
                                                  F64Prime P, int rounds) {
    __shared__ int32_t lds[4][16 * 17];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t line0 = wave * 16;
    if (line0 >= lines) return;                 // lines is a multiple of 64 (the host checks)
    v4i B[7];
#pragma unroll
    for (int s = 0; s < 7; s++) B[s] = btab[s * 64 + lane];
    const int m = lane & 15, g = lane >> 4;
    int4 v = *reinterpret_cast<const int4*>(in + (line0 + m) * 16 + 4 * g);
    int32_t* blk = lds[wv];
    double h[4];
#pragma unroll 1
    for (int rd = 0; rd < rounds; rd++) {
        v4i a;
        a.x = (int)(((uint32_t)v.x + 0x80808080u) ^ 0x80808080u);
        a.y = (int)(((uint32_t)v.y + 0x80808080u) ^ 0x80808080u);
        a.z = (int)(((uint32_t)v.z + 0x80808080u) ^ 0x80808080u);
        a.w = (int)(((uint32_t)v.w + 0x80808080u) ^ 0x80808080u);
        v4i D[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            const v4i z = {0, 0, 0, 0};
            D[s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, B[s], z, 0, 0, 0);
        }
        // element r of lane l: line 4 g + r of the block, output k = l % 16
#pragma unroll
        for (int r = 0; r < 4; r++) {
            double t = (double)D[6][r];
            t = __builtin_fma(t, 256.0, (double)D[5][r]);
            t = __builtin_fma(t, 256.0, (double)D[4][r]);
            t = __builtin_fma(t, 256.0, (double)D[3][r]);      // |t| < 2^45
            t = f64_reduce(t, P);
            t = __builtin_fma(t, 256.0, (double)D[2][r]);
            t = __builtin_fma(t, 256.0, (double)D[1][r]);      // |t| < 2^47
            t = f64_reduce(t, P);
            t = __builtin_fma(t, 256.0, (double)D[0][r]);
            h[r] = f64_reduce(t, P);                           // centered, |h| <= p/2 + 1
        }
        if (rd + 1 == rounds) break;
        // back to the A layout through LDS (the wave's own 16 lines)
#pragma unroll
        for (int r = 0; r < 4; r++) blk[(4 * g + r) * 17 + m] = (int32_t)h[r];
        __syncthreads();
        v.x = blk[m * 17 + 4 * g];
        v.y = blk[m * 17 + 4 * g + 1];
        v.z = blk[m * 17 + 4 * g + 2];
        v.w = blk[m * 17 + 4 * g + 3];
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; r++) out[(line0 + 4 * g + r) * 16 + m] = canon_u32(h[r], P);
}

__global__ void count_diff(const uint32_t* a, const uint32_t* b, int64_t n, unsigned long long* cnt) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (a[i] != b[i]) atomicAdd(cnt, 1ULL);
}

static uint64_t mulmod_h(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((unsigned __int128)a * b % p); }
static uint64_t powmod_h(uint64_t a, uint64_t e, uint64_t p) {
    uint64_t r = 1;
    a %= p;
    while (e) {
        if (e & 1) r = mulmod_h(r, a, p);
        a = mulmod_h(a, a, p);
        e >>= 1;
    }
    return r;
}
// balanced base-256 digits of a centered value |c| < 2^30, as signed bytes
static void balanced_digits(int64_t c, int8_t d[4]) {
    const uint32_t e = (uint32_t)((int32_t)c) + 0x80808080u;
    for (int a = 0; a < 4; a++) d[a] = (int8_t)((uint8_t)(e >> (8 * a)) ^ 0x80);
}

// rounds: DFT stages per line (with a stage boundary between).
// res[0]: VALU seconds, res[1]: MFMA seconds, res[2]: lines, res[3]: outputs
// that differ between the two, res[4]: w (the 16th root used); in_host (if
// not null) receives the first 16*nsample inputs and out_host the MFMA
// outputs of those lines.  Returns 0, or -1 on a HIP error.
extern "C" int bwp_mfma_dft16(int64_t lines, int reps, int rounds, double* res, int32_t* in_host,
                              uint32_t* out_host, int nsample) {
    if (lines <= 0 || lines % 256 || rounds < 1) return -1;
    const uint64_t p = 2113929217ULL;            // 63 * 2^25 + 1, primitive root 5
    const uint64_t w = powmod_h(5, (p - 1) / 16, p);
    const uint64_t w32 = powmod_h(5, (p - 1) / 32, p);
    F64PrimeDev pc;
    std::memset(&pc, 0, sizeof(pc));
    pc.P.p = (double)p;
    pc.P.pinv = 1.0 / (double)p;
    pc.P.cm = (double)(p - 1);
    for (int e = 0; e < 16; e++) {
        const uint64_t v = powmod_h(w32, e, p), iv = powmod_h(w32, (32 - e) % 32, p);
        pc.w32[e][0] = (double)v;
        pc.w32[e][1] = (double)v / (double)p;
        pc.iw32[e][0] = (double)iv;
        pc.iw32[e][1] = (double)iv / (double)p;
    }
    // B_s: lane l = 16 g + n, byte j: row (jj = 4 g + j / 4, a = j % 4), column n
    std::vector<int8_t> bt(7 * 64 * 16, 0);
    for (int l = 0; l < 64; l++) {
        const int n = l & 15, g = l >> 4;
        for (int j = 0; j < 16; j++) {
            const int jj = 4 * g + j / 4, a = j % 4;
            const uint64_t t = powmod_h(w, (uint64_t)jj * n, p);
            const int64_t c = t > p / 2 ? (int64_t)t - (int64_t)p : (int64_t)t;
            int8_t d[4];
            balanced_digits(c, d);
            for (int s = 0; s < 7; s++) {
                const int e = s - a;
                bt[(s * 64 + l) * 16 + j] = (e >= 0 && e < 4) ? d[e] : 0;
            }
        }
    }
    // inputs: centered residues from a fixed linear congruential sequence
    std::vector<int32_t> hin((size_t)lines * 16);
    uint64_t st = 0x9E3779B97F4A7C15ULL;
    for (auto& x : hin) {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        const uint64_t r = (st >> 16) % p;
        x = (int32_t)(r > p / 2 ? (int64_t)r - (int64_t)p : (int64_t)r);
    }
    int32_t* din = nullptr;
    uint32_t *dv = nullptr, *dm = nullptr;
    v4i* db = nullptr;
    unsigned long long* dcnt = nullptr;
    const size_t nb = (size_t)lines * 16 * 4;
    if (hipMalloc(&din, nb) != hipSuccess || hipMalloc(&dv, nb) != hipSuccess || hipMalloc(&dm, nb) != hipSuccess ||
        hipMalloc(&db, bt.size()) != hipSuccess || hipMalloc(&dcnt, 8) != hipSuccess)
        return -1;
    if (hipMemcpy(din, hin.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, bt.data(), bt.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(dcnt, 0, 8) != hipSuccess)
        return -1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
    const unsigned gv = (unsigned)((lines + 255) / 256);
    const unsigned gm = (unsigned)((lines / 16 * 64 + 255) / 256);
    double tv = 0.0, tm = 0.0;
    for (int r = 0; r <= reps; r++) {          // run 0 warms up
        float ms = 0.0f;
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(dft16_valu, dim3(gv), dim3(256), 0, 0, din, dv, lines, pc, rounds);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return -1;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) tv += ms * 1e-3;
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(dft16_mfma, dim3(gm), dim3(256), 0, 0, din, dm, lines, db, pc.P, rounds);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return -1;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) tm += ms * 1e-3;
    }
    hipLaunchKernelGGL(count_diff, dim3(1024), dim3(256), 0, 0, dv, dm, lines * 16, dcnt);
    unsigned long long diff = 0;
    if (hipMemcpy(&diff, dcnt, 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (in_host && out_host && nsample > 0) {
        const int64_t ns = nsample < lines ? nsample : lines;
        std::memcpy(in_host, hin.data(), (size_t)ns * 16 * 4);
        if (hipMemcpy(out_host, dm, (size_t)ns * 16 * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    }
    res[0] = tv / reps;
    res[1] = tm / reps;
    res[2] = (double)lines;
    res[3] = (double)diff;
    res[4] = (double)w;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(din); (void)hipFree(dv); (void)hipFree(dm); (void)hipFree(db); (void)hipFree(dcnt);
    return 0;
}
