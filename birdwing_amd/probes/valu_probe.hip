// valu_probe.hip -- arithmetic-rate probes for gfx950 (measurement only; not
// part of libbirdwing_hip.so).  Each kernel runs NCH independent dependency
// chains per lane for `iters` steps, so issue, not latency, bounds it; the
// host times it with hipEvents and reports operations per second over the
// whole device.
//
//   fma_f64        v_fma_f64 chains                         (1 op per step)
//   mul_lo_u32     v_mul_lo_u32 chains                      (1 op per step)
//   mul_hi_u32     v_mul_hi_u32 chains                      (1 op per step)
//   mad_u64_u32    64-bit products from v_mad_u64_u32        (1 op per step)
//   f64_modmul     the FP64 modular product of the 32-bit-digit kernels
//                  (q = rint(y*w/p) by the 1.5*2^52 trick, r = fma(y, w, -q(p-1)) - q)
//   shoup_u32      Shoup's precomputed-quotient product on 32-bit words
//                  (q = mulhi(y, w'), r = y*w - q*p, one conditional subtract)
//   mont_u64       the 64-bit Montgomery product of the 58-bit-prime kernels
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NCH 8

__global__ __launch_bounds__(256) void k_fma_f64(double* out, double a, double b, int iters) {
    double x[NCH];
    for (int k = 0; k < NCH; k++) x[k] = threadIdx.x + k;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) x[k] = __builtin_fma(x[k], a, b);
    double s = 0;
    for (int k = 0; k < NCH; k++) s += x[k];
    if (s == 1.2345) out[threadIdx.x] = s;        // never true for these inputs; keeps the chains live
}

__global__ __launch_bounds__(256) void k_mul_lo_u32(uint32_t* out, uint32_t a, int iters) {
    uint32_t x[NCH];
    for (int k = 0; k < NCH; k++) x[k] = threadIdx.x * 2 + 1 + k;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) x[k] = x[k] * a;
    uint32_t s = 0;
    for (int k = 0; k < NCH; k++) s ^= x[k];
    if (s == 0x12345u) out[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mul_hi_u32(uint32_t* out, uint32_t a, int iters) {
    uint32_t x[NCH];
    for (int k = 0; k < NCH; k++) x[k] = 0xF0000000u + threadIdx.x + k;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) x[k] = __umulhi(x[k], a);
    uint32_t s = 0;
    for (int k = 0; k < NCH; k++) s ^= x[k];
    if (s == 0x12345u) out[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mad_u64_u32(uint64_t* out, uint32_t a, int iters) {
    uint64_t x[NCH];
    for (int k = 0; k < NCH; k++) x[k] = threadIdx.x + k;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) x[k] = (uint64_t)(uint32_t)x[k] * a + (x[k] >> 32);
    uint64_t s = 0;
    for (int k = 0; k < NCH; k++) s ^= x[k];
    if (s == 0x12345u) out[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_f64_modmul(double* out, double w, double wp, double p, double cm, int iters) {
    const double M = 6755399441055744.0;
    double y[NCH];
    for (int k = 0; k < NCH; k++) y[k] = (double)(threadIdx.x * 977 + k * 131);
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) {
            const double q = __builtin_fma(y[k], wp, M) - M;
            y[k] = __builtin_fma(y[k], w, -(q * cm)) - q;
        }
    double s = 0;
    for (int k = 0; k < NCH; k++) s += y[k];
    if (s == 1.2345) out[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_shoup_u32(uint32_t* out, uint32_t w, uint32_t wq, uint32_t p, int iters) {
    uint32_t y[NCH];
    for (int k = 0; k < NCH; k++) y[k] = (threadIdx.x * 977 + k * 131) % p;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) {
            const uint32_t q = __umulhi(y[k], wq);
            uint32_t r = y[k] * w - q * p;
            y[k] = r >= p ? r - p : r;
        }
    uint32_t s = 0;
    for (int k = 0; k < NCH; k++) s ^= y[k];
    if (s == 0x12345u) out[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mont_u64(uint64_t* out, uint64_t b, uint64_t p, uint64_t pinv_neg, int iters) {
    uint64_t y[NCH];
    for (int k = 0; k < NCH; k++) y[k] = (threadIdx.x * 977ull + k * 131ull) % p;
    for (int i = 0; i < iters; i++)
#pragma unroll
        for (int k = 0; k < NCH; k++) {
            const uint64_t lo = y[k] * b, hi = __umul64hi(y[k], b);
            const uint64_t m = lo * pinv_neg;
            const uint64_t r = hi + __umul64hi(m, p) + (lo != 0 ? 1ull : 0ull);
            y[k] = r >= p ? r - p : r;
        }
    uint64_t s = 0;
    for (int k = 0; k < NCH; k++) s ^= y[k];
    if (s == 0x12345u) out[threadIdx.x] = s;
}

// ops/s of probe `which` (0..6, the order above): `blocks` workgroups of 256
// threads, `iters` steps of NCH chains.  Returns -1 on a HIP error.
extern "C" double bwp_valu_rate(int which, int blocks, int iters) {
    void* out = nullptr;
    if (hipMalloc(&out, 256 * 8) != hipSuccess) return -1.0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0;
    // 2113929217 = 63 * 2^25 + 1: the FP64 path's first prime; w a residue below it
    const double p = 2113929217.0, w = 1234567891.0;
    const uint32_t p32 = 2113929217u, w32 = 1234567891u;
    const uint32_t wq = (uint32_t)(((uint64_t)w32 << 32) / p32);
    const uint64_t p64 = 0x400002600000001ull;
    uint64_t inv = 1;
    for (int i = 0; i < 6; i++) inv *= 2 - p64 * inv;      // p^-1 mod 2^64 by Newton steps
    double secs = 0.0;
    for (int rep = 0; rep < 3; rep++) {                     // the first run warms the device
        (void)hipEventRecord(e0, 0);
        switch (which) {
        case 0: hipLaunchKernelGGL(k_fma_f64, dim3(blocks), dim3(256), 0, 0, (double*)out, 1.0000001, 1e-9, iters); break;
        case 1: hipLaunchKernelGGL(k_mul_lo_u32, dim3(blocks), dim3(256), 0, 0, (uint32_t*)out, 2654435761u, iters); break;
        case 2: hipLaunchKernelGGL(k_mul_hi_u32, dim3(blocks), dim3(256), 0, 0, (uint32_t*)out, 2654435761u, iters); break;
        case 3: hipLaunchKernelGGL(k_mad_u64_u32, dim3(blocks), dim3(256), 0, 0, (uint64_t*)out, 2654435761u, iters); break;
        case 4: hipLaunchKernelGGL(k_f64_modmul, dim3(blocks), dim3(256), 0, 0, (double*)out, w, w / p, p, p - 1.0, iters); break;
        case 5: hipLaunchKernelGGL(k_shoup_u32, dim3(blocks), dim3(256), 0, 0, (uint32_t*)out, w32, wq, p32, iters); break;
        case 6: hipLaunchKernelGGL(k_mont_u64, dim3(blocks), dim3(256), 0, 0, (uint64_t*)out, 0x123456789abcdefull % p64,
                                   p64, (uint64_t)0 - inv, iters); break;
        default: return -1.0;
        }
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) return -1.0;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && (secs == 0.0 || ms * 1e-3 < secs)) secs = ms * 1e-3;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return (double)blocks * 256.0 * NCH * (double)iters / secs;
}
