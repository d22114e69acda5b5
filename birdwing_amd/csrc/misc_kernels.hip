// misc_kernels.hip -- elementwise kernels behind the stage-level entry points.
#include "bw_internal.h"
#include "misc_kernels.h"

// src/zero_pad.cu:3-12: widen 32-bit limbs into the NTT word, zero [L, N); no reduction.
template <class Wd>
__global__ void zero_pad_kernel(const uint32_t* __restrict__ src, Wd* __restrict__ dst, int64_t L, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = i < L ? (Wd)src[i] : (Wd)0;
}

// src/zero_pad.cu:28-42: 64-bit limbs reduced mod p, zero [L, N).
__global__ void zero_pad_u64_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, int64_t L,
                                    int64_t N, U64Prime P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dst[i] = i < L ? u64_reduce(src[i], P) : 0;
}

// src/gpu_ntt.cu:216-238: C = A*B mod p for arbitrary words (full product, then reduce).
__global__ void pointwise_u32_kernel(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                                     uint32_t* __restrict__ c, int64_t N, F64Prime P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double br = f64_reduce((double)b[i], P);
    c[i] = f64_canon(f64_mulmod((double)a[i], br, br * P.pinv, P), P);
}

__global__ void pointwise_u64_kernel(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                     uint64_t* __restrict__ c, int64_t N, U64Prime P) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint64_t x = u64_reduce(a[i], P), y = u64_reduce(b[i], P);
    c[i] = u64_montmul(u64_montmul(x, y, P), P.r2, P);
}

// x[plane j][i] *= s_j mod p_j (planes of `plane` words, k planes).
__global__ void scale_u32_kernel(uint32_t* __restrict__ x, int64_t N, int64_t plane, const F64PrimeDev* pc,
                                 const double* pairs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= N) return;
    const F64Prime P = pc[j].P;
    uint32_t* xp = x + (int64_t)j * plane;
    xp[i] = f64_canon(f64_mulmod((double)xp[i], pairs[2 * j], pairs[2 * j + 1], P), P);
}

__global__ void scale_u64_kernel(uint64_t* __restrict__ x, int64_t N, int64_t plane, const U64PrimeDev* pc,
                                 const uint64_t* mont_pairs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= N) return;
    const U64Prime P = pc[j].P;
    uint64_t* xp = x + (int64_t)j * plane;
    // montmul(montmul(x, K), 1*2^64 -> ...) : K = s 2^128 gives montmul(montmul(x,K), 1) = x s
    xp[i] = u64_montmul(u64_montmul(xp[i], mont_pairs[2 * j], P), 1, P);
}

static dim3 grid1(int64_t n, unsigned y = 1) { return dim3((unsigned)((n + 255) / 256), y); }

hipError_t bw_launch_zero_pad(const uint32_t* src, void* dst, int word_bytes, int64_t L, int64_t N, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    if (word_bytes == 8)
        hipLaunchKernelGGL(zero_pad_kernel<uint64_t>, grid1(N), dim3(256), 0, s, src, (uint64_t*)dst, L, N);
    else
        hipLaunchKernelGGL(zero_pad_kernel<uint32_t>, grid1(N), dim3(256), 0, s, src, (uint32_t*)dst, L, N);
    return hipGetLastError();
}

hipError_t bw_launch_zero_pad_u64(const uint64_t* src, uint64_t* dst, int64_t L, int64_t N, const U64Prime& P,
                                  hipStream_t s) {
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_pad_u64_kernel, grid1(N), dim3(256), 0, s, src, dst, L, N, P);
    return hipGetLastError();
}

hipError_t bw_launch_pointwise(int word_bytes, const void* a, const void* b, void* c, int64_t N, uint64_t p,
                               hipStream_t s) {
    if (N <= 0) return hipSuccess;
    if (word_bytes == 8) {
        U64Prime P;
        P.p = p;
        uint64_t inv = 1;
        for (int it = 0; it < 7; it++) inv *= 2 - p * inv;
        P.pinv_neg = (uint64_t)0 - inv;
        typedef unsigned __int128 u128;
        const u128 r = (((u128)1) << 64) % p;
        P.r2 = (uint64_t)(r * r % p);
        P.floor_r = ~0ULL / p;
        hipLaunchKernelGGL(pointwise_u64_kernel, grid1(N), dim3(256), 0, s, (const uint64_t*)a, (const uint64_t*)b,
                           (uint64_t*)c, N, P);
    } else {
        F64Prime P;
        P.p = (double)p;
        P.pinv = 1.0 / (double)p;
        P.cm = (double)(p - 1);
        hipLaunchKernelGGL(pointwise_u32_kernel, grid1(N), dim3(256), 0, s, (const uint32_t*)a, (const uint32_t*)b,
                           (uint32_t*)c, N, P);
    }
    return hipGetLastError();
}

hipError_t bw_launch_scale(int limb_bits, void* x, int64_t N, int64_t plane, int k, const void* consts,
                           const void* pairs, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    if (limb_bits == 64)
        hipLaunchKernelGGL(scale_u64_kernel, grid1(N, k), dim3(256), 0, s, (uint64_t*)x, N, plane,
                           (const U64PrimeDev*)consts, (const uint64_t*)pairs);
    else
        hipLaunchKernelGGL(scale_u32_kernel, grid1(N, k), dim3(256), 0, s, (uint32_t*)x, N, plane,
                           (const F64PrimeDev*)consts, (const double*)pairs);
    return hipGetLastError();
}
