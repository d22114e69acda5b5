// The reference-API three-pass carry kernels from carry_prop.h, launched as
// the reference's host code launches them (intra, inter, fixup until no carry
// escapes a segment), against the library's single-pass carry on the same
// coefficients.  The all-ones case makes carries cross many segments.  Built for 32-bit (128-bit coefficients) and for
// NATIVE_HOST_LIMBS (160-bit coefficients).
#include "carry_prop.h"

#include <cstring>
#include <random>
#include <vector>

#define HIP_OK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            std::exit(2);                                                                \
        }                                                                                \
    } while (0)

int main() {
    int fails = 0;
    if (bw_abi_version() != BW_ABI_VERSION) {
        std::fprintf(stderr, "ABI mismatch: header %d, library %d (rebuild tests/cpp)\n", BW_ABI_VERSION,
                     bw_abi_version());
        return 3;
    }
    std::mt19937_64 rng(5);
    for (size_t N : {size_t(1), size_t(1000), size_t(1024), size_t(5000), size_t(70001)}) {
        for (int mode = 0; mode < 2; mode++) {  // 0: random, 1: all-ones words (long carry chains)
            std::vector<uint64_t> w0(N), w1(N);
            std::vector<uint32_t> w2(N);
            for (size_t i = 0; i < N; i++) {
                w0[i] = mode ? ~0ULL : rng();
#if defined(NATIVE_HOST_LIMBS)
                w1[i] = mode ? ~0ULL : rng();
#else
                // 128-bit coefficients stay below 2^88 (the product bound), so
                // the reference's int64 segment carries cannot overflow
                w1[i] = mode ? (1ULL << 24) - 1 : rng() >> 40;
#endif
                w2[i] = mode ? 0x7fffffffu : (uint32_t)(rng() >> 33);
            }
            const size_t n_out = N + (OUTPUT_LIMB_BITS == 64 ? 3 : 4);
            const size_t segs = (N + CARRY_SEG - 1) / CARRY_SEG;
            uint64_t *d0, *d1;
            uint32_t* d2;
            OutputLimbType *o1, *o2;
            int* esc;
            HIP_OK(hipMalloc(&d0, N * 8));
            HIP_OK(hipMalloc(&d1, N * 8));
            HIP_OK(hipMalloc(&d2, N * 4));
            HIP_OK(hipMalloc(&o1, n_out * sizeof(OutputLimbType)));
            HIP_OK(hipMalloc(&o2, n_out * sizeof(OutputLimbType)));
            HIP_OK(hipMalloc(&esc, sizeof(int)));
            HIP_OK(hipMemcpy(d0, w0.data(), N * 8, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(d1, w1.data(), N * 8, hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(d2, w2.data(), N * 4, hipMemcpyHostToDevice));
            HIP_OK(hipMemset(o1, 0, n_out * sizeof(OutputLimbType)));
#if defined(NATIVE_HOST_LIMBS)
            carry_propagate_gpu_u160(d0, d1, d2, N, o2, n_out);
            uint64_t *sl, *sm;
            uint32_t* sh;
            HIP_OK(hipMalloc(&sl, segs * 8));
            HIP_OK(hipMalloc(&sm, segs * 8));
            HIP_OK(hipMalloc(&sh, segs * 4));
            carry_intra_segment_kernel_u160<<<segs, 64>>>(d0, d1, d2, o1, sl, sm, sh, N);
            carry_inter_segment_kernel_u160<<<1, 64>>>(sl, sm, sh, segs);
            for (int it = 0; it < 64; it++) {
                HIP_OK(hipMemset(esc, 0, sizeof(int)));
                carry_fixup_kernel_u160<<<segs, 64>>>(o1, sl, sm, sh, N, segs, esc);
                int e = 0;
                HIP_OK(hipMemcpy(&e, esc, sizeof(int), hipMemcpyDeviceToHost));
                if (!e) break;
            }
            HIP_OK(hipFree(sl)); HIP_OK(hipFree(sm)); HIP_OK(hipFree(sh));
#else
            // 128-bit coefficients: hi = w1, lo = w0
            carry_propagate_gpu(d1, d0, N, o2, n_out);
            int64_t* sc;
            HIP_OK(hipMalloc(&sc, segs * 8));
            carry_intra_segment_kernel<<<segs, 64>>>(d1, d0, o1, sc, N);
            carry_inter_segment_kernel<<<1, 64>>>(sc, segs);
            for (int it = 0; it < 64; it++) {
                HIP_OK(hipMemset(esc, 0, sizeof(int)));
                carry_fixup_kernel<<<segs, 64>>>(o1, sc, N, segs, esc);
                int e = 0;
                HIP_OK(hipMemcpy(&e, esc, sizeof(int), hipMemcpyDeviceToHost));
                if (!e) break;
            }
            HIP_OK(hipFree(sc));
#endif
            HIP_OK(hipDeviceSynchronize());
            std::vector<OutputLimbType> a(n_out), b(n_out);
            HIP_OK(hipMemcpy(a.data(), o1, n_out * sizeof(OutputLimbType), hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(b.data(), o2, n_out * sizeof(OutputLimbType), hipMemcpyDeviceToHost));
            // the three-pass kernels write limbs [0, N) only (as in the reference)
            if (std::memcmp(a.data(), b.data(), N * sizeof(OutputLimbType)) != 0) {
                std::fprintf(stderr, "FAIL N=%zu mode=%d\n", N, mode);
                fails++;
            }
            HIP_OK(hipFree(d0)); HIP_OK(hipFree(d1)); HIP_OK(hipFree(d2));
            HIP_OK(hipFree(o1)); HIP_OK(hipFree(o2)); HIP_OK(hipFree(esc));
        }
    }
    std::printf("dropin_carry_kernels: %s (%d failures)\n", fails ? "FAIL" : "ok", fails);
    return fails ? 1 : 0;
}
