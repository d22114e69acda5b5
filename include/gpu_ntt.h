// gpu_ntt.h -- drop-in for the reference's include/gpu_ntt.h
// (src/gpu_ntt.cu:240-1019): precompute / upload / allocate / execute /
// cleanup, as header-only C++ over libbirdwing_hip.so's C ABI.
//
// Differences a caller can observe:
//   * the spectrum layout and twiddle tables are this library's (the GPU-NTT
//     Root<T>/Modulus<T> tables of the reference are not exposed);
//   * execute_ntt_multiply() runs on ctx.stream_a with no host synchronization
//     until the final copy; C_out receives L_A + L_B limbs followed by zeros up
//     to its size (callers presize it to >= N + 1 as in the reference);
//   * failures print to stderr and exit(1), as the reference's size checks do.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <string>
#include <vector>
#include <hip/hip_runtime_api.h>

#include "birdwing.h"
#include "carry_prop.h"
#include "config.h"
#include "crt_gpu.h"
#include "ntt_limits.h"

struct PrecomputeTiming {
    float factors_ms = 0.0f;
    float params_ms = 0.0f;
    float twiddle_host_ms = 0.0f;
    float garner_host_ms = 0.0f;
    float total_ms = 0.0f;
};

struct SetupUploadTiming {
    float twiddle_upload_ms = 0.0f;
    float mod_constants_ms = 0.0f;
    float garner_upload_ms = 0.0f;
    float total_ms = 0.0f;
};

struct NTTTiming {
    float ingress_fwd_ms = 0.0f;
    float h2d_ms = 0.0f;
    float fwd_pad_ntt_ms = 0.0f;
    float fwd_pad_ntt_a_ms = 0.0f;
    float fwd_pad_ntt_b_ms = 0.0f;
    float pointwise_mul_ms = 0.0f;
    float intt_ms = 0.0f;
    float crt_ms = 0.0f;
    float carry_ms = 0.0f;
    float d2h_ms = 0.0f;
    float total_ms = 0.0f;
};

struct NTTPrecomputed {
    size_t N = 0;
    int logN = 0;
    int prime_set = -1;
    std::vector<uint64_t> primes;
    CRTGarnerParams garner{};
    bool gpu_uploaded = false;
};

struct NTTContext {
    size_t N = 0;
    int logN = 0;
    InputLimbType* a_raw_dev = nullptr;
    InputLimbType* b_raw_dev = nullptr;
    size_t L_A = 0, L_B = 0;
    hipStream_t stream_a = nullptr, stream_b = nullptr;
#if defined(NATIVE_HOST_LIMBS)
    uint64_t* d_C_lo = nullptr;
    uint64_t* d_C_mid = nullptr;
    uint32_t* d_C_hi32 = nullptr;
#else
    uint64_t* d_C_hi = nullptr;
    uint64_t* d_C_lo = nullptr;
#endif
    OutputLimbType* d_out = nullptr;
    bw_plan* plan = nullptr;        // products: the library's automatic prime set
    bw_plan* crt_plan = nullptr;    // CRT early exit: the precompute's three-prime set (made on first use)
    int prime_set = 0;
};

namespace bw_dropin {
inline void die(const char* what) {
    std::fprintf(stderr, "[NTT] %s: %s\n", what, bw_last_error());
    std::exit(1);
}
inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "[NTT] %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}
}  // namespace bw_dropin

namespace bw_dropin {
inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace bw_dropin

// Host side of the precompute: the prime set and its Garner constants (the
// twiddle tables are built once per prime set by upload_ntt_precomputed).
inline NTTPrecomputed precompute_ntt(size_t N, PrecomputeTiming* timing_out = nullptr) {
    const double t0 = bw_dropin::now_ms();
    ensure_ntt_size_supported(N);
    NTTPrecomputed pre;
    pre.N = N;
    while ((size_t(1) << pre.logN) < N) pre.logN++;
#if LIMB_BITS == 64 && !defined(NATIVE_HOST_LIMBS)
    pre.prime_set = BW_SET_HYBRID;        // the hybrid pipeline's two 58-bit primes (src/gpu_ntt.cu:40-43)
#else
    pre.prime_set = bw_default_prime_set(BW_CFG_LIMB_BITS, N);
#endif
    int k = 0;
    uint64_t P[3] = {0, 0, 0};
    if (bw_prime_set_info(pre.prime_set, &k, P, nullptr) != BW_OK) bw_dropin::die("precompute_ntt");
    pre.primes.assign(P, P + k);
    const double t1 = bw_dropin::now_ms();
    if (k == NUM_MODULI) {
        std::vector<TestDataTypeUint> pv(P, P + k);
        pre.garner = compute_garner_params(pv);
    }
    const double t2 = bw_dropin::now_ms();
    if (timing_out) {
        *timing_out = PrecomputeTiming{};
        timing_out->factors_ms = (float)(t1 - t0);
        timing_out->garner_host_ms = (float)(t2 - t1);
        timing_out->total_ms = (float)(t2 - t0);
    }
    return pre;
}

// Precompute cache (docs/TIMING.md:24-26,98-100).  The host tables are per
// prime set, not per N: save_ntt_precomputed writes the images of the
// precompute's set and of every set the library picks for products to
// `dir`/bw_tables_<set>.bin; load_ntt_precomputed reads the images found there
// (in a later process, before upload_ntt_precomputed) and returns the
// precompute for N.  Both return false / exit on a missing or damaged image.
inline bool save_ntt_precomputed(const NTTPrecomputed& pre, const std::string& dir) {
    for (int set : {pre.prime_set, (int)BW_SET_D64}) {
        const std::string path = dir + "/bw_tables_" + std::to_string(set) + ".bin";
        if (bw_tables_save(set, path.c_str()) != BW_OK) {
            std::fprintf(stderr, "[NTT] save_ntt_precomputed: %s\n", bw_last_error());
            return false;
        }
    }
    return true;
}

inline NTTPrecomputed load_ntt_precomputed(size_t N, const std::string& dir, PrecomputeTiming* timing_out = nullptr) {
    for (int set = 0; set < 16; set++) {
        const std::string path = dir + "/bw_tables_" + std::to_string(set) + ".bin";
        if (FILE* f = std::fopen(path.c_str(), "rb")) {
            std::fclose(f);
            int got = -1;
            if (bw_tables_load(path.c_str(), &got) != BW_OK || got != set) bw_dropin::die("load_ntt_precomputed");
        }
    }
    return precompute_ntt(N, timing_out);
}

// Twiddle tables are built and uploaded once per prime set and device (the
// first call pays for it; later ones find them resident).
inline void upload_ntt_precomputed(NTTPrecomputed& pre, SetupUploadTiming* timing_out = nullptr) {
    if (pre.gpu_uploaded) return;
    const double t0 = bw_dropin::now_ms();
    int dev = 0;
    bw_dropin::hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (bw_init(dev) != BW_OK) bw_dropin::die("upload_ntt_precomputed");
    bw_dropin::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    pre.gpu_uploaded = true;
    if (timing_out) {
        *timing_out = SetupUploadTiming{};
        timing_out->twiddle_upload_ms = (float)(bw_dropin::now_ms() - t0);
        timing_out->total_ms = timing_out->twiddle_upload_ms;
    }
}

inline NTTContext allocate_ntt_context(const NTTPrecomputed& pre, size_t L_A, size_t L_B) {
    if (!pre.gpu_uploaded) {
        std::fprintf(stderr, "[NTT] allocate_ntt_context: call upload_ntt_precomputed first\n");
        std::exit(1);
    }
    NTTContext ctx;
    ctx.N = pre.N;
    ctx.logN = pre.logN;
    ctx.L_A = L_A;
    ctx.L_B = L_B;
    ctx.prime_set = pre.prime_set;
    // Products are the same under every prime set; the automatic choice (the
    // five-prime 64-bit-digit set where it fits) is the fastest.
    if (bw_plan_create(&ctx.plan, BW_CFG_LIMB_BITS, L_A, L_B, -1) != BW_OK)
        bw_dropin::die("allocate_ntt_context");
    using bw_dropin::hip_check;
    hip_check(hipMalloc((void**)&ctx.a_raw_dev, L_A * sizeof(InputLimbType)), "hipMalloc");
    hip_check(hipMalloc((void**)&ctx.b_raw_dev, L_B * sizeof(InputLimbType)), "hipMalloc");
    hip_check(hipMalloc((void**)&ctx.d_out, (L_A + L_B) * sizeof(OutputLimbType)), "hipMalloc");
    hip_check(hipStreamCreate(&ctx.stream_a), "hipStreamCreate");
    hip_check(hipStreamCreate(&ctx.stream_b), "hipStreamCreate");
    return ctx;
}

inline void execute_ntt_multiply(NTTContext& ctx, const InputLimbType* a_pinned, const InputLimbType* b_pinned,
                                 std::vector<OutputLimbType>& C_out, NTTTiming* timing_out = nullptr,
                                 std::vector<uint64_t>* crt_hi_out = nullptr, std::vector<uint64_t>* crt_lo_out = nullptr,
                                 std::vector<uint64_t>* crt_mid_out = nullptr,
                                 std::vector<uint32_t>* crt_hi32_out = nullptr) {
    using bw_dropin::hip_check;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
#if defined(NATIVE_HOST_LIMBS)
    const bool early = crt_lo_out && crt_mid_out && crt_hi32_out;
#else
    const bool early = crt_hi_out && crt_lo_out;
#endif
    (void)crt_hi_out;
    (void)crt_mid_out;
    (void)crt_hi32_out;
    // the CRT coefficients are those of the precompute's three-prime set
    if (early && !ctx.crt_plan &&
        bw_plan_create(&ctx.crt_plan, BW_CFG_LIMB_BITS, ctx.L_A, ctx.L_B, ctx.prime_set) != BW_OK)
        bw_dropin::die("execute_ntt_multiply");
    bw_plan* const run = early ? ctx.crt_plan : ctx.plan;
    if (timing_out) {
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventCreate(&e2), "hipEventCreate");
        hip_check(hipEventCreate(&e3), "hipEventCreate");
        hip_check(hipEventRecord(e0, ctx.stream_a), "hipEventRecord");
        bw_plan_set_timing(run, 1);
    }
    hip_check(hipMemcpyAsync(ctx.a_raw_dev, a_pinned, ctx.L_A * sizeof(InputLimbType), hipMemcpyHostToDevice,
                             ctx.stream_a), "H2D");
    hip_check(hipMemcpyAsync(ctx.b_raw_dev, b_pinned, ctx.L_B * sizeof(InputLimbType), hipMemcpyHostToDevice,
                             ctx.stream_a), "H2D");
    if (timing_out) hip_check(hipEventRecord(e1, ctx.stream_a), "hipEventRecord");

    bw_plan_info info;
    bw_plan_get_info(run, &info);
    const size_t Nx = info.N_exec;
    if (early) {
#if defined(NATIVE_HOST_LIMBS)
        if (!ctx.d_C_lo) {
            hip_check(hipMalloc((void**)&ctx.d_C_lo, Nx * 8), "hipMalloc");
            hip_check(hipMalloc((void**)&ctx.d_C_mid, Nx * 8), "hipMalloc");
            hip_check(hipMalloc((void**)&ctx.d_C_hi32, Nx * 4), "hipMalloc");
        }
        if (bw_execute_crt(run, ctx.a_raw_dev, ctx.b_raw_dev, ctx.d_C_lo, ctx.d_C_mid, ctx.d_C_hi32,
                           (void*)ctx.stream_a) != BW_OK)
            bw_dropin::die("execute_ntt_multiply");
        crt_lo_out->resize(ctx.N);
        crt_mid_out->resize(ctx.N);
        crt_hi32_out->resize(ctx.N);
        hip_check(hipMemcpyAsync(crt_lo_out->data(), ctx.d_C_lo, ctx.N * 8, hipMemcpyDeviceToHost, ctx.stream_a), "D2H");
        hip_check(hipMemcpyAsync(crt_mid_out->data(), ctx.d_C_mid, ctx.N * 8, hipMemcpyDeviceToHost, ctx.stream_a), "D2H");
        hip_check(hipMemcpyAsync(crt_hi32_out->data(), ctx.d_C_hi32, ctx.N * 4, hipMemcpyDeviceToHost, ctx.stream_a),
                  "D2H");
#else
        if (!ctx.d_C_lo) {
            hip_check(hipMalloc((void**)&ctx.d_C_lo, Nx * 8), "hipMalloc");
            hip_check(hipMalloc((void**)&ctx.d_C_hi, Nx * 8), "hipMalloc");
        }
        if (bw_execute_crt(run, ctx.a_raw_dev, ctx.b_raw_dev, ctx.d_C_lo, ctx.d_C_hi, nullptr,
                           (void*)ctx.stream_a) != BW_OK)
            bw_dropin::die("execute_ntt_multiply");
        crt_hi_out->resize(ctx.N);
        crt_lo_out->resize(ctx.N);
        hip_check(hipMemcpyAsync(crt_hi_out->data(), ctx.d_C_hi, ctx.N * 8, hipMemcpyDeviceToHost, ctx.stream_a), "D2H");
        hip_check(hipMemcpyAsync(crt_lo_out->data(), ctx.d_C_lo, ctx.N * 8, hipMemcpyDeviceToHost, ctx.stream_a), "D2H");
#endif
    } else {
        if (bw_execute(ctx.plan, ctx.a_raw_dev, ctx.b_raw_dev, ctx.d_out, (void*)ctx.stream_a) != BW_OK)
            bw_dropin::die("execute_ntt_multiply");
        if (timing_out) hip_check(hipEventRecord(e2, ctx.stream_a), "hipEventRecord");
        const size_t n = ctx.L_A + ctx.L_B;
        if (C_out.size() < n) C_out.resize(n);
        hip_check(hipMemcpyAsync(C_out.data(), ctx.d_out, n * sizeof(OutputLimbType), hipMemcpyDeviceToHost,
                                 ctx.stream_a), "D2H");
        hip_check(hipStreamSynchronize(ctx.stream_a), "hipStreamSynchronize");
        std::fill(C_out.begin() + n, C_out.end(), (OutputLimbType)0);
    }
    hip_check(hipStreamSynchronize(ctx.stream_a), "hipStreamSynchronize");
    if (timing_out) {
        hip_check(hipEventRecord(e3, ctx.stream_a), "hipEventRecord");
        hip_check(hipEventSynchronize(e3), "hipEventSynchronize");
        NTTTiming t{};
        hip_check(hipEventElapsedTime(&t.h2d_ms, e0, e1), "hipEventElapsedTime");
        float ms[64];
        int n = 0;
        bw_plan_timing(run, ms, 64, &n);
        bw_plan_set_timing(run, 0);
        // launch order (bw_plan_info.schedule): forward passes, the fused pass
        // (last forward + pointwise + first inverse), the remaining inverse
        // passes -- over all primes at once (pass-major), prime by prime
        // (prime-major, forward passes per operand), or forward over all primes
        // and the rest prime by prime (hybrid); then one Garner(+carry) launch
        const int P = info.npass, k = info.k;
        for (int i = 0; i + 1 < n; i++) {
            int kind = 1;                               // 0 forward, 1 fused, 2 inverse
            if (P > 1) {
                if (info.schedule == BW_SCHED_PRIME_MAJOR) {
                    const int pos = i % (3 * (P - 1) + 1);
                    kind = pos < 2 * (P - 1) ? 0 : (pos == 2 * (P - 1) ? 1 : 2);
                } else if (info.schedule == BW_SCHED_HYBRID) {
                    kind = i < P - 1 ? 0 : ((i - (P - 1)) % P == 0 ? 1 : 2);
                } else {
                    kind = i < P - 1 ? 0 : (i == P - 1 ? 1 : 2);
                }
            }
            (kind == 0 ? t.fwd_pad_ntt_ms : kind == 1 ? t.pointwise_mul_ms : t.intt_ms) += ms[i];
        }
        (void)k;
        if (n > 0) (early ? t.crt_ms : t.carry_ms) = ms[n - 1];
        t.fwd_pad_ntt_a_ms = t.fwd_pad_ntt_b_ms = t.fwd_pad_ntt_ms / 2;
        t.ingress_fwd_ms = t.h2d_ms + t.fwd_pad_ntt_ms;
        if (!early) hip_check(hipEventElapsedTime(&t.d2h_ms, e2, e3), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&t.total_ms, e0, e3), "hipEventElapsedTime");
        *timing_out = t;
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); (void)hipEventDestroy(e2); (void)hipEventDestroy(e3);
    }
}

inline void cleanup_ntt_context(NTTContext& ctx) {
    bw_plan_destroy(ctx.plan);
    bw_plan_destroy(ctx.crt_plan);
    ctx.plan = ctx.crt_plan = nullptr;
    (void)hipFree(ctx.a_raw_dev);
    (void)hipFree(ctx.b_raw_dev);
    (void)hipFree(ctx.d_out);
#if defined(NATIVE_HOST_LIMBS)
    (void)hipFree(ctx.d_C_lo); (void)hipFree(ctx.d_C_mid); (void)hipFree(ctx.d_C_hi32);
    ctx.d_C_lo = ctx.d_C_mid = nullptr; ctx.d_C_hi32 = nullptr;
#else
    (void)hipFree(ctx.d_C_hi); (void)hipFree(ctx.d_C_lo);
    ctx.d_C_hi = ctx.d_C_lo = nullptr;
#endif
    if (ctx.stream_a) (void)hipStreamDestroy(ctx.stream_a);
    if (ctx.stream_b) (void)hipStreamDestroy(ctx.stream_b);
    ctx.a_raw_dev = ctx.b_raw_dev = nullptr;
    ctx.d_out = nullptr;
    ctx.stream_a = ctx.stream_b = nullptr;
}

// Device tables stay with the library (shared per prime set); this only marks
// the precompute as no longer uploaded, as the reference does.
inline void cleanup_ntt_precomputed(NTTPrecomputed& pre) { pre.gpu_uploaded = false; }
