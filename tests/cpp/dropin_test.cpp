// Drop-in API test: the reference's C++ interface (config.h, ntt_limits.h,
// crt.h, crt_utils.h, crt_gpu.h, carry_prop.h, zero_pad.h, gpu_ntt.h,
// multiply.h) compiled against libbirdwing_hip.so, with products checked
// against GMP's mpz_mul.  Built three times: LIMB_BITS=32, LIMB_BITS=64
// (hybrid) and LIMB_BITS=64 NATIVE_HOST_LIMBS.
//
//   dropin_test          host-only checks (size policy, host CRT)
//   dropin_test --gpu    adds the device checks (needs a GPU)
#define BW_DEFINE_TABLES
#include "config.h"
#include "crt.h"
#include "crt_gpu.h"
#include "gpu_ntt.h"
#include "multiply.h"
#include "ntt_limits.h"
#include "zero_pad.h"

#include <gmp.h>

#include <cstring>
#include <unistd.h>
#include <random>
#include <string>

static int g_fail = 0;
#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);    \
            g_fail++;                                                            \
        }                                                                        \
    } while (0)

template <class T>
static std::vector<T> gmp_product(const std::vector<T>& A, const std::vector<T>& B) {
    mpz_t a, b, c;
    mpz_inits(a, b, c, nullptr);
    mpz_import(a, A.size(), -1, sizeof(T), 0, 0, A.data());
    mpz_import(b, B.size(), -1, sizeof(T), 0, 0, B.data());
    mpz_mul(c, a, b);
    std::vector<T> C(A.size() + B.size(), 0);
    size_t cnt = 0;
    mpz_export(C.data(), &cnt, -1, sizeof(T), 0, 0, c);
    mpz_clears(a, b, c, nullptr);
    return C;
}

template <class T>
static std::vector<T> random_limbs(size_t n, std::mt19937_64& rng) {
    std::vector<T> v(n);
    for (auto& x : v) x = (T)rng();
    return v;
}

#ifndef BW_EXTENDED_SIZES
// The reference's size-policy expectations for each configuration
// (tests/test_ntt_limits.cpp:140-311: configured limits, CRT bound, multiply
// limits, accepted and rejected N, the benchmark's L_arg convention), checked
// through the drop-in headers.  max_log: log2 of the largest N; crt_ok /
// crt_bad: log2 of square sizes inside / beyond the CRT bound; larg_ok /
// larg_bad: L_arg values (2^L limbs) inside / beyond the multiply limits.
struct RefPolicy { int max_log, crt_ok, crt_bad, larg_ok, larg_bad; };
#if defined(NATIVE_HOST_LIMBS)
static constexpr RefPolicy kRef = {32, 31, 50, 30, 32};
#elif LIMB_BITS == 64
static constexpr RefPolicy kRef = {33, 32, 53, 32, 33};
#else
static constexpr RefPolicy kRef = {23, 23, 24, 22, 23};
#endif

static void reference_policy_checks() {
    const size_t one = 1;
    CHECK(max_supported_logN() == kRef.max_log);
    CHECK(max_supported_N() == one << kRef.max_log);
    CHECK(max_supported_limb_count() == one << (kRef.max_log - 1));
    std::string why;
    CHECK(crt_coefficient_bound_satisfied(one << 22, one << 22, &why));
    CHECK(crt_coefficient_bound_satisfied(one << kRef.crt_ok, one << kRef.crt_ok, &why));
    why.clear();
    CHECK(!crt_coefficient_bound_satisfied(one << kRef.crt_bad, one << kRef.crt_bad, &why));
    CHECK(why.find("CRT modulus product") != std::string::npos);
    CHECK(crt_coefficient_bound_satisfied(one << kRef.crt_bad, 1, &why));   // the short operand bounds the terms
    CHECK(multiply_size_supported(one << 22, one << 22, &why));
    CHECK(!multiply_size_supported(one << kRef.max_log, one << kRef.max_log, &why));   // needs N = 2^(max_log+1)
    for (int logN = 1; logN <= kRef.max_log; logN++) CHECK(ntt_size_supported(one << logN, &why));
    CHECK(ntt_size_supported(padded_ntt_size(one << (kRef.max_log - 1), one << (kRef.max_log - 1)), &why));
    CHECK(!ntt_size_supported(0, &why));
    CHECK(!ntt_size_supported(3, &why));
    why.clear();
    CHECK(!ntt_size_supported(one << (kRef.max_log + 1), &why));
    CHECK(why.find(std::to_string(kRef.max_log + 1)) != std::string::npos);
    CHECK(!ntt_size_supported(padded_ntt_size(one << kRef.max_log, one << kRef.max_log), &why));
    CHECK(multiply_size_supported(one << kRef.larg_ok, one << kRef.larg_ok, &why));
    CHECK(!multiply_size_supported(one << kRef.larg_bad, one << kRef.larg_bad, &why));
}
#else
// BW_EXTENDED_SIZES: the library's automatic choice (32-bit limbs: three
// primes up to N = 2^26, limb pairs under the 58-bit primes beyond).
static void extended_policy_checks() {
    const size_t one = 1;
    std::string why;
    CHECK(max_supported_logN() == (BW_CFG_LIMB_BITS == 32 ? 33 : 32));
    CHECK(multiply_size_supported(one << 23, one << 23, &why));
    CHECK(multiply_size_supported(one << 25, one << 25, &why));
    CHECK(multiply_size_supported((one << 25) + 5, (one << 25) + 5, &why));
    CHECK(ntt_size_supported(one << 24, &why) && ntt_size_supported(one << 30, &why));
    CHECK(!multiply_size_supported(0, 5, &why));
}
#endif

static void host_checks() {
#ifndef BW_EXTENDED_SIZES
    reference_policy_checks();
#else
    extended_policy_checks();
#endif
    // per-prime limits of the configured moduli
    for (size_t i = 0; i < moduli.size(); i++) {
        const int v = max_logN_for_prime(moduli[i]);
        CHECK(v >= 23);
        const int r = max_root_logN(roots_of_unity_max[i], moduli[i]);
        CHECK(r >= 23 && r <= v);
    }
    CHECK(padded_ntt_size(1, 1) == 1);
    CHECK(padded_ntt_size(1024, 1024) == 2048);
    CHECK(padded_ntt_size(3, 5) == 8);
    CHECK(!ntt_size_supported(3));
    CHECK(ntt_size_supported(1024));
    std::string why;
    CHECK(!multiply_size_supported(0, 5, &why) && !why.empty());
    CHECK(multiply_size_supported(1000, 1000));
    CHECK(multiply_size_supported(max_supported_limb_count(), max_supported_limb_count()));
    CHECK(!multiply_size_supported(max_supported_limb_count() + 1, max_supported_limb_count() + 1));
    CHECK(max_supported_N() == (size_t(1) << max_supported_logN()));
    // host CRT
    CHECK(modinv_u64(3, 7) == 5);
    CHECK(modinv_u64(0x2d000001ULL % 0x23800001ULL, 0x23800001ULL) * (0x2d000001ULL % 0x23800001ULL) % 0x23800001ULL == 1);
    std::mt19937_64 rng(7);
    if (moduli.size() >= 2) {
        CRT2Params c2(moduli[0], moduli[1]);
        for (int t = 0; t < 1000; t++) {
            const unsigned __int128 x = (((unsigned __int128)rng() << 64) | rng()) % c2.modulus;
            const TestDataTypeUint a = (TestDataTypeUint)(x % moduli[0]), b = (TestDataTypeUint)(x % moduli[1]);
            CHECK(crt_combine_2(c2, a, b) == x);
            CHECK(crt_combine_many({moduli[0], moduli[1]}, {a, b}) == x);
        }
    }
    const CRTGarnerParams g = compute_garner_params(moduli);
    for (int i = 1; i < NUM_MODULI; i++)
        CHECK((unsigned __int128)g.inv[i] * g.prefix_M[i] % g.primes[i] == 1);
    // precompute cache: save the host tables, load them back (host only)
    {
        char dir[] = "/tmp/bw_dropin_XXXXXX";
        CHECK(mkdtemp(dir) != nullptr);
        const NTTPrecomputed pre = precompute_ntt(4096);
        CHECK(save_ntt_precomputed(pre, dir));
        PrecomputeTiming t{};
        const NTTPrecomputed back = load_ntt_precomputed(4096, dir, &t);
        CHECK(back.prime_set == pre.prime_set && back.primes == pre.primes && back.N == 4096);
        for (int set : {pre.prime_set, (int)BW_SET_D64}) {
            const std::string path = std::string(dir) + "/bw_tables_" + std::to_string(set) + ".bin";
            int got = -1;
            CHECK(bw_tables_load(path.c_str(), &got) == BW_OK && got == set);
            std::remove(path.c_str());
        }
        rmdir(dir);
    }
}

static void hip_ok(hipError_t e) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error %s\n", hipGetErrorString(e));
        std::exit(2);
    }
}

static void gpu_checks() {
    std::mt19937_64 rng(11);
    const size_t sizes[][2] = {{1, 1}, {3, 2}, {1000, 1003}, {4096, 4096}, {20000, 777}};
    for (auto& sz : sizes) {
        auto A = random_limbs<InputLimbType>(sz[0], rng), B = random_limbs<InputLimbType>(sz[1], rng);
        std::vector<OutputLimbType> C;
        std::chrono::duration<double, std::milli> d{};
#if defined(NATIVE_HOST_LIMBS)
        host_multiply_merge_native(A, B, C, d);
#else
        host_multiply_merge(A, B, C, d);
#endif
        CHECK(C == gmp_product(A, B));
        CHECK(d.count() >= 0.0);
    }
    // precompute / upload / allocate / execute, pinned host inputs
    {
        const size_t LA = 5000, LB = 3001, N = padded_ntt_size(LA, LB);
        auto A = random_limbs<InputLimbType>(LA, rng), B = random_limbs<InputLimbType>(LB, rng);
        InputLimbType *ha = nullptr, *hb = nullptr;
        hip_ok(hipHostMalloc((void**)&ha, LA * sizeof(InputLimbType), 0));
        hip_ok(hipHostMalloc((void**)&hb, LB * sizeof(InputLimbType), 0));
        std::memcpy(ha, A.data(), LA * sizeof(InputLimbType));
        std::memcpy(hb, B.data(), LB * sizeof(InputLimbType));
        PrecomputeTiming pt;
        NTTPrecomputed pre = precompute_ntt(N, &pt);
        CHECK(pre.N == N);
        SetupUploadTiming ut;
        upload_ntt_precomputed(pre, &ut);
        NTTContext ctx = allocate_ntt_context(pre, LA, LB);
        const auto want = gmp_product(A, B);
        std::vector<OutputLimbType> C(N + 1, 0xAB);
        execute_ntt_multiply(ctx, ha, hb, C);
        CHECK(C.size() == N + 1);
        CHECK(std::equal(want.begin(), want.end(), C.begin()));
        for (size_t i = want.size(); i < C.size(); i++) CHECK(C[i] == 0);
        NTTTiming t;
        std::vector<OutputLimbType> C2(N + 1);
        execute_ntt_multiply(ctx, ha, hb, C2, &t);
        CHECK(C2 == C);
        CHECK(t.total_ms > 0 && t.fwd_pad_ntt_ms > 0 && t.intt_ms >= 0 && t.carry_ms > 0);
        // early exit after Garner, then the carry wrapper on the coefficients
        uint64_t *d0 = nullptr, *d1 = nullptr;
        OutputLimbType* dout = nullptr;
        hip_ok(hipMalloc((void**)&d0, N * 8));
        hip_ok(hipMalloc((void**)&d1, N * 8));
        hip_ok(hipMalloc((void**)&dout, (LA + LB) * sizeof(OutputLimbType)));
        std::vector<OutputLimbType> C3(LA + LB);
#if defined(NATIVE_HOST_LIMBS)
        std::vector<uint64_t> lo, mid;
        std::vector<uint32_t> hi32;
        std::vector<OutputLimbType> dummy;
        execute_ntt_multiply(ctx, ha, hb, dummy, nullptr, nullptr, &lo, &mid, &hi32);
        CHECK(lo.size() == N && mid.size() == N && hi32.size() == N);
        uint32_t* d2 = nullptr;
        hip_ok(hipMalloc((void**)&d2, N * 4));
        hip_ok(hipMemcpy(d0, lo.data(), N * 8, hipMemcpyHostToDevice));
        hip_ok(hipMemcpy(d1, mid.data(), N * 8, hipMemcpyHostToDevice));
        hip_ok(hipMemcpy(d2, hi32.data(), N * 4, hipMemcpyHostToDevice));
        carry_propagate_gpu_u160(d0, d1, d2, N, dout, LA + LB);
        hip_ok(hipDeviceSynchronize());
        (void)hipFree(d2);
#else
        std::vector<uint64_t> hi, lo;
        std::vector<OutputLimbType> dummy;
        execute_ntt_multiply(ctx, ha, hb, dummy, nullptr, &hi, &lo);
        CHECK(hi.size() == N && lo.size() == N);
        hip_ok(hipMemcpy(d0, hi.data(), N * 8, hipMemcpyHostToDevice));
        hip_ok(hipMemcpy(d1, lo.data(), N * 8, hipMemcpyHostToDevice));
        carry_propagate_gpu(d0, d1, N, dout, LA + LB);
        hip_ok(hipDeviceSynchronize());
#endif
        hip_ok(hipMemcpy(C3.data(), dout, (LA + LB) * sizeof(OutputLimbType), hipMemcpyDeviceToHost));
        CHECK(C3 == want);
        (void)hipFree(d0); (void)hipFree(d1); (void)hipFree(dout);
        cleanup_ntt_context(ctx);
        cleanup_ntt_precomputed(pre);
        CHECK(!pre.gpu_uploaded);
        (void)hipHostFree(ha); (void)hipHostFree(hb);
    }
    // zero padding
    {
        const size_t L = 1000, N = 2048;
        std::vector<uint32_t> src(L);
        for (auto& x : src) x = (uint32_t)rng();
        uint32_t* ds = nullptr;
        TestDataTypeUint* dd = nullptr;
        hip_ok(hipMalloc((void**)&ds, L * 4));
        hip_ok(hipMalloc((void**)&dd, N * sizeof(TestDataTypeUint)));
        hip_ok(hipMemcpy(ds, src.data(), L * 4, hipMemcpyHostToDevice));
        zero_pad_gpu(ds, dd, L, N, nullptr);
        std::vector<TestDataTypeUint> out(N);
        hip_ok(hipMemcpy(out.data(), dd, N * sizeof(TestDataTypeUint), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) CHECK(out[i] == (i < L ? (TestDataTypeUint)src[i] : 0));
        (void)hipFree(ds); (void)hipFree(dd);
    }
    // Garner on the device, residues of random coefficients below the moduli product
    {
        const int N = 4097;
        std::vector<std::vector<TestDataTypeUint>> res(NUM_MODULI, std::vector<TestDataTypeUint>(N));
        std::vector<unsigned __int128> x128(N);
        std::vector<uint64_t> xl(N), xm(N);
        std::vector<uint32_t> xh(N);
#if defined(NATIVE_HOST_LIMBS)
        // coefficients below 2^160 (the U160 output width), residues derived from them
        std::vector<uint64_t> wl(N), wm(N);
        std::vector<uint32_t> wh(N);
        for (int i = 0; i < N; i++) {
            wl[i] = rng();
            wm[i] = rng();
            wh[i] = (uint32_t)rng();
            for (int j = 0; j < 3; j++) {
                const unsigned __int128 p = moduli[j];
                const unsigned __int128 t64 = ((unsigned __int128)1 << 64) % p;
                res[j][i] = (TestDataTypeUint)(((unsigned __int128)wl[i] % p + (unsigned __int128)wm[i] % p * t64 % p +
                                                (unsigned __int128)wh[i] * (t64 * t64 % p) % p) % p);
            }
        }
#else
        for (int i = 0; i < N; i++)
            for (int j = 0; j < NUM_MODULI; j++) res[j][i] = (TestDataTypeUint)(rng() % moduli[j]);
#endif
        std::vector<TestDataTypeUint*> dres(NUM_MODULI);
        for (int j = 0; j < NUM_MODULI; j++) {
            hip_ok(hipMalloc((void**)&dres[j], N * sizeof(TestDataTypeUint)));
            hip_ok(hipMemcpy(dres[j], res[j].data(), N * sizeof(TestDataTypeUint), hipMemcpyHostToDevice));
        }
        upload_garner_params(compute_garner_params(moduli));
        upload_residue_ptrs(dres);
        uint64_t *d0 = nullptr, *d1 = nullptr;
        hip_ok(hipMalloc((void**)&d0, N * 8));
        hip_ok(hipMalloc((void**)&d1, N * 8));
#if defined(NATIVE_HOST_LIMBS)
        uint32_t* d2 = nullptr;
        hip_ok(hipMalloc((void**)&d2, N * 4));
        crt_combine_gpu_u160(d0, d1, d2, N);
        hip_ok(hipMemcpy(xl.data(), d0, N * 8, hipMemcpyDeviceToHost));
        hip_ok(hipMemcpy(xm.data(), d1, N * 8, hipMemcpyDeviceToHost));
        hip_ok(hipMemcpy(xh.data(), d2, N * 4, hipMemcpyDeviceToHost));
        (void)hipFree(d2);
        for (int i = 0; i < N; i++) CHECK(xl[i] == wl[i] && xm[i] == wm[i] && xh[i] == wh[i]);
#else
        crt_combine_gpu(d1, d0, N);
        hip_ok(hipMemcpy(xl.data(), d0, N * 8, hipMemcpyDeviceToHost));
        hip_ok(hipMemcpy(xm.data(), d1, N * 8, hipMemcpyDeviceToHost));
        for (int i = 0; i < N; i++) {
            const unsigned __int128 x = ((unsigned __int128)xm[i] << 64) | xl[i];
            std::vector<TestDataTypeUint> r;
            for (int j = 0; j < NUM_MODULI; j++) r.push_back(res[j][i]);
            CHECK(x == crt_combine_many(moduli, r));
        }
#endif
        (void)hipFree(d0); (void)hipFree(d1);
        for (auto p : dres) (void)hipFree(p);
    }
}

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    if (bw_abi_version() != BW_ABI_VERSION) {   // built against another header than the library's
        std::fprintf(stderr, "ABI mismatch: header %d, library %d (rebuild tests/cpp)\n", BW_ABI_VERSION,
                     bw_abi_version());
        return 3;
    }
    host_checks();
    if (gpu) gpu_checks();
    std::printf("dropin_test LIMB_BITS=%d native=%d: %s (%d failures)\n", LIMB_BITS,
#if defined(NATIVE_HOST_LIMBS)
                1,
#else
                0,
#endif
                g_fail ? "FAIL" : "ok", g_fail);
    return g_fail ? 1 : 0;
}
