// bench_full_multiply.cpp -- the reference's full-multiply benchmark front end
// (bench/gpu_full_multiply_benchmark.cpp) over this library's drop-in C++ API.
//
// Same command line and CSV rows, so the reference's runner and plotting
// scripts read its output unchanged:
//
//   bench_full_multiply_<pipeline> [--warmup N] [--iters N] [--csv FILE] [--append] [L ...]
//
// L < 64 means 2^L limbs, larger values are limb counts, "a-b" is the inclusive
// range a, a+1, ..., b (bench/gpu_full_multiply_benchmark.cpp:243-278); the
// default sweep is 6..22.  Each L runs the reference's sequence per size --
// pinned host buffers, precompute_ntt, upload_ntt_precomputed,
// allocate_ntt_context, `warmup` untimed and `iters` timed
// execute_ntt_multiply calls, teardown -- and appends one CSV row with the
// execute statistics (mean, population standard deviation, min, max), the
// setup/teardown buckets and the per-stage means (column list:
// bench/gpu_full_multiply_benchmark.cpp:384-404).
//
// --list prints the resolved sizes (and which are unsupported) and exits
// without touching the GPU.
//
// Built three ways like the reference (config.h): pipeline "32" (32-bit limbs),
// "hybrid" (-DLIMB_BITS=64: 32-bit limbs), "64bit" (-DLIMB_BITS=64
// -DNATIVE_HOST_LIMBS: 64-bit limbs).
#define BW_DEFINE_TABLES
#include "config.h"
#include "gpu_ntt.h"
#include "ntt_limits.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

namespace {

const char* pipeline_name() {
#if defined(NATIVE_HOST_LIMBS)
    return "64bit";
#elif LIMB_BITS == 64
    return "hybrid";
#else
    return "32";
#endif
}

// Synthetic operands of the reference benchmark (gpu_full_multiply_benchmark.cpp:68-108):
// every eighth limb 0, the next 1, 64-bit limbs also all-ones, the rest random
// (below 2^30 on the 32-bit-limb pipelines).
std::vector<InputLimbType> bench_limbs(size_t n, uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::vector<InputLimbType> v(n);
    for (size_t i = 0; i < n; i++) {
        switch (i % 8) {
        case 0: v[i] = 0; break;
        case 1: v[i] = 1; break;
#if defined(NATIVE_HOST_LIMBS)
        case 2: v[i] = ~(uint64_t)0; break;
        default: v[i] = rng(); break;
#else
        default: v[i] = (uint32_t)(rng() % (1ULL << 30)); break;
#endif
        }
    }
    return v;
}

struct Stats {
    double mean = 0, sd = 0, min = 0, max = 0;
};

Stats stats_of(const std::vector<double>& x) {
    Stats s;
    if (x.empty()) return s;
    double sum = 0, sq = 0;
    s.min = s.max = x[0];
    for (double v : x) {
        sum += v;
        sq += v * v;
        s.min = std::min(s.min, v);
        s.max = std::max(s.max, v);
    }
    const double n = (double)x.size();
    s.mean = sum / n;
    s.sd = std::sqrt(std::max(0.0, sq / n - s.mean * s.mean));
    return s;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Row {
    size_t L_arg = 0, L = 0, N = 0;
    int logN = 0, warmup = 0, iters = 0;
    double pinned = 0, precompute = 0, upload = 0, alloc = 0;
    PrecomputeTiming pre{};
    SetupUploadTiming up{};
    double free_ctx = 0, free_pre = 0, free_pinned = 0;
    Stats total;
    // per-stage means
    double ingress_fwd = 0, h2d = 0, fwd = 0, fwd_a = 0, fwd_b = 0, mul = 0, intt = 0, crt = 0, carry = 0, d2h = 0;
};

size_t resolve(size_t L_arg) { return L_arg < 64 ? (size_t)1 << L_arg : L_arg; }

bool parse_count(const std::string& s, size_t& out) {
    if (s.empty() || s.find_first_not_of("0123456789") != std::string::npos) return false;
    out = (size_t)std::strtoull(s.c_str(), nullptr, 10);
    return out != 0;
}

// one L spec: a number or an inclusive range "lo-hi"
bool add_spec(const std::string& spec, std::vector<size_t>& out) {
    const size_t dash = spec.find('-');
    size_t lo = 0, hi = 0;
    if (dash == std::string::npos) {
        if (!parse_count(spec, lo)) return false;
        out.push_back(lo);
        return true;
    }
    if (!parse_count(spec.substr(0, dash), lo) || !parse_count(spec.substr(dash + 1), hi) || lo > hi) return false;
    for (size_t v = lo; v <= hi; v++) out.push_back(v);
    return true;
}

Row run_size(size_t L_arg, int warmup, int iters, uint64_t seed) {
    Row r;
    r.L_arg = L_arg;
    r.L = resolve(L_arg);
    r.N = padded_ntt_size(r.L, r.L);
    while (((size_t)1 << r.logN) < r.N) r.logN++;
    r.warmup = warmup;
    r.iters = iters;
    const std::vector<InputLimbType> A = bench_limbs(r.L, seed), B = bench_limbs(r.L, seed + 1);

    InputLimbType *a = nullptr, *b = nullptr;
    auto t = std::chrono::steady_clock::now();
    bw_dropin::hip_check(hipHostMalloc((void**)&a, r.L * sizeof(InputLimbType), hipHostMallocDefault), "hipHostMalloc");
    bw_dropin::hip_check(hipHostMalloc((void**)&b, r.L * sizeof(InputLimbType), hipHostMallocDefault), "hipHostMalloc");
    std::memcpy(a, A.data(), r.L * sizeof(InputLimbType));
    std::memcpy(b, B.data(), r.L * sizeof(InputLimbType));
    r.pinned = ms_since(t);

    NTTPrecomputed pre = precompute_ntt(r.N, &r.pre);
    r.precompute = r.pre.total_ms;
    upload_ntt_precomputed(pre, &r.up);
    r.upload = r.up.total_ms;
    t = std::chrono::steady_clock::now();
    NTTContext ctx = allocate_ntt_context(pre, r.L, r.L);
    r.alloc = ms_since(t);

    std::vector<OutputLimbType> C(r.N + 1, 0);
    for (int i = 0; i < warmup; i++) execute_ntt_multiply(ctx, a, b, C);
    std::vector<double> total, ingress, h2d, fwd, fwd_a, fwd_b, mul, intt, crt, carry, d2h;
    for (int i = 0; i < iters; i++) {
        NTTTiming tm{};
        execute_ntt_multiply(ctx, a, b, C, &tm);
        total.push_back(tm.total_ms);
        ingress.push_back(tm.ingress_fwd_ms);
        h2d.push_back(tm.h2d_ms);
        fwd.push_back(tm.fwd_pad_ntt_ms);
        fwd_a.push_back(tm.fwd_pad_ntt_a_ms);
        fwd_b.push_back(tm.fwd_pad_ntt_b_ms);
        mul.push_back(tm.pointwise_mul_ms);
        intt.push_back(tm.intt_ms);
        crt.push_back(tm.crt_ms);
        carry.push_back(tm.carry_ms);
        d2h.push_back(tm.d2h_ms);
    }
    r.total = stats_of(total);
    r.ingress_fwd = stats_of(ingress).mean;
    r.h2d = stats_of(h2d).mean;
    r.fwd = stats_of(fwd).mean;
    r.fwd_a = stats_of(fwd_a).mean;
    r.fwd_b = stats_of(fwd_b).mean;
    r.mul = stats_of(mul).mean;
    r.intt = stats_of(intt).mean;
    r.crt = stats_of(crt).mean;
    r.carry = stats_of(carry).mean;
    r.d2h = stats_of(d2h).mean;

    t = std::chrono::steady_clock::now();
    cleanup_ntt_context(ctx);
    r.free_ctx = ms_since(t);
    t = std::chrono::steady_clock::now();
    cleanup_ntt_precomputed(pre);
    r.free_pre = ms_since(t);
    t = std::chrono::steady_clock::now();
    (void)hipHostFree(a);
    (void)hipHostFree(b);
    r.free_pinned = ms_since(t);
    return r;
}

const char* kCsvHeader =
    "pipeline,host_limb_bits,operand_bits,L_arg,L,N,logN,warmup,iters,"
    "mean_ms,stddev_ms,min_ms,max_ms,"
    "setup_pinned_ms,setup_precompute_ms,setup_upload_ms,setup_alloc_ms,"
    "pre_factors_ms,pre_params_ms,pre_twiddle_host_ms,pre_garner_host_ms,"
    "upload_twiddle_ms,upload_mod_constants_ms,upload_garner_ms,"
    "teardown_free_ctx_ms,teardown_free_pre_ms,teardown_free_pinned_ms,"
    "ingress_fwd_mean_ms,h2d_mean_ms,fwd_pad_ntt_mean_ms,fwd_pad_ntt_a_mean_ms,fwd_pad_ntt_b_mean_ms,"
    "mul_mean_ms,intt_mean_ms,crt_mean_ms,carry_mean_ms,d2h_mean_ms\n";

bool write_csv(const std::string& path, const std::vector<Row>& rows, bool append) {
    bool empty = true;
    {
        std::ifstream probe(path);
        empty = !probe.good() || probe.peek() == std::ifstream::traits_type::eof();
    }
    FILE* f = std::fopen(path.c_str(), append ? "a" : "w");
    if (!f) return false;
    if (!append || empty) std::fputs(kCsvHeader, f);
    for (const Row& r : rows) {
        std::fprintf(f, "%s,%d,%zu,%zu,%zu,%zu,%d,%d,%d,", pipeline_name(), INPUT_LIMB_BITS,
                     r.L * (size_t)INPUT_LIMB_BITS, r.L_arg, r.L, r.N, r.logN, r.warmup, r.iters);
        const double v[] = {r.total.mean, r.total.sd, r.total.min, r.total.max,
                            r.pinned, r.precompute, r.upload, r.alloc,
                            r.pre.factors_ms, r.pre.params_ms, r.pre.twiddle_host_ms, r.pre.garner_host_ms,
                            r.up.twiddle_upload_ms, r.up.mod_constants_ms, r.up.garner_upload_ms,
                            r.free_ctx, r.free_pre, r.free_pinned,
                            r.ingress_fwd, r.h2d, r.fwd, r.fwd_a, r.fwd_b, r.mul, r.intt, r.crt, r.carry, r.d2h};
        const int n = (int)(sizeof(v) / sizeof(v[0]));
        for (int i = 0; i < n; i++) std::fprintf(f, "%.6f%s", v[i], i + 1 < n ? "," : "\n");
    }
    return std::fclose(f) == 0;
}

void print_row(const Row& r) {
    std::printf("L_arg=%3zu  L=%10zu  N=%10zu  logN=%2d\n", r.L_arg, r.L, r.N, r.logN);
    std::printf("  setup:   pinned=%8.3f  precompute=%8.3f  upload=%8.3f  alloc=%8.3f  (total %.3f ms)\n", r.pinned,
                r.precompute, r.upload, r.alloc, r.pinned + r.precompute + r.upload + r.alloc);
    std::printf("  execute: mean=%8.3f  stddev=%7.3f  min=%8.3f  max=%8.3f ms\n", r.total.mean, r.total.sd, r.total.min,
                r.total.max);
    std::printf("    ingress_fwd=%.3f (h2d=%.3f fwd=%.3f)  mul=%.3f  intt=%.3f  crt=%.3f  carry=%.3f  d2h=%.3f ms\n",
                r.ingress_fwd, r.h2d, r.fwd, r.mul, r.intt, r.crt, r.carry, r.d2h);
    std::printf("  teardown: free_ctx=%8.3f  free_pre=%8.3f  free_pinned=%8.3f  (total %.3f ms)\n", r.free_ctx,
                r.free_pre, r.free_pinned, r.free_ctx + r.free_pre + r.free_pinned);
}

void usage(const char* prog) {
    std::fprintf(stderr,
                 "Usage: %s [--warmup N] [--iters N] [--csv FILE] [--append] [--list] [L ...]\n"
                 "  --warmup N   warmup iterations per L (default 2)\n"
                 "  --iters N    timed iterations per L (default 20)\n"
                 "  --csv FILE   output CSV path (default gpu_multiply_bench.csv)\n"
                 "  --append     append rows to CSV instead of overwriting\n"
                 "  --list       print the resolved sizes and exit (no GPU work)\n"
                 "  L ...        limb spec: if L < 64, use 1<<L limbs; else L limbs\n"
                 "               ranges inclusive: 16-24 -> 16,17,...,24\n"
                 "               (default sweep: log2 sizes 6..22)\n",
                 prog);
}

}  // namespace

int main(int argc, char** argv) {
    int warmup = 2, iters = 20;
    std::string csv = "gpu_multiply_bench.csv";
    bool append = false, list_only = false;
    std::vector<size_t> L_args;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto value = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "Missing value for %s\n", what);
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--help" || a == "-h") { usage(argv[0]); return 0; }
        if (a == "--warmup") { warmup = std::atoi(value("--warmup")); continue; }
        if (a == "--iters") { iters = std::atoi(value("--iters")); continue; }
        if (a == "--csv") { csv = value("--csv"); continue; }
        if (a == "--append") { append = true; continue; }
        if (a == "--list") { list_only = true; continue; }
        if (!add_spec(a, L_args)) {
            std::fprintf(stderr, "Invalid limb spec: %s\n", a.c_str());
            usage(argv[0]);
            return 1;
        }
    }
    if (L_args.empty())
        for (size_t l = 6; l <= 22; l++) L_args.push_back(l);
    if (warmup < 0 || iters <= 0) {
        std::fprintf(stderr, "warmup must be >= 0 and iters must be > 0\n");
        return 1;
    }
    std::printf("GPU full multiply benchmark (pipeline=%s, host_limb_bits=%d, warmup=%d, iters=%d, "
                "max L_arg=%d / max L=%zu)\n",
                pipeline_name(), INPUT_LIMB_BITS, warmup, iters, max_supported_logN() - 1,
                max_supported_limb_count());
    std::printf("%s\n", std::string(72, '-').c_str());

    std::vector<size_t> run;
    for (size_t L_arg : L_args) {
        const size_t L = resolve(L_arg);
        std::string why;
        if (!multiply_size_supported(L, L, &why)) {
            std::fprintf(stderr, "Skipping L_arg=%zu (L=%zu, N=%zu): %s\n", L_arg, L, padded_ntt_size(L, L),
                         why.c_str());
            continue;
        }
        run.push_back(L_arg);
        if (list_only) std::printf("L_arg=%zu L=%zu N=%zu\n", L_arg, L, padded_ntt_size(L, L));
    }
    if (list_only) return 0;

    {   // the first pinned allocation of a process pays for runtime start-up
        void *h = nullptr, *d = nullptr;
        bw_dropin::hip_check(hipHostMalloc(&h, 4096, hipHostMallocDefault), "hipHostMalloc");
        bw_dropin::hip_check(hipMalloc(&d, 4096), "hipMalloc");
        (void)hipFree(d);
        (void)hipHostFree(h);
    }
    std::vector<Row> rows;
    uint64_t seed = 1234;
    for (size_t L_arg : run) {
        std::printf("Benchmarking L_arg=%zu (L=%zu) ... ", L_arg, resolve(L_arg));
        std::fflush(stdout);
        rows.push_back(run_size(L_arg, warmup, iters, seed));
        seed += 17;
        std::printf("done\n");
        print_row(rows.back());
    }
    if (!write_csv(csv, rows, append)) {
        std::fprintf(stderr, "Failed to open CSV: %s\n", csv.c_str());
        return 1;
    }
    std::printf("%s\nWrote %zu rows to %s\n", std::string(72, '-').c_str(), rows.size(), csv.c_str());
    return 0;
}
