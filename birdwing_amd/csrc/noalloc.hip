// noalloc.hip -- bw_multiply_device_noalloc: one product on caller-owned
// device buffers with no device allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation).
//
// It runs the same kernels as a plan of the five-prime 64-bit-digit set
// (forward passes, the fused pass, inverse passes, Garner + carry) but keeps
// everything a plan allocates -- the two residue buffers, the carry tile
// flags, the tile counter and error word, the N^-1 pairs -- in `scratch`, a
// file-scope device array of fixed size that each GPU holds until the process
// ends.  Its size bounds the products: transforms of at most
// 2^BW_NOALLOC_MAX_LOGN 64-bit digits.
//
// Host code never uses the array's name as an address: bw_noalloc_setup()
// (called by bw_init for the device) asks the runtime for the array's device
// address on that device with hipGetSymbolAddress and keeps it per device;
// every kernel argument and copy below uses that pointer.  Calls on one GPU
// share its array, so calls that overlap on one GPU (different streams, or
// different host threads) race; calls on one stream are ordered.
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "batch.h"
#include "engine.h"

#define BW_NOALLOC_MAX_LOGN 18          // 64-bit digits per transform
#define BW_NOALLOC_MAX_DEVICES 64

namespace {
constexpr int kPrimes = 5;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
// one operand's residue planes: five primes x 2^18 four-byte words
constexpr size_t kPlaneBytes = (size_t)kPrimes << BW_NOALLOC_MAX_LOGN << 2;
constexpr size_t kOffA = 0;
constexpr size_t kOffB = kPlaneBytes;
// carry tiles of the largest product: 2^18 + 1 output words
constexpr int64_t kMaxTiles = (((int64_t)1 << BW_NOALLOC_MAX_LOGN) + 1 + BW_CARRY_TILE - 1) / BW_CARRY_TILE;
constexpr size_t kOffFlags = 2 * kPlaneBytes;
constexpr size_t kOffCounter = align_up(kOffFlags + 8 * (size_t)kMaxTiles);   // [0] tile ids, [1] error word
// (N^-1, N^-1 / p) per prime for each log2 N in [0, BW_NOALLOC_MAX_LOGN]
constexpr size_t kNinvEntry = 16 * kPrimes;
constexpr size_t kOffNinv = align_up(kOffCounter + 2 * sizeof(int));
constexpr size_t kScratchBytes = align_up(kOffNinv + kNinvEntry * (BW_NOALLOC_MAX_LOGN + 1));
static_assert(kScratchBytes < (11u << 20), "scratch stays near 10 MiB");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_noalloc_mu;
static unsigned char* g_scratch_ptr[BW_NOALLOC_MAX_DEVICES];   // device address of `scratch` per device

namespace bw {

// Once per device, with that device current: find the device address of
// `scratch` and fill its N^-1 table (set-up; allocates nothing).
int noalloc_setup(int device) {
    if (device < 0 || device >= BW_NOALLOC_MAX_DEVICES) return fail(BW_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_scratch_ptr[device]) return BW_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(BW_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    const PrimeSet* s = prime_set(SET_D64);
    std::vector<double> tab(2 * kPrimes * (BW_NOALLOC_MAX_LOGN + 1), 0.0);
    for (int lg = 1; lg <= BW_NOALLOC_MAX_LOGN; lg++) {
        std::vector<uint64_t> n64;
        std::vector<double> nf;
        ninv_pairs(*s, (uint64_t)1 << lg, n64, nf);
        for (int i = 0; i < 2 * kPrimes; i++) tab[(size_t)lg * 2 * kPrimes + i] = nf[i];
    }
    unsigned char* base = static_cast<unsigned char*>(ptr);
    HIP_TRY(hipMemcpy(base + kOffNinv, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(base + kOffCounter, 0, 2 * sizeof(int)));
    g_scratch_ptr[device] = base;
    return BW_OK;
}

}  // namespace bw

using namespace bw;

extern "C" int bw_multiply_device_noalloc(int limb_bits, const void* d_a, uint64_t L_A, const void* d_b, uint64_t L_B,
                                          void* d_c, int* d_status, void* stream) {
    if (limb_bits != 32 && limb_bits != 64) return fail(BW_ERR_ARG, "limb_bits must be 32 or 64");
    if (L_A == 0 || L_B == 0) return fail(BW_ERR_UNSUPPORTED_SIZE, "operand limb count must be positive");
    if (!d_a || !d_b || !d_c) return fail(BW_ERR_ARG, "NULL device buffer");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        if (dev >= 0 && dev < BW_NOALLOC_MAX_DEVICES) base = g_scratch_ptr[dev];
    }
    if (!base) return fail(BW_ERR_STATE, "bw_init(%d) has not been called: the scratch of device %d is not set up", dev, dev);
    const PrimeSet* s = prime_set(SET_D64);
    std::string why;
    if (!multiply_size_supported(*s, limb_bits, L_A, L_B, &why))
        return fail(BW_ERR_UNSUPPORTED_SIZE, "unsupported multiply: %s", why.c_str());
    const uint64_t N = padded_size_for(*s, limb_bits, L_A, L_B);
    const uint64_t N_exec = N < 2 ? 2 : N;
    const int logN = valuation2(N_exec);
    if (logN > BW_NOALLOC_MAX_LOGN)
        return fail(BW_ERR_UNSUPPORTED_SIZE,
                    "product of %llu x %llu %d-bit limbs needs transforms of 2^%d digits; the allocation-free "
                    "entry point holds at most 2^%d (%zu bytes of scratch)",
                    (unsigned long long)L_A, (unsigned long long)L_B, limb_bits, logN, BW_NOALLOC_MAX_LOGN,
                    kScratchBytes);
    DevTables* tabs = nullptr;
    int rc = find_tables(SET_D64, &tabs);
    if (rc) return rc;

    Xform x;
    x.limb_bits = s->limb_bits;
    x.ingress = 2;
    x.set = s;
    x.tabs = tabs;
    x.logN = logN;
    x.N = (int64_t)N_exec;
    x.B = 1;
    TuneChoice c;
    if (tune_lookup(TuneKey{SET_D64, logN, 0}, &c)) {
        x.passes = c.passes;
        x.gshift = c.gshift;
        x.schedule = c.schedule;
    } else {
        default_passes(logN, x.passes);
    }
    x.bufA = base + kOffA;
    x.bufB = base + kOffB;
    x.ninv = base + kOffNinv + kNinvEntry * (size_t)logN;
    hipStream_t st = (hipStream_t)stream;
    const int64_t in_A = (int64_t)(L_A * (limb_bits / 32)), in_B = (int64_t)(L_B * (limb_bits / 32));
    if ((rc = run_transforms(x, d_a, d_b, in_A, in_B, st, nullptr))) return rc;

    const uint64_t n_words = limb_bits == 32 ? (L_A + L_B + 1) / 2 : L_A + L_B;
    unsigned long long* flags = reinterpret_cast<unsigned long long*>(base + kOffFlags);
    int* counter = reinterpret_cast<int*>(base + kOffCounter);
    int* err = d_status ? d_status : counter + 1;
    CarryParams cp;
    std::memset(&cp, 0, sizeof(cp));
    cp.src = BW_SRC_GARNER_D64;
    cp.gd = tabs->gd;
    cp.half_last = limb_bits == 32 && ((L_A + L_B) & 1);
    cp.a0 = x.bufA;
    cp.plane = (int64_t)N_exec;
    cp.N = (int64_t)N_exec;
    cp.n_out = (int64_t)n_words;
    cp.out = d_c;
    cp.g = tabs->garner;
    for (int i = 0; i < 3; i++) cp.fp[i] = tabs->fp[i];
    cp.flags = flags;
    cp.counter = counter;
    cp.err = err;
    cp.epoch = 1;                       // the flags are cleared before every launch
    cp.ntiles = bw_carry_tiles((int64_t)n_words);
    if (cp.ntiles > kMaxTiles) return fail(BW_ERR_STATE, "carry tiles %lld exceed the scratch", (long long)cp.ntiles);
    // flags, tile counter and error word start at zero: a fill kernel, no
    // memset node (the sequence may be captured into a graph)
    LAUNCH_TRY(bw_launch_carry_reset(flags, cp.ntiles, counter, st, err));
    LAUNCH_TRY(bw_launch_garner_carry(cp, st));
    return BW_OK;
}
