// engine.h -- host-side machinery shared by plans, batches and the server:
// errors, per-(device, prime set) tables, and the launch sequence of a
// stacked multi-pass transform.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "bw_internal.h"
#include "carry_kernels.h"
#include "host_math.h"
#include "ntt_kernels.h"

// 1: the passes of a product alternate the direction of their workgroup walk
// (PassParams::rev); 0: every pass walks forward.
#ifndef BW_ALTERNATE_WALK
#define BW_ALTERNATE_WALK 1
#endif

namespace bw {

// Records the message for bw_last_error() (thread-local) and returns `code`.
int fail(int code, const char* fmt, ...);

struct DevTables {
    SetTables t;
    GarnerDev garner;
    F64Prime fp[3];
    GarnerD64Dev gd;      // five-prime sets only
};
// Tables of prime set `set_id` on the current device (built on first use).
int get_tables(int set_id, DevTables** out);
// The same tables only if they are already built on the current device (no
// allocation, no copy); BW_ERR_STATE otherwise.
int find_tables(int set_id, DevTables** out);
// Whether a pass split can run: every pass within its kernel family (the
// register-staged kernels up to 2^10 points -- the only ones reading 64-bit
// digits --, the LDS kernels up to 2^13), and the last pass of a split of two
// or more may be a row-resident fused pass of 2^14 or 2^15 points on the FP64
// path.  arith_bits: 32 (FP64 kernels) or 64; ingress as in Xform.
bool split_supported(const std::vector<int>& passes, int arith_bits, int ingress, std::string* why);
// Host images of a prime set's tables saved to / loaded from a file (host only).
int tables_save(int set_id, const char* path);
int tables_load(const char* path, int* set_id);

// One transform job: B products of one padded size N = 2^logN stacked along
// the outer dimension.  Prime plane i of bufA / bufB holds product b at
// elements [i*B*N + b*N, i*B*N + (b+1)*N).  B = 1 is a single product.
struct Xform {
    int limb_bits = 32;
    const PrimeSet* set = nullptr;
    DevTables* tabs = nullptr;
    int logN = 0;
    int64_t N = 0;
    int64_t B = 1;
    std::vector<int> passes;
    int gshift = 0;                 // tile cap: lines per workgroup <= max >> gshift
    void* bufA = nullptr;
    void* bufB = nullptr;
    const void* ninv = nullptr;     // per-prime N^-1 pairs for this N
    int schedule = 0;               // BW_SCHED_*
    int ingress = 1;                // raw operands: 1 a limb per digit, 2 64-bit digits of 32-bit limb pairs,
                                    // 3 32-bit limbs into 64-bit residue words
    const BwIngress* bdesc = nullptr;   // stacked ingress (B > 1), device memory
};

struct PassPlan {
    int logT;
    int64_t S, O;
    int logM;
    int G;
    int64_t tiles;
};

PassPlan pass_geom(const Xform& x, int j);
PassParams base_params(const Xform& x, const PassPlan& g);
hipError_t launch_pass(const Xform& x, int mode, const PassPlan& g, const PassParams& prm, int nops, hipStream_t st);

// Optional per-launch hook (timing events).
struct LaunchMarker {
    virtual hipError_t mark() = 0;
    virtual ~LaunchMarker() {}
};

// Forward transforms of A and B (raw limbs in), pointwise product, inverse
// transform: canonical residues of the cyclic product land in bufA.  la / lb
// are the raw limb counts of a single product (ignored when x.bdesc is set).
int run_transforms(const Xform& x, const void* d_a, const void* d_b, int64_t la, int64_t lb, hipStream_t st,
                   LaunchMarker* mk);
int count_transform_launches(const Xform& x);

// Default split of 2^logN: one pass up to 2^10, [7, 4] at 2^11, two passes
// (the larger first) up to 2^16 -- the autotuner's choices for stacked
// batches --; above, passes of at most 2^9 points, as balanced as possible,
// the larger last ([8, 8, 8] at 2^24).
void default_passes(int logN, std::vector<int>& out);

}  // namespace bw

#define HIP_TRY(expr)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return bw::fail(BW_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_TRY(expr)                                                                                     \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return bw::fail(BW_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e_)); \
    } while (0)
