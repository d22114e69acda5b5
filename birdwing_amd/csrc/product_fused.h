// product_fused.h -- whole products in one workgroup, many products per launch.
//
// A product whose transforms fit in one CU's LDS (at most 2^11 64-bit digits,
// the five-prime set) runs start to finish inside one 1024-thread workgroup:
// operand digits are read once per prime straight from the caller's limbs,
// the forward transforms of A and B, the pointwise product (with N^-1), the
// inverse transform, Garner's reconstruction and the carry all work on LDS,
// and only the result limbs go back to HBM.  Shorter products run several to
// a workgroup (a bundle: up to 16 products of one transform length, as many
// as 120 KiB of LDS holds), so every barrier step has enough work.  The
// launch is persistent: each workgroup takes bundles from an atomic counter
// (largest transforms first) until none are left, so one launch covers any
// number of products of any mix of these sizes.
#pragma once
#include "bw_internal.h"

#define BW_FUSED_MAX_LOG 11        // largest transform: 2^11 64-bit digits
// Default limit of batches: the fused launch is faster than stacked
// transforms up to 2^8 digits (1.2x at 2^8, 2.3x at 2^7, 3-5x below on
// MI355X), slower at 2^9 and 2^10 (whose single register-staged pass of
// radix-32 codelets does fewer LDS round trips per butterfly), level at 2^11.
#define BW_FUSED_DEFAULT_LOG 8
#define BW_FUSED_THREADS 1024
#define BW_FUSED_MAX_BUNDLE 16
#define BW_FUSED_LDS_WORDS 30720   // dynamic LDS per workgroup (32-bit words, 120 KiB)

// One product: operand limbs at A + a_off, B + b_off (bytes), la / lb limbs of
// limb_bits; its la + lb result limbs at C + c_off (8-byte aligned); transform
// length 2^logN 64-bit digits (logN >= 1, 2^logN >= digits(A) + digits(B) - 1).
struct BwProductJob {
    int64_t a_off, b_off, c_off;
    int64_t la, lb;
    int32_t limb_bits;
    int32_t logN;
};

// Products one workgroup runs together: jobs[first, first + count), all with
// the same logN, count <= bw_fused_bundle_size(logN).
struct BwProductBundle {
    int32_t first;
    int32_t count;
};

// Products of transform length 2^logN that one workgroup holds at once.
// (LDS: five twiddle tables, then ten transforms per product, n words each)
__host__ __device__ constexpr int bw_fused_bundle_size(int logN) {
    return (BW_FUSED_LDS_WORDS - (5 << logN)) / (10 << logN) < BW_FUSED_MAX_BUNDLE
               ? (BW_FUSED_LDS_WORDS - (5 << logN)) / (10 << logN)
               : BW_FUSED_MAX_BUNDLE;
}

struct FusedParams {
    const unsigned char* A;
    const unsigned char* B;
    unsigned char* C;
    const BwProductJob* jobs;        // device memory
    const BwProductBundle* bundles;  // device memory, largest transforms first
    int nbundles;
    int* counter;                 // bundle allocator: zero before the launch, zero again after it
    const F64PrimeDev* consts;    // the five primes
    const double2* tw_stage;      // [5][2^BW_STAGE_LOG]: omega_{2^13}^e as (w, w/p)
    GarnerD64Dev gd;
};

// grid: workgroups (each loops over bundles).
hipError_t bw_launch_product_fused(const FusedParams& p, int grid, hipStream_t stream);
