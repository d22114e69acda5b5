/*
 * birdwing.h -- C ABI of libbirdwing_hip.so, the MI355X (gfx950) big-integer
 * multiplier.  Plain pointers and sizes only; every call returns a status
 * code (BW_OK == 0) instead of calling exit(), and bw_last_error() returns the
 * message of the calling thread's last failure.
 *
 * Product method (the reference's, src/gpu_ntt.cu:442-973): zero-pad the
 * little-endian limbs of A and B to N = next_pow2(L_A + L_B - 1), forward NTT
 * under each RNS prime, pointwise product, inverse NTT, Garner CRT, carry
 * propagation.  C has L_A + L_B limbs (src/host_multiply.cpp:30,47).
 *
 * Digit widths (a runtime argument here; compile-time LIMB_BITS in the
 * reference, include/config.h:17-53):
 *   limb_bits = 32: 32-bit limbs, three ~30-bit primes, 128-bit Garner.
 *   limb_bits = 64: 64-bit limbs, three 58-bit primes, 160-bit Garner.
 * Prime sets (src/gpu_ntt.cu:28-47 plus three of this library's):
 *   BW_SET_U32_REF  {0x2d000001, 0x23800001, 0x26800001}, N <= 2^23
 *   BW_SET_U32_WIDE {2013265921, 1811939329, 469762049},  N <= 2^26
 *   BW_SET_U64_REF  {0x400002600000001, 0x400004200000001, 0x400001100000001}
 *   BW_SET_HYBRID   the first two u64 primes on 32-bit limbs (the reference's hybrid
 *                   pipeline, include/config.h:26-30, src/gpu_ntt.cu:40-43): 64-bit
 *                   integer arithmetic, 128-bit two-prime Garner, N <= 2^33
 *   BW_SET_D64      {2113929217, 2013265921, 1811939329, 1711276033, 1107296257}:
 *                   64-bit digits (a pair of 32-bit limbs or one 64-bit limb),
 *                   five transforms of half the length, FP64 arithmetic,
 *                   digit transforms of up to 2^25 points.
 *   BW_SET_U64_PAIRS the BW_SET_U64_REF primes on 64-bit digits (pairs of
 *                   32-bit limbs): the reach of the reference's hybrid
 *                   pipeline for 32-bit limbs (include/config.h:26-30), up to
 *                   2^32 limbs per operand (memory permitting).
 * bw_default_prime_set() is the three-prime set of the reference pipeline for
 * a 32-bit-digit length N: BW_SET_U32_REF while N <= 2^23, BW_SET_U32_WIDE
 * above (u32), BW_SET_U64_REF (u64).  A plan created with prime_set = -1 runs
 * under BW_SET_D64 when that set supports the sizes (fewer butterflies and
 * residue bytes for the same product), else under the default set when that
 * does, else (32-bit limbs) under BW_SET_U64_PAIRS; every product is exact
 * either way, but only the reference's sets give its CRT coefficients
 * (bw_execute_crt).  bw_multiply_size_supported(-1, ...) answers for that
 * automatic choice.  Arrays receiving primes hold at least 5 entries.
 *
 * Streams are hipStream_t values passed as void* (NULL = default stream).
 * Device entry points are asynchronous on that stream; *_host entry points
 * synchronize before returning.  Plans are independent objects: distinct plans
 * may be used from distinct threads; one plan must not run two executions at
 * once (they share scratch buffers).
 */
#ifndef BIRDWING_H
#define BIRDWING_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BW_OK = 0,
    BW_ERR_ARG = -1,
    BW_ERR_UNSUPPORTED_SIZE = -2,
    BW_ERR_HIP = -3,
    BW_ERR_NOMEM = -4,
    BW_ERR_STATE = -5
};

/* Launch order of a plan's passes (bw_plan_set_schedule). */
enum { BW_SCHED_PASS_MAJOR = 0, BW_SCHED_PRIME_MAJOR = 1, BW_SCHED_HYBRID = 2 };

enum { BW_SET_DEFAULT = -1, BW_SET_U32_REF = 0, BW_SET_U32_WIDE = 1, BW_SET_U64_REF = 2, BW_SET_HYBRID = 3,
       BW_SET_D64 = 4, BW_SET_U64_PAIRS = 5 };

typedef struct bw_plan bw_plan;

typedef struct bw_plan_info {
    int limb_bits;
    int prime_set;
    int k;                 /* number of primes */
    uint64_t primes[5];
    uint64_t L_A, L_B;
    uint64_t N;            /* next_pow2(L_A + L_B - 1), the reference's padded size */
    uint64_t N_exec;       /* transform length actually run (>= 2) */
    int npass;             /* passes per transform */
    int pass_log[4];       /* log2 of each pass's points */
    int nlaunch;           /* kernel launches per bw_execute */
    uint64_t device_bytes; /* device memory held by the plan */
    int digit_bits;        /* 64 for the five-prime 64-bit-digit set, else limb_bits */
    int schedule;          /* launch order, BW_SCHED_* */
    int tile_shift;        /* tile cap: lines per workgroup at most the kernel's maximum >> tile_shift */
} bw_plan_info;

/* ---- library ---------------------------------------------------------- */
/* Layout version of the structs in this header; a caller built against one
 * version must not run against a library reporting another. */
#define BW_ABI_VERSION 4
int bw_abi_version(void);
const char* bw_last_error(void);
const char* bw_version(void);
/* hipSetDevice + table warm-up for the device; optional. */
int bw_init(int device);
/* Precompute cache (docs/TIMING.md:24-26,98-100: the host tables are the
 * disk-cache payload of precompute_ntt).  bw_tables_save writes the host
 * twiddle tables and constants of a prime set (BW_SET_*) to a file;
 * bw_tables_load reads one back (reporting its set) so that tables for that
 * set are uploaded from it, not rebuilt, on every device that has not yet
 * built them.  Files of another format or library, or damaged ones, are
 * rejected (BW_ERR_ARG).  Host only: no GPU call. */
int bw_tables_save(int set, const char* path);
int bw_tables_load(const char* path, int* set);

/* ---- size policy (src/ntt_limits.cpp) ---------------------------------- */
uint64_t bw_padded_ntt_size(uint64_t L_A, uint64_t L_B);                         /* :145-152 */
int bw_default_prime_set(int limb_bits, uint64_t N);
int bw_prime_set_info(int prime_set, int* k, uint64_t* primes, uint64_t* roots);  /* src/gpu_ntt.cu:28-47 */
int bw_max_supported_logN(int prime_set);                                         /* :125-135 */
uint64_t bw_max_limb_count_by_crt(int prime_set, int in_bits);                    /* :154-174 */
uint64_t bw_max_supported_limb_count(int prime_set, int in_bits);                 /* :176-184 */
/* The three predicates return 1/0 and fill `why` (if non-NULL) on rejection. */
int bw_ntt_size_supported(int prime_set, uint64_t N, char* why, size_t why_len);  /* :216-239 */
int bw_crt_bound_satisfied(int prime_set, int in_bits, uint64_t L_A, uint64_t L_B,
                           char* why, size_t why_len);                              /* :186-214 */
int bw_multiply_size_supported(int prime_set, int in_bits, uint64_t L_A, uint64_t L_B,
                               char* why, size_t why_len);                          /* :241-254 */

/* ---- plans: precompute_ntt + upload_ntt_precomputed + allocate_ntt_context
 *      (src/gpu_ntt.cu:240-440), cleanup_ntt_context + cleanup_ntt_precomputed
 *      (:975-1019).  Twiddle tables are built once per prime set and device and
 *      shared by all plans. */
int bw_plan_create(bw_plan** plan, int limb_bits, uint64_t L_A, uint64_t L_B, int prime_set);
/* Same, with an explicit pass split (sum of pass_log == log2 N_exec, each <= 13). */
int bw_plan_create_ex(bw_plan** plan, int limb_bits, uint64_t L_A, uint64_t L_B, int prime_set,
                      const int* pass_log, int npass);
void bw_plan_destroy(bw_plan* plan);
int bw_plan_get_info(const bw_plan* plan, bw_plan_info* info);

/* ---- execution: execute_ntt_multiply (src/gpu_ntt.cu:442-973) --------------
 * d_a: L_A limbs, d_b: L_B limbs, d_c: L_A + L_B limbs, all device memory.
 * Runs on `stream` with no host synchronization. */
int bw_execute(bw_plan* plan, const void* d_a, const void* d_b, void* d_c, void* stream);
/* Early exit after Garner (src/gpu_ntt.cu:843-884): N_exec coefficients as
 * 32-bit path: c0 = lo 64 bits, c1 = hi 64 bits (c2 unused, may be NULL);
 * 64-bit path: c0 = lo, c1 = mid, c2 = hi 32 bits (U160). */
int bw_execute_crt(bw_plan* plan, const void* d_a, const void* d_b, uint64_t* d_c0, uint64_t* d_c1,
                   uint32_t* d_c2, void* stream);
/* Launch order: BW_SCHED_PASS_MAJOR issues one launch per pass over all
 * primes; BW_SCHED_PRIME_MAJOR finishes one prime's chain of passes before the
 * next so intermediates are re-read from the Infinity Cache; BW_SCHED_HYBRID
 * runs the forward passes pass-major and the fused and inverse passes prime by
 * prime.  Plans start pass-major; bw_plan_tune also chooses the schedule. */
int bw_plan_set_schedule(bw_plan* plan, int schedule);
/* Autotuner: times candidate pass splits and tile caps of this plan's
 * transforms on the given operands (reps timed runs each, synchronizing),
 * keeps the fastest, and records it in the process-wide table that later
 * plans of the same limb width and size start from. */
int bw_plan_tune(bw_plan* plan, const void* d_a, const void* d_b, void* stream, int reps);
/* Synchronizes the device and reports BW_ERR_STATE if a carry look-back wait
 * of this plan exceeded its bound (a failure that would otherwise be a hang;
 * the products of the executions since the last check are then not valid).
 * The error word is sticky across bw_execute calls until this call reads and
 * clears it. */
int bw_plan_status(bw_plan* plan);
/* Host buffers; copies in, runs, copies out, synchronizes.  Returns
 * BW_ERR_STATE (and clears the error word) if this execution's carry
 * look-back wait exceeded its bound. */
int bw_execute_host(bw_plan* plan, const void* h_a, const void* h_b, void* h_c);
/* Per-launch timing (hipEvents recorded on the execution's stream between
 * launches).  bw_plan_set_timing(plan, 1) starts a recording; bw_plan_timing
 * synchronizes and returns, for each launch of an execution, its duration in ms
 * averaged over every execution recorded since then. */
int bw_plan_set_timing(bw_plan* plan, int enabled);
int bw_plan_timing(const bw_plan* plan, float* ms, int max, int* n);

/* ---- one-shot host multiply: host_multiply_merge / host_multiply_merge_native
 *      (src/host_multiply.cpp:27-134).  c gets L_A + L_B limbs; *ms (optional)
 *      gets the wall time from plan creation to plan release. */
int bw_multiply_host(int limb_bits, const void* h_a, uint64_t L_A, const void* h_b, uint64_t L_B, void* h_c,
                     double* ms);

/* ---- allocation-free product (no counterpart in the reference) -------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  One product
 * of device buffers (d_a: L_A limbs, d_b: L_B limbs, d_c: L_A + L_B limbs of
 * limb_bits 32 or 64), the same limbs as bw_execute, computed in a fixed
 * device scratch array of about 10 MiB that each GPU holds until the process
 * ends.  bw_init(device) must have been called for the current device: it
 * sets up that device's scratch address and builds the tables the call uses.
 * The call itself performs no hipMalloc, hipFree or other allocation.
 * Limit: transforms of at most 2^18 64-bit digits, i.e.
 * next_pow2(ceil(L_A/2) + ceil(L_B/2) - 1) <= 2^18 for 32-bit limbs and
 * next_pow2(L_A + L_B - 1) <= 2^18 for 64-bit limbs (two 2^18-limb or two
 * 2^17-limb operands); larger products return BW_ERR_UNSUPPORTED_SIZE.
 * Every call on a GPU shares that GPU's scratch, which keeps the last call's
 * contents: calls that overlap on one GPU (other streams or host threads)
 * race; calls on one stream are ordered.  d_status (optional, a device int)
 * is set to 0 by the call and to 1 if the carry look-back wait exceeded its
 * bound (the result is then not valid).  Asynchronous on `stream`. */
int bw_multiply_device_noalloc(int limb_bits, const void* d_a, uint64_t L_A, const void* d_b, uint64_t L_B,
                               void* d_c, int* d_status, void* stream);

/* ---- stage-level entry points ------------------------------------------- */
/* Forward transforms of raw limbs (L of them, zero-padded to N_exec) under every
 * prime of the plan: d_spec gets k planes of N_exec words (u32 words on the
 * 32-bit path, u64 on the 64-bit path) in the plan's spectrum order. */
int bw_forward(bw_plan* plan, const void* d_raw, uint64_t L, void* d_spec, void* stream);
/* Inverse transforms in place (k planes), including the N^-1 scale. */
int bw_inverse(bw_plan* plan, void* d_spec, void* stream);
/* order[pos] = frequency index held at spectrum position pos (N_exec entries). */
int bw_spectrum_order(const bw_plan* plan, uint64_t* order);
/* src/zero_pad.cu:14-24 -- widen 32-bit limbs to word_bytes (4 or 8), no reduction. */
int bw_zero_pad(const uint32_t* d_src, void* d_dst, int word_bytes, uint64_t L, uint64_t N, void* stream);
/* src/zero_pad.cu:44-56 -- 64-bit limbs reduced mod p. */
int bw_zero_pad_u64(const uint64_t* d_src, uint64_t* d_dst, uint64_t L, uint64_t N, uint64_t p, void* stream);
/* src/gpu_ntt.cu:216-238 -- c = a*b mod p elementwise (word_bytes 4 or 8). */
int bw_pointwise(int word_bytes, const void* d_a, const void* d_b, void* d_c, uint64_t N, uint64_t p,
                 void* stream);
/* src/crt_gpu.cu:278-302 -- Garner over the residue arrays of a prime set.
 * 32-bit sets (3 arrays): o0 = lo, o1 = hi (128-bit).  64-bit set (3 arrays):
 * o0 = lo, o1 = mid, o2 = hi32.  Hybrid set (2 arrays, r2 unused): o0 = lo,
 * o1 = hi. */
int bw_crt_combine(int prime_set, const void* d_r0, const void* d_r1, const void* d_r2, uint64_t N,
                   uint64_t* d_o0, uint64_t* d_o1, uint32_t* d_o2, void* stream);
/* src/carry_prop.cu (3-pass scheme) replaced by a single-pass carry scan.
 * 128-bit coefficients (hi, lo) -> n_out limbs of out_bits (32 or 64);
 * coefficients beyond N count as zero. */
int bw_carry128(const uint64_t* d_hi, const uint64_t* d_lo, uint64_t N, void* d_out, uint64_t n_out,
                int out_bits, void* stream);
/* U160 coefficients (lo, mid, hi32) -> n_out 64-bit limbs. */
int bw_carry160(const uint64_t* d_lo, const uint64_t* d_mid, const uint32_t* d_hi, uint64_t N, void* d_out,
                uint64_t n_out, void* stream);

/* ---- batched products (no counterpart in the reference, whose API runs one
 *      product per call: execute_ntt_multiply, src/gpu_ntt.cu:442-973) --------
 * n products of uneven sizes; limb_bits[i] is 32 or 64 per product.  Operands
 * and results are packed: product i reads L_A[i] limbs at d_a + off_a[i] and
 * L_B[i] limbs at d_b + off_b[i] (byte offsets from bw_batch_offsets, 256-byte
 * aligned) and writes L_A[i] + L_B[i] limbs at d_c + off_c[i].  Products
 * whose five-prime transform has at most 2^bw_set_fused_limit 64-bit digits
 * (default 2^8, at most 2^11) run whole in one persistent launch: a workgroup
 * runs a bundle of up to 16 products of one length, transforms, pointwise
 * product, Garner and carry in LDS, bundles taken largest first from a job
 * counter.  The other products with the same
 * limb width and padded size share every launch; one Garner+carry launch per
 * prime set covers them.  The batch owns its device workspace
 * (bw_batch_info.workspace_bytes). */
typedef struct bw_batch bw_batch;
typedef struct bw_batch_info {
    int n;                   /* products */
    int ngroups;             /* (limb width, padded size) groups */
    int nlaunch;             /* kernel launches per bw_batch_execute */
    int use_graph;
    uint64_t bytes_a, bytes_b, bytes_c;   /* packed buffer sizes */
    uint64_t workspace_bytes;
} bw_batch_info;
int bw_batch_create(bw_batch** batch, int n, const int* limb_bits, const uint64_t* L_A, const uint64_t* L_B);
void bw_batch_destroy(bw_batch* batch);
/* Host-only: device workspace and packed buffer bytes a batch of these sizes
 * needs (admission control; no device is touched). */
int bw_batch_footprint(int n, const int* limb_bits, const uint64_t* L_A, const uint64_t* L_B,
                       uint64_t* workspace_bytes, uint64_t* bytes_a, uint64_t* bytes_b, uint64_t* bytes_c);
int bw_batch_get_info(const bw_batch* batch, bw_batch_info* info);
int bw_batch_offsets(const bw_batch* batch, uint64_t* off_a, uint64_t* off_b, uint64_t* off_c);
/* Group gi: limb width, log2 N, products, stacked slots, pass split, tile cap. */
int bw_batch_group_info(const bw_batch* batch, int gi, int* limb_bits, int* logN, int* members, int* slots,
                        int* pass_log, int* npass, int* gshift);
/* Products the batch runs through the fused whole-product launch, and the
 * largest log2 transform length among them (0 when there are none). */
int bw_batch_fused_info(const bw_batch* batch, int* products, int* max_logN);
/* Largest log2 transform length (64-bit digits, 0 .. 11) of the products that
 * batches (and server batches) laid out from now on run through the fused
 * whole-product launch; 0 sends every product through stacked transforms.
 * Process-wide.  Returns the previous limit, or a negative status. */
int bw_set_fused_limit(int max_log);
/* Runs the batch on `stream` with no host synchronization. */
int bw_batch_execute(bw_batch* batch, const void* d_a, const void* d_b, void* d_c, void* stream);
/* enabled 1: capture the launch sequence into a hipGraph on the first execute
 * on a non-null stream and replay it while (d_a, d_b, d_c, stream) stay the
 * same; 2: the same with every size group's transforms on a branch of its own
 * (the groups run concurrently until the Garner+carry launches join them);
 * 0: plain launches in order.  On the default (null) stream the launches are
 * always issued in order without a graph. */
int bw_batch_set_graph(bw_batch* batch, int enabled);
/* Autotuner: times candidate pass splits and tile caps of every group on the
 * given operands (reps timed runs each, synchronizing), keeps the fastest in
 * the batch and in a process-wide table that later batches of the same
 * (limb width, size, group-size bucket) start from. */
int bw_batch_tune(bw_batch* batch, const void* d_a, const void* d_b, void* stream, int reps);
int bw_tune_count(void);
void bw_tune_clear(void);
/* Synchronizes the device; BW_ERR_STATE if a carry look-back wait exceeded its
 * bound since the last check (sticky across executions until read here, then
 * cleared). */
int bw_batch_status(bw_batch* batch);

/* ---- serving (no counterpart in the reference) -----------------------------
 * A request loop with continuous batching.  bw_server_submit() queues one
 * product of host buffers (L_A + L_B result limbs; the buffers stay owned by
 * the caller and must stay valid until the ticket completes; pinned memory,
 * e.g. hipHostMalloc, lets the copies overlap the running batches).  A worker
 * thread admits queued requests in arrival order into a batch while its device
 * footprint fits the free part of an arena reserved at creation, and keeps up
 * to `streams` batches in flight.  Every result is the same as the product run
 * alone. */
typedef struct bw_server bw_server;
typedef struct bw_server_config {
    uint64_t device_bytes;   /* device arena (0: 8 GiB) */
    int streams;             /* batches in flight (0: 3, at most 8) */
    int max_batch;           /* products per batch (0: 1024) */
} bw_server_config;
typedef struct bw_server_stats {
    int64_t completed, failed, batches, batched_products, pending;
    int max_concurrent_batches;
    uint64_t arena_bytes, arena_peak_bytes;
} bw_server_stats;
int bw_server_create(bw_server** server, const bw_server_config* config);
/* Returns a ticket (>= 0) or a negative status. */
int64_t bw_server_submit(bw_server* server, int limb_bits, const void* h_a, uint64_t L_A, const void* h_b,
                         uint64_t L_B, void* h_c);
/* Blocks until the ticket completes (timeout_ms < 0: no limit); releases the
 * ticket and reports its submit-to-completion latency.  Returns the ticket's
 * status: BW_ERR_STATE when the carry look-back wait of its batch exceeded its
 * bound (the result is then not valid). */
int bw_server_wait(bw_server* server, int64_t ticket, int timeout_ms, double* latency_ms);
/* 1 when the ticket has completed, 0 while it is queued or running. */
int bw_server_poll(bw_server* server, int64_t ticket);
int bw_server_get_stats(bw_server* server, bw_server_stats* stats);
/* Drains queued and running requests, then frees everything. */
void bw_server_destroy(bw_server* server);

#ifdef __cplusplus
}
#endif
#endif
