// batch.h -- many products of uneven sizes in one launch sequence.
//
// Products are grouped by (limb width, padded size N).  A group of B products
// is one stacked transform (engine.h: product b at offset b*N of each prime
// plane), so every pass of the group is one launch over all its products;
// ingress reads each product's limbs from packed input buffers through a
// per-product descriptor.  One Garner+carry launch per prime set then covers
// every product of the batch (tile -> product table, look-back inside a
// product only).  The launch sequence can be captured once into a hipGraph.
#pragma once
#include <map>
#include <mutex>
#include <vector>

#include "engine.h"
#include "product_fused.h"

namespace bw {

struct BatchGroup {
    int limb_bits = 32;             // arithmetic width of the set (32: FP64 kernels)
    int ingress = 1;                // 2: 64-bit digits read as 32-bit limb pairs
    int set_id = 0;
    int logN = 0;
    int64_t N = 0;
    int64_t B = 0;                  // stacked slots (members padded to the tile multiple)
    int64_t cap = 0;                // slots the workspace holds (B <= cap)
    std::vector<int> members;       // product indices, slot order
    std::vector<int> passes;
    int gshift = 0;
    // workspace offsets (bytes)
    size_t off_bufA = 0, off_bufB = 0, off_ninv = 0, off_ing = 0;
};

struct BatchCarry {
    int set_id = 0;
    int limb_bits = 32;
    int64_t ntiles = 0;
    int njobs = 0;
    size_t off_jobs = 0, off_tile_job = 0, off_flags = 0, off_counter = 0;
};

// Layout of a batch: computed on the host from the sizes alone.
struct BatchLayout {
    int n = 0;
    std::vector<int> bits;
    std::vector<uint64_t> LA, LB;
    std::vector<uint64_t> offA, offB, offC;   // byte offsets in the packed buffers
    uint64_t bytesA = 0, bytesB = 0, bytesC = 0;
    std::vector<BatchGroup> groups;
    std::vector<BatchCarry> carries;
    // products run whole by the fused kernel (product_fused.h), largest first
    std::vector<int> fused;
    std::vector<BwProductBundle> fbundles;    // indices into fused
    int fused_max_log = 0;
    size_t off_fjobs = 0, off_fbundles = 0, off_fcounter = 0;
    size_t off_desc = 0, desc_bytes = 0;      // descriptor block (uploaded as one copy)
    size_t off_epoch = 0, off_err = 0;
    size_t ws_bytes = 0;                      // device workspace
};

// Tuned (pass split, tile cap) per (limb width, log2 N, log2 B bucket).
struct TuneKey {
    int bits, logN, logB;           // bits: the prime set id
    bool operator<(const TuneKey& o) const {
        return bits != o.bits ? bits < o.bits : logN != o.logN ? logN < o.logN : logB < o.logB;
    }
};
struct TuneChoice {
    std::vector<int> passes;
    int gshift = 0;
    int schedule = 0;   // BW_SCHED_* (single-product plans)
    float ms = 0.0f;
};
bool tune_lookup(const TuneKey& k, TuneChoice* c);
void tune_store(const TuneKey& k, const TuneChoice& c);
int log2_bucket(int64_t B);
// Candidate (pass split, tile cap) pairs for transforms of 2^logN points.
void tune_candidates(int logN, std::vector<TuneChoice>& out);

// Largest transform (log2, 64-bit digits) that batches laid out from now on
// run through the fused whole-product kernel; 0 runs every product through
// stacked transforms.  Returns the previous value.
int set_fused_limit(int max_log);
int fused_limit();

// Builds the layout (and validates every size).  Returns BW_OK or an error.
int batch_layout(int n, const int* bits, const uint64_t* LA, const uint64_t* LB, BatchLayout& L);
// Host image of the descriptor block for a workspace at device address `ws`.
void batch_descriptors(const BatchLayout& L, void* ws, std::vector<unsigned char>& img);
// Side streams and events for issuing the groups of a batch concurrently:
// group i runs on streams[i % streams.size()] after a fork event recorded on
// the main stream, and the main stream waits for every group's join event
// before the Garner+carry launches.  Captured into a hipGraph this gives one
// branch per group.
struct BatchBranches {
    std::vector<hipStream_t> streams;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;   // one per group
};
// Issues the whole batch on `st`: transforms per group (on the branches when
// given), then Garner+carry.
int batch_issue(const BatchLayout& L, void* ws, const void* dA, const void* dB, void* dC, hipStream_t st,
                const BatchBranches* br = nullptr);
// Transform launches of one group only (autotuning).
int batch_issue_group(const BatchLayout& L, const BatchGroup& g, void* ws, const void* dA, const void* dB,
                      hipStream_t st);

}  // namespace bw
