// serve.cpp -- request loop with continuous batching (bw_server_* C ABI).
//
// Callers submit products (host operand and result buffers) from any thread.
// One worker thread owns the device: it admits pending requests in arrival
// order into a batch while the batch's device footprint fits the free part
// of a device arena reserved at start (no allocation per request), copies the
// operands in, runs the batch (batch.h: one launch sequence over all its
// products), copies every result straight into its caller's buffer and
// records an event.  Up to `streams` batches are in flight at once, each on
// its own stream, so requests that arrive while others run are admitted as
// soon as a stream and memory are free and their copies and kernels overlap
// the running batches.  A request completes when its batch's event has
// passed; its latency is submit -> completion.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/birdwing.h"
#include "batch.h"

using namespace bw;
using Clock = std::chrono::steady_clock;

namespace {

// First-fit allocator over [0, cap) in 256-byte units.
struct Arena {
    size_t cap = 0;
    std::map<size_t, size_t> free_;   // offset -> size
    size_t used = 0, peak = 0;
    void init(size_t c) {
        cap = c;
        free_.clear();
        free_[0] = c;
        used = peak = 0;
    }
    bool alloc(size_t n, size_t* off) {
        n = (n + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n;
                const size_t roff = it->first + n;
                free_.erase(it);
                if (rest) free_[roff] = rest;
                used += n;
                if (used > peak) peak = used;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) / 256 * 256;
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    int64_t id = 0;
    int bits = 32;
    const void* a = nullptr;
    const void* b = nullptr;
    void* c = nullptr;
    uint64_t la = 0, lb = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
    int status = BW_OK;
};

struct Flight {
    std::vector<Req*> reqs;
    size_t off = 0, bytes = 0;
    int slot = 0;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Device bytes a batch of these requests needs (workspace + packed buffers).
size_t flight_bytes(const BatchLayout& L) {
    return align256(L.ws_bytes) + align256(L.bytesA) + align256(L.bytesB) + align256(L.bytesC);
}

}  // namespace

struct bw_server {
    int device = 0;
    int nslots = 2;
    int max_batch = 1024;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    std::vector<Flight*> slot_flight;
    std::vector<unsigned char*> pinned;       // descriptor staging per slot
    std::vector<size_t> pinned_cap;
    std::vector<int*> errw;                   // pinned copy of each slot's look-back error word
    unsigned char* dmem = nullptr;
    Arena arena;

    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Req*> pending;
    std::unordered_map<int64_t, Req*> reqs;
    int64_t next_id = 0;
    bool stop = false;
    std::thread worker;

    // statistics
    int64_t completed = 0, batches = 0, batched_products = 0, failed = 0;
    int max_concurrent = 0;
    std::string worker_error;
};

static void complete(bw_server* s, Flight* f, int status) {
    const auto now = Clock::now();
    for (Req* r : f->reqs) {
        r->done = true;
        r->t_done = now;
        r->status = status;
        if (status == BW_OK) s->completed++;
        else s->failed++;
    }
    s->arena.release(f->off, f->bytes);
    s->slot_flight[f->slot] = nullptr;
    delete f;
}

// Issues one batch on its slot's stream (worker thread, lock not held).
static int launch_flight(bw_server* s, Flight* f, const BatchLayout& L) {
    hipStream_t st = s->streams[f->slot];
    unsigned char* ws = s->dmem + f->off;
    unsigned char* dA = ws + align256(L.ws_bytes);
    unsigned char* dB = dA + align256(L.bytesA);
    unsigned char* dC = dB + align256(L.bytesB);
    std::vector<unsigned char> img;
    batch_descriptors(L, ws, img);
    if (s->pinned_cap[f->slot] < img.size()) {
        if (s->pinned[f->slot]) HIP_TRY(hipHostFree(s->pinned[f->slot]));
        s->pinned[f->slot] = nullptr;
        const size_t cap = img.size() * 2;
        HIP_TRY(hipHostMalloc((void**)&s->pinned[f->slot], cap, 0));
        s->pinned_cap[f->slot] = cap;
    }
    std::memcpy(s->pinned[f->slot], img.data(), img.size());
    HIP_TRY(hipMemcpyAsync(ws + L.off_desc, s->pinned[f->slot], img.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(ws + L.off_err, 0, sizeof(int), st));
    for (int i = 0; i < L.n; i++) {
        const Req* r = f->reqs[i];
        const size_t w = (size_t)r->bits / 8;
        HIP_TRY(hipMemcpyAsync(dA + L.offA[i], r->a, r->la * w, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dB + L.offB[i], r->b, r->lb * w, hipMemcpyHostToDevice, st));
    }
    int rc = batch_issue(L, ws, dA, dB, dC, st);
    if (rc) return rc;
    for (int i = 0; i < L.n; i++) {
        const Req* r = f->reqs[i];
        const size_t w = (size_t)r->bits / 8;
        HIP_TRY(hipMemcpyAsync(r->c, dC + L.offC[i], (r->la + r->lb) * w, hipMemcpyDeviceToHost, st));
    }
    // the flight's carry look-back error word, read when its event completes
    HIP_TRY(hipMemcpyAsync(s->errw[f->slot], ws + L.off_err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(s->events[f->slot], st));
    return BW_OK;
}

// Takes requests from the head of the queue while the batch fits the largest
// free arena block (lock held).  Requests that cannot fit an empty arena fail.
static Flight* form_flight(bw_server* s, int slot, BatchLayout& L) {
    std::vector<Req*> take;
    std::vector<int> bits;
    std::vector<uint64_t> la, lb;
    size_t largest = 0;
    for (auto& kv : s->arena.free_) largest = std::max(largest, kv.second);
    size_t est = 0;
    // individual footprint estimate: 2 planes of k words per coefficient, packed io, slack
    auto estimate = [](const Req* r) {
        const size_t w = (size_t)r->bits / 8;
        const size_t N = (size_t)padded_ntt_size(r->la, r->lb);
        return 2 * 3 * (N < 2 ? 2 : N) * w + 2 * (r->la + r->lb) * w + 4096;
    };
    while (!s->pending.empty() && (int)take.size() < s->max_batch) {
        Req* r = s->pending.front();
        const size_t e = estimate(r);
        if (!take.empty() && est + e > largest) break;
        take.push_back(r);
        bits.push_back(r->bits);
        la.push_back(r->la);
        lb.push_back(r->lb);
        est += e;
        s->pending.pop_front();
    }
    while (!take.empty()) {
        int rc = batch_layout((int)take.size(), bits.data(), la.data(), lb.data(), L);
        if (rc) {
            // a request the layout rejects fails alone; the rest go back in order
            if (take.size() == 1) {
                take[0]->done = true;
                take[0]->status = rc;
                take[0]->t_done = Clock::now();
                s->failed++;
                return nullptr;
            }
            for (size_t i = take.size(); i-- > 1;) s->pending.push_front(take[i]);
            take.resize(1);
            bits.resize(1);
            la.resize(1);
            lb.resize(1);
            continue;
        }
        const size_t need = flight_bytes(L);
        size_t off = 0;
        if (s->arena.alloc(need, &off)) {
            Flight* f = new Flight();
            f->reqs = take;
            f->off = off;
            f->bytes = need;
            f->slot = slot;
            return f;
        }
        if (take.size() == 1) {
            if (need > s->arena.cap) {          // can never fit
                take[0]->done = true;
                take[0]->status = BW_ERR_NOMEM;
                take[0]->t_done = Clock::now();
                s->failed++;
            } else {
                s->pending.push_front(take[0]);  // wait for memory
            }
            return nullptr;
        }
        // shrink from the tail and retry
        const size_t keep = take.size() / 2;
        for (size_t i = take.size(); i-- > keep;) s->pending.push_front(take[i]);
        take.resize(keep);
        bits.resize(keep);
        la.resize(keep);
        lb.resize(keep);
    }
    return nullptr;
}

static void worker_main(bw_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // completions
        int busy = 0;
        for (int i = 0; i < s->nslots; i++) {
            Flight* f = s->slot_flight[i];
            if (!f) continue;
            const hipError_t q = hipEventQuery(s->events[i]);
            if (q == hipSuccess) {
                // a look-back wait past its bound leaves the flight's products invalid
                const bool bad = *s->errw[i] != 0;
                if (bad) s->worker_error = "carry look-back wait exceeded its bound";
                complete(s, f, bad ? BW_ERR_STATE : BW_OK);
                s->cv_done.notify_all();
            } else if (q != hipErrorNotReady) {
                s->worker_error = hipGetErrorString(q);
                complete(s, f, BW_ERR_HIP);
                s->cv_done.notify_all();
            } else {
                busy++;
            }
        }
        if (s->stop && s->pending.empty() && busy == 0) break;
        // launch into a free slot
        int slot = -1;
        for (int i = 0; i < s->nslots && slot < 0; i++)
            if (!s->slot_flight[i]) slot = i;
        if (slot >= 0 && !s->pending.empty()) {
            BatchLayout L;
            Flight* f = form_flight(s, slot, L);
            s->cv_done.notify_all();             // requests failed by admission are done
            if (f) {
                s->slot_flight[slot] = f;
                s->batches++;
                s->batched_products += (int64_t)f->reqs.size();
                if (busy + 1 > s->max_concurrent) s->max_concurrent = busy + 1;
                lk.unlock();
                const int rc = launch_flight(s, f, L);
                lk.lock();
                if (rc) {
                    s->worker_error = bw_last_error();
                    (void)hipStreamSynchronize(s->streams[slot]);
                    complete(s, f, rc);
                    s->cv_done.notify_all();
                }
                continue;
            }
        }
        if (busy == 0 && s->pending.empty()) {
            s->cv_submit.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        } else {
            // running batches: poll their events at a short interval, waking
            // early for new submissions
            s->cv_submit.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

extern "C" int bw_server_create(bw_server** out, const bw_server_config* cfg) {
    if (!out) return fail(BW_ERR_ARG, "server pointer is NULL");
    *out = nullptr;
    bw_server_config c;
    std::memset(&c, 0, sizeof(c));
    if (cfg) c = *cfg;
    if (c.device_bytes == 0) c.device_bytes = 8ULL << 30;
    if (c.streams <= 0) c.streams = 3;
    if (c.streams > 8) return fail(BW_ERR_ARG, "at most 8 streams");
    if (c.max_batch <= 0) c.max_batch = 1024;
    bw_server* s = new bw_server();
    s->nslots = c.streams;
    s->max_batch = c.max_batch;
    s->streams.assign(s->nslots, nullptr);
    s->events.assign(s->nslots, nullptr);
    s->slot_flight.assign(s->nslots, nullptr);
    s->pinned.assign(s->nslots, nullptr);
    s->pinned_cap.assign(s->nslots, 0);
    s->errw.assign(s->nslots, nullptr);
    int rc = BW_OK;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipMalloc((void**)&s->dmem, c.device_bytes);
    if (e != hipSuccess) rc = fail(BW_ERR_NOMEM, "server arena of %llu bytes: %s",
                                   (unsigned long long)c.device_bytes, hipGetErrorString(e));
    for (int i = 0; i < s->nslots && !rc; i++) {
        e = hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->events[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s->errw[i], sizeof(int), 0);
        if (e == hipSuccess) *s->errw[i] = 0;
        if (e != hipSuccess) rc = fail(BW_ERR_HIP, "server streams: %s", hipGetErrorString(e));
    }
    // tables of every prime set a request can need, before the first request
    for (int set : {SET_U32_REF, SET_U32_WIDE, SET_U64_REF, SET_D64, SET_U64_PAIRS}) {
        DevTables* t;
        if (!rc) rc = get_tables(set, &t);
    }
    if (rc) {
        bw_server_destroy(s);
        return rc;
    }
    s->arena.init(c.device_bytes);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return BW_OK;
}

extern "C" int64_t bw_server_submit(bw_server* s, int limb_bits, const void* h_a, uint64_t L_A, const void* h_b,
                                    uint64_t L_B, void* h_c) {
    if (!s) return fail(BW_ERR_ARG, "server is NULL");
    if (!h_a || !h_b || !h_c) return fail(BW_ERR_ARG, "NULL host buffer");
    if (limb_bits != 32 && limb_bits != 64) return fail(BW_ERR_ARG, "limb_bits must be 32 or 64");
    if (L_A == 0 || L_B == 0) return fail(BW_ERR_UNSUPPORTED_SIZE, "operand limb count must be positive");
    {
        const PrimeSet* ps = prime_set(auto_set(limb_bits, L_A, L_B));
        std::string why;
        if (!ps || !multiply_size_supported(*ps, limb_bits, L_A, L_B, &why))
            return fail(BW_ERR_UNSUPPORTED_SIZE, "unsupported multiply: %s", why.c_str());
    }
    Req* r = new Req();
    r->bits = limb_bits;
    r->a = h_a;
    r->b = h_b;
    r->c = h_c;
    r->la = L_A;
    r->lb = L_B;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete r;
            return fail(BW_ERR_STATE, "server is stopping");
        }
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->pending.push_back(r);
    }
    s->cv_submit.notify_one();
    return r->id;
}

extern "C" int bw_server_wait(bw_server* s, int64_t ticket, int timeout_ms, double* latency_ms) {
    if (!s) return fail(BW_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(BW_ERR_ARG, "unknown ticket %lld", (long long)ticket);
    Req* r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return fail(BW_ERR_STATE, "ticket %lld not done within %d ms", (long long)ticket, timeout_ms);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    const int st = r->status;
    s->reqs.erase(it);
    delete r;
    if (st != BW_OK)
        return fail(st, "request failed%s%s", s->worker_error.empty() ? "" : ": ", s->worker_error.c_str());
    return BW_OK;
}

extern "C" int bw_server_poll(bw_server* s, int64_t ticket) {
    if (!s) return fail(BW_ERR_ARG, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(BW_ERR_ARG, "unknown ticket %lld", (long long)ticket);
    return it->second->done ? 1 : 0;
}

extern "C" int bw_server_get_stats(bw_server* s, bw_server_stats* st) {
    if (!s || !st) return fail(BW_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->completed = s->completed;
    st->failed = s->failed;
    st->batches = s->batches;
    st->batched_products = s->batched_products;
    st->pending = (int64_t)s->pending.size();
    st->max_concurrent_batches = s->max_concurrent;
    st->arena_bytes = s->arena.cap;
    st->arena_peak_bytes = s->arena.peak;
    return BW_OK;
}

extern "C" void bw_server_destroy(bw_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_submit.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (auto& kv : s->reqs) delete kv.second;
    for (int i = 0; i < s->nslots; i++) {
        if (s->streams[i]) (void)hipStreamDestroy(s->streams[i]);
        if (s->events[i]) (void)hipEventDestroy(s->events[i]);
        if (s->pinned[i]) (void)hipHostFree(s->pinned[i]);
        if (s->errw[i]) (void)hipHostFree(s->errw[i]);
    }
    (void)hipFree(s->dmem);
    delete s;
}
