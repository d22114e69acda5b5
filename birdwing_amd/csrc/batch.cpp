// batch.cpp -- batched products (batch.h) and the bw_batch_* C ABI.
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>

#include "../../include/birdwing.h"

namespace bw {

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int log2_bucket(int64_t B) {
    int l = 0;
    while (((int64_t)1 << l) < B) l++;
    return l;
}

// ---------------------------------------------------------------------------
// tuning cache (process-wide)
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneChoice> g_tune;

bool tune_lookup(const TuneKey& k, TuneChoice* c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(k);
    if (it == g_tune.end()) return false;
    *c = it->second;
    return true;
}

void tune_store(const TuneKey& k, const TuneChoice& c) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[k] = c;
}

static std::atomic<int> g_fused_limit{BW_FUSED_DEFAULT_LOG};

int set_fused_limit(int max_log) { return g_fused_limit.exchange(max_log); }
int fused_limit() { return g_fused_limit.load(); }

// Workgroups of a fused launch: one per CU (its LDS and registers hold one).
static int fused_grid(int nbundles) {
    static const int cus = [] {
        int dev = 0, c = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0)
            return c;
        return 256;
    }();
    return std::min(nbundles, cus);
}

// Slots of a one-pass group are padded to the rows tile so that every
// workgroup holds whole products.
static int64_t rows_tile(int logN, int limb_bits, int gshift) {
    int g = bw_pass_max_G(logN, limb_bits, false) >> gshift;
    if (g < 1) g = 1;
    while (g > 1 && ((int64_t)g << logN) > 4096) g >>= 1;
    return g;
}

// Carry words of product i: 64-bit words for 64-bit digits (half a word
// more than needed when an odd number of 32-bit limbs), else its limbs.
static uint64_t carry_words(const BatchGroup& g, const BatchLayout& L, int i) {
    const uint64_t n = L.LA[i] + L.LB[i];
    return (g.ingress == 2 && L.bits[i] == 32) ? (n + 1) / 2 : n;
}

static void group_geometry(BatchGroup& g) {
    const int64_t m = (int64_t)g.members.size();
    g.B = m;
    if (g.passes.size() == 1) {
        const int64_t t = rows_tile(g.logN, g.limb_bits, g.gshift);
        g.B = (m + t - 1) / t * t;
    }
}

int batch_layout(int n, const int* bits, const uint64_t* LA, const uint64_t* LB, BatchLayout& L) {
    if (n <= 0) return fail(BW_ERR_ARG, "batch needs at least one product");
    if (!bits || !LA || !LB) return fail(BW_ERR_ARG, "NULL size array");
    L = BatchLayout();
    L.n = n;
    L.bits.assign(bits, bits + n);
    L.LA.assign(LA, LA + n);
    L.LB.assign(LB, LB + n);
    L.offA.resize(n);
    L.offB.resize(n);
    L.offC.resize(n);
    std::map<std::pair<int, int>, int> gidx;   // (prime set, logN) -> group
    const int flim = fused_limit();
    std::vector<std::pair<int, int>> fused;    // (logN, product)
    for (int i = 0; i < n; i++) {
        const int b = bits[i];
        if (b != 32 && b != 64) return fail(BW_ERR_ARG, "product %d: limb_bits must be 32 or 64", i);
        if (LA[i] == 0 || LB[i] == 0) return fail(BW_ERR_UNSUPPORTED_SIZE, "product %d: empty operand", i);
        const int set_id = auto_set(b, LA[i], LB[i]);
        const PrimeSet* s = prime_set(set_id);
        std::string why;
        if (!s || !multiply_size_supported(*s, b, LA[i], LB[i], &why))
            return fail(BW_ERR_UNSUPPORTED_SIZE, "product %d: unsupported multiply: %s", i, why.c_str());
        const uint64_t N = padded_size_for(*s, b, LA[i], LB[i]);
        const size_t w = (size_t)b / 8;
        L.offA[i] = L.bytesA;
        L.offB[i] = L.bytesB;
        L.offC[i] = L.bytesC;
        L.bytesA = align_up(L.bytesA + LA[i] * w, 256);
        L.bytesB = align_up(L.bytesB + LB[i] * w, 256);
        L.bytesC = align_up(L.bytesC + (LA[i] + LB[i]) * w, 256);
        const uint64_t Nx = N < 2 ? 2 : N;
        const int logN = valuation2(Nx);
        if (set_id == BW_SET_D64 && logN <= flim) {
            fused.push_back(std::make_pair(logN, i));
            continue;
        }
        auto key = std::make_pair(set_id, logN);
        auto it = gidx.find(key);
        if (it == gidx.end()) {
            BatchGroup g;
            g.limb_bits = s->limb_bits;
            g.ingress = s->digit_bits == 64 ? 2 : 1;
            g.logN = logN;
            g.N = (int64_t)Nx;
            g.set_id = set_id;
            gidx[key] = (int)L.groups.size();
            L.groups.push_back(g);
            it = gidx.find(key);
        }
        BatchGroup& g = L.groups[it->second];
        if (g.set_id != set_id) return fail(BW_ERR_STATE, "prime set differs inside a size group");
        g.members.push_back(i);
    }
    // fused products: largest transforms first (the job queue hands them out in order)
    std::stable_sort(fused.begin(), fused.end(),
                     [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first > b.first; });
    for (const auto& f : fused) {
        L.fused.push_back(f.second);
        L.fused_max_log = std::max(L.fused_max_log, f.first);
    }
    // bundles: runs of one transform length, as many products as a workgroup holds
    for (size_t f = 0; f < fused.size();) {
        const int lg = fused[f].first;
        size_t e = f;
        while (e < fused.size() && fused[e].first == lg && (int)(e - f) < bw_fused_bundle_size(lg)) e++;
        L.fbundles.push_back(BwProductBundle{(int32_t)f, (int32_t)(e - f)});
        f = e;
    }
    // pass split and tile cap: tuned choice if known, else the default split
    for (auto& g : L.groups) {
        TuneChoice c;
        if (tune_lookup(TuneKey{g.set_id, g.logN, log2_bucket((int64_t)g.members.size())}, &c)) {
            g.passes = c.passes;
            g.gshift = c.gshift;
        } else {
            default_passes(g.logN, g.passes);
            g.gshift = 0;
        }
        group_geometry(g);
        g.cap = g.B;
    }
    // carry launches: one per prime set
    std::map<int, int> cidx;
    for (auto& g : L.groups) {
        if (!cidx.count(g.set_id)) {
            BatchCarry c;
            c.set_id = g.set_id;
            c.limb_bits = g.limb_bits;
            cidx[g.set_id] = (int)L.carries.size();
            L.carries.push_back(c);
        }
        BatchCarry& c = L.carries[cidx[g.set_id]];
        for (int i : g.members) {
            c.njobs++;
            c.ntiles += bw_carry_tiles((int64_t)carry_words(g, L, i));
        }
    }
    // workspace: residue planes first, then the descriptor block, then flags
    size_t off = 0;
    for (auto& g : L.groups) {
        const size_t w = (size_t)g.limb_bits / 8;
        const size_t bytes = (size_t)prime_set(g.set_id)->k * (size_t)g.cap * (size_t)g.N * w;
        g.off_bufA = off;
        off = align_up(off + bytes, 256);
        g.off_bufB = off;
        off = align_up(off + bytes, 256);
    }
    L.off_desc = off;
    for (auto& g : L.groups) {
        g.off_ninv = off;
        off = align_up(off + 16 * (size_t)prime_set(g.set_id)->k, 64);   // (N^-1, aux) pair per prime
        g.off_ing = off;
        off = align_up(off + sizeof(BwIngress) * (size_t)g.cap, 64);
    }
    for (auto& c : L.carries) {
        c.off_jobs = off;
        off = align_up(off + sizeof(BwCarryJob) * (size_t)c.njobs, 64);
        c.off_tile_job = off;
        off = align_up(off + sizeof(int) * (size_t)c.ntiles, 64);
    }
    L.off_fjobs = off;
    off = align_up(off + sizeof(BwProductJob) * L.fused.size(), 64);
    L.off_fbundles = off;
    off = align_up(off + sizeof(BwProductBundle) * L.fbundles.size(), 64);
    // the fused job counter starts at zero with every descriptor upload (a
    // server reuses arena memory) and returns to zero after every launch
    L.off_fcounter = off;
    off = align_up(off + sizeof(int), 64);
    L.desc_bytes = off - L.off_desc;
    off = align_up(off, 256);
    for (auto& c : L.carries) {
        c.off_flags = off;
        off = align_up(off + 8 * (size_t)c.ntiles, 256);
        c.off_counter = off;
        off = align_up(off + 2 * sizeof(int), 64);
    }
    L.off_err = off;
    off = align_up(off + sizeof(int), 256);
    L.ws_bytes = off;
    return BW_OK;
}

void batch_descriptors(const BatchLayout& L, void* ws, std::vector<unsigned char>& img) {
    img.assign(L.desc_bytes, 0);
    auto at = [&](size_t ws_off) { return img.data() + (ws_off - L.off_desc); };
    unsigned char* dws = static_cast<unsigned char*>(ws);
    for (const auto& g : L.groups) {
        const PrimeSet& s = *prime_set(g.set_id);
        std::vector<uint64_t> n64;
        std::vector<double> nf;
        ninv_pairs(s, (uint64_t)g.N, n64, nf);
        std::memcpy(at(g.off_ninv), g.limb_bits == 64 ? (const void*)n64.data() : (const void*)nf.data(),
                    16 * (size_t)s.k);
        BwIngress* ing = reinterpret_cast<BwIngress*>(at(g.off_ing));
        for (int64_t slot = 0; slot < g.B; slot++) {
            BwIngress d = {0, 0, 0, 0};
            if (slot < (int64_t)g.members.size()) {
                const int i = g.members[slot];
                // raw units: 32-bit limbs for 64-bit digits, else the product's limbs
                const size_t u = g.ingress == 2 ? 4 : (size_t)L.bits[i] / 8;
                const uint64_t per = g.ingress == 2 ? (uint64_t)L.bits[i] / 32 : 1;
                d.a_off = (int64_t)(L.offA[i] / u);
                d.b_off = (int64_t)(L.offB[i] / u);
                d.a_len = (int64_t)(L.LA[i] * per);
                d.b_len = (int64_t)(L.LB[i] * per);
            }
            ing[slot] = d;
        }
    }
    BwProductJob* fj = reinterpret_cast<BwProductJob*>(at(L.off_fjobs));
    for (size_t f = 0; f < L.fused.size(); f++) {
        const int i = L.fused[f];
        const PrimeSet& s = *prime_set(BW_SET_D64);
        const uint64_t N = padded_size_for(s, L.bits[i], L.LA[i], L.LB[i]);
        BwProductJob J;
        J.a_off = (int64_t)L.offA[i];
        J.b_off = (int64_t)L.offB[i];
        J.c_off = (int64_t)L.offC[i];
        J.la = (int64_t)L.LA[i];
        J.lb = (int64_t)L.LB[i];
        J.limb_bits = L.bits[i];
        J.logN = valuation2(N < 2 ? 2 : N);
        fj[f] = J;
    }
    std::memcpy(at(L.off_fbundles), L.fbundles.data(), sizeof(BwProductBundle) * L.fbundles.size());
    for (const auto& c : L.carries) {
        BwCarryJob* jobs = reinterpret_cast<BwCarryJob*>(at(c.off_jobs));
        int* tile_job = reinterpret_cast<int*>(at(c.off_tile_job));
        int j = 0;
        int64_t t = 0;
        for (const auto& g : L.groups) {
            if (g.set_id != c.set_id) continue;
            const size_t w = (size_t)g.limb_bits / 8;
            for (size_t slot = 0; slot < g.members.size(); slot++) {
                const int i = g.members[slot];
                BwCarryJob J;
                J.res = dws + g.off_bufA + slot * (size_t)g.N * w;
                J.plane = g.B * g.N;
                J.N = g.N;
                J.n_out = (int64_t)carry_words(g, L, i);
                // output words: 64-bit for 64-bit digits, else the product's limbs
                J.out_off = (int64_t)(L.offC[i] / (g.ingress == 2 ? 8 : (size_t)L.bits[i] / 8));
                J.tile0 = t;
                J.half_last = g.ingress == 2 && L.bits[i] == 32 && ((L.LA[i] + L.LB[i]) & 1);
                const int64_t nt = bw_carry_tiles(J.n_out);
                for (int64_t q = 0; q < nt; q++) tile_job[t + q] = j;
                t += nt;
                jobs[j++] = J;
            }
        }
    }
}

static Xform group_xform(const BatchGroup& g, void* ws) {
    unsigned char* d = static_cast<unsigned char*>(ws);
    Xform x;
    x.limb_bits = g.limb_bits;
    x.set = prime_set(g.set_id);
    x.logN = g.logN;
    x.N = g.N;
    x.B = g.B;
    x.passes = g.passes;
    x.gshift = g.gshift;
    x.bufA = d + g.off_bufA;
    x.bufB = d + g.off_bufB;
    x.ninv = d + g.off_ninv;
    x.schedule = BW_SCHED_PASS_MAJOR;
    x.ingress = g.ingress;
    x.bdesc = reinterpret_cast<const BwIngress*>(d + g.off_ing);
    return x;
}

int batch_issue_group(const BatchLayout& L, const BatchGroup& g, void* ws, const void* dA, const void* dB,
                      hipStream_t st) {
    (void)L;
    Xform x = group_xform(g, ws);
    int rc = get_tables(g.set_id, &x.tabs);
    if (rc) return rc;
    return run_transforms(x, dA, dB, 0, 0, st, nullptr);
}

static int issue_fused(const BatchLayout& L, void* ws, const void* dA, const void* dB, void* dC, hipStream_t st) {
    if (L.fused.empty()) return BW_OK;
    unsigned char* d = static_cast<unsigned char*>(ws);
    DevTables* tabs;
    int rc = get_tables(BW_SET_D64, &tabs);
    if (rc) return rc;
    FusedParams fp;
    std::memset(&fp, 0, sizeof(fp));
    fp.A = static_cast<const unsigned char*>(dA);
    fp.B = static_cast<const unsigned char*>(dB);
    fp.C = static_cast<unsigned char*>(dC);
    fp.jobs = reinterpret_cast<const BwProductJob*>(d + L.off_fjobs);
    fp.bundles = reinterpret_cast<const BwProductBundle*>(d + L.off_fbundles);
    fp.nbundles = (int)L.fbundles.size();
    fp.counter = reinterpret_cast<int*>(d + L.off_fcounter);
    fp.consts = static_cast<const F64PrimeDev*>(tabs->t.consts);
    fp.tw_stage = static_cast<const double2*>(tabs->t.tw_stage);
    fp.gd = tabs->gd;
    LAUNCH_TRY(bw_launch_product_fused(fp, fused_grid(fp.nbundles), st));
    return BW_OK;
}

int batch_issue(const BatchLayout& L, void* ws, const void* dA, const void* dB, void* dC, hipStream_t st,
                const BatchBranches* br) {
    unsigned char* d = static_cast<unsigned char*>(ws);
    int rc;
    if (br && !br->streams.empty() && L.groups.size() > 1) {
        // groups are independent until Garner+carry: one branch each, the
        // whole small products on the main stream beside them
        HIP_TRY(hipEventRecord(br->fork, st));
        if ((rc = issue_fused(L, ws, dA, dB, dC, st))) return rc;
        for (size_t i = 0; i < L.groups.size(); i++) {
            hipStream_t s = br->streams[i % br->streams.size()];
            HIP_TRY(hipStreamWaitEvent(s, br->fork, 0));
            if ((rc = batch_issue_group(L, L.groups[i], ws, dA, dB, s))) return rc;
            HIP_TRY(hipEventRecord(br->joins[i], s));
        }
        for (size_t i = 0; i < L.groups.size(); i++) HIP_TRY(hipStreamWaitEvent(st, br->joins[i], 0));
    } else {
        if ((rc = issue_fused(L, ws, dA, dB, dC, st))) return rc;
        for (const auto& g : L.groups)
            if ((rc = batch_issue_group(L, g, ws, dA, dB, st))) return rc;
    }
    for (const auto& c : L.carries) {
        DevTables* tabs;
        if ((rc = get_tables(c.set_id, &tabs))) return rc;
        unsigned long long* flags = reinterpret_cast<unsigned long long*>(d + c.off_flags);
        int* counter = reinterpret_cast<int*>(d + c.off_counter);
        LAUNCH_TRY(bw_launch_carry_reset(flags, c.ntiles, counter, st));
        CarryParams cp;
        std::memset(&cp, 0, sizeof(cp));
        cp.src = prime_set(c.set_id)->k == 5 ? BW_SRC_GARNER_D64
                 : (prime_set(c.set_id)->limb_bits == 64 ? BW_SRC_GARNER_U64 : BW_SRC_GARNER_U32);
        cp.gd = tabs->gd;
        cp.out = dC;
        cp.g = tabs->garner;
        for (int i = 0; i < 3; i++) cp.fp[i] = tabs->fp[i];
        cp.flags = flags;
        cp.counter = counter;
        cp.err = reinterpret_cast<int*>(d + L.off_err);
        cp.epoch = 1;                       // flags are cleared before every launch
        cp.ntiles = c.ntiles;
        cp.jobs = reinterpret_cast<const BwCarryJob*>(d + c.off_jobs);
        cp.tile_job = reinterpret_cast<const int*>(d + c.off_tile_job);
        LAUNCH_TRY(bw_launch_garner_carry(cp, st));
    }
    return BW_OK;
}

}  // namespace bw

using namespace bw;

// ---------------------------------------------------------------------------
// C ABI
struct bw_batch {
    BatchLayout L;
    void* ws = nullptr;
    // hipGraph of the launch sequence, valid for (A, B, C, stream)
    bool use_graph = false;
    bool branches = false;          // groups on concurrent streams (graph: one branch per group)
    BatchBranches br;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    const void* gA = nullptr;
    const void* gB = nullptr;
    void* gC = nullptr;
    hipStream_t gst = nullptr;
};

static void drop_graph(bw_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static void drop_branches(bw_batch* b) {
    for (auto s : b->br.streams) (void)hipStreamDestroy(s);
    for (auto e : b->br.joins) (void)hipEventDestroy(e);
    if (b->br.fork) (void)hipEventDestroy(b->br.fork);
    b->br = BatchBranches();
}

// Side streams (default priority) and events for concurrent groups, made once.
static int make_branches(bw_batch* b) {
    if (b->br.fork) return BW_OK;
    const size_t ng = b->L.groups.size();
    const size_t ns = ng < 8 ? ng : 8;
    HIP_TRY(hipEventCreateWithFlags(&b->br.fork, hipEventDisableTiming));
    for (size_t i = 0; i < ns; i++) {
        hipStream_t s;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->br.streams.push_back(s);
    }
    for (size_t i = 0; i < ng; i++) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->br.joins.push_back(e);
    }
    return BW_OK;
}

static int upload_descriptors(bw_batch* b) {
    std::vector<unsigned char> img;
    batch_descriptors(b->L, b->ws, img);
    HIP_TRY(hipMemcpy(static_cast<unsigned char*>(b->ws) + b->L.off_desc, img.data(), img.size(),
                      hipMemcpyHostToDevice));
    return BW_OK;
}

extern "C" int bw_batch_create(bw_batch** out, int n, const int* limb_bits, const uint64_t* L_A, const uint64_t* L_B) {
    if (!out) return fail(BW_ERR_ARG, "batch pointer is NULL");
    *out = nullptr;
    bw_batch* b = new bw_batch();
    int rc = batch_layout(n, limb_bits, L_A, L_B, b->L);
    if (rc) { delete b; return rc; }
    // tables are built here, never inside a captured launch sequence
    for (const auto& g : b->L.groups) {
        DevTables* t;
        if ((rc = get_tables(g.set_id, &t))) { delete b; return rc; }
    }
    if (!b->L.fused.empty()) {
        DevTables* t;
        if ((rc = get_tables(BW_SET_D64, &t))) { delete b; return rc; }
    }
    hipError_t e = hipMalloc(&b->ws, b->L.ws_bytes);
    if (e != hipSuccess) {
        delete b;
        return fail(BW_ERR_NOMEM, "hipMalloc(%zu) for the batch workspace failed: %s", b->L.ws_bytes,
                    hipGetErrorString(e));
    }
    // flags, counters and the error word start at zero
    e = hipMemset(static_cast<unsigned char*>(b->ws) + b->L.off_desc, 0, b->L.ws_bytes - b->L.off_desc);
    if (e != hipSuccess || (rc = upload_descriptors(b))) {
        (void)hipFree(b->ws);
        delete b;
        return rc ? rc : fail(BW_ERR_HIP, "hipMemset: %s", hipGetErrorString(e));
    }
    *out = b;
    return BW_OK;
}

extern "C" int bw_batch_footprint(int n, const int* limb_bits, const uint64_t* L_A, const uint64_t* L_B,
                                  uint64_t* workspace_bytes, uint64_t* bytes_a, uint64_t* bytes_b, uint64_t* bytes_c) {
    BatchLayout L;
    int rc = batch_layout(n, limb_bits, L_A, L_B, L);
    if (rc) return rc;
    if (workspace_bytes) *workspace_bytes = L.ws_bytes;
    if (bytes_a) *bytes_a = L.bytesA;
    if (bytes_b) *bytes_b = L.bytesB;
    if (bytes_c) *bytes_c = L.bytesC;
    return BW_OK;
}

extern "C" void bw_batch_destroy(bw_batch* b) {
    if (!b) return;
    drop_graph(b);
    for (auto s : b->br.streams) (void)hipStreamSynchronize(s);
    drop_branches(b);
    (void)hipFree(b->ws);
    delete b;
}

extern "C" int bw_batch_get_info(const bw_batch* b, bw_batch_info* info) {
    if (!b || !info) return fail(BW_ERR_ARG, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    info->n = b->L.n;
    info->ngroups = (int)b->L.groups.size();
    info->bytes_a = b->L.bytesA;
    info->bytes_b = b->L.bytesB;
    info->bytes_c = b->L.bytesC;
    info->workspace_bytes = b->L.ws_bytes;
    int nl = 0;
    for (const auto& g : b->L.groups) {
        Xform x;
        x.passes = g.passes;
        x.set = prime_set(g.set_id);
        x.schedule = BW_SCHED_PASS_MAJOR;
        nl += count_transform_launches(x);
    }
    info->nlaunch = nl + 2 * (int)b->L.carries.size() + (b->L.fused.empty() ? 0 : 1);
    info->use_graph = b->use_graph ? 1 : 0;
    return BW_OK;
}

extern "C" int bw_batch_offsets(const bw_batch* b, uint64_t* off_a, uint64_t* off_b, uint64_t* off_c) {
    if (!b) return fail(BW_ERR_ARG, "batch is NULL");
    for (int i = 0; i < b->L.n; i++) {
        if (off_a) off_a[i] = b->L.offA[i];
        if (off_b) off_b[i] = b->L.offB[i];
        if (off_c) off_c[i] = b->L.offC[i];
    }
    return BW_OK;
}

extern "C" int bw_batch_group_info(const bw_batch* b, int gi, int* limb_bits, int* logN, int* members, int* slots,
                                   int* pass_log, int* npass, int* gshift) {
    if (!b || gi < 0 || gi >= (int)b->L.groups.size()) return fail(BW_ERR_ARG, "bad group index");
    const BatchGroup& g = b->L.groups[gi];
    if (limb_bits) *limb_bits = g.limb_bits;
    if (logN) *logN = g.logN;
    if (members) *members = (int)g.members.size();
    if (slots) *slots = (int)g.B;
    if (npass) *npass = (int)g.passes.size();
    if (pass_log)
        for (size_t j = 0; j < g.passes.size() && j < 4; j++) pass_log[j] = g.passes[j];
    if (gshift) *gshift = g.gshift;
    return BW_OK;
}

extern "C" int bw_batch_fused_info(const bw_batch* b, int* products, int* max_logN) {
    if (!b) return fail(BW_ERR_ARG, "batch is NULL");
    if (products) *products = (int)b->L.fused.size();
    if (max_logN) *max_logN = b->L.fused_max_log;
    return BW_OK;
}

extern "C" int bw_set_fused_limit(int max_log) {
    if (max_log < 0 || max_log > BW_FUSED_MAX_LOG)
        return fail(BW_ERR_ARG, "fused limit %d (0 .. %d)", max_log, BW_FUSED_MAX_LOG);
    return set_fused_limit(max_log);
}

extern "C" int bw_batch_set_graph(bw_batch* b, int enabled) {
    if (!b) return fail(BW_ERR_ARG, "batch is NULL");
    if (enabled < 0 || enabled > 2) return fail(BW_ERR_ARG, "graph mode %d (0, 1 or 2)", enabled);
    const bool branches = enabled == 2;
    if (branches != b->branches || enabled == 0) drop_graph(b);
    b->use_graph = enabled != 0;
    b->branches = branches;
    if (branches) return make_branches(b);
    return BW_OK;
}

extern "C" int bw_batch_execute(bw_batch* b, const void* d_a, const void* d_b, void* d_c, void* stream) {
    if (!b) return fail(BW_ERR_ARG, "batch is NULL");
    if (!d_a || !d_b || !d_c) return fail(BW_ERR_ARG, "NULL device buffer");
    hipStream_t st = (hipStream_t)stream;
    const BatchBranches* br = b->branches ? &b->br : nullptr;
    if (!b->use_graph || st == nullptr) return batch_issue(b->L, b->ws, d_a, d_b, d_c, st);
    if (!b->exec || b->gA != d_a || b->gB != d_b || b->gC != d_c || b->gst != st) {
        drop_graph(b);
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = batch_issue(b->L, b->ws, d_a, d_b, d_c, st, br);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(BW_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(b);
            return fail(BW_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
        b->gA = d_a;
        b->gB = d_b;
        b->gC = d_c;
        b->gst = st;
    }
    HIP_TRY(hipGraphLaunch(b->exec, st));
    return BW_OK;
}

extern "C" int bw_batch_status(bw_batch* b) {
    if (!b) return fail(BW_ERR_ARG, "batch is NULL");
    int err = 0;
    int* word = reinterpret_cast<int*>(static_cast<unsigned char*>(b->ws) + b->L.off_err);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(&err, word, sizeof(int), hipMemcpyDeviceToHost));
    if (!err) return BW_OK;
    // reported once: the next execution starts with a clear word
    HIP_TRY(hipMemset(word, 0, sizeof(int)));
    return fail(BW_ERR_STATE, "carry look-back wait exceeded its bound: the batch's products are not valid");
}

namespace bw {
// Candidate (pass split, tile cap) pairs for transforms of 2^logN points.
void tune_candidates(int logN, std::vector<TuneChoice>& out) {
    out.clear();
    auto add = [&](std::vector<int> p) {
        for (int gs = 0; gs <= 2; gs++) {
            TuneChoice c;
            c.passes = p;
            c.gshift = gs;
            out.push_back(c);
        }
    };
    if (logN <= 10) add({logN});
    for (int a = 2; a <= 10; a++) {
        const int r = logN - a;
        if (r >= 2 && r <= 10 && std::abs(a - r) <= 4) add({a, r});
    }
    // a row-resident last pass of 2^14 or 2^15 points (FP64 path; filtered by
    // split_supported for the others)
    for (int t = BW_ROW_MIN_LOG; t <= BW_ROW_MAX_LOG; t++)
        if (logN - t >= 2 && logN - t <= 10) add({logN - t, t});
    if (logN >= 15) {
        for (int a = 4; a <= 10; a++)
            for (int c2 = 4; c2 <= 10; c2++) {
                const int r = logN - a - c2;
                if (r >= 4 && r <= 10 && std::max(a, std::max(c2, r)) - std::min(a, std::min(c2, r)) <= 2)
                    add({a, c2, r});
            }
    }
}
}  // namespace bw

extern "C" int bw_batch_tune(bw_batch* b, const void* d_a, const void* d_b, void* stream, int reps) {
    if (!b) return fail(BW_ERR_ARG, "batch is NULL");
    if (!d_a || !d_b) return fail(BW_ERR_ARG, "NULL device buffer");
    if (reps < 1) reps = 3;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = BW_OK;
    BatchLayout& L = b->L;
    std::vector<TuneChoice> cands;
    for (auto& g : L.groups) {
        tune_candidates(g.logN, cands);
        TuneChoice best;
        best.ms = -1.0f;
        const int64_t m = (int64_t)g.members.size();
        for (auto& c : cands) {
            if (!split_supported(c.passes, g.limb_bits, g.ingress, nullptr)) continue;
            BatchGroup t = g;
            t.passes = c.passes;
            t.gshift = c.gshift;
            group_geometry(t);
            if (t.B > g.cap) continue;           // more slots than the workspace holds
            float ms_min = -1.0f;
            for (int r = 0; r < reps + 1 && !rc; r++) {
                if ((rc = (hipEventRecord(e0, st) == hipSuccess ? BW_OK : fail(BW_ERR_HIP, "hipEventRecord"))))
                    break;
                if ((rc = batch_issue_group(L, t, b->ws, d_a, d_b, st))) break;
                if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = fail(BW_ERR_HIP, "tuning event");
                    break;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (r > 0 && (ms_min < 0 || ms < ms_min)) ms_min = ms;   // first run warms up
            }
            if (rc) break;
            if (best.ms < 0 || ms_min < best.ms) {
                best = c;
                best.ms = ms_min;
            }
        }
        if (rc) break;
        if (best.ms >= 0) {
            g.passes = best.passes;
            g.gshift = best.gshift;
            group_geometry(g);
            tune_store(TuneKey{g.set_id, g.logN, log2_bucket(m)}, best);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graph(b);
    // slot counts may have changed: descriptors follow the tuned geometry
    if (!rc) {
        HIP_TRY(hipStreamSynchronize(st));
        rc = upload_descriptors(b);
    }
    return rc;
}

extern "C" int bw_tune_count(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return (int)g_tune.size();
}

extern "C" void bw_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}
