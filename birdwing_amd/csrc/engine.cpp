// engine.cpp -- errors, table registry and the stacked transform sequence.
#include "engine.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "../../include/birdwing.h"

namespace bw {

// ---------------------------------------------------------------------------
// errors
static thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// per (device, set) tables, built from one host image per set (made here, or
// loaded from a saved image)
static std::mutex g_tab_mu;
static std::map<std::pair<int, int>, std::unique_ptr<DevTables>> g_tabs;
static std::map<int, HostTables> g_host;     // guarded by g_tab_mu

static const HostTables& host_tables_locked(const PrimeSet& s) {
    auto it = g_host.find(s.id);
    if (it == g_host.end()) {
        HostTables h;
        build_tables(s, h);
        it = g_host.emplace(s.id, std::move(h)).first;
    }
    return it->second;
}

int tables_save(int set_id, const char* path) {
    const PrimeSet* s = prime_set(set_id);
    if (!s) return fail(BW_ERR_ARG, "unknown prime set %d", set_id);
    if (!path) return fail(BW_ERR_ARG, "NULL path");
    std::lock_guard<std::mutex> lk(g_tab_mu);
    std::string why;
    if (!write_table_image(path, set_id, host_tables_locked(*s), &why)) return fail(BW_ERR_ARG, "%s", why.c_str());
    return BW_OK;
}

int tables_load(const char* path, int* set_id) {
    if (!path) return fail(BW_ERR_ARG, "NULL path");
    HostTables h;
    int id = -1;
    std::string why;
    if (!read_table_image(path, &id, h, &why)) return fail(BW_ERR_ARG, "table image %s: %s", path, why.c_str());
    std::lock_guard<std::mutex> lk(g_tab_mu);
    g_host[id] = std::move(h);
    if (set_id) *set_id = id;
    return BW_OK;
}

int find_tables(int set_id, DevTables** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tabs.find(std::make_pair(dev, set_id));
    if (it == g_tabs.end())
        return fail(BW_ERR_STATE, "tables of prime set %d are not built on device %d (call bw_init first)", set_id, dev);
    *out = it->second.get();
    return BW_OK;
}

int get_tables(int set_id, DevTables** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto key = std::make_pair(dev, set_id);
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) { *out = it->second.get(); return BW_OK; }
    const PrimeSet* s = prime_set(set_id);
    if (!s) return fail(BW_ERR_ARG, "unknown prime set %d", set_id);
    const HostTables& h = host_tables_locked(*s);
    std::unique_ptr<DevTables> d(new DevTables());
    d->t.limb_bits = s->limb_bits;
    d->t.k = s->k;
    d->t.table_log = h.table_log;
    void** ptrs[4] = {&d->t.consts, &d->t.tw_stage, &d->t.tw_lo, &d->t.tw_hi};
    const std::vector<unsigned char>* srcs[4] = {&h.consts, &h.tw_stage, &h.tw_lo, &h.tw_hi};
    for (int i = 0; i < 4; i++) {
        HIP_TRY(hipMalloc(ptrs[i], srcs[i]->size()));
        HIP_TRY(hipMemcpy(*ptrs[i], srcs[i]->data(), srcs[i]->size(), hipMemcpyHostToDevice));
    }
    d->garner = garner_constants(*s);
    if (s->k == 5) d->gd = garner_d64_constants(*s);
    for (int i = 0; i < 3; i++) f64_prime(i < s->k ? s->p[i] : 3, d->fp[i]);
    *out = d.get();
    g_tabs[key] = std::move(d);
    return BW_OK;
}

// ---------------------------------------------------------------------------
// pass geometry
void default_passes(int logN, std::vector<int>& out) {
    out.clear();
    // Up to 2^16 the autotuner's choices for stacked batches (measured on the
    // 4096-product mix): one register-staged pass up to 2^10, [7, 4] at 2^11,
    // two passes with the larger first from 2^12 to 2^16.
    if (logN <= BW_REG2_MAX_LOG) {
        out.push_back(logN < 1 ? 1 : logN);
        return;
    }
    if (logN == 11) {
        out = {7, 4};
        return;
    }
    if (logN <= 16) {
        out = {(logN + 1) / 2, logN / 2};
        return;
    }
    // Above 2^16, passes of at most 2^9 points; the odd points go to the
    // middle passes first, then the first: a last (row, fused) pass of 2^9
    // points measured 1.7x slower than one of 2^8 (482 against 278 us at
    // 2^24), a middle column pass of 2^9 only ~8% slower per point.
    const int P = (logN + 8) / 9 < 1 ? 1 : (logN + 8) / 9;
    const int base = logN / P;
    int extra = logN % P;
    out.assign(P, base);
    for (int j = 1; j < P - 1 && extra > 0; j++, extra--) out[j]++;
    for (int j = 0; j < P - 1 && extra > 0; j++, extra--) out[j]++;
    if (extra > 0) out[P - 1]++;
}

PassPlan pass_geom(const Xform& x, int j) {
    PassPlan g;
    int prefix = 0;
    for (int i = 0; i < j; i++) prefix += x.passes[i];
    g.logT = x.passes[j];
    g.logM = x.logN - prefix;
    g.S = (int64_t)1 << (x.logN - prefix - g.logT);
    g.O = x.B << prefix;
    int gmax = bw_pass_max_G(g.logT, x.limb_bits, g.S > 1) >> x.gshift;
    if (gmax < 1) gmax = 1;
    if (g.S > 1) {
        g.G = (int)(g.S < gmax ? g.S : gmax);
    } else {
        // rows: G * T <= 4096 keeps the fused kernel's per-thread registers small
        while (gmax > 1 && ((int64_t)gmax << g.logT) > 4096) gmax >>= 1;
        g.G = (int)(g.O < gmax ? g.O : gmax);
        while (g.G > 1 && g.O % g.G) g.G >>= 1;     // whole tiles only
    }
    g.tiles = g.O * g.S / g.G;
    return g;
}

PassParams base_params(const Xform& x, const PassPlan& g) {
    PassParams prm;
    std::memset(&prm, 0, sizeof(prm));
    prm.N = x.N;
    prm.plane = x.N * x.B;
    prm.S = g.S;
    prm.O = g.O;
    prm.logM = g.logM;
    prm.table_log = x.tabs->t.table_log;
    prm.consts = x.tabs->t.consts;
    prm.tw_stage = x.tabs->t.tw_stage;
    prm.tw_lo = x.tabs->t.tw_lo;
    prm.tw_hi = x.tabs->t.tw_hi;
    prm.ninv = x.ninv;
    prm.canon_out = 1;
    return prm;
}

hipError_t launch_pass(const Xform& x, int mode, const PassPlan& g, const PassParams& prm, int nops,
                       hipStream_t st) {
    PassLaunch L;
    L.mode = mode;
    L.logT = g.logT;
    L.G = g.G;
    L.limb_bits = x.limb_bits;
    L.nprimes = x.set->k;
    L.nops = nops;
    L.tiles = g.tiles;
    return bw_launch_pass(L, prm, st);
}

bool split_supported(const std::vector<int>& passes, int arith_bits, int ingress, std::string* why) {
    const int P = (int)passes.size();
    if (P < 1 || P > 4) {
        if (why) *why = "1 to 4 passes";
        return false;
    }
    const int maxlog = (ingress == 2 || ingress == 3) ? BW_REG2_MAX_LOG : BW_MAX_PASS_LOG;
    for (int i = 0; i < P; i++) {
        const int t = passes[i];
        const bool row = i == P - 1 && P >= 2 && arith_bits == 32 && t >= BW_ROW_MIN_LOG && t <= BW_ROW_MAX_LOG;
        if (t < 1 || (t > maxlog && !row)) {
            if (why) *why = "pass " + std::to_string(i) + " of 2^" + std::to_string(t) + " points is outside its kernels";
            return false;
        }
    }
    return true;
}

int count_transform_launches(const Xform& x) {
    const int P = (int)x.passes.size();
    if (P == 1) return 1;
    const int k = x.set->k;
    if (x.schedule == BW_SCHED_PRIME_MAJOR) return k * (2 * (P - 1) + 1 + (P - 1));
    if (x.schedule == BW_SCHED_HYBRID) return (P - 1) + k * (1 + (P - 1));
    return (P - 1) + 1 + (P - 1);
}

#define MARK()                                  \
    do {                                        \
        if (mk) LAUNCH_TRY(mk->mark());         \
    } while (0)

// BW_SCHED_PASS_MAJOR: one launch per pass over every (prime, operand).
// BW_SCHED_PRIME_MAJOR: the chain of one prime -- forward passes of A_p, of
// B_p, the fused pass, the inverse passes -- runs before the next prime
// starts, so each intermediate plane is re-read while the 256 MB Infinity
// Cache still holds it.
// BW_SCHED_HYBRID: forward passes pass-major (every prime and both operands in
// one launch), then the fused pass and the inverse passes prime by prime: the
// product plane of one prime (written by the fused pass) is re-read by the
// inverse passes from the Infinity Cache.
int run_transforms(const Xform& x, const void* d_a, const void* d_b, int64_t la, int64_t lb, hipStream_t st,
                   LaunchMarker* mk) {
    const int P = (int)x.passes.size();
    // the final inverse pass stores canonical residues for the three-prime
    // Garner kernels; the five-prime one reads centered residues as they are
    const int final_canon = x.ingress == 2 ? 0 : 1;
    if (P == 1) {
        PassPlan g = pass_geom(x, 0);
        PassParams prm = base_params(x, g);
        prm.canon_out = final_canon;
        prm.ingress = x.ingress;
        prm.in[0] = d_a;
        prm.in[1] = d_b;
        prm.in_len[0] = la;
        prm.in_len[1] = lb;
        prm.bdesc = x.bdesc;
        prm.out[0] = x.bufA;
        LAUNCH_TRY(launch_pass(x, BW_MODE_FUSED, g, prm, 1, st));
        MARK();
        return BW_OK;
    }
    // stacked ingress names the operand by launch slot (A in 0, B in 1)
    if (x.bdesc && x.schedule == BW_SCHED_PRIME_MAJOR)
        return fail(BW_ERR_ARG, "stacked ingress needs a pass-major forward schedule");
    const bool folded = pass_geom(x, P - 1).logT <= BW_REG2_MAX_LOG && pass_geom(x, P - 2).logT <= BW_REG2_MAX_LOG;
    const int k = x.set->k;
    const void* raw[2] = {d_a, d_b};
    void* buf[2] = {x.bufA, x.bufB};
    const int64_t len[2] = {la, lb};
    // Each launch walks its workgroups in the direction opposite to the
    // launch before it, so it first reads what that launch wrote last: a
    // pass over 2^24-point planes then takes part of its input from the
    // 256 MB Infinity Cache (measured on a memory-only copy of the second
    // forward pass: 270 instead of 290 us).
    int walk = 0;
    auto next_rev = [&]() { return BW_ALTERNATE_WALK ? (walk++ & 1) : 0; };
    // forward passes 1 .. P-1 of primes [prime0, prime0 + nper): one operand
    // (op 0 or 1) or both at once (op < 0)
    auto forward = [&](int prime0, int nper, int op) -> int {
        for (int j = 0; j < P - 1; j++) {
            PassPlan g = pass_geom(x, j);
            PassParams prm = base_params(x, g);
            prm.prime0 = prime0;
            prm.canon_out = 0;
            const int nops = op < 0 ? 2 : 1;
            for (int q = 0; q < nops; q++) {
                const int which = op < 0 ? q : op;
                prm.in[q] = j == 0 ? raw[which] : (const void*)buf[which];
                prm.in_len[q] = len[which];
                prm.out[q] = buf[which];
            }
            if (j == 0) prm.bdesc = x.bdesc;
            prm.ingress = j == 0 ? x.ingress : 0;
            prm.nops = nops;
            prm.rev = next_rev();
            PassLaunch L = {BW_MODE_FWD, g.logT, g.G, x.limb_bits, nper, nops, g.tiles};
            LAUNCH_TRY(bw_launch_pass(L, prm, st));
            MARK();
        }
        return BW_OK;
    };
    // fused pass and inverse passes of primes [prime0, prime0 + nper)
    auto backward = [&](int prime0, int nper) -> int {
        {
            PassPlan g = pass_geom(x, P - 1);
            PassParams prm = base_params(x, g);
            prm.prime0 = prime0;
            prm.in[0] = x.bufA;
            prm.in[1] = x.bufB;
            prm.out[0] = x.bufA;
            prm.canon_out = 0;
            prm.fold_ninv = folded;
            prm.rev = next_rev();
            PassLaunch L = {BW_MODE_FUSED, g.logT, g.G, x.limb_bits, nper, 1, g.tiles};
            LAUNCH_TRY(bw_launch_pass(L, prm, st));
            MARK();
        }
        for (int j = P - 2; j >= 0; j--) {
            PassPlan g = pass_geom(x, j);
            PassParams prm = base_params(x, g);
            prm.prime0 = prime0;
            prm.in[0] = x.bufA;
            prm.out[0] = x.bufA;
            prm.canon_out = j == 0 ? final_canon : 0;
            prm.fold_ninv = (j == P - 2) && folded;
            prm.rev = next_rev();
            PassLaunch L = {BW_MODE_INV, g.logT, g.G, x.limb_bits, nper, 1, g.tiles};
            LAUNCH_TRY(bw_launch_pass(L, prm, st));
            MARK();
        }
        return BW_OK;
    };
    int rc;
    if (x.schedule == BW_SCHED_PRIME_MAJOR) {
        for (int c = 0; c < k; c++) {
            for (int op = 0; op < 2; op++)
                if ((rc = forward(c, 1, op))) return rc;
            if ((rc = backward(c, 1))) return rc;
        }
    } else if (x.schedule == BW_SCHED_HYBRID) {
        if ((rc = forward(0, k, -1))) return rc;
        for (int c = 0; c < k; c++)
            if ((rc = backward(c, 1))) return rc;
    } else {
        if ((rc = forward(0, k, -1))) return rc;
        if ((rc = backward(0, k))) return rc;
    }
    return BW_OK;
}

}  // namespace bw
