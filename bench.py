#!/usr/bin/env python3
"""bench.py -- MI355X big-integer multiply benchmark (one JSON line on stdout).

Default workload (`value`): the plain single product of two 2^24-limb
operands of 32-bit limbs (uniform random, top limb nonzero), inputs resident
in HBM.  It runs under the library's automatic method: five primes below 2^31
on 64-bit digits (pairs of 32-bit limbs), N = 2^24 digits per prime, exact
integer residues in FP64 arithmetic (`config.method`).  A step is one product;
K steps are timed between device synchronizations.

Beside it, on the same GPU in the same run (`legs`): the same product with
64-bit limbs (five primes, N = 2^24), and the reference pipelines' configured
three-prime methods at N = 2^25 -- the wide ~30-bit set on 32-bit limbs and
the reference 58-bit set on 64-bit limbs -- each with per-kernel GB/s and the
whole product against the staged byte model; the PCIe-inclusive plain time
(host limb arrays in and out, `plain_host`); the batched and served workloads.

Reported beside products/s:
  * p50/p99 latency of the timed products (hipEvents per step),
  * roofline of the dominant kernel: its algorithmic bytes per launch (the
    staged one-pass byte model: each stage reads its logical inputs once and
    writes its outputs once) over its average launch duration, measured live
    with hipEvents on the launch stream during the timed region,
  * the whole product's algorithmic bytes under the staged model of the
    reference method (136 B per coefficient on 32-bit limbs, 248 B on 64-bit)
    over the time per product,
  * bit-exactness of the GPU product against GMP mpz_mul,
  * CPU baselines on this host (core count and CPU model recorded): the
    oracle's NTT/Garner/carry pipeline (a port of the reference's CPU path)
    and GMP mpz_mul on one bounded sample of the same operands, and for the
    batched and served workloads both as thread pools (one product per
    thread).

`--gpus N` with N > 1 exits: this build measures one GPU.
`--valu-peak`, `--mfma-dft16` and `--hbm-peak` run the rate probes instead.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def synthetic_limbs(n, seed, bits):
    """Uniform limbs from a seeded generator, top limb forced nonzero."""
    rng = np.random.default_rng(seed)
    dt = np.uint64 if bits == 64 else np.uint32
    a = rng.integers(0, 2**bits, n, dtype=np.uint64, endpoint=False).astype(dt)
    if a[-1] == 0:
        a[-1] = 1
    return a


def launch_names(plan, schedule):
    """(name, operand-planes) per launch, in issue order."""
    P = len(plan.passes)
    logs = plan.passes
    k = plan.k
    if P == 1:
        return [("fused_pass(T=2^%d)" % logs[0], k), ("garner_carry", k)]
    out = []

    def forward(tag, kk, ops):
        for o in ops:
            for j in range(P - 1):
                out.append(("fwd_pass%d(T=2^%d)%s%s" % (j + 1, logs[j], tag, o), kk * (2 if o == "AB" else 1)))

    def backward(tag, kk):
        out.append(("fused_pass%d(T=2^%d)%s" % (P, logs[P - 1], tag), kk))
        for j in range(P - 2, -1, -1):
            out.append(("inv_pass%d(T=2^%d)%s" % (j + 1, logs[j], tag), kk))

    if schedule == 1:          # prime-major
        for c in range(k):
            forward("[p%d]" % c, 1, ["A", "B"])
            backward("[p%d]" % c, 1)
    elif schedule == 2:        # hybrid: forward pass-major, the rest per prime
        forward("", k, ["AB"])
        for c in range(k):
            backward("[p%d]" % c, 1)
    else:
        forward("", k, ["AB"])
        backward("", k)
    out.append(("garner_carry", k))
    return out


def launch_bytes(plan, bits, names):
    """Algorithmic bytes per launch (staged one-pass model) for what the launch
    processes (k primes of N_exec residue words; 4-byte words for 32-bit limbs
    and for 64-bit digits, 8-byte words on the 64-bit-prime path).

    fwd pass 1 (ingress):  read A,B limbs, write 2k spectra      = S1
    fwd pass j > 1:        read + write 2k planes
    fused last pass:       read 2k planes, write k planes          = S2
    inv pass j > 1:        read + write k planes
    inv pass 1:            read + write k planes                   = S3
    garner_carry:          read k planes, write L_A+L_B limbs      = S4+S5 without the
                           wide-coefficient round trip the fusion removes
    """
    wl = bits // 8
    w = 4 if (bits == 32 or plan.digit_bits == 64) else 8
    N = plan.N_exec
    L = plan.L_A + plan.L_B
    P = len(plan.passes)
    out = []
    for name, planes in names:
        if name.startswith("garner"):
            out.append(plan.k * N * w + L * wl)
        elif name.startswith("fused"):
            out.append((L * wl + planes * N * w) if P == 1 else 3 * planes * N * w)
        elif name.startswith("fwd_pass1"):
            nop = 2 if name.endswith("AB") else 1
            raw = L * wl if nop == 2 else (plan.L_A if name.endswith("A") else plan.L_B) * wl
            out.append(raw + planes * N * w)
        else:
            out.append(2 * planes * N * w)
    return out


def model_bytes(plan, bits):
    """Whole-product bytes of the staged model of the reference method (fixed,
    whatever this library runs): three primes on digits of the limb width,
    N = next_pow2(L_A + L_B - 1), (L_A+L_B)(w_in+w_out) + 8kNw_r + 2Nw_c with
    (w_in, w_r, w_c, w_out) = (4, 4, 16, 4) B for 32-bit limbs, (8, 8, 20, 8) B for 64-bit."""
    import birdwing_amd as bw
    w = bits // 8
    wc = 16 if bits == 32 else 20
    N = bw.padded_ntt_size(plan.L_A, plan.L_B)
    return (plan.L_A + plan.L_B) * 2 * w + 8 * 3 * N * w + 2 * N * wc


def load_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` (a bench label) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, profiles/traffic_<workload>.json
    (counter values scaled by the copy-kernel calibration recorded there), or
    None when that kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return k["hbm_bytes_per_launch"] if k else None


def cpu_baselines(A, B, bits, gpu_out, sample_oracle, min_seconds=10.0):
    """Oracle pipeline (port of the reference's CPU path: NTT per prime on one
    thread each, Garner, carry) and GMP mpz_mul on this host, each repeated on
    the same operands until about min_seconds of CPU work (at least once)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    res = {}
    ref = None
    if sample_oracle:
        n, t0 = 0, time.perf_counter()
        while n == 0 or time.perf_counter() - t0 < min_seconds:
            ref = pyoracle.multiply(A, B, bits)
            n += 1
        t = time.perf_counter() - t0
        res["oracle"] = {"value": n / t, "unit": "products/s", "cores": 3, "kind": "port",
                         "sample": "%d product(s) of the same %dx%d-limb operands (%d-bit limbs; one prime per "
                                   "thread), %.1f s" % (n, len(A), len(B), bits, t)}
    exact = None
    if pyoracle.gmp_available():
        n, tg, g = 0, 0.0, None
        while n == 0 or tg < min_seconds:
            g, dt = pyoracle.gmp_mul(A, B, bits)
            tg += dt
            n += 1
        res["gmp"] = {"value": n / tg, "unit": "products/s", "cores": 1, "kind": "library",
                      "sample": "%d mpz_mul of the same operands, %.2f s (import/export excluded)" % (n, tg)}
        exact = bool(np.array_equal(g, gpu_out))
        if ref is not None:
            exact = exact and bool(np.array_equal(ref, gpu_out))
    elif ref is not None:
        exact = bool(np.array_equal(ref, gpu_out))
    return res, exact


def product_model_bytes(LA, LB, bits, k=3):
    """Staged-model bytes of one product (32-bit: 136 B per coefficient + limbs)."""
    import birdwing_amd as bw
    w = bits // 8
    wc = 16 if bits == 32 else 20
    N = bw.padded_ntt_size(int(LA), int(LB))
    return (int(LA) + int(LB)) * 2 * w + 8 * k * N * w + 2 * N * wc


def bench_batch(n, steps, warmup, seed=4096, host_sample=128, tune=True, compare=True, pool_seconds=8.0,
                graph_mode=1):
    """Batched workload: n products, limb counts per operand log-uniform in
    [2^8, 2^16] (seeded), uniform random 32-bit limbs (top limb nonzero);
    autotuned, hipGraph-captured; against the plain per-product path on the
    same products (device-resident Plan.execute, plans made beforehand) and
    the host one-shot path (plan, copies, execute, release) on a sample.
    tune=False keeps the default splits and compare=False skips both plain
    paths (a profiled run then traces the batch launches alone)."""
    import torch
    import birdwing_amd as bw
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    rng = np.random.default_rng(seed)
    LA = np.floor(2.0 ** rng.uniform(8, 16, n)).astype(np.int64)
    LB = np.floor(2.0 ** rng.uniform(8, 16, n)).astype(np.int64)
    b = bw.Batch(32, LA, LB)
    pa = rng.integers(0, 2**32, b.bytes_a // 4, dtype=np.uint64).astype(np.uint32)
    pb = rng.integers(0, 2**32, b.bytes_b // 4, dtype=np.uint64).astype(np.uint32)
    for i in range(n):                  # top limbs nonzero
        pa[(b.off_a[i] >> 2) + LA[i] - 1] |= 1
        pb[(b.off_b[i] >> 2) + LB[i] - 1] |= 1
    da = torch.from_numpy(pa.view(np.uint8)).cuda()
    db = torch.from_numpy(pb.view(np.uint8)).cuda()
    dc = torch.zeros(b.bytes_c, dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    t_tune = time.perf_counter()
    if tune:
        with torch.cuda.stream(s):
            b.tune(da, db, stream=s, reps=2)
    t_tune = time.perf_counter() - t_tune
    def timed(mode):
        b.set_graph(mode)
        with torch.cuda.stream(s):
            for _ in range(warmup):
                b.execute(da, db, dc, stream=s)
            s.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
            t0 = time.perf_counter()
            ev[0].record(s)
            for i in range(steps):
                b.execute(da, db, dc, stream=s)
                ev[i + 1].record(s)
            s.synchronize()
            t1 = time.perf_counter()
        b.status()
        return (t1 - t0) / steps, sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))

    # the other graph form first, for comparison; the reported one last
    other = 1 if graph_mode == 2 else 2
    other_s, _ = timed(other)
    batch_s, lat = timed(graph_mode)
    out = dc.cpu().numpy()
    got = b.unpack(out)
    mbytes = sum(product_model_bytes(x, y, 32) for x, y in zip(LA, LB))
    groups = b.groups()
    res = {
        "workload": "batch_%d_logu_2^8-2^16_u32" % n,
        "products": n,
        "steps": steps,
        "products_per_s": n / batch_s,
        "ms_per_batch": batch_s * 1e3,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "alg_bytes": mbytes,
        "achieved": mbytes / batch_s / 1e9,
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": mbytes / batch_s / 1e9 / HBM_PEAK_GBPS,
        "graph": ["none", "one chain", "a branch per size group"][graph_mode],
        "ms_per_batch_other_graph": {"graph": ["none", "one chain", "a branch per size group"][other],
                                     "ms": other_s * 1e3},
        "launches": int(b.info.nlaunch),
        "fused_products": b.fused()[0],
        "autotuned": tune,
        "autotune_s": t_tune,
        "groups": [{"logN": g["logN"], "members": g["members"], "passes": g["passes"], "gshift": g["gshift"]}
                   for g in groups],
    }
    if not compare:
        return res

    # plain per-product path on the same products, same device buffers
    plans = [bw.Plan(32, int(LA[i]), int(LB[i])) for i in range(n)]
    av = [da[b.off_a[i]:b.off_a[i] + 4 * int(LA[i])] for i in range(n)]
    bv = [db[b.off_b[i]:b.off_b[i] + 4 * int(LB[i])] for i in range(n)]
    pc = torch.zeros(2 * (1 << 16) + 8, dtype=torch.int32, device="cuda")
    with torch.cuda.stream(s):
        for i in range(min(n, 64)):
            plans[i].execute(av[i], bv[i], pc, stream=s)
        s.synchronize()
        same = True
        t2 = time.perf_counter()
        for i in range(n):
            plans[i].execute(av[i], bv[i], pc, stream=s)
        s.synchronize()
        t3 = time.perf_counter()
        for i in range(0, n, max(1, n // 64)):      # the plain path agrees with the batch
            plans[i].execute(av[i], bv[i], pc, stream=s)
            s.synchronize()
            m = int(LA[i] + LB[i])
            same = same and bool(np.array_equal(pc[:m].cpu().numpy().view(np.uint32), got[i]))
    plain_s = (t3 - t2) / n
    del plans
    if pool_seconds > 0:           # CPU thread pools over the same products
        prods = [(32, pa[(b.off_a[i] >> 2):(b.off_a[i] >> 2) + LA[i]], pb[(b.off_b[i] >> 2):(b.off_b[i] >> 2) + LB[i]])
                 for i in range(n)]
        res["cpu_pool"] = cpu_pool(prods, pool_seconds, res["workload"])
    # host one-shot path (the reference's host_multiply_merge shape) on a sample
    idx = list(range(0, n, max(1, n // host_sample)))[:host_sample]
    t4 = time.perf_counter()
    exact = True
    import pyoracle
    for j, i in enumerate(idx):
        A = pa[(b.off_a[i] >> 2):(b.off_a[i] >> 2) + LA[i]]
        B = pb[(b.off_b[i] >> 2):(b.off_b[i] >> 2) + LB[i]]
        r = bw.multiply(A, B, 32)
        if j < 16 and pyoracle.gmp_available():
            g, _ = pyoracle.gmp_mul(A, B, 32)
            exact = exact and bool(np.array_equal(g, r)) and bool(np.array_equal(g, got[i]))
    host_s = (time.perf_counter() - t4) / len(idx)
    res.update({
        "plain_device_products_per_s": 1.0 / plain_s,
        "plain_host_products_per_s": 1.0 / host_s,
        "vs_plain_device": plain_s / batch_s * n,
        "vs_plain_host": host_s / batch_s * n,
        "bit_exact_vs_gmp_sample": exact,
        "same_as_plain_sample": same,
    })
    return res


def bench_fused_classes(steps=10):
    """Batches of equal products, one size class each (32-bit limbs, L x L),
    through the fused whole-product launch (limit 11) and through stacked
    transforms (limit 0, autotuned splits), hipGraph replays; both outputs
    compared limb for limb."""
    import torch
    import birdwing_amd as bw
    out = []
    for L, n in ((2048, 1024), (1024, 2048), (512, 2048), (256, 4096), (128, 4096), (64, 4096), (16, 8192)):
        row = {"limbs": L, "products": n}
        outs = []
        for lim in (11, 0):
            old = bw.set_fused_limit(lim)
            b = bw.Batch(32, [L] * n, [L] * n)
            bw.set_fused_limit(old)
            rng = np.random.default_rng(L)
            da = torch.from_numpy(rng.integers(0, 256, b.bytes_a, dtype=np.uint8)).cuda()
            db = torch.from_numpy(rng.integers(0, 256, b.bytes_b, dtype=np.uint8)).cuda()
            dc = torch.zeros(b.bytes_c, dtype=torch.uint8, device="cuda")
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                if lim == 0:
                    b.tune(da, db, stream=s, reps=2)
                b.set_graph(1)
                for _ in range(2):
                    b.execute(da, db, dc, stream=s)
                s.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    b.execute(da, db, dc, stream=s)
                s.synchronize()
                t = (time.perf_counter() - t0) / steps
            b.status()
            nf, lg = b.fused()
            row["logN"] = lg if nf else [g["logN"] for g in b.groups()][0]
            row["fused_ms" if lim else "stacked_ms"] = t * 1e3
            if not lim:
                row["stacked_passes"] = [g["passes"] for g in b.groups()][0]
            outs.append(dc.cpu().numpy())
        row["fused_vs_stacked"] = row["stacked_ms"] / row["fused_ms"]
        row["same_limbs"] = bool(np.array_equal(outs[0], outs[1]))
        out.append(row)
        progress("fused classes: %d x %d limbs done" % (L, L))
    return {"probe": "fused_classes", "unit": "ms per batch", "steps": steps, "classes": out}


class _ServeMix:
    """Serving workload: limb counts log-uniform in [2^6, 2^22] per operand,
    80% 32-bit / 20% 64-bit digits, uniform random limbs taken as slices of
    pinned random pools; results land in a pinned ring (sampled requests get
    their own arrays for checking)."""

    def __init__(self, seed, ring_bytes=1 << 30):
        import torch
        self.rng = np.random.default_rng(seed)
        words = 1 << 23
        self.pool = {}
        for bits, dt, tdt in ((32, np.uint32, torch.int32), (64, np.uint64, torch.int64)):
            pa = torch.empty(words, dtype=tdt, pin_memory=True).numpy().view(dt)
            pb = torch.empty(words, dtype=tdt, pin_memory=True).numpy().view(dt)
            pa[:] = self.rng.integers(0, 2**bits, words, dtype=np.uint64, endpoint=False).astype(dt)
            pb[:] = self.rng.integers(0, 2**bits, words, dtype=np.uint64, endpoint=False).astype(dt)
            pa[pa == 0] = 1
            pb[pb == 0] = 1
            self.pool[bits] = (pa, pb)
        self.ring = torch.empty(ring_bytes, dtype=torch.uint8, pin_memory=True).numpy()
        self.cur = 0

    def next(self):
        bits = 64 if self.rng.random() < 0.2 else 32
        la = int(2.0 ** self.rng.uniform(6, 22))
        lb = int(2.0 ** self.rng.uniform(6, 22))
        pa, pb = self.pool[bits]
        oa = int(self.rng.integers(0, len(pa) - la))
        ob = int(self.rng.integers(0, len(pb) - lb))
        return bits, pa[oa:oa + la], pb[ob:ob + lb]

    def out(self, n, bits):
        nb = n * bits // 8
        if self.cur + nb > len(self.ring):
            self.cur = 0
        v = self.ring[self.cur:self.cur + nb].view(np.uint64 if bits == 64 else np.uint32)
        self.cur = (self.cur + nb + 255) // 256 * 256
        return v


def bench_serve(seconds, seed=7, sat_seconds=10.0, plain_seconds=10.0, pool_seconds=8.0, sweep_mults=(2.0, 4.0, 8.0),
                sweep_seconds=6.0):
    """Poisson request stream through the server at 80% of the plain path's
    sustained rate, a saturated run, and the plain per-request path (host
    one-shot multiply: plan, copies, execute, release per request)."""
    import threading
    import birdwing_amd as bw
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    mix = _ServeMix(seed)
    # plain per-request path, closed loop
    lat_plain = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < plain_seconds:
        bits, A, B = mix.next()
        t = time.perf_counter()
        bw.multiply(A, B, bits)
        lat_plain.append((time.perf_counter() - t) * 1e3)
    plain_rate = len(lat_plain) / (time.perf_counter() - t0)

    def run(rate, duration, max_outstanding=None, check_every=0):
        srv = bw.Server(device_bytes=32 << 30, streams=3, max_batch=1024)
        tickets = []
        checks = []
        lock = threading.Lock()
        done = threading.Event()
        lats = []

        def producer():
            t_start = time.perf_counter()
            nxt = t_start
            i = 0
            rng = np.random.default_rng(seed + 1)
            while True:
                now = time.perf_counter()
                if now - t_start >= duration:
                    break
                if rate is not None:
                    if now < nxt:
                        time.sleep(min(nxt - now, 0.001))
                        continue
                    nxt += rng.exponential(1.0 / rate)
                elif max_outstanding is not None:
                    with lock:
                        outstanding = len(tickets) - len(lats)
                    if outstanding >= max_outstanding:
                        time.sleep(0.0002)
                        continue
                bits, A, B = mix.next()
                if check_every and i % check_every == 0:
                    out = np.zeros(len(A) + len(B), dtype=A.dtype)
                    checks.append((bits, A.copy(), B.copy(), out))
                else:
                    out = mix.out(len(A) + len(B), bits)
                t = srv.submit(A, B, out, bits)
                with lock:
                    tickets.append(t)
                i += 1
            done.set()

        def waiter():
            j = 0
            while True:
                with lock:
                    have = j < len(tickets)
                    t = tickets[j] if have else None
                if not have:
                    if done.is_set():
                        with lock:
                            if j >= len(tickets):
                                return
                    time.sleep(0.0001)
                    continue
                lat = srv.wait(t)
                with lock:
                    lats.append(lat)
                j += 1

        th_p = threading.Thread(target=producer)
        th_w = threading.Thread(target=waiter)
        t_begin = time.perf_counter()
        th_p.start()
        th_w.start()
        th_p.join()
        th_w.join()
        elapsed = time.perf_counter() - t_begin
        st = srv.stats()
        srv.close()
        return len(lats) / elapsed, lats, st, checks

    progress("serve: plain per-request path %.0f products/s" % plain_rate)
    served_rate, lat_s, st_s, checks = run(0.8 * plain_rate, seconds, check_every=50)
    progress("serve: Poisson run done")
    sat_rate, lat_sat, st_sat, _ = run(None, sat_seconds, max_outstanding=256)
    progress("serve: saturated run done")
    # arrival rates above what the plain path sustains: continuous batching
    # absorbs them (batch size grows with the queue)
    sweep = []
    for mult in sweep_mults:
        r, l, st, _ = run(mult * plain_rate, sweep_seconds)
        sweep.append({"arrival_rate": mult * plain_rate, "x_plain": mult, "products_per_s": r,
                      "latency_ms": {"p50": float(np.percentile(l, 50)), "p99": float(np.percentile(l, 99))},
                      "mean_batch": st["batched_products"] / max(1, st["batches"])})
    if sweep_mults:
        progress("serve: arrival-rate sweep done")
    # sampled results of the Poisson run equal the same products alone
    import pyoracle
    exact = True
    for bits, A, B, out in checks:
        alone = bw.multiply(A, B, bits)
        exact = exact and bool(np.array_equal(alone, out))
        if len(A) + len(B) <= (1 << 17) and pyoracle.gmp_available():
            g, _ = pyoracle.gmp_mul(A, B, bits)
            exact = exact and bool(np.array_equal(g, out))
    pct = lambda v, q: float(np.percentile(v, q)) if v else None
    pool = None
    if pool_seconds > 0:           # CPU thread pools over requests of the same mix
        pmix = _ServeMix(seed + 2, ring_bytes=1 << 20)
        pool = cpu_pool([pmix.next() for _ in range(4096)], pool_seconds, "serve_logu_2^6-2^22_80u32_20u64")
    return {
        "workload": "serve_poisson_logu_2^6-2^22_80u32_20u64",
        "seconds": seconds,
        "arrival_rate": 0.8 * plain_rate,
        "products_per_s": served_rate,
        "latency_ms": {"p50": pct(lat_s, 50), "p99": pct(lat_s, 99)},
        "requests": len(lat_s),
        "batches": int(st_s["batches"]),
        "mean_batch": st_s["batched_products"] / max(1, st_s["batches"]),
        "max_concurrent_batches": int(st_s["max_concurrent_batches"]),
        "saturated_products_per_s": sat_rate,
        "saturated_latency_ms": {"p50": pct(lat_sat, 50), "p99": pct(lat_sat, 99)},
        "saturated_mean_batch": st_sat["batched_products"] / max(1, st_sat["batches"]),
        "plain_products_per_s": plain_rate,
        "plain_latency_ms": {"p50": pct(lat_plain, 50), "p99": pct(lat_plain, 99)},
        "saturated_vs_plain": sat_rate / plain_rate,
        "sweep": sweep,
        "same_as_alone_sample": exact,
        "checked": len(checks),
        "cpu_pool": pool,
    }


def progress(msg):
    """One line on stderr when a stage of the run ends (stdout keeps the JSON line)."""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def host_info():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def pool_cores():
    """Threads for the CPU thread pools: the CPUs this process may use, at
    most 16 (a GPU box's CPU share per GPU)."""
    return max(1, min(16, host_info()["cpus_available"] or 1))


def cpu_pool(products, seconds, label):
    """GMP mpz_mul and the oracle pipeline (one thread per product, no
    threads inside a product) over `products` -- a list of (bits, A, B) --
    on a thread pool of pool_cores() threads, each until about `seconds` of
    wall time.  ctypes releases the GIL inside the C calls, so the threads
    run in parallel.  Returns {"gmp": {...}, "oracle": {...}}."""
    import concurrent.futures as cf
    import threading
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cores = pool_cores()
    out = {}
    fns = [("oracle", lambda b, x, y: pyoracle.multiply(x, y, b))]
    if pyoracle.gmp_available():
        fns.insert(0, ("gmp", lambda b, x, y: pyoracle.gmp_mul(x, y, b)))
    pyoracle.set_prime_threads(False)
    try:
        for name, fn in fns:
            lock = threading.Lock()
            nxt = [0]
            done = [0]
            t0 = time.perf_counter()

            def worker():
                while time.perf_counter() - t0 < seconds:
                    with lock:
                        i = nxt[0]
                        nxt[0] += 1
                    b, x, y = products[i % len(products)]
                    fn(b, x, y)
                    with lock:
                        done[0] += 1

            with cf.ThreadPoolExecutor(max_workers=cores) as ex:
                for f in [ex.submit(worker) for _ in range(cores)]:
                    f.result()
            t = time.perf_counter() - t0
            out[name] = {"value": done[0] / t, "unit": "products/s", "cores": cores,
                         "kind": "library" if name == "gmp" else "port",
                         "sample": "%d products of the %s workload in %.1f s, one product per thread" % (
                             done[0], label, t)}
    finally:
        pyoracle.set_prime_threads(True)
    out["host"] = host_info()
    return out


def bench_plain(A, B, bits, pset=-1, passes=None, schedule=-1, steps=20, warmup=3, tune=True, workload=None):
    """One product shape, inputs resident in HBM: autotune (unless told not
    to), W warm-up products, then K timed products between synchronizations
    (one hipEvent between products for the latency percentiles), then the
    same K products again with a hipEvent after every launch for the
    per-launch times.  The per-launch events stay out of the first run: on a
    product of a few launches of a few microseconds each they add more time
    than the launches take.  Returns (fields, plan facts, the product limbs)."""
    import re
    import torch
    import birdwing_amd as bw
    LA, LB = len(A), len(B)
    tdt = torch.int64 if bits == 64 else torch.int32
    ndt = np.int64 if bits == 64 else np.int32
    da = torch.from_numpy(A.view(ndt)).cuda()
    db = torch.from_numpy(B.view(ndt)).cuda()
    dc = torch.empty(LA + LB, dtype=tdt, device="cuda")
    plan = bw.Plan(bits, LA, LB, pset, passes)
    stream = torch.cuda.current_stream()
    tuned = False
    if not passes and tune:
        plan.tune(da, db, stream, reps=3)      # pass split / tile cap / launch order, before any timing
        tuned = True
    if schedule >= 0:
        plan.set_schedule(schedule)
    sched = int(plan.info.schedule)
    for _ in range(warmup):
        plan.execute(da, db, dc, stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        plan.execute(da, db, dc, stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan.set_timing(True)
    for i in range(steps):
        plan.execute(da, db, dc, stream)
    torch.cuda.synchronize()
    plan.set_timing(False)
    per_launch = plan.timing()
    plan.status()
    ms_step = (t1 - t0) * 1e3 / steps
    lat = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))
    nl = launch_names(plan, sched)
    names = [n for n, _ in nl]
    lbytes = launch_bytes(plan, bits, nl)
    agg = {}                                   # launches of one kernel type, across primes and operands
    for n, t, by in zip(names, per_launch, lbytes):
        key = re.sub(r"\[p\d\]|(?<=\))(A|B|AB)$", "", n)
        a = agg.setdefault(key, [0.0, 0, 0])
        a[0] += t
        a[1] += by
        a[2] += 1
    dom_name = max(agg, key=lambda k: agg[k][0])
    dom_ms, dom_bytes, dom_n = agg[dom_name]
    dom_gbps = dom_bytes / (dom_ms * 1e-3) / 1e9
    workload = workload or "plain_%dx%d_u%d" % (LA, LB, bits)
    mbytes = model_bytes(plan, bits)
    arith = "u64" if plan.prime_set in (bw.SET_U64_REF, bw.SET_HYBRID, bw.SET_U64_PAIRS) else "f64"
    fields = {
        "workload": workload,
        "products_per_s": steps / (t1 - t0),
        "ms_per_product": ms_step,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "dtype": arith,
        "config": {"workload": workload, "mode": "plain", "limbs_a": LA, "limbs_b": LB, "limb_bits": bits,
                   "prime_set": plan.prime_set, "primes": plan.primes, "digit_bits": plan.digit_bits,
                   "N": plan.N, "N_exec": plan.N_exec, "passes": plan.passes,
                   "schedule": ["pass-major", "prime-major", "hybrid"][sched],
                   "tile_shift": int(plan.info.tile_shift), "autotuned": tuned},
        "kernels": [{"name": k, "launches": v[2], "ms_per_product": v[0], "alg_bytes": v[1],
                     "alg_GBps": v[1] / (v[0] * 1e-3) / 1e9} for k, v in agg.items()],
        "roofline": {"bound": "hbm", "kernel": dom_name, "launches_per_product": dom_n,
                     "avg_launch_ms": dom_ms / dom_n, "alg_bytes_per_launch": dom_bytes / dom_n,
                     "achieved": dom_gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": dom_gbps / HBM_PEAK_GBPS, "traffic": load_traffic(workload, dom_name)},
        "product_roofline": {"alg_bytes": mbytes, "achieved": mbytes / (ms_step * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                             "frac": mbytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                             "target_ms_at_half_peak": mbytes / (0.5 * HBM_PEAK_GBPS * 1e9) * 1e3},
    }
    out = dc.cpu().numpy().view(np.uint64 if bits == 64 else np.uint32).copy()
    plan.close()
    del da, db, dc
    torch.cuda.empty_cache()
    return fields, out


METHOD_TEXT = {
    "f64": "primes < 2^31 (%d of them) on %d-bit digits, N = 2^%d per prime, exact integer residues in FP64 "
           "arithmetic",
    "u64": "58-bit primes (%d of them) on %d-bit digits, N = 2^%d per prime, 64-bit integer arithmetic",
}


def method_of(f):
    c = f["config"]
    return METHOD_TEXT[f["dtype"]] % (len(c["primes"]), c["digit_bits"], int(c["N_exec"]).bit_length() - 1)


def plain_host(A, B, bits, reps=5):
    """The PCIe-inclusive plain path: host limb arrays in and out.
    execute_host: a plan made beforehand, then copy in, run, copy out
    (the reference's execute_ntt_multiply bucket, src/gpu_ntt.cu:468-473,
    932-933); one-shot: plan creation, copies, run, release
    (host_multiply_merge, src/host_multiply.cpp:27-86)."""
    import birdwing_amd as bw
    plan = bw.Plan(bits, len(A), len(B))
    plan.execute_host(A, B)
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = plan.execute_host(A, B)
        t.append((time.perf_counter() - t0) * 1e3)
    plan.close()
    t1 = []
    for _ in range(max(2, reps // 2)):
        t0 = time.perf_counter()
        bw.multiply(A, B, bits)
        t1.append((time.perf_counter() - t0) * 1e3)
    return {"workload": "plain_%dx%d_u%d_host_buffers" % (len(A), len(B), bits),
            "execute_host_ms": {"p50": float(np.percentile(t, 50)), "min": min(t)},
            "one_shot_ms": {"p50": float(np.percentile(t1, 50)), "min": min(t1)},
            "products_per_s": 1e3 / float(np.percentile(t, 50)),
            "host_memory": "pageable numpy arrays"}, out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--limbs", type=int, default=1 << 24, help="limbs per operand")
    ap.add_argument("--bits", type=int, default=32, choices=(32, 64))
    ap.add_argument("--prime-set", type=int, default=-1, help="prime set of the main product (-1: automatic)")
    ap.add_argument("--passes", type=str, default="", help="pass split, e.g. 12,13")
    ap.add_argument("--schedule", type=int, default=-1,
                    help="launch order: 0 pass-major, 1 prime-major, 2 hybrid (default: the autotuner's choice)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-tune", action="store_true", help="keep the default pass split (no autotuning)")
    ap.add_argument("--no-oracle", action="store_true", help="GMP only as CPU baseline")
    ap.add_argument("--no-legs", action="store_true", help="skip the 64-bit-limb and three-prime legs")
    ap.add_argument("--leg-steps", type=int, default=10)
    ap.add_argument("--no-host", action="store_true", help="skip the PCIe-inclusive plain path")
    ap.add_argument("--pool-seconds", type=float, default=8.0,
                    help="wall time of each CPU thread-pool baseline (0: skip)")
    ap.add_argument("--batch-n", type=int, default=4096, help="products in the batched workload (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-no-tune", action="store_true", help="batched workload with the default splits")
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="skip the plain-path comparisons of the batched workload (profiling)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="Poisson serving run length (0: skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (HBM counter calibration) and exit")
    ap.add_argument("--valu-peak", action="store_true",
                    help="only measure the VALU rates of FP64 FMA and 32-bit integer multiplies and exit")
    ap.add_argument("--mfma-dft16", action="store_true",
                    help="only time a 16-point DFT stage on int8 MFMA against the VALU codelet and exit")
    ap.add_argument("--hbm-peak", action="store_true",
                    help="only measure HBM copy/read/write streaming rates and a column pass's memory pattern "
                         "without arithmetic, and exit")
    ap.add_argument("--fused-classes", action="store_true",
                    help="only time batches of one size class through the fused whole-product launch "
                         "against stacked transforms and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (--gpus 1); multi-GPU runs are not part of it",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import birdwing_amd as bw

    if args.calibrate_copy > 0:
        n = args.calibrate_copy << 20
        x = torch.zeros(n // 4, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):           # one elementwise kernel: reads n bytes, writes n bytes
            torch.bitwise_not(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": n, "bytes_written": n, "copies": 3}))
        return
    if args.fused_classes:
        print(json.dumps(bench_fused_classes()))
        return
    if args.valu_peak or args.mfma_dft16 or args.hbm_peak:
        from birdwing_amd import probes
        if args.hbm_peak:
            print(json.dumps(probes.hbm_rates()))
        else:
            print(json.dumps(probes.valu_rates() if args.valu_peak else probes.mfma_dft16()))
        return

    L = args.limbs
    bits = args.bits
    passes = [int(x) for x in args.passes.split(",")] if args.passes else None
    A = synthetic_limbs(L, 1, bits)
    B = synthetic_limbs(L, 2, bits)
    f, gpu_out = bench_plain(A, B, bits, args.prime_set, passes, args.schedule, args.steps, args.warmup,
                             not args.no_tune)
    progress("plain product timed (%.3f ms)" % f["ms_per_product"])
    cpu, exact = ({}, None) if args.no_cpu else cpu_baselines(A, B, bits, gpu_out, not args.no_oracle)
    progress("CPU baselines done")
    cfg = dict(f["config"])
    cfg["method"] = method_of(f)
    line = {
        "metric": "products/s batched & served vs plain (bit-exact); p50/p99 latency; % HBM roofline",
        "value": f["products_per_s"],
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f["ms_per_product"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f["dtype"],
        "data": "synthetic",
        "config": cfg,
        "latency_ms": f["latency_ms"],
        "bit_exact_vs_gmp": exact,
        "kernels": f["kernels"],
        "roofline": f["roofline"],
        "product_roofline": f["product_roofline"],
    }
    if "oracle" in cpu:
        line["cpu_baseline"] = dict(cpu["oracle"], **host_info())
    elif "gmp" in cpu:
        line["cpu_baseline"] = dict(cpu["gmp"], **host_info())
    if "gmp" in cpu:
        line["gmp_baseline"] = dict(cpu["gmp"], **host_info())

    if not args.no_host:
        ph, hout = plain_host(A, B, bits)
        ph["same_as_device"] = bool(np.array_equal(hout, gpu_out))
        line["plain_host"] = ph
        progress("host-buffer path timed")

    if not args.no_legs and L == 1 << 24 and bits == 32 and args.prime_set < 0:
        # the same operands under the reference pipeline's configured method
        # (three ~30-bit primes at N = 2^25); its product must equal the main one
        legs = {}
        f3, out3 = bench_plain(A, B, 32, bw.SET_U32_WIDE, None, -1, args.leg_steps, 2, not args.no_tune,
                               "plain_%dx%d_u32_3prime" % (L, L))
        f3["method"] = method_of(f3)
        f3["same_as_main"] = bool(np.array_equal(out3, gpu_out))
        legs["three_prime_u32"] = f3
        del out3
        progress("three-prime leg (32-bit limbs) timed")
        # 64-bit limbs: the automatic (five-prime) method and the reference 58-bit set
        A64 = synthetic_limbs(L, 3, 64)
        B64 = synthetic_limbs(L, 4, 64)
        f64a, o64a = bench_plain(A64, B64, 64, -1, None, -1, args.leg_steps, 2, not args.no_tune)
        f64a["method"] = method_of(f64a)
        f64b, o64b = bench_plain(A64, B64, 64, bw.SET_U64_REF, None, -1, args.leg_steps, 2, not args.no_tune,
                                 "plain_%dx%d_u64_3prime" % (L, L))
        f64b["method"] = method_of(f64b)
        same = bool(np.array_equal(o64a, o64b))
        f64a["same_as_three_prime"] = f64b["same_as_five_prime"] = same
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle
            if pyoracle.gmp_available():
                g, _ = pyoracle.gmp_mul(A64, B64, 64)
                f64a["bit_exact_vs_gmp"] = bool(np.array_equal(g, o64a))
                f64b["bit_exact_vs_gmp"] = bool(np.array_equal(g, o64b))
                del g
        legs["u64"] = f64a
        legs["three_prime_u64"] = f64b
        progress("64-bit-limb legs timed")
        del A64, B64, o64a, o64b
        # the small plain product (2048 x 2048 limbs, the reference's 32-bit
        # prime set, N = 4096: latency-bound, so p50/p99 per product) and the
        # unbalanced carry-stress shape (2^22 x 2^14 limbs)
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        for name, la, lb, ps, steps in (("small_2048x2048_u32_ref", 2048, 2048, bw.SET_U32_REF, 500),
                                        ("small_2048x2048_u32", 2048, 2048, -1, 500),
                                        ("stress_4194304x16384_u32", 1 << 22, 1 << 14, -1, 20)):
            As = synthetic_limbs(la, 5, 32)
            Bs = synthetic_limbs(lb, 6, 32)
            fs, outs = bench_plain(As, Bs, 32, ps, None, -1, steps, 20, not args.no_tune, name)
            fs["method"] = method_of(fs)
            if pyoracle.gmp_available():
                g, _ = pyoracle.gmp_mul(As, Bs, 32)
                fs["bit_exact_vs_gmp"] = bool(np.array_equal(g, outs))
            legs[name] = fs
            progress("%s timed (p50 %.3f ms)" % (name, fs["latency_ms"]["p50"]))
        line["legs"] = legs
    torch.cuda.empty_cache()
    if args.batch_n > 0:
        line["batched"] = bench_batch(args.batch_n, args.batch_steps, 2, tune=not args.batch_no_tune,
                                      compare=not args.batch_no_plain, pool_seconds=args.pool_seconds)
        torch.cuda.empty_cache()
        progress("batched workload done")
    if args.serve_seconds > 0:
        line["served"] = bench_serve(args.serve_seconds, pool_seconds=args.pool_seconds)
        progress("served workload done")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
