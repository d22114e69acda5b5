#!/usr/bin/env python3
"""bench.py -- MI355X big-integer multiply benchmark (one JSON line on stdout).

Default workload: the plain single product of two 2^24-limb (32-bit limbs)
operands, N = 2^25 per prime under the three ~30-bit primes
{2013265921, 1811939329, 469762049}; inputs resident in HBM.  A step is one
product; K steps are timed between device synchronizations.

Reported beside products/s:
  * p50/p99 latency of the timed products (hipEvents per step),
  * roofline of the dominant kernel: its algorithmic bytes per launch (the
    staged one-pass byte model: each stage reads its logical inputs once and
    writes its outputs once) over its average launch duration, measured live
    with hipEvents on the launch stream during the timed region,
  * the whole product's algorithmic bytes (136 B per coefficient on the 32-bit
    path) over the time per product,
  * bit-exactness of the GPU product against GMP mpz_mul,
  * CPU baselines on this host: the oracle's NTT/Garner/carry pipeline (a port
    of the reference's CPU path) and GMP mpz_mul, both on one bounded sample.

`--gpus N` with N > 1 exits: this build measures one GPU.
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
    chunks = range(k) if schedule == 1 else [None]
    for c in chunks:
        tag = "" if c is None else "[p%d]" % c
        kk = 1 if c is not None else k
        ops = ["A", "B"] if c is not None else ["AB"]
        for o in ops:
            for j in range(P - 1):
                out.append(("fwd_pass%d(T=2^%d)%s%s" % (j + 1, logs[j], tag, o), kk * (2 if o == "AB" else 1)))
        out.append(("fused_pass%d(T=2^%d)%s" % (P, logs[P - 1], tag), kk))
        for j in range(P - 2, -1, -1):
            out.append(("inv_pass%d(T=2^%d)%s" % (j + 1, logs[j], tag), kk))
    out.append(("garner_carry", k))
    return out


def launch_bytes(plan, bits, names):
    """Algorithmic bytes per launch (staged one-pass model).

    fwd pass 1 (ingress):  read A,B limbs, write 2k spectra      = S1
    fwd pass j > 1:        read + write 2k planes
    fused last pass:       read 2k planes, write k planes          = S2
    inv pass j > 1:        read + write k planes
    inv pass 1:            read + write k planes                   = S3
    garner_carry:          read k planes, write L_A+L_B limbs      = S4+S5 without the
                           128/160-bit coefficient round trip the fusion removes
    """
    w = bits // 8
    N = plan.N_exec
    L = plan.L_A + plan.L_B
    P = len(plan.passes)
    out = []
    for name, planes in names:
        if name.startswith("garner"):
            out.append(plan.k * N * w + L * w)
        elif name.startswith("fused"):
            out.append((L * w + planes * N * w) if P == 1 else 3 * planes * N * w)
        elif name.startswith("fwd_pass1"):
            nop = 2 if name.endswith("AB") else 1
            raw = L * w if nop == 2 else (plan.L_A if name.endswith("A") else plan.L_B) * w
            out.append(raw + planes * N * w)
        else:
            out.append(2 * planes * N * w)
    return out


def model_bytes(plan, bits):
    """Whole-product bytes of the staged model: (L_A+L_B)(w_in+w_out) + 8kNw_r + 2Nw_c."""
    w = bits // 8
    wc = 16 if bits == 32 else 20
    return (plan.L_A + plan.L_B) * 2 * w + 8 * plan.k * plan.N * w + 2 * plan.N * wc


def load_traffic(workload):
    """HBM bytes per launch of the dominant kernel from the committed rocprof
    PMC summary (profiles/traffic_<workload>.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baselines(A, B, bits, gpu_out, sample_oracle):
    """Oracle pipeline (port of the reference's CPU path) and GMP on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    res = {}
    t0 = time.perf_counter()
    ref = None
    if sample_oracle:
        ref = pyoracle.multiply(A, B, bits)
        t = time.perf_counter() - t0
        res["oracle"] = {"value": 1.0 / t, "unit": "products/s", "cores": 3, "kind": "port",
                         "sample": "one %dx%d-limb product (%d-bit limbs, one prime per thread), %.1f s"
                                   % (len(A), len(B), bits, t)}
    exact = None
    if pyoracle.gmp_available():
        g, tg = pyoracle.gmp_mul(A, B, bits)
        res["gmp"] = {"value": 1.0 / tg, "unit": "products/s", "cores": 1, "kind": "library",
                      "sample": "one mpz_mul of the same operands, %.2f s (import/export excluded)" % tg}
        exact = bool(np.array_equal(g, gpu_out))
        if ref is not None:
            exact = exact and bool(np.array_equal(ref, gpu_out))
    elif ref is not None:
        exact = bool(np.array_equal(ref, gpu_out))
    return res, exact


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--limbs", type=int, default=1 << 24, help="limbs per operand")
    ap.add_argument("--bits", type=int, default=32, choices=(32, 64))
    ap.add_argument("--passes", type=str, default="", help="pass split, e.g. 12,13")
    ap.add_argument("--schedule", type=int, default=1, help="0 pass-major, 1 prime-major launches")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-oracle", action="store_true", help="GMP only as CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (--gpus 1); multi-GPU runs are not part of it",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import birdwing_amd as bw

    L = args.limbs
    bits = args.bits
    passes = [int(x) for x in args.passes.split(",")] if args.passes else None
    A = synthetic_limbs(L, 1, bits)
    B = synthetic_limbs(L, 2, bits)
    tdt = torch.int64 if bits == 64 else torch.int32
    ndt = np.int64 if bits == 64 else np.int32
    da = torch.from_numpy(A.view(ndt)).cuda()
    db = torch.from_numpy(B.view(ndt)).cuda()
    dc = torch.empty(2 * L, dtype=tdt, device="cuda")
    plan = bw.Plan(bits, L, L, bw.SET_DEFAULT, passes)
    plan.set_schedule(args.schedule)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        plan.execute(da, db, dc, stream)
    torch.cuda.synchronize()

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        plan.execute(da, db, dc, stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    plan.set_timing(False)
    per_launch = plan.timing()
    plan.status()

    elapsed = t1 - t0
    ms_step = elapsed * 1e3 / args.steps
    lat = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps))
    p50 = float(np.percentile(lat, 50))
    p99 = float(np.percentile(lat, 99))
    nl = launch_names(plan, args.schedule)
    names = [n for n, _ in nl]
    lbytes = launch_bytes(plan, bits, nl)
    # aggregate launches of one kernel type (across primes and operands)
    import re
    agg = {}
    for n, t, by in zip(names, per_launch, lbytes):
        key = re.sub(r"\[p\d\]|(?<=\))(A|B|AB)$", "", n)
        a = agg.setdefault(key, [0.0, 0, 0])
        a[0] += t
        a[1] += by
        a[2] += 1
    dom_name = max(agg, key=lambda k: agg[k][0])
    dom_ms, dom_bytes, dom_n = agg[dom_name]
    dom_gbps = dom_bytes / (dom_ms * 1e-3) / 1e9
    workload = "plain_%dx%d_u%d" % (L, L, bits)
    traffic = load_traffic(workload)
    mbytes = model_bytes(plan, bits)

    gpu_out = dc.cpu().numpy().view(np.uint64 if bits == 64 else np.uint32)
    cpu, exact = ({}, None) if args.no_cpu else cpu_baselines(A, B, bits, gpu_out, not args.no_oracle)

    line = {
        "metric": "products/s batched & served vs plain (bit-exact); p50/p99 latency; % HBM roofline",
        "value": args.steps / elapsed,
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload, "mode": "plain", "limbs_a": L, "limbs_b": L, "limb_bits": bits,
                   "N": plan.N, "primes": plan.primes, "passes": plan.passes},
        "latency_ms": {"p50": p50, "p99": p99},
        "bit_exact_vs_gmp": exact,
        "kernels": [{"name": k, "launches": v[2], "ms_per_product": v[0], "alg_bytes": v[1],
                     "alg_GBps": v[1] / (v[0] * 1e-3) / 1e9} for k, v in agg.items()],
        "roofline": {"bound": "hbm", "kernel": dom_name, "launches_per_product": dom_n,
                     "avg_launch_ms": dom_ms / dom_n, "alg_bytes_per_launch": dom_bytes / dom_n,
                     "achieved": dom_gbps, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": dom_gbps / HBM_PEAK_GBPS,
                     "traffic": traffic.get("hbm_bytes_per_launch") if traffic else None},
        "product_roofline": {"alg_bytes": mbytes, "achieved": mbytes / (ms_step * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                             "frac": mbytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBPS},
    }
    if "oracle" in cpu:
        line["cpu_baseline"] = cpu["oracle"]
    elif "gmp" in cpu:
        line["cpu_baseline"] = cpu["gmp"]
    if "gmp" in cpu:
        line["gmp_baseline"] = cpu["gmp"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
