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


def launch_names(plan):
    P = len(plan.passes)
    logs = plan.passes
    if P == 1:
        return ["fused_pass(T=2^%d)" % logs[0], "garner_carry"]
    names = ["fwd_pass%d(T=2^%d)" % (j + 1, logs[j]) for j in range(P - 1)]
    names.append("fused_pass%d(T=2^%d)" % (P, logs[P - 1]))
    names += ["inv_pass%d(T=2^%d)" % (j + 1, logs[j]) for j in range(P - 2, -1, -1)]
    names.append("garner_carry")
    return names


def launch_bytes(plan, bits):
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
    k, N = plan.k, plan.N_exec
    L = plan.L_A + plan.L_B
    P = len(plan.passes)
    out = []
    if P == 1:
        out.append(L * w + k * N * w)                # raw limbs in, k product planes out
    else:
        out.append(L * w + 2 * k * N * w)
        out += [4 * k * N * w] * (P - 2)
        out.append(3 * k * N * w)
        out += [2 * k * N * w] * (P - 1)
    out.append(k * N * w + L * w)
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
        res["oracle"] = {"value": 1.0 / t, "unit": "products/s", "cores": 1, "kind": "port",
                         "sample": "one %dx%d-limb product (%d-bit limbs), %.1f s" % (len(A), len(B), bits, t)}
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
    names = launch_names(plan)
    lbytes = launch_bytes(plan, bits)
    dom = int(np.argmax(per_launch))
    dom_gbps = lbytes[dom] / (per_launch[dom] * 1e-3) / 1e9
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
        "kernels": [{"name": n, "avg_ms": t, "alg_bytes": b, "alg_GBps": b / (t * 1e-3) / 1e9}
                    for n, t, b in zip(names, per_launch, lbytes)],
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": dom_gbps, "peak": HBM_PEAK_GBPS,
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
