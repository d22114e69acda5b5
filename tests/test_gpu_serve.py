"""Serving loop on the GPU: requests submitted while others run are batched
and overlapped, and every result is bit-identical to the same product alone
(GMP mpz_mul and a Plan on its own) -- tolerance 0 bits.

The serving path has no counterpart in the reference; its parity bar is the
reference's end-to-end one (tests/test_full_multiply.cpp: products vs GMP).
"""
import threading
import time

import numpy as np
import pytest
import torch

from devbuf import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _requests(oracle, n, seed, lo_log=6, hi_log=14, frac64=0.2):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        bits = 64 if rng.random() < frac64 else 32
        la = int(2.0 ** rng.uniform(lo_log, hi_log))
        lb = int(2.0 ** rng.uniform(lo_log, hi_log))
        A = oracle.uniform_limbs(la, seed * 10007 + 2 * i + 1, bits)
        B = oracle.uniform_limbs(lb, seed * 10007 + 2 * i + 2, bits)
        reqs.append((bits, A, B))
    return reqs


def test_serve_mix_equals_alone(gpu, oracle):
    reqs = _requests(oracle, 240, 1)
    srv = gpu.Server(device_bytes=1 << 30, streams=3, max_batch=64)
    outs = [np.zeros(len(A) + len(B), dtype=A.dtype) for _, A, B in reqs]
    tickets = [None] * len(reqs)
    rng = np.random.default_rng(5)

    def producer():                      # arrivals spread over time, overlapping running batches
        for i, (bits, A, B) in enumerate(reqs):
            tickets[i] = srv.submit(A, B, outs[i], bits)
            if rng.random() < 0.3:
                time.sleep(rng.uniform(0, 0.002))

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    lats = [srv.wait(t, timeout_ms=60000) for t in tickets]
    st = srv.stats()
    print("serve stats:", st, "p50 %.3f ms p99 %.3f ms" % (np.percentile(lats, 50), np.percentile(lats, 99)),
          flush=True)
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batched_products"] == len(reqs)
    assert st["batches"] < len(reqs)          # requests were batched together
    for i, (bits, A, B) in enumerate(reqs):
        want, _ = oracle.gmp_mul(A, B, bits)
        assert np.array_equal(outs[i], want), i
    # the same products alone, through a plan each
    for i in range(0, len(reqs), 12):
        bits, A, B = reqs[i]
        p = gpu.Plan(bits, len(A), len(B))
        c = torch.zeros(len(A) + len(B), dtype=torch.int64 if bits == 64 else torch.int32, device="cuda")
        p.execute(to_dev(A), to_dev(B), c)
        assert np.array_equal(from_dev(c), outs[i]), i


def test_serve_memory_admission(gpu, oracle):
    """A small arena: requests wait for memory; one that can never fit fails
    alone with BW_ERR_NOMEM while the rest complete exactly."""
    srv = gpu.Server(device_bytes=48 << 20, streams=2, max_batch=16)
    reqs = _requests(oracle, 30, 2, lo_log=10, hi_log=17, frac64=0.0)
    big = oracle.uniform_limbs(1 << 20, 77, 32)          # needs ~ 2 x 3 x 2^21 x 4 B > 48 MiB
    outs = [np.zeros(len(A) + len(B), dtype=A.dtype) for _, A, B in reqs]
    tickets = []
    for i, (bits, A, B) in enumerate(reqs):
        tickets.append(srv.submit(A, B, outs[i], bits))
        if i == 10:
            big_out = np.zeros(2 * len(big), dtype=np.uint32)
            big_t = srv.submit(big, big, big_out, 32)
    for t in tickets:
        srv.wait(t, timeout_ms=60000)
    with pytest.raises(gpu.BWError) as ei:
        srv.wait(big_t, timeout_ms=60000)
    assert ei.value.code == gpu.BW_ERR_NOMEM
    st = srv.stats()
    srv.close()
    assert st["failed"] == 1 and st["completed"] == len(reqs)
    assert st["arena_peak_bytes"] <= 48 << 20
    for i, (bits, A, B) in enumerate(reqs):
        want, _ = oracle.gmp_mul(A, B, bits)
        assert np.array_equal(outs[i], want), i


def test_serve_rejects_bad_requests(gpu):
    srv = gpu.Server(device_bytes=16 << 20, streams=1)
    a = np.ones(4, dtype=np.uint32)
    with pytest.raises(gpu.BWError):
        srv.submit(a[:0], a, np.zeros(4, dtype=np.uint32), 32)
    with pytest.raises(gpu.BWError):
        srv.submit(a, a, np.zeros(8, dtype=np.uint32), 48)
    t = srv.submit(a, a, np.zeros(8, dtype=np.uint32), 32)
    srv.wait(t, timeout_ms=60000)
    srv.close()
