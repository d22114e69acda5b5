"""Batched products on the GPU: every product of a mixed batch bit-exact
against GMP mpz_mul (and identical to the same product run alone through a
Plan), with and without hipGraph replay, before and after autotuning.

The batch path has no counterpart in the reference (which multiplies one
product per execute_ntt_multiply call, src/gpu_ntt.cu:442-973); its parity bar
is the reference's end-to-end one (tests/test_full_multiply.cpp: products vs
GMP), applied to every product of the batch.
"""
import numpy as np
import pytest
import torch

from devbuf import from_dev, to_dev

pytestmark = pytest.mark.gpu


def _log_uniform(rng, n, lo_log, hi_log):
    return np.floor(2.0 ** rng.uniform(lo_log, hi_log, n)).astype(np.int64)


def _operands(oracle, bits, LA, LB, seed):
    As, Bs = [], []
    for i, (la, lb) in enumerate(zip(LA, LB)):
        As.append(oracle.uniform_limbs(int(la), seed * 7919 + 2 * i + 1, bits))
        Bs.append(oracle.uniform_limbs(int(lb), seed * 7919 + 2 * i + 2, bits))
    return As, Bs


def _bytes_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.uint8)).cuda()


def _run(batch, pa, pb, graph=False, reps=1):
    da, db = _bytes_dev(pa), _bytes_dev(pb)
    dc = torch.zeros(max(batch.bytes_c, 1), dtype=torch.uint8, device="cuda")
    batch.set_graph(graph)
    s = torch.cuda.Stream()
    outs = []
    with torch.cuda.stream(s):
        for _ in range(reps):
            dc.fill_(0xA5)                  # poison: every limb must be rewritten
            batch.execute(da, db, dc, stream=s)
            s.synchronize()
            outs.append(batch.unpack(dc.cpu().numpy()))
    batch.status()
    return outs, (da, db)


def _check_vs_gmp(oracle, bits, As, Bs, got, sample=None):
    idx = range(len(As)) if sample is None else sample
    for i in idx:
        want, _ = oracle.gmp_mul(As[i], Bs[i], bits[i] if isinstance(bits, list) else bits)
        assert np.array_equal(got[i], want), (i, len(As[i]), len(Bs[i]))


@pytest.mark.parametrize("bits", [32, 64])
def test_batch_mixed_sizes_vs_gmp(gpu, oracle, bits):
    rng = np.random.default_rng(100 + bits)
    n = 97
    LA = _log_uniform(rng, n, 0, 13)
    LB = _log_uniform(rng, n, 0, 13)
    As, Bs = _operands(oracle, bits, LA, LB, 1)
    b = gpu.Batch(bits, LA, LB)
    pa, pb = b.pack(As, Bs)
    print("batch bits=%d n=%d groups=%s" % (bits, n, b.groups()), flush=True)
    (got,), _ = _run(b, pa, pb)
    _check_vs_gmp(oracle, bits, As, Bs, got)


def test_batch_mixed_widths_and_edges(gpu, oracle):
    """32- and 64-bit products together; one-limb operands; L_A + L_B - 1 = N
    exactly; unbalanced; all-ones operands (longest carry chains)."""
    specs = [(32, 1, 1), (64, 1, 1), (32, 1, 2), (32, 2048, 2049), (64, 1024, 1025), (32, 3000, 7),
             (32, 5000, 5000), (64, 4096, 17), (32, 513, 512), (64, 33, 31)]
    bits = [s[0] for s in specs]
    LA = [s[1] for s in specs]
    LB = [s[2] for s in specs]
    As = [oracle.uniform_limbs(la, 900 + i, bt) for i, (bt, la) in enumerate(zip(bits, LA))]
    Bs = [oracle.uniform_limbs(lb, 950 + i, bt) for i, (bt, lb) in enumerate(zip(bits, LB))]
    # all-ones operands for two of them
    for i in (6, 7):
        As[i][:] = np.iinfo(As[i].dtype).max
        Bs[i][:] = np.iinfo(Bs[i].dtype).max
    b = gpu.Batch(bits, LA, LB)
    pa, pb = b.pack(As, Bs)
    (got,), _ = _run(b, pa, pb)
    _check_vs_gmp(oracle, bits, As, Bs, got)


def test_batch_equals_plan_alone(gpu, oracle):
    """Each product of a batch is identical to the same product run alone."""
    rng = np.random.default_rng(7)
    n = 40
    LA = _log_uniform(rng, n, 4, 14)
    LB = _log_uniform(rng, n, 4, 14)
    As, Bs = _operands(oracle, 32, LA, LB, 3)
    b = gpu.Batch(32, LA, LB)
    pa, pb = b.pack(As, Bs)
    (got,), _ = _run(b, pa, pb)
    for i in range(n):
        p = gpu.Plan(32, int(LA[i]), int(LB[i]))
        c = torch.zeros(int(LA[i] + LB[i]), dtype=torch.int32, device="cuda")
        p.execute(to_dev(As[i]), to_dev(Bs[i]), c)
        alone = from_dev(c)
        assert np.array_equal(alone, got[i]), i


def test_batch_graph_replays(gpu, oracle):
    """hipGraph replays (with eager work between them) equal the eager run."""
    rng = np.random.default_rng(11)
    n = 64
    LA = _log_uniform(rng, n, 6, 15)
    LB = _log_uniform(rng, n, 6, 15)
    As, Bs = _operands(oracle, 32, LA, LB, 4)
    b = gpu.Batch(32, LA, LB)
    pa, pb = b.pack(As, Bs)
    (eager,), _ = _run(b, pa, pb, graph=False)
    reps, _ = _run(b, pa, pb, graph=True, reps=4)
    for r in reps:
        for i in range(n):
            assert np.array_equal(r[i], eager[i]), i
    _check_vs_gmp(oracle, 32, As, Bs, eager, sample=range(0, n, 7))


@pytest.mark.parametrize("bits", [32, 64])
def test_batch_graph_branches(gpu, oracle, bits):
    """A graph with one concurrent branch per size group (mode 2), replayed
    four times with eager work between, equals the eager in-order run; the
    groups include products large enough for the long-stride inverse tiles."""
    rng = np.random.default_rng(17 + bits)
    n = 80
    LA = _log_uniform(rng, n, 4, 18)
    LB = _log_uniform(rng, n, 4, 18)
    As, Bs = _operands(oracle, bits, LA, LB, 6)
    b = gpu.Batch(bits, LA, LB)
    pa, pb = b.pack(As, Bs)
    print("branches: groups", [(g["logN"], g["members"]) for g in b.groups()], flush=True)
    (eager,), _ = _run(b, pa, pb, graph=0)
    reps, _ = _run(b, pa, pb, graph=2, reps=4)
    for r in reps:
        for i in range(n):
            assert np.array_equal(r[i], eager[i]), i
    _check_vs_gmp(oracle, bits, As, Bs, eager, sample=range(0, n, 9))


def test_batch_autotune_keeps_results(gpu, oracle):
    rng = np.random.default_rng(13)
    n = 48
    LA = _log_uniform(rng, n, 8, 16)
    LB = _log_uniform(rng, n, 8, 16)
    As, Bs = _operands(oracle, 32, LA, LB, 5)
    b = gpu.Batch(32, LA, LB)
    pa, pb = b.pack(As, Bs)
    (before,), (da, db) = _run(b, pa, pb)
    s = torch.cuda.Stream()
    b.tune(da, db, stream=s, reps=2)
    assert gpu.tune_count() >= 1
    groups = b.groups()
    print("tuned groups:", groups, flush=True)
    (after,), _ = _run(b, pa, pb)
    for i in range(n):
        assert np.array_equal(before[i], after[i]), i
    # a new batch of the same sizes starts from the tuned choices
    b2 = gpu.Batch(32, LA, LB)
    assert [g["passes"] for g in b2.groups()] == [g["passes"] for g in groups]
    _check_vs_gmp(oracle, 32, As, Bs, after, sample=range(0, n, 5))


def _mod_q(limbs, bits, q, pw):
    """Value of a little-endian limb array mod q (q < 2^31), vectorized."""
    x = np.asarray(limbs).astype(np.uint64)
    if bits == 64:
        lo = x & np.uint64(0xFFFFFFFF)
        hi = x >> np.uint64(32)
        x = np.empty(2 * len(lo), dtype=np.uint64)
        x[0::2], x[1::2] = lo, hi
    x %= np.uint64(q)
    return int(((x * pw[:len(x)]) % np.uint64(q)).sum() % q)


def test_batch_4096_uneven_checksums(gpu, oracle):
    """4096 products, limb counts log-uniform in [2^8, 2^16]: every product
    checked by residues mod two primes (the product of residues), a sample
    against GMP."""
    rng = np.random.default_rng(2024)
    n = 4096
    LA = _log_uniform(rng, n, 8, 16)
    LB = _log_uniform(rng, n, 8, 16)
    b = gpu.Batch(32, LA, LB)
    # uniform random packed operands straight from the generator
    pa = rng.integers(0, 256, b.bytes_a, dtype=np.uint8)
    pb = rng.integers(0, 256, b.bytes_b, dtype=np.uint8)
    print("4096 batch: groups=%d workspace=%d MB" % (len(b.groups()), b.info.workspace_bytes >> 20), flush=True)
    (got,), _ = _run(b, pa, pb, graph=True)
    maxlen = 2 * (1 << 17) + 8
    for q in (2147483647, 2147483629):
        pw = np.ones(maxlen, dtype=np.uint64)
        base = (1 << 32) % q
        for i in range(1, maxlen):
            pw[i] = pw[i - 1] * np.uint64(base) % np.uint64(q)
        for i in range(n):
            a = pa[b.off_a[i]:b.off_a[i] + 4 * int(LA[i])].view(np.uint32)
            bb = pb[b.off_b[i]:b.off_b[i] + 4 * int(LB[i])].view(np.uint32)
            lhs = _mod_q(a, 32, q, pw) * _mod_q(bb, 32, q, pw) % q
            assert _mod_q(got[i], 32, q, pw) == lhs, (q, i)
    for i in range(0, n, 256):
        a = pa[b.off_a[i]:b.off_a[i] + 4 * int(LA[i])].view(np.uint32)
        bb = pb[b.off_b[i]:b.off_b[i] + 4 * int(LB[i])].view(np.uint32)
        want, _ = oracle.gmp_mul(a, bb, 32)
        assert np.array_equal(got[i], want), i


# ---------------------------------------------------------------------------
# Whole products in one workgroup (the fused launch of products up to 2^11
# 64-bit digits, several of one length to a workgroup): every transform
# length, the largest operands of each, all-ones operands, odd limb counts,
# against GMP and against the same products run through stacked transforms.

def _fused_specs(bits):
    """(L_A, L_B) per transform length 2^L, L = 1 .. 11: the largest product of
    that length (digits(A) + digits(B) - 1 = 2^L), an odd total of limbs, an
    unbalanced one."""
    per = 2 if bits == 32 else 1
    specs = []
    for L in range(1, 12):
        n = 1 << L
        da, db = n // 2 + 1, n // 2
        specs.append((per * da, per * db))
        if bits == 32:
            specs.append((2 * da - 1, 2 * db))      # odd limb total: half of the top word
        specs.append((per * (n - 1) + (1 if L > 1 else 0), 1))
    return specs


@pytest.mark.parametrize("bits", [32, 64])
def test_fused_products_vs_gmp(gpu, oracle, bits):
    specs = _fused_specs(bits)
    LA = [s[0] for s in specs]
    LB = [s[1] for s in specs]
    As, Bs = _operands(oracle, bits, LA, LB, 21)
    # all-ones operands for the largest product of every length: coefficients
    # at the CRT bound's worst case for that length, carries through every limb
    per_len = len(specs) // 11           # specs per transform length, the largest first
    for i in range(0, len(specs), per_len):
        As[i][:] = np.iinfo(As[i].dtype).max
        Bs[i][:] = np.iinfo(Bs[i].dtype).max
    old = gpu.set_fused_limit(11)
    try:
        b = gpu.Batch(bits, LA, LB)
    finally:
        gpu.set_fused_limit(old)
    nf, mx = b.fused()
    print("fused bits=%d products=%d fused=%d max_logN=%d groups=%s" % (bits, len(specs), nf, mx, b.groups()),
          flush=True)
    assert nf == len(specs) and mx == 11 and not b.groups()
    pa, pb = b.pack(As, Bs)
    (got,), _ = _run(b, pa, pb)
    _check_vs_gmp(oracle, bits, As, Bs, got)
    # a batch of fused products only, captured and replayed
    reps, _ = _run(b, pa, pb, graph=True, reps=2)
    for r in reps:
        for i in range(len(specs)):
            assert np.array_equal(r[i], got[i]), i


@pytest.mark.parametrize("bits", [32, 64])
def test_fused_equals_stacked(gpu, oracle, bits):
    """The same mixed batch with the fused launch on and off: identical limbs
    (the larger products run stacked in both)."""
    rng = np.random.default_rng(31 + bits)
    n = 160
    LA = _log_uniform(rng, n, 0, 14)
    LB = _log_uniform(rng, n, 0, 14)
    As, Bs = _operands(oracle, bits, LA, LB, 8)
    outs = []
    for lim in (11, 0):
        old = gpu.set_fused_limit(lim)
        try:
            b = gpu.Batch(bits, LA, LB)
        finally:
            gpu.set_fused_limit(old)
        nf, _ = b.fused()
        assert (nf > 0) == (lim > 0)
        pa, pb = b.pack(As, Bs)
        (got,), _ = _run(b, pa, pb)
        outs.append(got)
    for i in range(n):
        assert np.array_equal(outs[0][i], outs[1][i]), (i, LA[i], LB[i])
    _check_vs_gmp(oracle, bits, As, Bs, outs[0], sample=range(0, n, 5))


def test_fused_graph_replays(gpu, oracle):
    """The fused launch's job counter returns to zero after every launch:
    graph replays and eager runs (both graph forms) give the same limbs."""
    rng = np.random.default_rng(41)
    n = 600
    LA = _log_uniform(rng, n, 0, 12.5)
    LB = _log_uniform(rng, n, 0, 12.5)
    As, Bs = _operands(oracle, 32, LA, LB, 9)
    b = gpu.Batch(32, LA, LB)
    print("fused replays: fused=%s groups=%s" % (b.fused(), [(g["logN"], g["members"]) for g in b.groups()]),
          flush=True)
    pa, pb = b.pack(As, Bs)
    (eager,), _ = _run(b, pa, pb, graph=0)
    for mode in (1, 2):
        reps, _ = _run(b, pa, pb, graph=mode, reps=3)
        for r in reps:
            for i in range(n):
                assert np.array_equal(r[i], eager[i]), (mode, i)
    (again,), _ = _run(b, pa, pb, graph=0, reps=1)
    for i in range(n):
        assert np.array_equal(again[i], eager[i]), i
    _check_vs_gmp(oracle, 32, As, Bs, eager, sample=range(0, n, 11))


def test_fused_mixed_widths_in_bundles(gpu, oracle):
    """32- and 64-bit products of the same transform lengths share bundles of
    the fused launch (several products per workgroup, each with its own limb
    width, lengths and output offset): every product against GMP."""
    rng = np.random.default_rng(57)
    specs = []
    for L in range(3, 10):
        n = 1 << L
        for _ in range(6):
            bits = int(rng.choice([32, 64]))
            per = 2 if bits == 32 else 1
            da = int(rng.integers(1, n // 2 + 1))
            db = int(rng.integers(1, n - da + 2))
            specs.append((bits, per * da - int(rng.integers(0, per)), per * db))
    bits = [s[0] for s in specs]
    LA = [max(1, s[1]) for s in specs]
    LB = [max(1, s[2]) for s in specs]
    As = [oracle.uniform_limbs(la, 700 + i, bt) for i, (bt, la) in enumerate(zip(bits, LA))]
    Bs = [oracle.uniform_limbs(lb, 800 + i, bt) for i, (bt, lb) in enumerate(zip(bits, LB))]
    old = gpu.set_fused_limit(11)
    try:
        b = gpu.Batch(bits, LA, LB)
    finally:
        gpu.set_fused_limit(old)
    nf, _ = b.fused()
    print("fused mixed widths: %d products, %d fused, groups %s" % (len(specs), nf, b.groups()), flush=True)
    assert nf == len(specs)
    pa, pb = b.pack(As, Bs)
    (got,), _ = _run(b, pa, pb)
    _check_vs_gmp(oracle, bits, As, Bs, got)
