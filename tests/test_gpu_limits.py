"""Products at the limits the product path runs at, on the GPU.

* The automatic (five-prime, 64-bit-digit) set at its CRT limit: the product of
  its primes is 2^153.36 and the largest coefficient of 2^25 x 2^25 32-bit
  limbs (2^24 64-bit digits) or 2^24 x 2^24 64-bit limbs is 2^24 (2^64 - 1)^2,
  about 2^152 -- a margin of 2.6x, reached only by operands of all-ones
  limbs, the inputs of the reference's carry stress
  (tests/test_full_multiply.cpp:173-237; the bound of src/ntt_limits.cpp:186-214).
* The reference pipelines' three-prime methods at N = 2^25: the wide 30-bit
  set on 2^24 x 2^24 32-bit limbs and the reference 58-bit set
  (src/gpu_ntt.cu:30-34) on 2^24 x 2^24 64-bit limbs.
* Batches and the serving loop with large products (2^19 - 2^22 limbs), whose
  stacked groups run the long-stride inverse tiles.

All-ones operands have a closed form, (2^b - 1)^2 = 2^(2b) - 2^(b+1) + 1,
checked limb for limb beside GMP mpz_mul.
"""
import numpy as np
import pytest
import torch

from devbuf import empty, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _ones(n, bits, top_bits=None):
    """n limbs of all ones; the top limb keeps top_bits bits when given."""
    dt = np.uint64 if bits == 64 else np.uint32
    x = np.full(n, np.iinfo(dt).max, dtype=dt)
    if top_bits:
        x[-1] = (1 << top_bits) - 1
    return x


def _square_of_ones(n, bits, b):
    """Limbs (2n of `bits`) of (2^b - 1)^2 = 2^(2b) - 2^(b+1) + 1."""
    dt = np.uint64 if bits == 64 else np.uint32
    if b == bits * n:
        out = np.zeros(2 * n, dtype=dt)
        out[0] = 1
        out[n] = np.iinfo(dt).max - 1
        out[n + 1:] = np.iinfo(dt).max
        return out
    v = (1 << (2 * b)) - (1 << (b + 1)) + 1
    return np.frombuffer(v.to_bytes(2 * n * bits // 8, "little"), dtype=dt).copy()


def _run(gpu, A, B, bits, pset=-1):
    plan = gpu.Plan(bits, len(A), len(B), pset)
    c = empty(len(A) + len(B), bits)
    print("  plan set=%d k=%d N_exec=2^%d passes=%s" % (plan.prime_set, plan.k, plan.N_exec.bit_length() - 1,
                                                          plan.passes), flush=True)
    plan.execute(to_dev(A), to_dev(B), c)
    got = from_dev(c, len(A) + len(B))
    plan.status()
    info = (plan.prime_set, plan.k)
    plan.close()
    del c
    torch.cuda.empty_cache()
    return got, info


@pytest.mark.parametrize("bits,n,b", [(32, 1 << 25, 32 << 25), (64, 1 << 24, 64 << 24),
                                      (32, 1 << 25, (32 << 25) - 7)])
def test_five_prime_crt_limit_all_ones(gpu, oracle, bits, n, b):
    """(2^b - 1)^2 under the automatic set at its CRT limit, vs the closed
    form and GMP."""
    top = b % bits or None
    X = _ones(n, bits, top)
    print("all-ones %d-bit limbs n=%d b=%d" % (bits, n, b), flush=True)
    got, (pset, k) = _run(gpu, X, X, bits)
    assert pset == gpu.SET_D64 and k == 5
    want = _square_of_ones(n, bits, b)
    assert np.array_equal(got, want)
    g, _ = oracle.gmp_mul(X, X, bits)
    assert np.array_equal(got, g)


@pytest.mark.parametrize("bits", [32, 64])
def test_three_prime_full_size_uniform(gpu, oracle, bits):
    """2^24 x 2^24 limbs under the reference pipeline's three-prime method at
    N = 2^25: the wide 30-bit set (32-bit limbs) or the reference 58-bit set
    (64-bit limbs), vs GMP."""
    L = 1 << 24
    pset = gpu.SET_U32_WIDE if bits == 32 else gpu.SET_U64_REF
    A = oracle.uniform_limbs(L, 241, bits)
    B = oracle.uniform_limbs(L, 242, bits)
    print("three-prime 2^24 x 2^24 u%d" % bits, flush=True)
    got, (s, k) = _run(gpu, A, B, bits, pset)
    assert s == pset and k == 3
    want, _ = oracle.gmp_mul(A, B, bits)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bits", [32, 64])
def test_three_prime_full_size_all_ones(gpu, oracle, bits):
    """All-ones 2^24 x 2^24 limbs under the same three-prime methods: the
    largest coefficients they can meet, vs the closed form."""
    L = 1 << 24
    pset = gpu.SET_U32_WIDE if bits == 32 else gpu.SET_U64_REF
    X = _ones(L, bits)
    print("three-prime all-ones 2^24 u%d" % bits, flush=True)
    got, (s, k) = _run(gpu, X, X, bits, pset)
    assert s == pset and k == 3
    assert np.array_equal(got, _square_of_ones(L, bits, bits * L))


@pytest.mark.parametrize("bits,la,lb", [(32, 1, 1 << 25), (32, 1 << 25, 3), (64, 1 << 24, 1),
                                        (64, 2, (1 << 24) - 1)])
def test_lopsided_full_size_all_ones(gpu, oracle, bits, la, lb):
    """The most unbalanced shapes at the automatic set's largest operand (a
    few limbs times 2^25 32-bit or 2^24 64-bit limbs, all ones: every
    coefficient at its largest for these lengths, carries the full length),
    vs GMP and the closed form (2^a - 1)(2^b - 1)."""
    A = _ones(la, bits)
    B = _ones(lb, bits)
    print("lopsided all-ones %d x %d u%d" % (la, lb, bits), flush=True)
    got, (pset, k) = _run(gpu, A, B, bits)
    assert pset == gpu.SET_D64 and k == 5
    g, _ = oracle.gmp_mul(A, B, bits)
    assert np.array_equal(got, g)
    a, b = bits * la, bits * lb
    v = (1 << (a + b)) - (1 << a) - (1 << b) + 1
    dt = np.uint64 if bits == 64 else np.uint32
    want = np.frombuffer(v.to_bytes((la + lb) * bits // 8, "little"), dtype=dt)
    assert np.array_equal(got, want)


# ---- batches and serving with large products ---------------------------------
def _value_mod(limbs, q, block=1 << 14):
    """Value mod q (q < 2^31) of little-endian limbs (32- or 64-bit), by
    blocks: sum_i a_i 2^(32 i) = sum_h 2^(32 B h) sum_l a_(hB+l) 2^(32 l)."""
    a = np.ascontiguousarray(limbs).view(np.uint32)
    n = -(-len(a) // block) * block
    x = np.zeros(n, dtype=np.uint64)
    x[:len(a)] = a
    x %= np.uint64(q)
    base = pow(2, 32, q)
    pw = np.empty(block, dtype=np.uint64)
    acc = 1
    for i in range(block):
        pw[i] = acc
        acc = acc * base % q
    inner = ((x.reshape(-1, block) * pw) % np.uint64(q)).sum(axis=1) % np.uint64(q)
    step = pow(2, 32 * block, q)
    total, w = 0, 1
    for v in inner.tolist():
        total = (total + v * w) % q
        w = w * step % q
    return total


def _check_product(oracle, A, B, bits, got, gmp_limit=1 << 21):
    """Against GMP while L_A + L_B <= gmp_limit; beyond, by residues mod three
    primes near 2^31 and the exact low 1024 limbs."""
    if len(A) + len(B) <= gmp_limit:
        want, _ = oracle.gmp_mul(A, B, bits)
        assert np.array_equal(got, want), (len(A), len(B), bits)
        return
    for q in (2147483647, 2147483629, 2147483587):
        assert _value_mod(got, q) == _value_mod(A, q) * _value_mod(B, q) % q, (q, len(A), len(B), bits)
    m = 1024
    low = lambda v: int.from_bytes(np.ascontiguousarray(v[:m]).tobytes(), "little")
    assert low(got) == (low(A) * low(B)) % (1 << (bits * m))


def test_batch_large_products_vs_gmp(gpu, oracle):
    """Products of 2^19 - 2^20 limbs stacked in groups whose first inverse pass
    has a row stride of 2^12 or more elements (the 64-column, 1024-thread
    inverse tiles), with 32- and 64-bit limbs and all-ones operands, vs GMP."""
    specs = [(32, 1 << 19, 1 << 19), (32, (1 << 19) + 3, (1 << 19) - 5), (32, 1 << 20, 1 << 19),
             (32, 1 << 19, 1 << 19), (64, 1 << 19, 1 << 19), (64, 3 << 17, 1 << 19), (32, 1 << 20, 1 << 20)]
    bits = [s[0] for s in specs]
    LA = [s[1] for s in specs]
    LB = [s[2] for s in specs]
    As, Bs = [], []
    for i, (bt, la, lb) in enumerate(specs):
        if i == 3:
            As.append(_ones(la, bt))
            Bs.append(_ones(lb, bt))
        else:
            As.append(oracle.uniform_limbs(la, 7000 + 2 * i, bt))
            Bs.append(oracle.uniform_limbs(lb, 7001 + 2 * i, bt))
    b = gpu.Batch(bits, LA, LB)
    groups = b.groups()
    print("large batch groups:", groups, flush=True)
    pa, pb = b.pack(As, Bs)
    da = torch.from_numpy(pa).cuda()
    db = torch.from_numpy(pb).cuda()
    dc = torch.zeros(max(b.bytes_c, 1), dtype=torch.uint8, device="cuda")
    for graph in (False, True):
        b.set_graph(graph)
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            dc.fill_(0xA5)
            b.execute(da, db, dc, stream=s)
            s.synchronize()
        b.status()
        got = b.unpack(dc.cpu().numpy())
        for i in range(len(specs)):
            want, _ = oracle.gmp_mul(As[i], Bs[i], bits[i])
            assert np.array_equal(got[i], want), (graph, i, specs[i])
    b.close()


def test_serve_full_range_equals_alone(gpu, oracle):
    """The serving configuration's range: limb counts log-uniform in
    [2^6, 2^22], 80% 32-bit / 20% 64-bit limbs, submitted while earlier ones
    run.  Every result equals the same product alone (a Plan of its own) and
    GMP, or for the largest its residues and low limbs."""
    import threading
    import time
    rng = np.random.default_rng(2022)
    reqs = []
    for i in range(40):
        bt = 64 if rng.random() < 0.2 else 32
        la = int(2.0 ** rng.uniform(6, 22))
        lb = int(2.0 ** rng.uniform(6, 22))
        reqs.append((bt, oracle.uniform_limbs(la, 3100 + 2 * i, bt), oracle.uniform_limbs(lb, 3101 + 2 * i, bt)))
    # the widest requests the range allows, in both widths
    reqs.append((32, oracle.uniform_limbs(1 << 22, 3998, 32), oracle.uniform_limbs(1 << 22, 3999, 32)))
    reqs.append((64, oracle.uniform_limbs(1 << 22, 3996, 64), oracle.uniform_limbs((1 << 22) - 1, 3997, 64)))
    srv = gpu.Server(device_bytes=6 << 30, streams=3, max_batch=64)
    outs = [np.zeros(len(A) + len(B), dtype=A.dtype) for _, A, B in reqs]
    tickets = [None] * len(reqs)
    prng = np.random.default_rng(9)

    def producer():
        for i, (bt, A, B) in enumerate(reqs):
            tickets[i] = srv.submit(A, B, outs[i], bt)
            if prng.random() < 0.3:
                time.sleep(prng.uniform(0, 0.002))

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    for t in tickets:
        srv.wait(t, timeout_ms=110000)
    st = srv.stats()
    print("serve full range stats:", st, flush=True)
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    for i, (bt, A, B) in enumerate(reqs):
        p = gpu.Plan(bt, len(A), len(B))
        c = empty(len(A) + len(B), bt)
        p.execute(to_dev(A), to_dev(B), c)
        alone = from_dev(c, len(A) + len(B))
        p.close()
        assert np.array_equal(alone, outs[i]), (i, bt, len(A), len(B))
        _check_product(oracle, A, B, bt, outs[i])
