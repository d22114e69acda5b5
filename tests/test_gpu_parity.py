"""GPU parity: every stage of the HIP path through the C ABI against the oracle.

Bar: bit-exact (integer arithmetic throughout).  Inputs are seeded
(mt19937_64); sizes are ones the CPU oracle finishes in seconds.  Reference
tests mirrored: tests/test_zero_pad.cu, tests/test_crt_gpu.cu,
tests/test_carry_prop.cu, tests/test_full_multiply.cpp,
tests/test_pipeline_crt.cpp.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U32_REF, U32_WIDE, U64_REF = 0, 1, 2


def dev():
    import devbuf
    return devbuf


# ---------------------------------------------------------------- transforms
@pytest.mark.parametrize("pset,logN", [(U32_REF, l) for l in (1, 2, 3, 4, 5, 6, 8, 11, 12, 13, 14, 16, 18)] +
                         [(U32_WIDE, l) for l in (4, 13, 17)] + [(U64_REF, l) for l in (1, 4, 7, 12, 13, 15)])
def test_forward_ntt_matches_oracle(gpu, oracle, pset, logN):
    d = dev()
    bits = 64 if pset == U64_REF else 32
    N = 1 << logN
    L = max(1, N // 2)
    plan = gpu.Plan(bits, L, L, pset)
    assert plan.N_exec == N
    raw = oracle.uniform_limbs(L, 1000 + logN, bits)
    spec = d.empty(plan.k * N, bits)
    plan.forward(d.to_dev(raw), L, spec)
    got = d.from_dev(spec).reshape(plan.k, N)
    order = plan.spectrum_order()
    assert sorted(order.tolist()) == list(range(N))
    P, R = oracle.prime_set(pset)
    for j in range(plan.k):
        x = np.zeros(N, np.uint64)
        x[:L] = raw.astype(np.uint64) % P[j]
        w = oracle.root_for_N(P[j], R[j], logN)
        want = oracle.ntt_forward(x, P[j], w)
        nat = np.zeros(N, np.uint64)
        nat[order.astype(np.int64)] = got[j].astype(np.uint64)
        assert np.array_equal(nat, want), "prime %d" % j


@pytest.mark.parametrize("pset,logN", [(U32_REF, 3), (U32_REF, 12), (U32_REF, 15), (U32_WIDE, 19), (U64_REF, 9),
                                       (U64_REF, 14)])
def test_inverse_roundtrip(gpu, oracle, pset, logN):
    d = dev()
    bits = 64 if pset == U64_REF else 32
    N = 1 << logN
    L = N // 2
    plan = gpu.Plan(bits, L, L, pset)
    raw = oracle.uniform_limbs(L, 77, bits)
    spec = d.empty(plan.k * N, bits)
    plan.forward(d.to_dev(raw), L, spec)
    plan.inverse(spec)
    got = d.from_dev(spec).reshape(plan.k, N)
    for j, p in enumerate(plan.primes):
        want = np.zeros(N, np.uint64)
        want[:L] = raw.astype(np.uint64) % p
        assert np.array_equal(got[j].astype(np.uint64), want)


# ---------------------------------------------------------------- zero pad / pointwise
def test_zero_pad_u32_reference_cases(gpu):
    """tests/test_zero_pad.cu:37-103 (32-bit build): widen without reduction."""
    import torch
    d = dev()
    cases = [([1, 2, 3, 4], 8), ([10, 20, 30, 40], 4), ([42], 8), ([5, 0, 7, 0], 8),
             ([0xFFFFFFFF, 0xDEADBEEF, 0x12345678], 4)]
    for src, N in cases:
        s = np.array(src, np.uint32)
        for wb, bits in ((4, 32), (8, 64)):
            out = d.empty(N, bits)
            gpu.zero_pad(d.to_dev(s), out, len(s), N)
            got = d.from_dev(out)
            want = np.zeros(N, np.uint64)
            want[:len(s)] = s
            assert np.array_equal(got.astype(np.uint64), want)
    L, N = 1 << 20, 1 << 23
    s = (np.arange(L, dtype=np.uint64) + 1).astype(np.uint32)
    out = torch.full((N,), -1, dtype=torch.int32, device="cuda")
    gpu.zero_pad(d.to_dev(s), out, L, N)
    got = d.from_dev(out)
    assert np.array_equal(got[:L], s) and not got[L:].any()


def test_zero_pad_u64_reference_cases(gpu):
    """tests/test_zero_pad.cu:126-145: 64-bit limbs reduced mod p."""
    d = dev()
    p = 0x400002600000001
    for src, N in (([42, 1 << 40, 2**64 - 1], 8), ([2**64 - 1, 2**64 - 2, 1], 4)):
        s = np.array(src, np.uint64)
        out = d.empty(N, 64)
        gpu.zero_pad_u64(d.to_dev(s), out, len(s), N, p)
        got = d.from_dev(out)
        want = [int(x) % p for x in src] + [0] * (N - len(src))
        assert [int(x) for x in got] == want


@pytest.mark.parametrize("p,bits", [(0x2d000001, 32), (0x23800001, 32), (2013265921, 32), (469762049, 32),
                                    (0x400002600000001, 64), (0x400001100000001, 64)])
def test_pointwise_full_range(gpu, oracle, p, bits):
    """src/gpu_ntt.cu:216-238: C = A*B mod p for any words (tests/test_pointwise_mul.cu inputs)."""
    d = dev()
    rng = np.random.default_rng(1234)
    n = 4096
    mx = 2**bits - 1
    a = rng.integers(0, mx, n, dtype=np.uint64, endpoint=True).astype(np.uint64 if bits == 64 else np.uint32)
    b = rng.integers(0, mx, n, dtype=np.uint64, endpoint=True).astype(a.dtype)
    a[:5] = [0, 1, mx, mx // 2, p - 1]
    b[:5] = [0, 1, mx, mx // 2, p - 1]
    c = d.empty(n, bits)
    gpu.pointwise(d.to_dev(a), d.to_dev(b), c, n, p)
    got = d.from_dev(c)
    want = [(int(x) * int(y)) % p for x, y in zip(a, b)]
    assert [int(x) for x in got] == want


# ---------------------------------------------------------------- CRT
def _crt_run(gpu, pset, residues):
    d = dev()
    bits = 64 if pset == U64_REF else 32
    N = residues.shape[1]
    r = [d.to_dev(residues[j].astype(np.uint64 if bits == 64 else np.uint32)) for j in range(3)]
    o0, o1 = d.empty(N, 64), d.empty(N, 64)
    o2 = d.empty(N, 32) if bits == 64 else None
    gpu.crt_combine(pset, r[0], r[1], r[2], N, o0, o1, o2)
    lo, hi = d.from_dev(o0, N), d.from_dev(o1, N)
    if bits == 64:
        h2 = d.from_dev(o2, N)
        return [int(a) | (int(b) << 64) | (int(c) << 128) for a, b, c in zip(lo, hi, h2)]
    return [int(a) | (int(b) << 64) for a, b in zip(lo, hi)]


@pytest.mark.parametrize("pset", [U32_REF, U32_WIDE])
def test_crt_reference_kats_u32(gpu, oracle, pset):
    """tests/test_crt_gpu.cu:159-265 values, plus values up to prod(p)-1."""
    P, _ = oracle.prime_set(pset)
    M = P[0] * P[1] * P[2]
    vals = [12, 0, 0xFFFFFFFF, 0, 1, 12, 255, 65535, 0xFFFFFFFF, M - 1, M // 2, 2**87]
    if 2**89 < M:
        vals.append(2**89)
    res = np.array([[v % p for v in vals] for p in P], dtype=np.uint64)
    assert _crt_run(gpu, pset, res) == vals
    n = 64                                   # synthetic (i*7 + j*13 + 1) mod p_j, :243-265
    res = np.array([[(i * 7 + j * 13 + 1) % P[j] for i in range(n)] for j in range(3)], dtype=np.uint64)
    hi, lo = oracle.garner128(P, res)
    assert _crt_run(gpu, pset, res) == [int(l) | (int(h) << 64) for h, l in zip(hi, lo)]


def test_crt_reference_kats_u160(gpu, oracle):
    """tests/test_crt_gpu.cu:403-485 (U160 with the 58-bit set)."""
    P, _ = oracle.prime_set(U64_REF)
    vals = [0, 0, 0, 0, 12, 0, 1, 12, 2**64 - 1, 2**63 - 1, 2**128, 2**160 - 1, (1 << 159) + 12345]
    res = np.array([[v % p for v in vals] for p in P], dtype=np.uint64)
    assert _crt_run(gpu, U64_REF, res) == vals
    n = 64
    res = np.array([[(i * 7 + j * 13 + 1) % P[j] for i in range(n)] for j in range(3)], dtype=np.uint64)
    lo, mid, hi = oracle.garner160(P, res)
    want = [int(a) | (int(b) << 64) | (int(c) << 128) for a, b, c in zip(lo, mid, hi)]
    assert _crt_run(gpu, U64_REF, res) == want


# ---------------------------------------------------------------- carry
def _carry_cases_128():
    """tests/test_carry_prop.cu:331-390."""
    M32 = 0xFFFFFFFF
    cases = []
    cases.append(("all zeros", [0] * 64))
    cases.append(("small", [i + 1 for i in range(64)]))
    v = [0] * 64; v[0] = M32 + 1; cases.append(("single overflow", v))
    cases.append(("max limbs 1024", [M32] * 1024))
    v = [0] * 16; v[0] = (1 << 70) + 12345; cases.append(("wide", v))
    v = [0] * 2048; v[1023] = M32 + 1; cases.append(("cross segment", v))
    big = []
    for i in range(1 << 16):
        if i % 4 == 0: big.append(i * 12345 + 1)
        elif i % 4 == 1: big.append((1 << 70) + (i % 1000))
        elif i % 4 == 2: big.append(M32)
        else: big.append(((i % 17) << 64) + i * 99991)
    cases.append(("large random 2^16", big))
    cases.append(("all-ones 2^128-1 x 5000", [2**128 - 1] * 5000))
    cases.append(("ripple 300k", [M32] * 300000 + [1 << 40]))
    return cases


@pytest.mark.parametrize("case", range(9))
def test_carry128_reference_kats(gpu, oracle, case):
    d = dev()
    name, vals = _carry_cases_128()[case]
    N = len(vals)
    hi = np.array([v >> 64 for v in vals], np.uint64)
    lo = np.array([v & (2**64 - 1) for v in vals], np.uint64)
    for out_bits in (32, 64):
        n_out = N + (4 if out_bits == 32 else 2)
        out = d.empty(n_out, out_bits)
        gpu.carry128(d.to_dev(hi), d.to_dev(lo), N, out, n_out, out_bits)
        got = d.from_dev(out, n_out)
        want = oracle.carry128(hi, lo, n_out, out_bits)
        assert np.array_equal(got, want), name
        total = sum(v << (out_bits * i) for i, v in enumerate(vals))
        assert oracle.to_int(got, out_bits) == total, name


def _carry_cases_160():
    """tests/test_carry_prop.cu:230-327."""
    U64 = 2**64 - 1
    cases = [("zeros", [0] * 64), ("small", [i + 1 for i in range(64)]), ("single overflow", [1 << 64] + [0] * 63),
             ("max limbs", [U64] * 1024), ("wide", [12345 + (1 << 159)] + [0] * 15),
             ("160 max", [2**160 - 1] * 8), ("128 boundary", [1 + (1 << 64)] + [1] * 15)]
    v = [0] * 2048; v[1023] = U64 + (1 << 64); cases.append(("cross segment", v))
    big = []
    for i in range(8192):
        m = i % 5
        big.append([i * 12345 + 1, (1 << 120) + (i % 1000), U64, ((i % 17) << 64) + i * 99991, (1 << 155) + i][m])
    cases.append(("multi-segment 8k", big))
    return cases


@pytest.mark.parametrize("case", range(9))
def test_carry160_reference_kats(gpu, oracle, case):
    d = dev()
    name, vals = _carry_cases_160()[case]
    N = len(vals)
    lo = np.array([v & (2**64 - 1) for v in vals], np.uint64)
    mid = np.array([(v >> 64) & (2**64 - 1) for v in vals], np.uint64)
    hi = np.array([v >> 128 for v in vals], np.uint32)
    n_out = N + 3
    out = d.empty(n_out, 64)
    gpu.carry160(d.to_dev(lo), d.to_dev(mid), d.to_dev(hi), N, out, n_out)
    got = d.from_dev(out, n_out)
    assert np.array_equal(got, oracle.carry160(lo, mid, hi, n_out)), name
    assert oracle.to_int(got, 64) == sum(v << (64 * i) for i, v in enumerate(vals)), name


# ---------------------------------------------------------------- whole product
def _product_check(gpu, oracle, A, B, bits, pset=-1, passes=None):
    d = dev()
    plan = gpu.Plan(bits, len(A), len(B), pset, passes)
    c = d.empty(len(A) + len(B), bits)
    plan.execute(d.to_dev(A), d.to_dev(B), c)
    got = d.from_dev(c, len(A) + len(B))
    want, _ = oracle.gmp_mul(A, B, bits)
    assert np.array_equal(got, want), (len(A), len(B), bits, plan.passes)
    return plan


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("LA,LB", [(1, 1), (1, 2), (2, 2), (3, 5), (4, 4), (7, 9), (8, 8), (16, 16), (33, 31),
                                   (64, 64), (100, 3), (128, 128), (1000, 1000), (2048, 2048), (4096, 4096),
                                   (4097, 4095), (5000, 17), (10000, 10000), (1 << 15, 1 << 15),
                                   (70000, 1000), (1 << 17, 1 << 17)])
def test_product_vs_gmp(gpu, oracle, LA, LB, bits):
    A = oracle.uniform_limbs(LA, LA * 31 + 1, bits)
    B = oracle.uniform_limbs(LB, LB * 17 + 2, bits)
    _product_check(gpu, oracle, A, B, bits)


# Long-stride inverse column passes run 64-column tiles in 1024-thread
# workgroups (row stride >= 2^12 elements, tiles of at most 256 points on the
# FP64 path); the first pass of each split below qualifies: 256-point tiles,
# 128-point tiles (half the threads idle in stage 1) and 64-point tiles.
@pytest.mark.parametrize("bits,LA,LB,passes", [
    (32, 1 << 20, 1 << 20, [8, 6, 6]),          # N = 2^20 64-bit digits, first-pass stride 2^12
    (32, 1 << 19, (1 << 19) - 3, [7, 6, 6]),    # N = 2^19, stride 2^12, 128-point tiles
    (32, 1 << 19, 1 << 19, [6, 7, 6]),          # N = 2^19, stride 2^13, 64-point tiles
    (64, 1 << 18, (1 << 18) + 1, [7, 6, 6]),    # 64-bit limbs, N = 2^19, stride 2^12
])
def test_long_stride_inverse_tiles(gpu, oracle, bits, LA, LB, passes):
    A = oracle.uniform_limbs(LA, LA + 91, bits)
    B = oracle.uniform_limbs(LB, LB + 92, bits)
    print("long-stride inverse", bits, LA, LB, passes, flush=True)
    plan = _product_check(gpu, oracle, A, B, bits, -1, passes)
    assert plan.passes == passes


def test_product_reference_simple_cases(gpu, oracle):
    """tests/test_full_multiply.cpp:271-351: 1x1, 0x42, root probe, identities."""
    one = np.array([1, 0, 0, 0], np.uint32)
    _product_check(gpu, oracle, one, one, 32)
    _product_check(gpu, oracle, np.zeros(4, np.uint32), np.array([42, 0, 0, 0], np.uint32), 32)
    A = np.zeros(8, np.uint32); A[1] = 1
    B = np.zeros(8, np.uint32); B[0] = 1
    _product_check(gpu, oracle, A, B, 32)
    L = 1 << 16
    Z, O = np.zeros(L, np.uint32), np.ones(L, np.uint32)
    _product_check(gpu, oracle, Z, Z, 32)
    _product_check(gpu, oracle, O, Z, 32)
    _product_check(gpu, oracle, O, O, 32)


@pytest.mark.parametrize("L", [4, 8, 16, 64, 128, 2048, 10000, 1 << 15])
def test_product_reference_random_limbs(gpu, oracle, L):
    """tests/test_full_multiply.cpp:240-269 inputs (seeds 1234/5678) vs schoolbook/GMP."""
    A = oracle.ref_random_limbs(L, 1234)
    B = oracle.ref_random_limbs(L, 5678)
    _product_check(gpu, oracle, A, B, 32)


@pytest.mark.parametrize("L", [4, 16, 64, 256, 1024, 4096, 1 << 16])
def test_product_reference_u64_limbs(gpu, oracle, L):
    """tests/test_full_multiply.cpp:488-504 (64-bit pipeline, seeds 1234/5678)."""
    A = oracle.ref_random_limbs(L, 1234, 64)
    B = oracle.ref_random_limbs(L, 5678, 64)
    _product_check(gpu, oracle, A, B, 64)


@pytest.mark.parametrize("bits", [32, 64])
def test_product_carry_stress(gpu, oracle, bits):
    """Edge operands: all-ones limbs (maximal carry chains), sparse limbs, 2^b-1 squared."""
    mx = 2**bits - 1
    dt = np.uint64 if bits == 64 else np.uint32
    for L in (1, 2, 1023, 1024, 4096, 1 << 14):
        A = np.full(L, mx, dt)
        _product_check(gpu, oracle, A, A, bits)
    rng = np.random.default_rng(5)
    L = 1 << 15
    A = np.where(rng.random(L) < 0.01, rng.integers(1, mx, L, endpoint=True, dtype=np.uint64), 0).astype(dt)
    B = np.where(rng.random(L) < 0.01, rng.integers(1, mx, L, endpoint=True, dtype=np.uint64), 0).astype(dt)
    A[-1] |= 1
    _product_check(gpu, oracle, A, B, bits)
    # (2^b - 1)^2 with b not a multiple of the limb size
    b = bits * 3000 - 7
    x = (1 << b) - 1
    n = (b + bits - 1) // bits
    A = oracle.from_int(x, n, bits)
    _product_check(gpu, oracle, A, A, bits)


@pytest.mark.parametrize("passes", [[1, 12], [6, 7], [7, 6], [12, 1], [4, 4, 5], [2, 3, 8], [1, 1, 11], [11, 1, 1], [3, 9, 1]])
def test_product_pass_splits(gpu, oracle, passes):
    """Every pass decomposition gives the same product (N = 2^13)."""
    A = oracle.uniform_limbs(4096, 11)
    B = oracle.uniform_limbs(4096, 12)
    _product_check(gpu, oracle, A, B, 32, gpu.default_prime_set(32, 8192), passes)
    A = oracle.uniform_limbs(4096, 13, 64)
    B = oracle.uniform_limbs(4096, 14, 64)
    _product_check(gpu, oracle, A, B, 64, gpu.default_prime_set(64, 8192), passes)


@pytest.mark.parametrize("passes", [[6, 6], [4, 8], [8, 4], [2, 10], [10, 2], [3, 4, 5], [1, 1, 10]])
def test_d64_pass_splits(gpu, oracle, passes):
    """Every pass decomposition of the 64-bit-digit transform (N = 2^12 digits)."""
    A = oracle.uniform_limbs(4096, 21)
    B = oracle.uniform_limbs(4095, 22)
    _product_check(gpu, oracle, A, B, 32, gpu.SET_D64, passes)
    A = oracle.uniform_limbs(2048, 23, 64)
    B = oracle.uniform_limbs(2047, 24, 64)
    _product_check(gpu, oracle, A, B, 64, gpu.SET_D64, passes)


@pytest.mark.parametrize("schedule", [0, 1, 2])
@pytest.mark.parametrize("bits,L", [(32, 3000), (32, 1 << 18), (64, 1 << 16)])
def test_both_schedules(gpu, oracle, bits, L, schedule):
    """Pass-major, prime-major and hybrid launch orders give the same product."""
    d = dev()
    A = oracle.uniform_limbs(L, 5, bits)
    B = oracle.uniform_limbs(L - 7, 6, bits)
    plan = gpu.Plan(bits, len(A), len(B))
    plan.set_schedule(schedule)
    c = d.empty(len(A) + len(B), bits)
    plan.execute(d.to_dev(A), d.to_dev(B), c)
    want, _ = oracle.gmp_mul(A, B, bits)
    assert np.array_equal(d.from_dev(c, len(A) + len(B)), want)
    plan.status()


@pytest.mark.parametrize("bits,L", [(32, 4), (32, 256), (32, 1 << 15), (64, 64), (64, 4096)])
def test_execute_crt_matches_oracle(gpu, oracle, bits, L):
    """Early exit after Garner (src/gpu_ntt.cu:843-884; tests/test_pipeline_crt.cpp:209-259)."""
    d = dev()
    A = oracle.ref_pipeline_limbs(L, 100 + L) if bits == 32 else oracle.uniform_limbs(L, 100 + L, 64)
    B = oracle.ref_pipeline_limbs(L, 101 + L) if bits == 32 else oracle.uniform_limbs(L, 101 + L, 64)
    plan = gpu.Plan(bits, L, L, gpu.default_prime_set(bits, gpu.padded_ntt_size(L, L)))
    N = plan.N_exec
    c0, c1 = d.empty(N, 64), d.empty(N, 64)
    c2 = d.empty(N, 32) if bits == 64 else None
    plan.execute_crt(d.to_dev(A), d.to_dev(B), c0, c1, c2)
    w0, w1, w2 = oracle.multiply_crt(A, B, bits)
    assert np.array_equal(d.from_dev(c0, N), w0[:N])
    assert np.array_equal(d.from_dev(c1, N), w1[:N])
    if bits == 64:
        assert np.array_equal(d.from_dev(c2, N), w2[:N])


def test_host_entry_points(gpu, oracle):
    A = oracle.uniform_limbs(3000, 1)
    B = oracle.uniform_limbs(2000, 2)
    want, _ = oracle.gmp_mul(A, B)
    assert np.array_equal(gpu.multiply(A, B), want)
    plan = gpu.Plan(32, 3000, 2000)
    for _ in range(3):
        assert np.array_equal(plan.execute_host(A, B), want)


def test_unsupported_sizes_raise(gpu):
    # beyond every set (limb pairs reach 2^32 limbs per operand); rejected
    # before any allocation
    with pytest.raises(gpu.BWError) as e:
        gpu.Plan(32, (1 << 32) + 2, (1 << 32) + 2)
    assert e.value.code == gpu.BW_ERR_UNSUPPORTED_SIZE
    with pytest.raises(gpu.BWError) as e:
        gpu.Plan(32, 1 << 26, 1 << 26, gpu.SET_U32_WIDE)
    assert e.value.code == gpu.BW_ERR_UNSUPPORTED_SIZE
    with pytest.raises(gpu.BWError):
        gpu.Plan(32, 0, 5)


@pytest.mark.parametrize("bits,L", [(32, 1 << 15), (64, 5000)])
def test_plan_autotune_keeps_product(gpu, oracle, bits, L):
    """The autotuned pass split computes the same product (vs GMP)."""
    A = oracle.uniform_limbs(L, 501, bits)
    B = oracle.uniform_limbs(L + 3, 502, bits)
    d = dev()
    plan = gpu.Plan(bits, len(A), len(B))
    da, db = d.to_dev(A), d.to_dev(B)
    plan.tune(da, db, reps=1)
    c = d.empty(len(A) + len(B), bits)
    plan.execute(da, db, c)
    want, _ = oracle.gmp_mul(A, B, bits)
    assert np.array_equal(d.from_dev(c, len(A) + len(B)), want), plan.passes
    # the tuned tile cap is one of the candidates (full, half or quarter width)
    assert 0 <= plan.info.tile_shift <= 2
    # a new plan of the same shape starts from the tuned split and tile cap
    again = gpu.Plan(bits, len(A), len(B))
    assert again.passes == plan.passes and again.info.tile_shift == plan.info.tile_shift


@pytest.mark.parametrize("bits", [32, 64])
def test_random_shapes_vs_gmp(gpu, oracle, bits):
    """Forty seeded shapes, limb counts log-uniform in [1, 2^18] per operand
    (odd, unbalanced, every transform length from 2^1 to 2^18 and the splits
    the plans pick for them), under the automatic set, each against GMP."""
    rng = np.random.default_rng(20260 + bits)
    for i in range(40):
        la, lb = (int(2.0 ** rng.uniform(0, 18)) for _ in range(2))
        A = oracle.uniform_limbs(la, 1000 + 2 * i, bits)
        B = oracle.uniform_limbs(lb, 1001 + 2 * i, bits)
        _product_check(gpu, oracle, A, B, bits)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("LA,LB", [(1, 1), (1, 2), (2, 3), (3, 5), (7, 7), (2047, 2049), (4096, 4096),
                                   (10001, 9999), (1 << 16, (1 << 16) + 1), (100000, 3), (1 << 18, 1 << 18)])
def test_d64_products_vs_gmp(gpu, oracle, bits, LA, LB):
    """64-bit digits under five primes (the automatic choice) against GMP:
    odd and even 32-bit limb counts, unbalanced operands."""
    A = oracle.uniform_limbs(LA, LA * 7 + 1, bits)
    B = oracle.uniform_limbs(LB, LB * 11 + 2, bits)
    plan = _product_check(gpu, oracle, A, B, bits)
    assert plan.prime_set == gpu.SET_D64 and plan.digit_bits == 64


@pytest.mark.parametrize("bits", [32, 64])
def test_d64_carry_stress(gpu, oracle, bits):
    """All-ones operands (longest carry chains) and (2^b - 1)^2 under 64-bit digits."""
    for L in (4097, 30001):
        A = np.full(L, np.iinfo(np.uint64 if bits == 64 else np.uint32).max,
                    dtype=np.uint64 if bits == 64 else np.uint32)
        plan = _product_check(gpu, oracle, A, A, bits)
        assert plan.prime_set == gpu.SET_D64
    b = bits * 5001 - 3
    x = (1 << b) - 1
    n = (b + bits - 1) // bits
    A = oracle.from_int(x, n, bits)
    _product_check(gpu, oracle, A, A, bits)


def test_d64_execute_crt_refused(gpu, oracle):
    plan = gpu.Plan(32, 1000, 1000)
    assert plan.prime_set == gpu.SET_D64
    d = dev()
    c0, c1 = d.empty(plan.N_exec, 64), d.empty(plan.N_exec, 64)
    A = oracle.uniform_limbs(1000, 5, 32)
    with pytest.raises(gpu.BWError):
        plan.execute_crt(d.to_dev(A), d.to_dev(A), c0, c1)


# ------------------------------------------------ the measured configurations
@pytest.mark.parametrize("pset", [-1, "ref3"])
def test_config_plain_2048(gpu, oracle, pset):
    """2048 x 2048 32-bit limbs, seeded uniform, top limb nonzero: under the
    reference's 32-bit prime set (N = 4096) and under the automatic choice."""
    A = oracle.uniform_limbs(2048, 1, 32)
    B = oracle.uniform_limbs(2048, 2, 32)
    s = gpu.SET_U32_REF if pset == "ref3" else -1
    plan = _product_check(gpu, oracle, A, B, 32, s)
    if pset == "ref3":
        assert plan.N == 4096 and plan.primes == [0x2d000001, 0x23800001, 0x26800001]


@pytest.mark.parametrize("pset", [-1, "ref3"])
def test_config_carry_edge_stress(gpu, oracle, pset):
    """Carry/edge stress at full size (32-bit limbs): unbalanced 2^22 x 2^14,
    all-ones 2^22-limb operands, 99%-zero operands, (2^b - 1)^2 for
    b = 32 * 2^22 - 7 and b = 32 * 2^22; under the reference pipeline's
    three-prime set and under the automatic (five-prime) choice."""
    L = 1 << 22

    def set_for(la, lb):
        return gpu.default_prime_set(32, gpu.padded_ntt_size(la, lb)) if pset == "ref3" else -1

    A = oracle.uniform_limbs(L, 31, 32)
    B = oracle.uniform_limbs(1 << 14, 32, 32)
    print("2^22 x 2^14", flush=True)
    _product_check(gpu, oracle, A, B, 32, set_for(L, 1 << 14))
    O = np.full(L, 0xFFFFFFFF, np.uint32)
    print("all-ones 2^22", flush=True)
    _product_check(gpu, oracle, O, O, 32, set_for(L, L))
    rng = np.random.default_rng(77)
    Z1 = np.where(rng.random(L) < 0.01, rng.integers(1, 2**32, L, dtype=np.uint64), 0).astype(np.uint32)
    Z2 = np.where(rng.random(L) < 0.01, rng.integers(1, 2**32, L, dtype=np.uint64), 0).astype(np.uint32)
    Z1[-1] |= 1
    Z2[-1] |= 1
    print("99% zero 2^22", flush=True)
    _product_check(gpu, oracle, Z1, Z2, 32, set_for(L, L))
    for b in (32 * L - 7, 32 * L):
        n = (b + 31) // 32
        X = np.full(n, 0xFFFFFFFF, np.uint32)
        if b % 32:
            X[-1] = (1 << (b % 32)) - 1
        print("(2^b-1)^2 b=%d" % b, flush=True)
        _product_check(gpu, oracle, X, X, 32, set_for(n, n))


@pytest.mark.parametrize("bits", [32, 64])
def test_config_plain_large(gpu, oracle, bits):
    """2^24 x 2^24 limbs (N = 2^25 32-bit digits), seeded uniform, vs GMP."""
    L = 1 << 24
    A = oracle.uniform_limbs(L, 41, bits)
    B = oracle.uniform_limbs(L, 42, bits)
    print("2^24 x 2^24 u%d" % bits, flush=True)
    _product_check(gpu, oracle, A, B, bits)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("LA,LB", [(1, 1), (3, 5), (1000, 1003), (1500, 2500), (2048, 2048), (4097, 4095),
                                   (1 << 15, 1 << 15), (70000, 1000), (1 << 18, (1 << 18) - 1)])
def test_three_prime_products_vs_gmp(gpu, oracle, bits, LA, LB):
    """The reference pipelines' three-prime sets (chosen explicitly) against GMP
    (up to 2048 output limbs the carry runs one tile of sixteen waves of two
    rows, up to 4096 one of four rows, beyond that tiles of the default shape)."""
    A = oracle.uniform_limbs(LA, LA * 3 + 7, bits)
    B = oracle.uniform_limbs(LB, LB * 5 + 9, bits)
    s = gpu.default_prime_set(bits, gpu.padded_ntt_size(LA, LB))
    plan = _product_check(gpu, oracle, A, B, bits, s)
    assert plan.k == 3


# ---- 32-bit limbs as limb pairs under the 58-bit primes ----------------------
@pytest.mark.parametrize("bits,LA,LB", [(32, 1, 1), (32, 3, 2), (32, 1000, 999), (32, 4097, 3000),
                                        (32, 1 << 16, (1 << 16) - 1), (32, 1 << 20, 1 << 20),
                                        (64, 2000, 1999), (64, 1 << 17, 5)])
def test_limb_pairs_vs_gmp(gpu, oracle, bits, LA, LB):
    """BW_SET_U64_PAIRS: 64-bit digits (odd 32-bit limb counts leave a half
    digit) in 64-bit integer arithmetic, against GMP."""
    A = oracle.uniform_limbs(LA, 500 + LA, bits)
    B = oracle.uniform_limbs(LB, 501 + LB, bits)
    plan = _product_check(gpu, oracle, A, B, bits, gpu.SET_U64_PAIRS)
    assert plan.info.prime_set == gpu.SET_U64_PAIRS and plan.info.digit_bits == 64


def _value_mod(limbs, q, block=1 << 14):
    """A little-endian 32-bit limb array's value mod q (q < 2^31), vectorized
    over blocks: sum_i a_i 2^(32 i) = sum_h 2^(32 B h) sum_l a_(hB+l) 2^(32 l)."""
    a = np.asarray(limbs, dtype=np.uint32)
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


def test_limb_pairs_auto_beyond_five_primes(gpu, oracle):
    """2^25 + 5 limbs per operand: past the five-prime set (2^25 limbs) and the
    32-bit-digit sets (N <= 2^26), so the automatic choice is the limb-pair set.
    Checked by residues mod three primes and by the exact low 2048 limbs."""
    L = (1 << 25) + 5
    rng = np.random.default_rng(77)
    A = rng.integers(0, 2 ** 32, L, dtype=np.uint64).astype(np.uint32)
    B = rng.integers(0, 2 ** 32, L - 3, dtype=np.uint64).astype(np.uint32)
    A[-1] |= 1
    B[-1] |= 1
    d = dev()
    plan = gpu.Plan(32, len(A), len(B))
    assert plan.info.prime_set == gpu.SET_U64_PAIRS
    print("limb pairs: %d x %d limbs, N = 2^%d digits" % (len(A), len(B), plan.info.N_exec.bit_length() - 1),
          flush=True)
    c = d.empty(len(A) + len(B), 32)
    plan.execute(d.to_dev(A), d.to_dev(B), c)
    got = d.from_dev(c, len(A) + len(B))
    plan.status()
    plan.close()
    for q in (2147483647, 2147483629, 2147483587):
        assert _value_mod(got, q) == _value_mod(A, q) * _value_mod(B, q) % q, q
    m = 2048
    low = lambda v: int.from_bytes(np.ascontiguousarray(v[:m]).tobytes(), "little")
    want_low = (low(A) * low(B)) % (1 << (32 * m))
    assert low(got) == want_low


# ---- row-resident fused pass (2^14 / 2^15-point rows) -----------------------
@pytest.mark.parametrize("bits,LA,LB,pset,passes", [
    (32, 1 << 15, (1 << 15) - 3, -1, [1, 14]),        # five primes, N = 2^15 digits
    (32, 1 << 16, 1 << 16, -1, [1, 15]),              # N = 2^16
    (32, 70000, 60001, -1, [2, 14]),                  # N = 2^16
    (64, 1 << 16, 1 << 15, -1, [2, 15]),              # 64-bit limbs, N = 2^17
    (32, 1 << 19, 1 << 19, 1, [6, 14]),               # three primes (wide set), 32-bit digits, N = 2^20
    (32, 1 << 19, (1 << 19) - 7, 1, [5, 15]),
    (32, 1 << 21, 1 << 21, -1, [6, 15]),              # N = 2^21
])
def test_row_fused_pass_vs_gmp(gpu, oracle, bits, LA, LB, pset, passes):
    """The last forward pass, the pointwise product and the first inverse pass
    on whole rows resident in LDS, against GMP."""
    A = oracle.uniform_limbs(LA, 900 + LA % 97, bits)
    B = oracle.uniform_limbs(LB, 901 + LB % 89, bits)
    plan = _product_check(gpu, oracle, A, B, bits, pset, passes)
    assert plan.passes == passes


def test_row_fused_full_size(gpu, oracle):
    """2^24 x 2^24 limbs with the [9, 15] and [10, 14] splits (three launches)."""
    A = oracle.uniform_limbs(1 << 24, 41, 32)
    B = oracle.uniform_limbs(1 << 24, 42, 32)
    want, _ = oracle.gmp_mul(A, B, 32)
    d = dev()
    da, db = d.to_dev(A), d.to_dev(B)
    for passes in ([9, 15], [10, 14]):
        print("row split", passes, flush=True)
        plan = gpu.Plan(32, len(A), len(B), gpu.SET_DEFAULT, passes)
        c = d.empty(len(A) + len(B), 32)
        plan.execute(da, db, c)
        assert np.array_equal(d.from_dev(c, len(A) + len(B)), want), passes
        plan.close()


def test_row_fused_split_rules(gpu):
    """Row passes are last passes of the FP64 path only."""
    with pytest.raises(gpu.BWError):
        gpu.Plan(32, 1 << 14, 1 << 14, gpu.SET_DEFAULT, [15])          # a single pass
    with pytest.raises(gpu.BWError):
        gpu.Plan(32, 1 << 20, 1 << 20, gpu.SET_DEFAULT, [15, 5])       # not the last pass
    with pytest.raises(gpu.BWError):
        gpu.Plan(64, 1 << 16, 1 << 16, gpu.SET_U64_REF, [3, 14])       # 64-bit arithmetic


# ---- the hybrid pipeline: 32-bit limbs under the two 58-bit primes -----------
HYBRID = 3


@pytest.mark.parametrize("logN", [1, 3, 6, 9, 12, 14, 17])
def test_hybrid_forward_and_roundtrip(gpu, oracle, logN):
    """BW_SET_HYBRID (src/gpu_ntt.cu:40-43): 32-bit limbs enter 64-bit residue
    words; the forward NTT against the oracle's per prime, then the inverse
    returns the zero-padded limbs."""
    d = dev()
    N = 1 << logN
    L = max(1, N // 2)
    plan = gpu.Plan(32, L, L, HYBRID)
    assert plan.N_exec == N and plan.k == 2 and plan.prime_set == HYBRID
    raw = oracle.uniform_limbs(L, 2000 + logN, 32)
    spec = d.empty(plan.k * N, 64)
    plan.forward(d.to_dev(raw), L, spec)
    got = d.from_dev(spec).reshape(plan.k, N)
    order = plan.spectrum_order()
    P, R = oracle.prime_set(HYBRID)
    assert list(P) == plan.primes
    for j in range(plan.k):
        x = np.zeros(N, np.uint64)
        x[:L] = raw.astype(np.uint64) % P[j]
        want = oracle.ntt_forward(x, P[j], oracle.root_for_N(P[j], R[j], logN))
        nat = np.zeros(N, np.uint64)
        nat[order.astype(np.int64)] = got[j].astype(np.uint64)
        assert np.array_equal(nat, want), "prime %d" % j
    plan.inverse(spec)
    back = d.from_dev(spec).reshape(plan.k, N)
    for j in range(plan.k):
        want = np.zeros(N, np.uint64)
        want[:L] = raw.astype(np.uint64) % P[j]
        assert np.array_equal(back[j].astype(np.uint64), want)


@pytest.mark.parametrize("LA,LB", [(1, 1), (1, 2), (3, 5), (100, 3), (1000, 1003), (4097, 4095), (5000, 17),
                                   (1 << 15, 1 << 15), (70000, 1000), (1 << 20, (1 << 20) - 1)])
def test_hybrid_products_vs_gmp(gpu, oracle, LA, LB):
    A = oracle.uniform_limbs(LA, 300 + LA, 32)
    B = oracle.uniform_limbs(LB, 301 + LB, 32)
    plan = _product_check(gpu, oracle, A, B, 32, HYBRID)
    assert plan.k == 2 and plan.info.digit_bits == 32


def test_hybrid_all_ones_and_splits(gpu, oracle):
    """All-ones operands (the largest coefficients) and several pass splits."""
    L = 1 << 16
    X = np.full(L, 0xFFFFFFFF, np.uint32)
    _product_check(gpu, oracle, X, X, 32, HYBRID)
    A = oracle.uniform_limbs(3001, 5, 32)
    B = oracle.uniform_limbs(2999, 6, 32)
    for passes in ([6, 7], [4, 9], [3, 4, 6], [10, 3]):
        _product_check(gpu, oracle, A, B, 32, HYBRID, passes)


@pytest.mark.parametrize("L", [4, 256, 1 << 15])
def test_hybrid_execute_crt_matches_oracle(gpu, oracle, L):
    """The hybrid pipeline's early exit: 128-bit coefficients (lo, hi) of its
    two primes, equal to the oracle's Garner (src/crt_gpu.cu:157-203)."""
    d = dev()
    A = oracle.ref_pipeline_limbs(L, 700 + L)
    B = oracle.ref_pipeline_limbs(L, 701 + L)
    plan = gpu.Plan(32, L, L, HYBRID)
    N = plan.N_exec
    c0, c1 = d.empty(N, 64), d.empty(N, 64)
    plan.execute_crt(d.to_dev(A), d.to_dev(B), c0, c1)
    w0, w1, _ = oracle.multiply_crt(A, B, 32, HYBRID)
    assert np.array_equal(d.from_dev(c0, N), w0[:N])
    assert np.array_equal(d.from_dev(c1, N), w1[:N])
    # and the standalone Garner over the same residues
    rng = np.random.default_rng(3)
    P, _ = oracle.prime_set(HYBRID)
    x = [int(v) for v in rng.integers(0, 2**62, 64, dtype=np.uint64)]
    x = [(a * (2**52) + 12345) % (P[0] * P[1]) for a in x]
    r = [d.to_dev(np.array([v % p for v in x], dtype=np.uint64)) for p in P]
    o0, o1 = d.empty(64, 64), d.empty(64, 64)
    gpu.crt_combine(HYBRID, r[0], r[1], r[0], 64, o0, o1)
    lo, hi = d.from_dev(o0, 64), d.from_dev(o1, 64)
    assert [int(h) << 64 | int(l) for l, h in zip(lo, hi)] == x
