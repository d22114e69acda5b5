"""Pins the CPU oracle (oracle/) before it is trusted as the GPU path's checker.

Each oracle function is checked against something that does not share its
code: the golden fixtures (pure Python, tests/golden/make_golden.py), the
reference's own known answers, the reference's own CPU CRT compiled from its
sources (oracle/_ref), g++'s std::mt19937_64, GMP, and mathematical
identities chosen so that a dropped term, wrong sign/index or swapped operand
fails at least one of them.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


SETS = {"u32_ref": 0, "u32_wide": 1, "u64_ref": 2}


# ------------------------------------------------------------------ MT19937-64
def test_mt64_standard_known_answer(oracle):
    """C++ [rand.predef]: the 10000th output of default-seeded mt19937_64."""
    m = oracle.MT64(5489)
    for _ in range(9999):
        m()
    assert m() == 9981545732273789042


def test_mt64_matches_std_library(oracle):
    """Same stream as g++'s std::mt19937_64 for the seeds the reference tests use."""
    src = r'''
#include <random>
#include <cstdio>
int main(){ for (unsigned long long s : {1234ULL, 5678ULL, 100ULL, 1ULL}) { std::mt19937_64 r(s);
  for (int i=0;i<5;i++) printf("%llu\n", (unsigned long long)r()); } }'''
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "m.cpp"), "w") as f:
            f.write(src)
        subprocess.run(["g++", "-O1", "-o", os.path.join(d, "m"), os.path.join(d, "m.cpp")], check=True)
        out = [int(x) for x in subprocess.run([os.path.join(d, "m")], capture_output=True, text=True,
                                              check=True).stdout.split()]
    mine = []
    for s in (1234, 5678, 100, 1):
        m = oracle.MT64(s)
        mine += [m() for _ in range(5)]
    assert mine == out


def test_reference_limb_patterns(oracle):
    """tests/test_full_multiply.cpp:61-73 pattern: i%8==0 -> 0, ==1 -> 1, else rng()%2^30
    (rng advanced only for the random slots)."""
    a = oracle.ref_random_limbs(24, 1234)
    m = oracle.MT64(1234)
    want = []
    for i in range(24):
        want.append(0 if i % 8 == 0 else 1 if i % 8 == 1 else (m() & 0xFFFFFFFF) % (1 << 30))
    assert a.tolist() == want


# ------------------------------------------------------------------ primes / size policy
def _is_prime(n):
    if n < 2:
        return False
    for sp in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % sp == 0:
            return n == sp
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


@pytest.mark.parametrize("s", [0, 1, 2])
def test_prime_sets(oracle, s):
    P, R = oracle.prime_set(s)
    for p, r in zip(P, R):
        assert _is_prime(p)
        L = oracle.lib().orc_max_root_logN(r, p)
        # root has exact order 2^L: r^(2^(L-1)) == -1
        assert pow(r, 1 << (L - 1), p) == p - 1 and pow(r, 1 << L, p) == 1
    assert oracle.lib().orc_max_supported_logN(s) == {0: 23, 1: 26, 2: 32}[s]


def test_size_policy_reference_expectations(oracle):
    """tests/test_ntt_limits.cpp:63-323, restated for the three configurations."""
    L = oracle.lib()
    assert [L.orc_max_logN_for_prime(p) for p in (0x2d000001, 0x23800001, 0x26800001)] == [24, 23, 23]
    assert [L.orc_max_logN_for_prime(p) for p in (0x400002600000001, 0x400004200000001, 0x400001100000001)] == [33, 33, 32]
    assert L.orc_max_root_logN(663, 0x2d000001) == 23
    assert L.orc_max_root_logN(721, 0x23800001) == 23
    assert L.orc_max_root_logN(19, 0x26800001) == 23
    assert L.orc_max_root_logN(273765203699653965, 0x400002600000001) == 33
    assert L.orc_max_root_logN(26231613454922890, 0x400004200000001) == 33
    assert L.orc_max_root_logN(144261359744151243, 0x400001100000001) == 32
    assert [L.orc_padded_ntt_size(a, b) for a, b in ((1, 1), (4, 4), (5, 5), (1 << 22, 1 << 22), (1 << 23, 1 << 23))] == \
        [1, 8, 16, 1 << 23, 1 << 24]
    # 32-bit configuration (u32 reference set, 32-bit inputs)
    assert L.orc_max_supported_logN(0) == 23
    assert L.orc_max_supported_limb_count(0, 32) == 1 << 22
    assert L.orc_crt_bound_ok(0, 32, 1 << 22, 1 << 22)
    assert L.orc_crt_bound_ok(0, 32, 1 << 23, 1 << 23)
    assert not L.orc_crt_bound_ok(0, 32, 1 << 24, 1 << 24)
    assert L.orc_crt_bound_ok(0, 32, 1 << 24, 1)          # rectangular: short operand keeps the bound
    assert L.orc_multiply_size_supported(0, 32, 1 << 22, 1 << 22)
    assert not L.orc_multiply_size_supported(0, 32, 1 << 23, 1 << 23)
    for logN in range(1, 24):
        assert L.orc_ntt_size_supported(0, 1 << logN)
    assert not L.orc_ntt_size_supported(0, 0) and not L.orc_ntt_size_supported(0, 3)
    assert not L.orc_ntt_size_supported(0, 1 << 24)
    # hybrid configuration (2 x 58-bit primes, 32-bit inputs)
    assert L.orc_max_supported_logN(3) == 33
    assert L.orc_max_supported_limb_count(3, 32) == 1 << 32
    assert L.orc_crt_bound_ok(3, 32, 1 << 32, 1 << 32)
    assert not L.orc_crt_bound_ok(3, 32, 1 << 53, 1 << 53)
    assert not L.orc_multiply_size_supported(3, 32, 1 << 33, 1 << 33)
    # 64-bit configuration (3 x 58-bit primes, 64-bit inputs)
    assert L.orc_max_supported_logN(2) == 32
    assert L.orc_max_supported_limb_count(2, 64) == 1 << 31
    assert L.orc_crt_bound_ok(2, 64, 1 << 31, 1 << 31)
    assert not L.orc_crt_bound_ok(2, 64, 1 << 50, 1 << 50)
    assert not L.orc_multiply_size_supported(2, 64, 1 << 32, 1 << 32)
    # wide u32 set reaches the 2^24-limb product
    assert L.orc_multiply_size_supported(1, 32, 1 << 24, 1 << 24)
    assert not L.orc_multiply_size_supported(0, 32, 1 << 24, 1 << 24)


@pytest.mark.parametrize("binary", ["test_ntt_limits_32", "test_ntt_limits_hybrid", "test_ntt_limits_64",
                                    "test_crt_32", "test_crt_64"])
def test_reference_cpu_test_binaries(binary):
    """The reference's own CPU tests, built from its sources into oracle/_ref."""
    path = os.path.join(ROOT, "oracle", "_ref", binary)
    if not os.path.exists(path):
        pytest.skip("oracle/_ref not built (reference tree absent)")
    r = subprocess.run([path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "FAIL" not in r.stdout


# ------------------------------------------------------------------ transforms
def test_ntt_golden_dft(oracle):
    """Oracle NTT == O(N^2) definition (pure Python fixture) for N <= 64, all sets."""
    n = 0
    for c in load("ntt_small.json"):
        p, logN = c["p"], c["logN"]
        w = oracle.root_for_N(p, c["root_max"], logN)
        assert w == c["w"]
        x = np.array(c["x"], dtype=np.uint64)
        assert oracle.ntt_forward(x, p, w).tolist() == c["X"]
        assert oracle.dft_bruteforce(x, p, w).tolist() == c["X"]
        assert oracle.ntt_inverse(np.array(c["X"], dtype=np.uint64), p, w).tolist() == c["x"]
        n += 1
    assert n == 3 * 3 * 7


@pytest.mark.parametrize("s,logN", [(0, 10), (1, 12), (2, 11), (0, 16)])
def test_ntt_convolution_theorem(oracle, s, logN):
    """INTT(NTT(a) * NTT(b)) == cyclic convolution, computed by schoolbook mod p."""
    P, R = oracle.prime_set(s)
    N = 1 << logN
    rng = np.random.default_rng(logN)
    for p, r in zip(P, R):
        w = oracle.root_for_N(p, r, logN)
        La, Lb = N // 2, N // 2 - 3
        a = rng.integers(0, p, La, dtype=np.uint64)
        b = rng.integers(0, p, Lb, dtype=np.uint64)
        A = np.zeros(N, np.uint64); A[:La] = a
        B = np.zeros(N, np.uint64); B[:Lb] = b
        c = oracle.ntt_inverse(oracle.pointwise(oracle.ntt_forward(A, p, w), oracle.ntt_forward(B, p, w), p), p, w)
        if logN <= 12:
            assert np.array_equal(c, oracle.conv_schoolbook_mod(a, b, N, p))
        else:   # spot-check coefficients against the definition
            for i in (0, 1, La - 1, N // 2, N - 5):
                want = sum(int(a[j]) * int(b[i - j]) for j in range(max(0, i - Lb + 1), min(i, La - 1) + 1)) % p
                assert int(c[i]) == want


def test_planted_mistake_is_caught(oracle):
    """A transform with the wrong root (w^-1) must fail the DFT fixture."""
    c = [x for x in load("ntt_small.json") if x["logN"] == 4][0]
    p, w = c["p"], c["w"]
    wrong = oracle.ntt_forward(np.array(c["x"], np.uint64), p, pow(w, -1, p))
    assert wrong.tolist() != c["X"]


# ------------------------------------------------------------------ Garner
def test_garner_golden_kats(oracle):
    """tests/test_crt_gpu.cu and tests/test_crt.cpp known answers (exact x)."""
    for c in load("crt_kats.json"):
        P = c["primes"]
        res = np.array(c["residues"], dtype=np.uint64).T
        if len(P) == 3 and c["set"] == "u64_ref":
            lo, mid, hi = oracle.garner160(P, res)
            got = [int(a) | (int(b) << 64) | (int(h) << 128) for a, b, h in zip(lo, mid, hi)]
            want = [x % 2**160 for x in c["x"]]        # U160 holds x mod 2^160 (include/wide_int.h:6-11)
        else:
            hi, lo = oracle.garner128(P, res)
            got = [int(l) | (int(h) << 64) for h, l in zip(hi, lo)]
            want = c["x"]
        assert got == want, c["set"]


@pytest.mark.parametrize("bits", [32, 64])
def test_garner_vs_reference_crt_combine_many(oracle, bits):
    """oracle/_ref: the reference's crt_combine_many (src/crt.cpp:42-87), valid
    while prod(p) < 2^128: the 32-bit sets (k=3) and the hybrid set (k=2)."""
    L = oracle.ref_crt_lib(bits)
    if L is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    sets = [0, 1] if bits == 32 else [3]
    rng = np.random.default_rng(bits)
    for s in sets:
        P, _ = oracle.prime_set(s)
        res = np.array([rng.integers(0, p, 300, dtype=np.uint64) for p in P])
        hi, lo = oracle.garner128(P, res)
        for i in range(res.shape[1]):
            want = oracle.ref_crt_combine_many(L, P, [int(r) for r in res[:, i]])
            assert want == int(lo[i]) | (int(hi[i]) << 64)


def test_garner_uniqueness_random(oracle):
    """x == r_j mod p_j and x < prod(p) for random residues; the U160 layout
    holds x mod 2^160 (x itself may reach prod(p) ~ 2^174)."""
    P, _ = oracle.prime_set(2)
    M = P[0] * P[1] * P[2]
    rng = np.random.default_rng(3)
    res = np.array([rng.integers(0, p, 200, dtype=np.uint64) for p in P])
    lo, mid, hi = oracle.garner160(P, res)
    for i in range(200):
        r = [int(res[j, i]) for j in range(3)]
        x = sum(rj * (M // p) * pow(M // p, -1, p) for rj, p in zip(r, P)) % M
        got = int(lo[i]) | (int(mid[i]) << 64) | (int(hi[i]) << 128)
        assert got == x % 2**160
    P32, _ = oracle.prime_set(1)
    res = np.array([rng.integers(0, p, 200, dtype=np.uint64) for p in P32])
    hi2, lo2 = oracle.garner128(P32, res)
    for i in range(200):
        x = int(lo2[i]) | (int(hi2[i]) << 64)
        assert x < P32[0] * P32[1] * P32[2] and all(x % p == int(res[j, i]) for j, p in enumerate(P32))


# ------------------------------------------------------------------ carry
def test_carry_golden_kats(oracle):
    """tests/test_carry_prop.cu inputs; expected limbs from exact Python sums."""
    for c in load("carry_kats.json"):
        vals = [int(v) for v in c["coeff"]]
        want = [int(v) for v in c["limbs"]]
        if c["out_bits"] == 32:
            hi = np.array([v >> 64 for v in vals], np.uint64)
            lo = np.array([v & (2**64 - 1) for v in vals], np.uint64)
            got = oracle.carry128(hi, lo, len(want), 32)
        else:
            lo = np.array([v & (2**64 - 1) for v in vals], np.uint64)
            mid = np.array([(v >> 64) & (2**64 - 1) for v in vals], np.uint64)
            hi = np.array([v >> 128 for v in vals], np.uint32)
            got = oracle.carry160(lo, mid, hi, len(want))
        assert [int(x) for x in got] == want, c["name"]


# ------------------------------------------------------------------ whole product
def test_products_golden(oracle):
    for c in load("products_small.json"):
        bits = c["bits"]
        dt = np.uint64 if bits == 64 else np.uint32
        A = np.array([int(x) for x in c["A"]], dtype=dt)
        B = np.array([int(x) for x in c["B"]], dtype=dt)
        want = [int(x) for x in c["C"]]
        assert [int(x) for x in oracle.multiply(A, B, bits)] == want
        assert [int(x) for x in oracle.schoolbook(A, B, bits)] == want
        if oracle.gmp_available():
            assert [int(x) for x in oracle.gmp_mul(A, B, bits)[0]] == want


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("LA,LB", [(1, 1), (8, 8), (1000, 999), (4096, 4096), (1 << 15, 3), (30000, 30000)])
def test_products_vs_python_int(oracle, bits, LA, LB):
    A = oracle.uniform_limbs(LA, LA + 1, bits)
    B = oracle.uniform_limbs(LB, LB + 2, bits)
    want = oracle.to_int(A, bits) * oracle.to_int(B, bits)
    assert oracle.to_int(oracle.multiply(A, B, bits), bits) == want
    if LA * LB <= 10**7:
        assert oracle.to_int(oracle.schoolbook(A, B, bits), bits) == want


def test_products_crt_coefficients(oracle):
    """multiply_crt coefficients equal the linear convolution (exact)."""
    A = oracle.ref_pipeline_limbs(64, 300)
    B = oracle.ref_pipeline_limbs(64, 301)
    c0, c1, _ = oracle.multiply_crt(A, B, 32)
    conv = [sum(int(A[j]) * int(B[i - j]) for j in range(max(0, i - 63), min(i, 63) + 1)) for i in range(128)]
    assert [int(a) | (int(b) << 64) for a, b in zip(c0, c1)] == conv
