"""CPU checks of the C ABI library: it loads, exports every entry point that
include/birdwing.h declares, and its host-only size policy agrees with the
oracle and the reference's own tests/test_ntt_limits.cpp expectations.  No
compute calls (no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "birdwing.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bw_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import birdwing_amd as bw
    L = bw.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", bw.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (bw_\w+)", out))
    assert set(names) <= exported
    assert set(names) == set(bw.SIGNATURES), "Python bindings out of sync with the header"


def test_version_and_error_string():
    import birdwing_amd as bw
    assert b"gfx950" in bw.lib().bw_version()
    assert isinstance(bw.lib().bw_last_error(), bytes)


def test_size_policy_matches_oracle(oracle):
    import birdwing_amd as bw
    O = oracle.lib()
    for s in (0, 1, 2, 3):
        assert bw.max_supported_logN(s) == O.orc_max_supported_logN(s)
        for bits in (32, 64):
            assert bw.max_limb_count_by_crt(s, bits) == O.orc_max_limb_count_by_crt(s, bits)
            assert bw.max_supported_limb_count(s, bits) == O.orc_max_supported_limb_count(s, bits)
    for L1, L2 in ((1, 1), (4, 4), (5, 5), (1 << 22, 1 << 22), (1 << 23, 1 << 23), (3, 1 << 20)):
        assert bw.padded_ntt_size(L1, L2) == O.orc_padded_ntt_size(L1, L2)
        for s, bits in ((0, 32), (1, 32), (2, 64), (3, 32)):
            assert bw.multiply_size_supported(s, bits, L1, L2)[0] == bool(O.orc_multiply_size_supported(s, bits, L1, L2))
            assert bw.crt_bound_satisfied(s, bits, L1, L2)[0] == bool(O.orc_crt_bound_ok(s, bits, L1, L2))
    for logN in range(0, 36):
        for s in (0, 1, 2, 3):
            assert bw.ntt_size_supported(s, 1 << logN)[0] == bool(O.orc_ntt_size_supported(s, 1 << logN))


def test_rejection_messages():
    """tests/test_ntt_limits.cpp:186-191,277-281: messages name the failing bound."""
    import birdwing_amd as bw
    ok, why = bw.crt_bound_satisfied(0, 32, 1 << 24, 1 << 24)
    assert not ok and "CRT modulus product" in why
    ok, why = bw.ntt_size_supported(0, 1 << 24)
    assert not ok and "24" in why
    ok, why = bw.ntt_size_supported(0, 3)
    assert not ok and "power of two" in why
    ok, why = bw.multiply_size_supported(0, 32, 0, 5)
    assert not ok and "positive" in why


def test_default_prime_sets():
    import birdwing_amd as bw
    assert bw.default_prime_set(32, 4096) == 0
    assert bw.default_prime_set(32, 1 << 23) == 0
    assert bw.default_prime_set(32, 1 << 25) == 1
    assert bw.default_prime_set(64, 1 << 25) == 2
    assert bw.prime_set_info(1)[0] == [2013265921, 1811939329, 469762049]


def _is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_d64_prime_set():
    """The five-prime 64-bit-digit set: primes c*2^m+1 < 2^31 with c < 128 and
    m >= 25, roots of exact order 2^m, product above the coefficient bound of
    2^24-digit operands (2^24 * (2^64-1)^2)."""
    import birdwing_amd as bw
    primes, roots = bw.prime_set_info(bw.SET_D64)
    assert len(primes) == 5
    prod = 1
    for p, r in zip(primes, roots):
        assert _is_prime(p) and p < 2**31
        m = ((p - 1) & -(p - 1)).bit_length() - 1
        assert m >= 25 and (p - 1) >> m < 128
        assert pow(r, 1 << m, p) == 1 and pow(r, 1 << (m - 1), p) == p - 1
        prod *= p
    assert prod > (1 << 24) * (2**64 - 1) ** 2
    assert bw.max_supported_logN(bw.SET_D64) == 25
    # 32-bit limbs pair into digits: 2^25 limbs per operand fit, one more does not
    assert bw.multiply_size_supported(bw.SET_D64, 32, 1 << 25, 1 << 25)[0]
    assert not bw.multiply_size_supported(bw.SET_D64, 32, (1 << 25) + 1, (1 << 25) + 1)[0]
    assert bw.multiply_size_supported(bw.SET_D64, 64, 1 << 24, 1 << 24)[0]
    assert not bw.multiply_size_supported(bw.SET_D64, 64, (1 << 24) + 1, (1 << 24) + 1)[0]
    assert bw.max_supported_limb_count(bw.SET_D64, 32) == 1 << 25


def test_limb_pair_prime_set():
    """32-bit limbs beyond the 32-bit-digit sets run as limb pairs under the
    reference's 58-bit primes (the reach of its hybrid pipeline,
    include/config.h:26-30): coefficients < N * 2^128 below their product,
    transforms up to 2^32 digits (2-adic valuation of the third prime)."""
    import birdwing_amd as bw
    primes, roots = bw.prime_set_info(bw.SET_U64_PAIRS)
    assert (primes, roots) == bw.prime_set_info(bw.SET_U64_REF)
    prod = primes[0] * primes[1] * primes[2]
    assert prod > (1 << 32) * (2**64 - 1) ** 2
    assert bw.max_supported_logN(bw.SET_U64_PAIRS) == 32
    assert bw.max_supported_limb_count(bw.SET_U64_PAIRS, 32) == 1 << 32
    assert bw.multiply_size_supported(bw.SET_U64_PAIRS, 32, 1 << 32, 1 << 32)[0]
    assert not bw.multiply_size_supported(bw.SET_U64_PAIRS, 32, (1 << 32) + 2, (1 << 32) + 2)[0]
    # the automatic choice (-1): the five-prime set through 2^25 limbs per
    # operand, limb pairs beyond it, nothing past 2^32
    for L, ok in ((1 << 25, True), ((1 << 25) + 1, True), (1 << 30, True), ((1 << 32) + 2, False)):
        assert bw.multiply_size_supported(-1, 32, L, L)[0] == ok, L
    assert not bw.multiply_size_supported(bw.SET_U32_WIDE, 32, (1 << 25) + 1, (1 << 25) + 1)[0]


def test_abi_version_matches_header():
    import re
    import birdwing_amd as bw
    hdr = open(os.path.join(ROOT, "include", "birdwing.h")).read()
    v = int(re.search(r"#define BW_ABI_VERSION (\d+)", hdr).group(1))
    assert bw.lib().bw_abi_version() == v


def test_fused_limit_setter():
    """bw_set_fused_limit: 0 .. 11 accepted (returns the previous limit, 8
    at start), anything else refused with BW_ERR_ARG; host only."""
    import birdwing_amd as bw
    L = bw.lib()
    first = L.bw_set_fused_limit(5)
    assert first == 8
    assert L.bw_set_fused_limit(11) == 5
    assert L.bw_set_fused_limit(12) == bw.BW_ERR_ARG
    assert L.bw_set_fused_limit(-1) == bw.BW_ERR_ARG
    assert "fused limit" in L.bw_last_error().decode()
    assert L.bw_set_fused_limit(0) == 11
    assert bw.set_fused_limit(first) == 0
    with pytest.raises(bw.BWError):
        bw.set_fused_limit(40)
