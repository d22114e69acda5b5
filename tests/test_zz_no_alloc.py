"""The allocation-free entry point, bw_multiply_device_noalloc, through the C
ABI: the same limbs as the plan entry point (bw_execute) on the same
operands, for two calls in a row on one GPU (the second smaller, so it runs
in scratch the first one left behind), and a clear refusal of products beyond
its scratch.  It has no counterpart in the reference; the bar is the plan's
result, which the other GPU tests hold to GMP.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from devbuf import empty, from_dev, to_dev

pytestmark = pytest.mark.gpu


def _noalloc(gpu, bits, da, LA, db, LB, dc, status):
    s = torch.cuda.current_stream()
    return gpu.lib().bw_multiply_device_noalloc(bits, C.c_void_p(da.data_ptr()), LA, C.c_void_p(db.data_ptr()), LB,
                                                C.c_void_p(dc.data_ptr()),
                                                C.c_void_p(status.data_ptr()) if status is not None else None,
                                                C.c_void_p(s.cuda_stream))


def _plan_product(gpu, bits, da, db, LA, LB):
    plan = gpu.Plan(bits, LA, LB)
    c = empty(LA + LB, bits)
    plan.execute(da, db, c)
    got = from_dev(c, LA + LB)
    plan.status()
    plan.close()
    return got


@pytest.mark.parametrize("bits,sizes", [
    (32, [(1 << 18, 1 << 18), (1000, 3001)]),
    (64, [(1 << 17, (1 << 17) - 3), (5, 7)]),
    (32, [(70001, 65536), (1, 1)]),
    (64, [(40000, 9), (3, 1)]),
])
def test_noalloc_matches_plan(gpu, oracle, bits, sizes):
    dev = torch.cuda.current_device()
    assert gpu.lib().bw_init(dev) == gpu.BW_OK, gpu.lib().bw_last_error()
    for call, (LA, LB) in enumerate(sizes):
        A = oracle.uniform_limbs(LA, 9100 + 7 * call + LA, bits)
        B = oracle.uniform_limbs(LB, 9200 + 7 * call + LB, bits)
        if call == 0:
            A[:64] = np.iinfo(A.dtype).max          # a carry chain across the first words
        da, db = to_dev(A), to_dev(B)
        want = _plan_product(gpu, bits, da, db, LA, LB)
        dc = empty(LA + LB, bits)
        dc.fill_(0x5A5A)                            # poison: every limb must be written
        status = torch.full((1,), 7, dtype=torch.int32, device="cuda")
        print("noalloc call %d: %d x %d %d-bit limbs" % (call, LA, LB, bits), flush=True)
        rc = _noalloc(gpu, bits, da, LA, db, LB, dc, status)
        assert rc == gpu.BW_OK, gpu.lib().bw_last_error()
        got = from_dev(dc, LA + LB)
        assert int(status.cpu()[0]) == 0
        assert np.array_equal(got, want), (call, LA, LB, bits)
        if call == 0:
            g, _ = oracle.gmp_mul(A, B, bits)
            assert np.array_equal(got, g)


def test_noalloc_without_status_word(gpu, oracle):
    assert gpu.lib().bw_init(torch.cuda.current_device()) == gpu.BW_OK
    A = oracle.uniform_limbs(3000, 1, 32)
    B = oracle.uniform_limbs(2000, 2, 32)
    da, db = to_dev(A), to_dev(B)
    dc = empty(5000, 32)
    assert _noalloc(gpu, 32, da, 3000, db, 2000, dc, None) == gpu.BW_OK
    g, _ = oracle.gmp_mul(A, B, 32)
    assert np.array_equal(from_dev(dc, 5000), g)


def test_noalloc_refuses_beyond_scratch(gpu, oracle):
    assert gpu.lib().bw_init(torch.cuda.current_device()) == gpu.BW_OK
    for bits, LA, LB in ((32, (1 << 18) + 1, (1 << 18) + 1), (64, (1 << 17) + 1, (1 << 17) + 1), (32, 0, 5), (48, 4, 4)):
        n = max(LA + LB, 1)
        da, db, dc = empty(max(LA, 1), 64), empty(max(LB, 1), 64), empty(n, 64)
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        rc = _noalloc(gpu, bits, da, LA, db, LB, dc, status)
        assert rc in (gpu.BW_ERR_UNSUPPORTED_SIZE, gpu.BW_ERR_ARG), (bits, LA, LB, rc)
        if bits in (32, 64) and LA and LB:
            assert rc == gpu.BW_ERR_UNSUPPORTED_SIZE
            assert b"allocation-free" in gpu.lib().bw_last_error()
