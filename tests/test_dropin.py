"""The reference's C++ interface (include/*.h) over libbirdwing_hip.so.

tests/cpp/dropin_test.cpp includes config.h, ntt_limits.h, crt.h,
crt_utils.h, crt_gpu.h, carry_prop.h, zero_pad.h, gpu_ntt.h and multiply.h as
a reference caller would (src/main.cpp, tests/test_host_multiply.cpp), in the
three pipeline configurations of include/config.h, and checks products
against GMP.  tests/cpp/dropin_carry_kernels.hip launches the three-pass
carry kernels of carry_prop.h (src/carry_prop.cu) and compares them with the
library's single-pass carry.  The binaries are built by tests/cpp/Makefile
(run by __graft_entry__.build()).
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
ROOT = os.path.dirname(HERE)
CONFIGS = {
    "u32": ["-DLIMB_BITS=32"],
    "hybrid": ["-DLIMB_BITS=64"],
    "u64": ["-DLIMB_BITS=64", "-DNATIVE_HOST_LIMBS"],
    "u32ext": ["-DLIMB_BITS=32", "-DBW_EXTENDED_SIZES"],
}


def _bin(name):
    path = os.path.join(CPP, name)
    if not os.path.exists(path):
        pytest.fail("%s is not built (run __graft_entry__.build() or make -C tests/cpp)" % path)
    return path


@pytest.mark.parametrize("cfg", sorted(CONFIGS))
def test_headers_compile(cfg):
    """Every drop-in header compiles with the host compiler in each configuration."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = os.path.join(CPP, "dropin_test.cpp")
    cmd = [gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-function",
           "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           "-I/opt/conda/include"] + CONFIGS[cfg] + [src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.parametrize("cfg", sorted(CONFIGS))
def test_dropin_host(cfg):
    """Size policy and host CRT through the drop-in headers (no device calls)."""
    r = subprocess.run([_bin("dropin_test_" + cfg)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failures)" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", sorted(CONFIGS))
def test_dropin_gpu(cfg):
    """host_multiply_merge*, execute_ntt_multiply (full and early-exit + carry),
    zero_pad_gpu and crt_combine_gpu* against GMP / host CRT."""
    print("dropin_test_%s --gpu" % cfg, flush=True)
    r = subprocess.run([_bin("dropin_test_" + cfg), "--gpu"], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failures)" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", ["u32", "u64"])
def test_reference_carry_kernels(cfg):
    """carry_intra/inter/fixup kernels (and _u160) equal the single-pass carry."""
    print("dropin_carry_kernels_%s" % cfg, flush=True)
    r = subprocess.run([_bin("dropin_carry_kernels_" + cfg)], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failures)" in r.stdout
