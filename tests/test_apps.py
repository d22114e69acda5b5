"""Command-line front ends: the benchmark harness with the reference's CSV
schema and size convention (bench/gpu_full_multiply_benchmark.cpp), the
limb-file CLI (main.cpp) and its file format (scripts/rng.py), and the
benchmark runner (scripts/run_gpu_bench.py).

Host-only checks run everywhere; the products (checked against GMP) and the
CSV rows of real runs need the GPU.  The binaries are built by
birdwing_amd/apps/Makefile (run by __graft_entry__.build()).
"""
import csv
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(ROOT, "birdwing_amd", "apps", "bin")
PIPES = ("32", "hybrid", "64bit")
LIMB_BITS = {"32": 32, "hybrid": 32, "64bit": 64}

# bench/gpu_full_multiply_benchmark.cpp:384-404 (header line of write_csv)
REFERENCE_CSV_COLUMNS = (
    "pipeline,host_limb_bits,operand_bits,L_arg,L,N,logN,warmup,iters,"
    "mean_ms,stddev_ms,min_ms,max_ms,"
    "setup_pinned_ms,setup_precompute_ms,setup_upload_ms,setup_alloc_ms,"
    "pre_factors_ms,pre_params_ms,pre_twiddle_host_ms,pre_garner_host_ms,"
    "upload_twiddle_ms,upload_mod_constants_ms,upload_garner_ms,"
    "teardown_free_ctx_ms,teardown_free_pre_ms,teardown_free_pinned_ms,"
    "ingress_fwd_mean_ms,h2d_mean_ms,fwd_pad_ntt_mean_ms,fwd_pad_ntt_a_mean_ms,fwd_pad_ntt_b_mean_ms,"
    "mul_mean_ms,intt_mean_ms,crt_mean_ms,carry_mean_ms,d2h_mean_ms"
).split(",")


def _bin(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail("%s is not built (run __graft_entry__.build() or make -C birdwing_amd/apps)" % path)
    return path


def _run(args, timeout=60):
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout)


def _listed(out):
    """(L_arg, L, N) triples printed by --list."""
    rows = []
    for line in out.splitlines():
        if line.startswith("L_arg="):
            rows.append(tuple(int(tok.split("=")[1]) for tok in line.split()))
    return rows


# ---------------------------------------------------------------- host only
@pytest.mark.parametrize("pipe", PIPES)
def test_bench_size_spec(pipe):
    """L < 64 means 2^L limbs, larger values are counts, a-b is inclusive
    (bench/gpu_full_multiply_benchmark.cpp:243-278, 280-285)."""
    r = _run([_bin("bench_full_multiply_" + pipe), "--list", "6-8", "100", "12"])
    assert r.returncode == 0, r.stderr
    got = _listed(r.stdout)
    import birdwing_amd as bw
    want = [(a, (1 << a) if a < 64 else a, bw.padded_ntt_size((1 << a) if a < 64 else a, (1 << a) if a < 64 else a))
            for a in (6, 7, 8, 100, 12)]
    assert got == want
    assert ("pipeline=%s," % pipe) in r.stdout


def test_bench_rejects_bad_spec_and_skips_unsupported():
    exe = _bin("bench_full_multiply_32")
    for bad in ("8-6", "x", "0", "5-"):
        r = _run([exe, "--list", bad])
        assert r.returncode == 1 and "Invalid limb spec" in r.stderr, bad
    # 2^31 limbs run as limb pairs under the 58-bit primes; 2^33 is beyond every set
    r = _run([exe, "--list", "10", "31", "33"])
    assert r.returncode == 0
    assert [t[0] for t in _listed(r.stdout)] == [10, 31]
    assert "Skipping L_arg=33" in r.stderr
    r = _run([exe, "--iters", "0", "--list"])
    assert r.returncode == 1


def test_bench_default_sweep():
    """No sizes: the reference's default sweep, log2 sizes 6..22."""
    r = _run([_bin("bench_full_multiply_32"), "--list"])
    assert r.returncode == 0
    assert [t[0] for t in _listed(r.stdout)] == list(range(6, 23))


def test_cli_usage_and_bad_file(tmp_path):
    exe = _bin("birdwing_32")
    r = _run([exe])
    assert r.returncode == 1 and "Usage" in r.stderr
    odd = tmp_path / "odd.bin"
    odd.write_bytes(b"\x01\x02\x03\x04\x05\x06")       # not a multiple of 2 limbs
    r = _run([exe, "merge", str(odd)])
    assert r.returncode == 1 and "not divisible" in r.stderr
    r = _run([exe, "merge", str(tmp_path / "missing.bin")])
    assert r.returncode == 1 and "Failed to open" in r.stderr


@pytest.mark.parametrize("bits", [32, 64])
def test_limb_file_round_trip(tmp_path, bits):
    """A then B, equal lengths, little-endian (scripts/rng.py:18-25)."""
    from birdwing_amd import rng as limbrng
    path = tmp_path / "in.bin"
    limbrng.main(["--limbs", "37", "--limb-bits", str(bits), "--output", str(path), "--seed", "5"])
    assert path.stat().st_size == 2 * 37 * bits // 8
    a, b = limbrng.read_limb_file(str(path), bits)
    raw = path.read_bytes()
    w = bits // 8
    assert int.from_bytes(raw[:w], "little") == int(a[0])
    assert int.from_bytes(raw[37 * w:38 * w], "little") == int(b[0])
    # same seed, same limbs; another seed, other limbs
    limbrng.main(["--limbs", "37", "--limb-bits", str(bits), "--output", str(tmp_path / "again.bin"), "--seed", "5"])
    assert (tmp_path / "again.bin").read_bytes() == raw
    limbrng.main(["--limbs", "37", "--limb-bits", str(bits), "--output", str(tmp_path / "other.bin"), "--seed", "6"])
    assert (tmp_path / "other.bin").read_bytes() != raw
    with pytest.raises(ValueError):
        limbrng.write_limb_file(str(tmp_path / "bad.bin"), np.zeros(3, np.uint32), np.zeros(4, np.uint32))


def test_runner_pipeline_names():
    """scripts/run_gpu_bench.py:12-25: pipelines, 'both', 'all' and the old aliases."""
    from birdwing_amd import run_gpu_bench as runner
    assert runner.selected("all") == ["32", "hybrid", "64bit"]
    assert runner.selected("both") == ["32", "hybrid"]
    assert runner.selected("64") == ["hybrid"]
    assert runner.selected("64native") == ["64bit"]
    assert os.path.basename(runner.binary("hybrid")) == "bench_full_multiply_hybrid"


# ---------------------------------------------------------------- on the GPU
@pytest.mark.gpu
def test_bench_csv_rows(gpu, tmp_path):
    """Rows of the reference's schema; --append writes the header once."""
    out = tmp_path / "bench.csv"
    print("bench_full_multiply_32 7 300 --iters 2 --warmup 1", flush=True)
    r = _run([_bin("bench_full_multiply_32"), "--csv", str(out), "--iters", "2", "--warmup", "1", "7", "300"], 110)
    assert r.returncode == 0, r.stdout + r.stderr
    print("bench_full_multiply_64bit 9 --append", flush=True)
    r = _run([_bin("bench_full_multiply_64bit"), "--csv", str(out), "--append", "--iters", "2", "--warmup", "0", "9"],
             110)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(out) as f:
        rows = list(csv.reader(f))
    assert rows[0] == REFERENCE_CSV_COLUMNS
    assert len(rows) == 4 and all(len(r) == len(REFERENCE_CSV_COLUMNS) for r in rows)
    recs = [dict(zip(rows[0], r)) for r in rows[1:]]
    assert [(d["pipeline"], d["L_arg"], d["L"]) for d in recs] == [("32", "7", "128"), ("32", "300", "300"),
                                                                   ("64bit", "9", "512")]
    assert [d["host_limb_bits"] for d in recs] == ["32", "32", "64"]
    assert recs[2]["operand_bits"] == str(512 * 64)
    for d in recs:
        assert float(d["min_ms"]) <= float(d["mean_ms"]) <= float(d["max_ms"])
        assert float(d["mean_ms"]) > 0 and float(d["carry_mean_ms"]) > 0
        # a one-pass transform runs as the fused launch alone (the mul bucket)
        stages = [float(d[c]) for c in ("fwd_pad_ntt_mean_ms", "mul_mean_ms", "intt_mean_ms")]
        assert min(stages) >= 0 and stages[1] > 0


@pytest.mark.gpu
@pytest.mark.parametrize("pipe,L", [("32", 3001), ("hybrid", 1500), ("64bit", 2500)])
def test_cli_product_vs_gmp(gpu, oracle, tmp_path, pipe, L):
    """merge on a limb file writes the product bit-exact against GMP."""
    from birdwing_amd import rng as limbrng
    bits = LIMB_BITS[pipe]
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    a = oracle.uniform_limbs(L, 70 + L, bits)
    b = oracle.uniform_limbs(L, 71 + L, bits)
    limbrng.write_limb_file(str(src), a, b)
    print("birdwing_%s merge (%d limbs)" % (pipe, L), flush=True)
    r = _run([_bin("birdwing_" + pipe), "merge", str(src), "--out", str(dst)], 110)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(str(dst), dtype=np.uint64 if bits == 64 else np.uint32)
    want, _ = oracle.gmp_mul(a, b, bits)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_cli_default_example_and_benchmark(gpu):
    """Without a file: the reference's example operands {1,2,3,4} x {5,6,7,8}."""
    print("birdwing_32 merge (example)", flush=True)
    r = _run([_bin("birdwing_32"), "merge"], 110)
    assert r.returncode == 0, r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("Result (merge):")][0]
    assert [int(t) for t in line.split(":")[1].split()] == [5, 16, 34, 60, 61, 52, 32, 0]
    print("birdwing_64bit benchmark (example)", flush=True)
    r = _run([_bin("birdwing_64bit"), "benchmark"], 110)
    assert r.returncode == 0 and "[OK] Results match!" in r.stdout, r.stdout + r.stderr
