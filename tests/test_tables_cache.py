"""Precompute cache (docs/TIMING.md:24-26,98-100): the host tables of a prime
set written to a file and loaded back, with damaged or foreign files rejected.

The image layout checked here is this library's own: a header (magic,
version, set id, k, table_log, five primes, four blob sizes), the blobs, and
an FNV-1a 64 checksum of all preceding bytes.
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import birdwing_amd as bw

HEADER = struct.Struct("<8sIiii5Q4Q")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _fnv1a(data):
    h = 0xcbf29ce484222325
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.parametrize("pset", [bw.SET_U32_REF, bw.SET_U32_WIDE, bw.SET_U64_REF, bw.SET_D64])
def test_save_load_round_trip(tmp_path, pset):
    path = tmp_path / "t.bin"
    bw.tables_save(pset, str(path))
    raw = path.read_bytes()
    magic, version, set_id, k, table_log, *rest = HEADER.unpack_from(raw)
    primes, sizes = rest[:5], rest[5:]
    assert magic == b"BWTABLES" and version == 1 and set_id == pset
    want_primes = bw.prime_set_info(pset)[0]
    assert k == len(want_primes) and list(primes[:k]) == want_primes and not any(primes[k:])
    assert len(raw) == HEADER.size + sum(sizes) + 8
    assert struct.unpack_from("<Q", raw, len(raw) - 8)[0] == _fnv1a(raw[:-8])
    assert bw.tables_load(str(path)) == pset
    # saving again gives the same bytes (deterministic tables)
    bw.tables_save(pset, str(tmp_path / "again.bin"))
    assert (tmp_path / "again.bin").read_bytes() == raw


def test_damaged_and_foreign_images_rejected(tmp_path):
    good = tmp_path / "good.bin"
    bw.tables_save(bw.SET_D64, str(good))
    raw = bytearray(good.read_bytes())

    def rejected(data, what):
        p = tmp_path / ("bad_%s.bin" % what)
        p.write_bytes(bytes(data))
        with pytest.raises(bw.BWError):
            bw.tables_load(str(p))

    flipped = bytearray(raw)
    flipped[HEADER.size + 100] ^= 1                 # one table bit
    rejected(flipped, "bit")
    rejected(raw[:-9], "truncated")
    rejected(raw + b"\0", "trailing")
    wrong_magic = bytearray(raw)
    wrong_magic[0:8] = b"NOTTABLE"
    rejected(wrong_magic, "magic")
    # a valid checksum over another prime set's header: the primes do not match
    foreign = bytearray(raw)
    hdr = list(HEADER.unpack_from(foreign))
    hdr[5] += 2                                     # first prime
    HEADER.pack_into(foreign, 0, *hdr)
    struct.pack_into("<Q", foreign, len(foreign) - 8, _fnv1a(bytes(foreign[:-8])))
    rejected(foreign, "foreign")
    with pytest.raises(bw.BWError):
        bw.tables_load(str(tmp_path / "missing.bin"))
    with pytest.raises(bw.BWError):
        bw.tables_save(99, str(tmp_path / "x.bin"))


@pytest.mark.gpu
def test_products_from_loaded_tables(gpu, oracle, tmp_path):
    """A fresh process loads the saved images before any table is built and
    multiplies with them: products stay bit-exact against GMP."""
    for s in (bw.SET_D64, bw.SET_U32_WIDE, bw.SET_U64_REF):
        bw.tables_save(s, str(tmp_path / ("set%d.bin" % s)))
    cases = [(32, 5000, bw.SET_D64), (32, 3000, bw.SET_U32_WIDE), (64, 2000, bw.SET_U64_REF)]
    for bits, L, _ in cases:
        a = oracle.uniform_limbs(L, 300 + L, bits)
        b = oracle.uniform_limbs(L - 1, 301 + L, bits)
        np.save(tmp_path / ("a%d_%d.npy" % (bits, L)), a)
        np.save(tmp_path / ("b%d_%d.npy" % (bits, L)), b)
    script = """
import sys, numpy as np
sys.path.insert(0, %r)
import birdwing_amd as bw
d = %r
for s in (bw.SET_D64, bw.SET_U32_WIDE, bw.SET_U64_REF):
    assert bw.tables_load(d + "/set%%d.bin" %% s) == s
for bits, L, s in %r:
    a = np.load(d + "/a%%d_%%d.npy" %% (bits, L)); b = np.load(d + "/b%%d_%%d.npy" %% (bits, L))
    p = bw.Plan(bits, len(a), len(b), s)
    np.save(d + "/c%%d_%%d.npy" %% (bits, L), p.execute_host(a, b))
    p.close()
print("ok")
""" % (ROOT, str(tmp_path), cases)
    print("fresh process: load tables, multiply", flush=True)
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr
    for bits, L, _ in cases:
        a = np.load(tmp_path / ("a%d_%d.npy" % (bits, L)))
        b = np.load(tmp_path / ("b%d_%d.npy" % (bits, L)))
        got = np.load(tmp_path / ("c%d_%d.npy" % (bits, L)))
        want, _ = oracle.gmp_mul(a, b, bits)
        assert np.array_equal(got, want)
