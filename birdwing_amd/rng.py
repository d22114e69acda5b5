"""Write two random big integers as a limb file, the input format of the
command-line front end (`birdwing_amd/apps/bin/birdwing_<pipeline>`).

The file is operand A then operand B, the same number of little-endian limbs
each, 32- or 64-bit (the format scripts/rng.py:18-25 writes and main.cpp:16-46
reads).  Limbs are uniform random; --seed makes them reproducible.

    python -m birdwing_amd.rng --limbs 64 --output input.bin
    python -m birdwing_amd.rng --limbs 64 --limb-bits 64 --output input_64bit.bin
"""
import argparse
import secrets

import numpy as np


def random_limbs(n, limb_bits, rng):
    """n uniform limbs of limb_bits (32 or 64) as a little-endian numpy array."""
    dt = np.dtype("<u8") if limb_bits == 64 else np.dtype("<u4")
    return rng.integers(0, 2 ** limb_bits, n, dtype=np.uint64, endpoint=False).astype(dt)


def write_limb_file(path, a, b):
    """A then B, back to back; both must hold the same number of limbs of one width."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype or a.dtype.itemsize not in (4, 8):
        raise ValueError("operands must be equal-length arrays of 32- or 64-bit limbs")
    dt = np.dtype("<u%d" % a.dtype.itemsize)
    with open(path, "wb") as f:
        f.write(a.astype(dt).tobytes())
        f.write(b.astype(dt).tobytes())


def read_limb_file(path, limb_bits):
    """(A, B) from a limb file; its size must be a multiple of two limbs."""
    dt = np.dtype("<u8") if limb_bits == 64 else np.dtype("<u4")
    data = np.fromfile(path, dtype=dt)
    if data.size == 0 or data.size % 2:
        raise ValueError("file size is not a positive multiple of 2 * %d bytes" % dt.itemsize)
    n = data.size // 2
    return data[:n].copy(), data[n:].copy()


def main(argv=None):
    ap = argparse.ArgumentParser(description="Generate two random big integers as limb arrays")
    ap.add_argument("--limbs", type=int, default=32, help="number of limbs per number")
    ap.add_argument("--limb-bits", type=int, choices=(32, 64), default=32, help="limb width in bits (default 32)")
    ap.add_argument("--output", default="bigints.bin", help="output binary file")
    ap.add_argument("--seed", type=int, default=None, help="RNG seed (default: fresh entropy)")
    args = ap.parse_args(argv)
    if args.limbs < 1:
        ap.error("--limbs must be positive")
    rng = np.random.default_rng(args.seed if args.seed is not None else secrets.randbits(64))
    a = random_limbs(args.limbs, args.limb_bits, rng)
    b = random_limbs(args.limbs, args.limb_bits, rng)
    write_limb_file(args.output, a, b)
    print("Two random numbers (%d x %d-bit limbs each) written to %s." % (args.limbs, args.limb_bits, args.output))


if __name__ == "__main__":
    main()
