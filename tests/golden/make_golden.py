#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ in pure Python.

Independent of oracle/ (no C code, no reference code): Python's big integers
and pow() are the arithmetic.  The fixtures hold
  ntt_small.json     -- X[k] = sum_n x[n] w^(nk) mod p by the O(N^2) definition,
                        N in {1..64}, every prime set, roots w = root_max^(2^(maxlog-logN))
                        (the reference's root derivation, src/gpu_ntt.cu:171-192)
  crt_kats.json      -- the reference's CRT known answers (tests/test_crt_gpu.cu:159-265,
                        :403-485; tests/test_crt.cpp:19-103) with their exact values
  carry_kats.json    -- the reference's carry inputs (tests/test_carry_prop.cu:331-390,
                        :230-327) and the exact limb outputs (Python sum + base conversion)
  products_small.json-- small products (tests/test_full_multiply.cpp:271-351 cases plus
                        seeded ones) by Python multiplication
Run: python tests/golden/make_golden.py
"""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))

SETS = {
    "u32_ref": ([0x2d000001, 0x23800001, 0x26800001], [663, 721, 19]),
    "u32_wide": ([2013265921, 1811939329, 469762049], [440564289, 72705542, 2187]),
    "u64_ref": ([0x400002600000001, 0x400004200000001, 0x400001100000001],
                [273765203699653965, 26231613454922890, 144261359744151243]),
}


def v2(n):
    k = 0
    while n % 2 == 0:
        n //= 2
        k += 1
    return k


def order_log(r, p):
    k = 0
    while r != 1 and k < 70:
        r = r * r % p
        k += 1
    return k


def root_for(p, rmax, logN):
    maxlog = min(v2(p - 1), order_log(rmax, p))
    w = rmax % p
    for _ in range(maxlog - logN):
        w = w * w % p
    return w


def ntt_small():
    out = []
    rnd = random.Random(20240601)
    for name, (P, R) in SETS.items():
        for p, r in zip(P, R):
            for logN in range(0, 7):
                N = 1 << logN
                w = root_for(p, r, logN)
                x = [rnd.randrange(p) for _ in range(N)]
                X = [sum(x[n] * pow(w, n * k, p) for n in range(N)) % p for k in range(N)]
                out.append({"set": name, "p": p, "root_max": r, "logN": logN, "w": w, "x": x, "X": X})
    return out


def crt_kats():
    kats = []
    # tests/test_crt_gpu.cu (32-bit build): x = 12, 0, 2^32-1, 2^89 (skipped when >= M), batch list
    for name in ("u32_ref", "u32_wide"):
        P = SETS[name][0]
        M = P[0] * P[1] * P[2]
        vals = [12, 0, 0xFFFFFFFF, 1, 255, 65535] + ([2**89] if 2**89 < M else []) + [M - 1]
        kats.append({"set": name, "primes": P, "x": vals, "residues": [[v % p for p in P] for v in vals]})
        syn = [[(i * 7 + j * 13 + 1) % P[j] for j in range(3)] for i in range(64)]
        xs = [crt_int(r, P) for r in syn]
        kats.append({"set": name, "primes": P, "x": xs, "residues": syn, "synthetic": True})
    # U160 cases (tests/test_crt_gpu.cu:403-485)
    P = SETS["u64_ref"][0]
    vals = [0, 12, 2**64 - 1, 2**63 - 1, 2**128, 2**160 - 1, (1 << 159) + 12345]
    kats.append({"set": "u64_ref", "primes": P, "x": vals, "residues": [[v % p for p in P] for v in vals]})
    syn = [[(i * 7 + j * 13 + 1) % P[j] for j in range(3)] for i in range(64)]
    kats.append({"set": "u64_ref", "primes": P, "x": [crt_int(r, P) for r in syn], "residues": syn,
                 "synthetic": True})
    # tests/test_crt.cpp: 2-prime and 3-prime cases with {2013265921, 1811939329, 469762049}
    p1, p2, p3 = 2013265921, 1811939329, 469762049
    x2 = 12345678901234567890 % (p1 * p2)
    x3 = ((123456789123 << 64) ^ 987654321) % (p1 * p2 * p3)
    kats.append({"set": "test_crt_2", "primes": [p1, p2], "x": [x2], "residues": [[x2 % p1, x2 % p2]]})
    kats.append({"set": "test_crt_3", "primes": [p1, p2, p3], "x": [x3],
                 "residues": [[x3 % p1, x3 % p2, x3 % p3]]})
    return kats


def crt_int(res, P):
    """Textbook CRT: x = sum r_i M_i (M_i^-1 mod p_i) mod M."""
    M = 1
    for p in P:
        M *= p
    x = 0
    for r, p in zip(res, P):
        Mi = M // p
        x += r * Mi * pow(Mi, -1, p)
    return x % M


def carry_kats():
    M32, U64 = 0xFFFFFFFF, 2**64 - 1
    c128 = [("all zeros", [0] * 64), ("small values no carry", [i + 1 for i in range(64)]),
            ("single overflow", [M32 + 1] + [0] * 63), ("max limb values", [M32] * 1024),
            ("wide 128-bit coefficient", [(1 << 70) + 12345] + [0] * 15)]
    v = [0] * 2048
    v[1023] = M32 + 1
    c128.append(("cross-segment carry", v))
    c160 = [("u160 all zeros", [0] * 64), ("u160 small", [i + 1 for i in range(64)]),
            ("u160 single overflow", [1 << 64] + [0] * 63), ("u160 max limb values", [U64] * 1024),
            ("u160 wide", [12345 + (1 << 159)] + [0] * 15), ("u160 max", [2**160 - 1] * 8),
            ("u160 128-bit boundary", [1 + (1 << 64)] + [1] * 15)]
    v = [0] * 2048
    v[1023] = U64 + (1 << 64)
    c160.append(("u160 cross-segment", v))
    out = []
    for name, vals in c128:
        total = sum(c << (32 * i) for i, c in enumerate(vals))
        n_out = len(vals) + 4
        limbs = [(total >> (32 * i)) & M32 for i in range(n_out)]
        out.append({"name": name, "coeff": [str(c) for c in vals], "out_bits": 32, "limbs": limbs})
    for name, vals in c160:
        total = sum(c << (64 * i) for i, c in enumerate(vals))
        n_out = len(vals) + 3
        limbs = [str((total >> (64 * i)) & U64) for i in range(n_out)]
        out.append({"name": name, "coeff": [str(c) for c in vals], "out_bits": 64, "limbs": limbs})
    return out


def products_small():
    out = []
    rnd = random.Random(7)
    cases = [([1, 0, 0, 0], [1, 0, 0, 0]), ([0, 0, 0, 0], [42, 0, 0, 0]),
             ([0, 1, 0, 0, 0, 0, 0, 0], [1, 0, 0, 0, 0, 0, 0, 0]), ([1, 2, 3, 4], [5, 6, 7, 8])]
    for bits in (32, 64):
        for a, b in cases:
            out.append({"bits": bits, "A": a, "B": b})
        for la, lb in ((1, 1), (3, 5), (17, 9), (64, 64), (100, 7)):
            A = [rnd.getrandbits(bits) for _ in range(la)]
            B = [rnd.getrandbits(bits) for _ in range(lb)]
            out.append({"bits": bits, "A": A, "B": B})
        mx = 2**bits - 1
        out.append({"bits": bits, "A": [mx] * 50, "B": [mx] * 50})
    for c in out:
        bits = c["bits"]
        a = sum(x << (bits * i) for i, x in enumerate(c["A"]))
        b = sum(x << (bits * i) for i, x in enumerate(c["B"]))
        n = len(c["A"]) + len(c["B"])
        prod = a * b
        c["C"] = [str((prod >> (bits * i)) & (2**bits - 1)) for i in range(n)]
        c["A"] = [str(x) for x in c["A"]]
        c["B"] = [str(x) for x in c["B"]]
    return out


def main():
    for name, fn in (("ntt_small.json", ntt_small), ("crt_kats.json", crt_kats),
                     ("carry_kats.json", carry_kats), ("products_small.json", products_small)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()
