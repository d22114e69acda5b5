"""Arithmetic-rate probes on the GPU (measurement only).

libbirdwing_probe.so (birdwing_amd/probes/*.hip, built beside the product
library) holds small kernels that measure what the product kernels' design
rests on: the VALU rates of FP64 FMA against 32-bit integer multiplies, and
of the three modular products (FP64, 32-bit Shoup, 64-bit Montgomery).
bench.py --valu-peak prints valu_rates() as one JSON line; bench.py
--mfma-dft16 prints mfma_dft16(): one 16-point DFT stage mod p on int8 MFMA
against the product's VALU codelet, on the same lines, timed and checked.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE_PATH = os.path.join(HERE, "libbirdwing_probe.so")

# MI355X published vector peak: 78.6 TFLOP/s FP64 = 39.3 T FMA/s
FP64_FMA_PEAK = 39.3e12
VALU_PROBES = ["fma_f64", "mul_lo_u32", "mul_hi_u32", "mad_u64_u32", "f64_modmul", "shoup_u32_modmul",
               "mont_u64_modmul"]

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(PROBE_PATH):
            raise ImportError("%s is missing; build it with `make -C %s`" % (PROBE_PATH, HERE))
        L = C.CDLL(PROBE_PATH)
        L.bwp_valu_rate.restype = C.c_double
        L.bwp_valu_rate.argtypes = [C.c_int, C.c_int, C.c_int]
        L.bwp_hbm_rate.restype = C.c_double
        L.bwp_hbm_rate.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_int]
        L.bwp_mfma_dft16.restype = C.c_int
        L.bwp_mfma_dft16.argtypes = [C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                     C.c_int]
        _LIB = L
    return _LIB


def valu_rates(blocks=4096, iters=1 << 16):
    """Operations per second of each probe over the whole device (8
    independent chains per lane, blocks x 256 threads, best of two timed runs)."""
    import torch
    torch.cuda.init()
    rates = {}
    for i, name in enumerate(VALU_PROBES):
        r = float(lib().bwp_valu_rate(i, blocks, iters))
        if r < 0:
            raise RuntimeError("probe %s failed" % name)
        rates[name] = r
    return {
        "probe": "valu_rates",
        "unit": "operations/s (device)",
        "rates": rates,
        "fp64_fma_peak": FP64_FMA_PEAK,
        "fma_f64_frac_of_peak": rates["fma_f64"] / FP64_FMA_PEAK,
        "f64_modmul_vs_shoup_u32": rates["f64_modmul"] / rates["shoup_u32_modmul"],
        "f64_modmul_vs_mont_u64": rates["f64_modmul"] / rates["mont_u64_modmul"],
        "config": {"blocks": blocks, "threads": 256, "chains_per_lane": 8, "iters": iters},
    }


def _dft16_rounds(x, p, w, rounds):
    for _ in range(rounds):
        x = [sum(x[j] * pow(w, j * k, p) for j in range(16)) % p for k in range(16)]
    return x


def mfma_dft16(lines=1 << 22, reps=5, rounds=(1, 8), nsample=32):
    """Lines of 16 residues mod p = 2113929217 through the VALU radix-16
    codelet and through int8 MFMA (see birdwing_amd/probes/mfma_probe.hip),
    `rounds` DFT stages per line with a stage boundary (LDS round trip)
    between; both store canonical residues in natural order.  The two outputs
    are compared on the device (every element) and a sample of lines against
    direct DFTs computed here.  One round is bound by HBM; eight rounds per
    load compare the arithmetic."""
    import numpy as np
    import torch
    torch.cuda.init()
    out = {"probe": "mfma_dft16", "prime": 2113929217, "lines": lines, "runs": []}
    for rnd in rounds:
        res = (C.c_double * 5)()
        hin = np.zeros(nsample * 16, dtype=np.int32)
        hout = np.zeros(nsample * 16, dtype=np.uint32)
        rc = lib().bwp_mfma_dft16(lines, reps, rnd, res, hin.ctypes.data_as(C.c_void_p),
                                  hout.ctypes.data_as(C.c_void_p), nsample)
        if rc != 0:
            raise RuntimeError("mfma_dft16 probe failed")
        p, w = 2113929217, int(res[4])
        ok = 0
        for i in range(nsample):
            x = [int(v) % p for v in hin[16 * i:16 * i + 16]]
            ok += int(_dft16_rounds(x, p, w, rnd) == [int(v) for v in hout[16 * i:16 * i + 16]])
        tv, tm, n = res[0], res[1], res[2]
        out["runs"].append({
            "rounds": rnd,
            "valu_codelet": {"seconds": tv, "dft16_per_s": rnd * n / tv, "elements_per_s": 16 * rnd * n / tv},
            "mfma_i8": {"seconds": tm, "dft16_per_s": rnd * n / tm, "elements_per_s": 16 * rnd * n / tm},
            "mfma_vs_valu": tv / tm,
            "outputs_differing": int(res[3]),
            "sample_lines_matching_direct_dft": "%d/%d" % (ok, nsample),
        })
    out["config"] = {"reps": reps, "lines_per_mfma_block": 16, "mfma": "v_mfma_i32_16x16x64_i8",
                     "mfmas_per_block_and_round": 7}
    return out


HBM_PEAK = 8.0e12   # MI355X published HBM3E bandwidth, bytes/s


def hbm_rates(mib=640, reps=10):
    """Device streaming rates in GB/s (best of `reps` runs): copy, read and
    write of `mib` MiB at three grid sizes, and a 256-point column pass of
    stride 2^8 over ten 2^24-word planes (the memory pattern of the middle
    forward pass, no arithmetic) walked forward and in reverse right after
    a pass of stride 2^16 over the same planes; the stride-2^16 pass timed
    after one of stride 2^8; both over planes padded with one 128-byte line
    after every 2^16 words."""
    import torch
    torch.cuda.init()
    out = {}
    for kind, name in ((0, "copy"), (1, "read"), (2, "write")):
        for grid in (1024, 4096, 16384):
            r = float(lib().bwp_hbm_rate(kind, mib, grid, reps))
            if r < 0:
                raise RuntimeError("hbm probe %s failed" % name)
            out["%s_grid%d" % (name, grid)] = r
    for kind, name in ((3, "colpass_s256_forward_walk"), (4, "colpass_s256_reverse_walk"),
                       (5, "colpass_s65536_forward_walk"), (6, "colpass_s65536_padded_planes"),
                       (7, "colpass_s256_padded_planes")):
        r = float(lib().bwp_hbm_rate(kind, 0, 1, reps))
        if r < 0:
            raise RuntimeError("hbm probe %s failed" % name)
        out[name] = r
    best = {k: max(v for kk, v in out.items() if kk.startswith(k + "_grid")) for k in ("copy", "read", "write")}
    return {
        "probe": "hbm_rates",
        "unit": "GB/s (1e9 bytes moved per second: read + written)",
        "rates": out,
        "best": best,
        "peak_GBps": HBM_PEAK / 1e9,
        "best_frac_of_peak": {k: v / (HBM_PEAK / 1e9) for k, v in best.items()},
        "config": {"MiB": mib, "threads": 256, "reps": reps, "colpass_bytes": 2 * 10 * (1 << 24) * 4},
    }
