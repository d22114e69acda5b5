"""Arithmetic-rate probes on the GPU (measurement only).

libbirdwing_probe.so (birdwing_amd/probes/*.hip, built beside the product
library) holds small kernels that measure what the product kernels' design
rests on: the VALU rates of FP64 FMA against 32-bit integer multiplies, and
of the three modular products (FP64, 32-bit Shoup, 64-bit Montgomery).
bench.py --valu-peak prints valu_rates() as one JSON line.
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
