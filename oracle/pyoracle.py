"""ctypes binding for the CPU oracle (oracle/liboracle.so) and oracle/_ref.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package birdwing_amd.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SET_U32_REF, SET_U32_WIDE, SET_U64_REF, SET_HYBRID = 0, 1, 2, 3

_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)


def build():
    """Compile oracle/ (and oracle/_ref when the reference tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sig = {
            "orc_mt64_seed": (None, [C.c_void_p, C.c_uint64]),
            "orc_mt64_next": (C.c_uint64, [C.c_void_p]),
            "orc_ref_random_limbs_u32": (None, [C.c_void_p, C.c_size_t, C.c_uint64]),
            "orc_ref_random_limbs_u64": (None, [C.c_void_p, C.c_size_t, C.c_uint64]),
            "orc_ref_pipeline_limbs_u32": (None, [C.c_void_p, C.c_size_t, C.c_uint64]),
            "orc_uniform_limbs": (None, [C.c_void_p, C.c_size_t, C.c_int, C.c_uint64]),
            "orc_prime_set": (C.c_int, [C.c_int, _u64p, _u64p]),
            "orc_valuation2": (C.c_int, [C.c_uint64]),
            "orc_max_logN_for_prime": (C.c_int, [C.c_uint64]),
            "orc_max_root_logN": (C.c_int, [C.c_uint64, C.c_uint64]),
            "orc_max_supported_logN": (C.c_int, [C.c_int]),
            "orc_padded_ntt_size": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_crt_bound_ok": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint64]),
            "orc_ntt_size_supported": (C.c_int, [C.c_int, C.c_uint64]),
            "orc_multiply_size_supported": (C.c_int, [C.c_int, C.c_int, C.c_uint64, C.c_uint64]),
            "orc_max_limb_count_by_crt": (C.c_uint64, [C.c_int, C.c_int]),
            "orc_max_supported_limb_count": (C.c_uint64, [C.c_int, C.c_int]),
            "orc_mulmod": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_powmod": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_modinv": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_root_for_N": (C.c_uint64, [C.c_uint64, C.c_uint64, C.c_int]),
            "orc_ntt_forward": (None, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_ntt_inverse": (None, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_dft_bruteforce": (None, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_pointwise": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64]),
            "orc_conv_schoolbook_mod": (None, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                                C.c_void_p, C.c_uint64, C.c_uint64]),
            "orc_garner_params": (None, [C.c_void_p, _u64p, C.c_int]),
            "orc_garner128": (None, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
            "orc_garner160": (None, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]),
            "orc_carry128": (None, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int]),
            "orc_carry160": (None, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]),
            "orc_choose_set": (C.c_int, [C.c_int, C.c_uint64]),
            "orc_multiply": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
                                       C.c_int, C.c_int]),
            "orc_multiply_crt": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_int,
                                           C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
            "orc_schoolbook": (None, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]),
            "orc_gmp_available": (C.c_int, []),
            "orc_set_prime_threads": (None, [C.c_int]),
            "orc_gmp_mul": (C.c_double, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _dt(bits):
    return np.uint64 if bits == 64 else np.uint32


class MT64:
    """std::mt19937_64 restated in C."""

    def __init__(self, seed):
        self._buf = C.create_string_buffer(312 * 8 + 16)
        lib().orc_mt64_seed(self._buf, seed)

    def __call__(self):
        return int(lib().orc_mt64_next(self._buf))


def ref_random_limbs(n, seed, bits=32):
    out = np.zeros(n, dtype=_dt(bits))
    if bits == 64:
        lib().orc_ref_random_limbs_u64(_p(out), n, seed)
    else:
        lib().orc_ref_random_limbs_u32(_p(out), n, seed)
    return out


def ref_pipeline_limbs(n, seed):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_ref_pipeline_limbs_u32(_p(out), n, seed)
    return out


def uniform_limbs(n, seed, bits=32):
    out = np.zeros(n, dtype=_dt(bits))
    lib().orc_uniform_limbs(_p(out), n, bits, seed)
    return out


def prime_set(s):
    P = (C.c_uint64 * 3)()
    R = (C.c_uint64 * 3)()
    k = lib().orc_prime_set(s, P, R)
    return [int(P[i]) for i in range(k)], [int(R[i]) for i in range(k)]


def root_for_N(p, root_max, logN):
    return int(lib().orc_root_for_N(p, root_max, logN))


def ntt_forward(x, p, w):
    y = np.ascontiguousarray(x, dtype=np.uint64).copy()
    lib().orc_ntt_forward(_p(y), len(y), p, w)
    return y


def ntt_inverse(x, p, w):
    y = np.ascontiguousarray(x, dtype=np.uint64).copy()
    lib().orc_ntt_inverse(_p(y), len(y), p, w)
    return y


def dft_bruteforce(x, p, w):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    X = np.zeros_like(x)
    lib().orc_dft_bruteforce(_p(x), _p(X), len(x), p, w)
    return X


def pointwise(a, b, p):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.zeros_like(a)
    lib().orc_pointwise(_p(a), _p(b), _p(c), len(a), p)
    return c


def conv_schoolbook_mod(a, b, N, p):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.zeros(N, dtype=np.uint64)
    lib().orc_conv_schoolbook_mod(_p(a), len(a), _p(b), len(b), _p(c), N, p)
    return c


class Garner(C.Structure):
    _fields_ = [("k", C.c_int), ("primes", C.c_uint64 * 3), ("inv", C.c_uint64 * 3),
                ("prefix_M", C.c_uint64 * 3), ("M_mod_table", (C.c_uint64 * 3) * 3),
                ("barrett_m", C.c_uint64 * 3)]


def garner_params(primes):
    g = Garner()
    P = (C.c_uint64 * 3)(*primes)
    lib().orc_garner_params(C.byref(g), P, len(primes))
    return g


def garner128(primes, residues):
    """residues: (k, N) array -> (hi, lo) uint64 arrays."""
    g = garner_params(primes)
    r = np.ascontiguousarray(residues, dtype=np.uint64)
    N = r.shape[1]
    hi = np.zeros(N, np.uint64)
    lo = np.zeros(N, np.uint64)
    lib().orc_garner128(C.byref(g), _p(r), N, _p(hi), _p(lo))
    return hi, lo


def garner160(primes, residues):
    g = garner_params(primes)
    r = np.ascontiguousarray(residues, dtype=np.uint64)
    N = r.shape[1]
    lo = np.zeros(N, np.uint64)
    mid = np.zeros(N, np.uint64)
    hi = np.zeros(N, np.uint32)
    lib().orc_garner160(C.byref(g), _p(r), N, _p(lo), _p(mid), _p(hi))
    return lo, mid, hi


def carry128(hi, lo, n_out=None, out_bits=32):
    hi = np.ascontiguousarray(hi, np.uint64)
    lo = np.ascontiguousarray(lo, np.uint64)
    n_out = len(lo) if n_out is None else n_out
    out = np.zeros(n_out, dtype=_dt(out_bits))
    lib().orc_carry128(_p(hi), _p(lo), len(lo), _p(out), n_out, out_bits)
    return out


def carry160(lo, mid, hi, n_out=None):
    lo = np.ascontiguousarray(lo, np.uint64)
    mid = np.ascontiguousarray(mid, np.uint64)
    hi = np.ascontiguousarray(hi, np.uint32)
    n_out = len(lo) if n_out is None else n_out
    out = np.zeros(n_out, dtype=np.uint64)
    lib().orc_carry160(_p(lo), _p(mid), _p(hi), len(lo), _p(out), n_out)
    return out


def choose_set(bits, N):
    return int(lib().orc_choose_set(bits, N))


def multiply(A, B, bits=32, pset=None):
    A = np.ascontiguousarray(A, dtype=_dt(bits))
    B = np.ascontiguousarray(B, dtype=_dt(bits))
    N = int(lib().orc_padded_ntt_size(len(A), len(B)))
    if pset is None:
        pset = choose_set(bits, N)
    out = np.zeros(len(A) + len(B), dtype=_dt(bits))
    rc = lib().orc_multiply(_p(A), len(A), _p(B), len(B), _p(out), bits, pset)
    if rc != 0:
        raise ValueError("unsupported multiply size")
    return out


def multiply_crt(A, B, bits=32, pset=None):
    A = np.ascontiguousarray(A, dtype=_dt(bits))
    B = np.ascontiguousarray(B, dtype=_dt(bits))
    N = int(lib().orc_padded_ntt_size(len(A), len(B)))
    if pset is None:
        pset = choose_set(bits, N)
    c0 = np.zeros(N, np.uint64)
    c1 = np.zeros(N, np.uint64)
    c2 = np.zeros(N, np.uint32)
    rc = lib().orc_multiply_crt(_p(A), len(A), _p(B), len(B), bits, pset, _p(c0), _p(c1), _p(c2))
    if rc != 0:
        raise ValueError("unsupported multiply size")
    return c0, c1, c2


def schoolbook(A, B, bits=32):
    A = np.ascontiguousarray(A, dtype=_dt(bits))
    B = np.ascontiguousarray(B, dtype=_dt(bits))
    out = np.zeros(len(A) + len(B), dtype=_dt(bits))
    lib().orc_schoolbook(_p(A), len(A), _p(B), len(B), _p(out), bits)
    return out


def set_prime_threads(on):
    """True: the primes of one multiply() run on their own threads (the
    default); False: one thread per product (for thread pools of products)."""
    lib().orc_set_prime_threads(1 if on else 0)


def gmp_available():
    return bool(lib().orc_gmp_available())


def gmp_mul(A, B, bits=32):
    """Returns (product limbs, mpz_mul seconds)."""
    A = np.ascontiguousarray(A, dtype=_dt(bits))
    B = np.ascontiguousarray(B, dtype=_dt(bits))
    out = np.zeros(len(A) + len(B), dtype=_dt(bits))
    t = lib().orc_gmp_mul(_p(A), len(A), _p(B), len(B), _p(out), bits)
    if t < 0:
        raise RuntimeError("GMP not available in this oracle build")
    return out, t


# ---- conversions --------------------------------------------------------
def to_int(limbs, bits=32):
    """Little-endian limbs -> Python int."""
    a = np.ascontiguousarray(limbs, dtype=_dt(bits))
    return int.from_bytes(a.tobytes(), "little")


def from_int(x, n, bits=32):
    return np.frombuffer(x.to_bytes(n * bits // 8, "little"), dtype=_dt(bits)).copy()


# ---- oracle/_ref (the reference's own crt.cpp / crt_utils.cpp) ------------
def ref_crt_lib(limb_bits):
    path = os.path.join(HERE, "_ref", "libref_crt%d.so" % limb_bits)
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_crt_combine_many.argtypes = [_u64p, _u64p, C.c_int, _u64p, _u64p]
    L.ref_crt_combine_many.restype = None
    L.ref_crt_combine_2.argtypes = [C.c_uint64] * 4 + [_u64p, _u64p]
    L.ref_crt_combine_2.restype = None
    L.ref_modinv_u64.argtypes = [C.c_uint64, C.c_uint64]
    L.ref_modinv_u64.restype = C.c_uint64
    L.ref_limb_bits.restype = C.c_int
    return L


def ref_crt_combine_many(L, primes, residues):
    k = len(primes)
    P = (C.c_uint64 * k)(*primes)
    R = (C.c_uint64 * k)(*residues)
    lo = C.c_uint64()
    hi = C.c_uint64()
    L.ref_crt_combine_many(P, R, k, C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value
