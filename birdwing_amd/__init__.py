"""birdwing_amd -- MI355X (gfx950) NTT big-integer multiplier.

Python binding of the C ABI in include/birdwing.h (libbirdwing_hip.so, built
in-tree by `make -C birdwing_amd`).  The library is required: every entry
point raises if it is missing, there is no CPU fallback.

Host-side mirror of the reference interface (nikitaborisov/birdwing):
  * multiply(A, B)            -- host_multiply_merge / host_multiply_merge_native
                                 (src/host_multiply.cpp:27-134): little-endian
                                 limb arrays in, L_A + L_B limbs out.
  * Plan                      -- precompute_ntt + upload_ntt_precomputed +
                                 allocate_ntt_context (src/gpu_ntt.cu:240-440) and
                                 execute_ntt_multiply (:442-973) on device buffers.
  * size policy functions     -- src/ntt_limits.cpp.
  * crt_combine / carry       -- src/crt_gpu.cu, src/carry_prop.cu entry points.
Device buffers are torch tensors on a HIP device (torch is plumbing for device
memory and streams only).
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "lib", "Plan", "multiply", "BWError", "SET_DEFAULT", "SET_U32_REF", "SET_U32_WIDE", "SET_U64_REF", "SET_D64",
    "padded_ntt_size", "multiply_size_supported", "ntt_size_supported", "crt_bound_satisfied",
    "max_supported_logN", "max_supported_limb_count", "max_limb_count_by_crt", "default_prime_set",
    "prime_set_info", "crt_combine", "carry128", "carry160", "zero_pad", "zero_pad_u64", "pointwise",
    "Batch", "batch_footprint", "set_fused_limit", "tune_count", "tune_clear", "Server",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbirdwing_hip.so")

SET_DEFAULT, SET_U32_REF, SET_U32_WIDE, SET_U64_REF, SET_HYBRID, SET_D64, SET_U64_PAIRS = -1, 0, 1, 2, 3, 4, 5
SCHED_PASS_MAJOR, SCHED_PRIME_MAJOR, SCHED_HYBRID = 0, 1, 2
BW_OK, BW_ERR_ARG, BW_ERR_UNSUPPORTED_SIZE, BW_ERR_HIP, BW_ERR_NOMEM, BW_ERR_STATE = 0, -1, -2, -3, -4, -5


class BWError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("birdwing error %d: %s" % (code, msg))
        self.code = code


class PlanInfo(C.Structure):
    _fields_ = [("limb_bits", C.c_int), ("prime_set", C.c_int), ("k", C.c_int),
                ("primes", C.c_uint64 * 5), ("L_A", C.c_uint64), ("L_B", C.c_uint64),
                ("N", C.c_uint64), ("N_exec", C.c_uint64), ("npass", C.c_int),
                ("pass_log", C.c_int * 4), ("nlaunch", C.c_int), ("device_bytes", C.c_uint64),
                ("digit_bits", C.c_int), ("schedule", C.c_int), ("tile_shift", C.c_int)]


class BatchInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("ngroups", C.c_int), ("nlaunch", C.c_int), ("use_graph", C.c_int),
                ("bytes_a", C.c_uint64), ("bytes_b", C.c_uint64), ("bytes_c", C.c_uint64),
                ("workspace_bytes", C.c_uint64)]


class ServerConfig(C.Structure):
    _fields_ = [("device_bytes", C.c_uint64), ("streams", C.c_int), ("max_batch", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("failed", C.c_int64), ("batches", C.c_int64),
                ("batched_products", C.c_int64), ("pending", C.c_int64), ("max_concurrent_batches", C.c_int),
                ("arena_bytes", C.c_uint64), ("arena_peak_bytes", C.c_uint64)]


_vp, _u64, _i, _sz, _cp = C.c_void_p, C.c_uint64, C.c_int, C.c_size_t, C.c_char_p
_pi, _pu64 = C.POINTER(C.c_int), C.POINTER(C.c_uint64)
# name -> (restype, argtypes); every symbol declared in include/birdwing.h
SIGNATURES = {
    "bw_tables_save": (_i, [_i, C.c_char_p]),
    "bw_tables_load": (_i, [C.c_char_p, C.POINTER(C.c_int)]),
    "bw_abi_version": (_i, []),
    "bw_last_error": (_cp, []),
    "bw_version": (_cp, []),
    "bw_init": (_i, [_i]),
    "bw_padded_ntt_size": (_u64, [_u64, _u64]),
    "bw_default_prime_set": (_i, [_i, _u64]),
    "bw_prime_set_info": (_i, [_i, C.POINTER(_i), C.POINTER(_u64), C.POINTER(_u64)]),
    "bw_max_supported_logN": (_i, [_i]),
    "bw_max_limb_count_by_crt": (_u64, [_i, _i]),
    "bw_max_supported_limb_count": (_u64, [_i, _i]),
    "bw_ntt_size_supported": (_i, [_i, _u64, _cp, _sz]),
    "bw_crt_bound_satisfied": (_i, [_i, _i, _u64, _u64, _cp, _sz]),
    "bw_multiply_size_supported": (_i, [_i, _i, _u64, _u64, _cp, _sz]),
    "bw_plan_create": (_i, [C.POINTER(_vp), _i, _u64, _u64, _i]),
    "bw_plan_create_ex": (_i, [C.POINTER(_vp), _i, _u64, _u64, _i, C.POINTER(_i), _i]),
    "bw_plan_destroy": (None, [_vp]),
    "bw_plan_get_info": (_i, [_vp, C.POINTER(PlanInfo)]),
    "bw_execute": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "bw_execute_crt": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "bw_execute_host": (_i, [_vp, _vp, _vp, _vp]),
    "bw_plan_status": (_i, [_vp]),
    "bw_plan_tune": (_i, [_vp, _vp, _vp, _vp, _i]),
    "bw_plan_set_schedule": (_i, [_vp, _i]),
    "bw_plan_set_timing": (_i, [_vp, _i]),
    "bw_plan_timing": (_i, [_vp, C.POINTER(C.c_float), _i, C.POINTER(_i)]),
    "bw_multiply_host": (_i, [_i, _vp, _u64, _vp, _u64, _vp, C.POINTER(C.c_double)]),
    "bw_multiply_device_noalloc": (_i, [_i, _vp, _u64, _vp, _u64, _vp, _vp, _vp]),
    "bw_forward": (_i, [_vp, _vp, _u64, _vp, _vp]),
    "bw_inverse": (_i, [_vp, _vp, _vp]),
    "bw_spectrum_order": (_i, [_vp, C.POINTER(_u64)]),
    "bw_zero_pad": (_i, [_vp, _vp, _i, _u64, _u64, _vp]),
    "bw_zero_pad_u64": (_i, [_vp, _vp, _u64, _u64, _u64, _vp]),
    "bw_pointwise": (_i, [_i, _vp, _vp, _vp, _u64, _u64, _vp]),
    "bw_crt_combine": (_i, [_i, _vp, _vp, _vp, _u64, _vp, _vp, _vp, _vp]),
    "bw_carry128": (_i, [_vp, _vp, _u64, _vp, _u64, _i, _vp]),
    "bw_carry160": (_i, [_vp, _vp, _vp, _u64, _vp, _u64, _vp]),
    "bw_batch_create": (_i, [C.POINTER(_vp), _i, _pi, _pu64, _pu64]),
    "bw_batch_destroy": (None, [_vp]),
    "bw_batch_footprint": (_i, [_i, _pi, _pu64, _pu64, _pu64, _pu64, _pu64, _pu64]),
    "bw_batch_get_info": (_i, [_vp, C.POINTER(BatchInfo)]),
    "bw_batch_offsets": (_i, [_vp, _pu64, _pu64, _pu64]),
    "bw_batch_group_info": (_i, [_vp, _i, _pi, _pi, _pi, _pi, _pi, _pi, _pi]),
    "bw_batch_execute": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "bw_batch_fused_info": (_i, [_vp, _pi, _pi]),
    "bw_set_fused_limit": (_i, [_i]),
    "bw_batch_set_graph": (_i, [_vp, _i]),
    "bw_batch_tune": (_i, [_vp, _vp, _vp, _vp, _i]),
    "bw_tune_count": (_i, []),
    "bw_tune_clear": (None, []),
    "bw_batch_status": (_i, [_vp]),
    "bw_server_create": (_i, [C.POINTER(_vp), C.POINTER(ServerConfig)]),
    "bw_server_submit": (C.c_int64, [_vp, _i, _vp, _u64, _vp, _u64, _vp]),
    "bw_server_wait": (_i, [_vp, C.c_int64, _i, C.POINTER(C.c_double)]),
    "bw_server_poll": (_i, [_vp, C.c_int64]),
    "bw_server_get_stats": (_i, [_vp, C.POINTER(ServerStats)]),
    "bw_server_destroy": (None, [_vp]),
}

_LIB = None


def lib():
    """Load libbirdwing_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("birdwing_amd: %s is missing; build it with `make -C %s`" % (LIB_PATH, HERE))
        # torch ships its own HIP runtime under the same soname: load it first so
        # that the library binds to it (one runtime per process, whichever of
        # the two touches the GPU first)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != BW_OK:
        raise BWError(rc, lib().bw_last_error().decode())
    return rc


def _dt(bits):
    return np.uint64 if bits == 64 else np.uint32


def _ptr(x):
    """Device/host address of a torch tensor or numpy array."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return C.c_void_p(x.data_ptr())
    return x.ctypes.data_as(C.c_void_p)


def _dev(x, nbytes, what):
    """Address of a device buffer that holds at least `nbytes` bytes.

    Device entry points take torch tensors on a HIP device, contiguous, large
    enough for what the call reads or writes; anything else (a host array, a
    strided view, a short tensor) would be an out-of-bounds device access, so
    it is refused here, before the C ABI is called."""
    if x is None or not hasattr(x, "data_ptr") or not getattr(x, "is_cuda", False):
        raise BWError(BW_ERR_ARG, "%s must be a torch tensor on a HIP device" % what)
    if not x.is_contiguous():
        raise BWError(BW_ERR_ARG, "%s must be contiguous" % what)
    have = x.numel() * x.element_size()
    if have < nbytes:
        raise BWError(BW_ERR_ARG, "%s holds %d bytes, the call needs %d" % (what, have, nbytes))
    return C.c_void_p(x.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


# ---- size policy (src/ntt_limits.cpp) --------------------------------------
def padded_ntt_size(L_A, L_B):
    return int(lib().bw_padded_ntt_size(L_A, L_B))


def tables_save(prime_set, path):
    """Write the host tables of a prime set (SET_*) to `path` (precompute cache)."""
    _check(lib().bw_tables_save(prime_set, os.fsencode(path)))


def tables_load(path):
    """Load a saved table image; later table builds for its set use it. Returns the set."""
    s = C.c_int(-1)
    _check(lib().bw_tables_load(os.fsencode(path), C.byref(s)))
    return int(s.value)


def default_prime_set(limb_bits, N):
    return int(lib().bw_default_prime_set(limb_bits, N))


def prime_set_info(s):
    k = C.c_int()
    P = (C.c_uint64 * 5)()
    R = (C.c_uint64 * 5)()
    _check(lib().bw_prime_set_info(s, C.byref(k), P, R))
    return [int(P[i]) for i in range(k.value)], [int(R[i]) for i in range(k.value)]


def max_supported_logN(s):
    return int(lib().bw_max_supported_logN(s))


def max_supported_limb_count(s, in_bits):
    return int(lib().bw_max_supported_limb_count(s, in_bits))


def max_limb_count_by_crt(s, in_bits):
    return int(lib().bw_max_limb_count_by_crt(s, in_bits))


def _why_call(fn, *args):
    buf = C.create_string_buffer(512)
    ok = fn(*args, buf, 512)
    return bool(ok), buf.value.decode()


def ntt_size_supported(s, N):
    return _why_call(lib().bw_ntt_size_supported, s, N)


def crt_bound_satisfied(s, in_bits, L_A, L_B):
    return _why_call(lib().bw_crt_bound_satisfied, s, in_bits, L_A, L_B)


def multiply_size_supported(s, in_bits, L_A, L_B):
    return _why_call(lib().bw_multiply_size_supported, s, in_bits, L_A, L_B)


# ---- plans ----------------------------------------------------------------
class Plan:
    """A product shape (limb width, L_A, L_B): tables, buffers and launches."""

    def __init__(self, limb_bits, L_A, L_B, prime_set=SET_DEFAULT, passes=None):
        self._p = C.c_void_p()
        if passes:
            arr = (C.c_int * len(passes))(*passes)
            _check(lib().bw_plan_create_ex(C.byref(self._p), limb_bits, L_A, L_B, prime_set, arr, len(passes)))
        else:
            _check(lib().bw_plan_create(C.byref(self._p), limb_bits, L_A, L_B, prime_set))
        self.limb_bits, self.L_A, self.L_B = limb_bits, L_A, L_B
        info = PlanInfo()
        _check(lib().bw_plan_get_info(self._p, C.byref(info)))
        self.info = info
        self.N, self.N_exec, self.k = int(info.N), int(info.N_exec), int(info.k)
        self.primes = [int(info.primes[i]) for i in range(self.k)]
        self.passes = [int(info.pass_log[i]) for i in range(info.npass)]
        self.prime_set = int(info.prime_set)
        self.digit_bits = int(info.digit_bits)

    def close(self):
        if self._p:
            lib().bw_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, a, b, c, stream=None):
        """Device tensors: a (L_A limbs), b (L_B limbs) -> c (L_A+L_B limbs)."""
        wb = self.limb_bits // 8
        _check(lib().bw_execute(self._p, _dev(a, self.L_A * wb, "a"), _dev(b, self.L_B * wb, "b"),
                                _dev(c, (self.L_A + self.L_B) * wb, "c"), _stream(stream)))

    def execute_crt(self, a, b, c0, c1, c2=None, stream=None):
        """Early exit after Garner: N_exec coefficients (see bw_execute_crt)."""
        wb = self.limb_bits // 8
        n = self.N_exec
        _check(lib().bw_execute_crt(self._p, _dev(a, self.L_A * wb, "a"), _dev(b, self.L_B * wb, "b"),
                                    _dev(c0, 8 * n, "c0"), _dev(c1, 8 * n, "c1"),
                                    _dev(c2, 4 * n, "c2") if (c2 is not None or self.limb_bits == 64) else None,
                                    _stream(stream)))

    def execute_host(self, A, B):
        dt = _dt(self.limb_bits)
        A = np.ascontiguousarray(A, dtype=dt)
        B = np.ascontiguousarray(B, dtype=dt)
        out = np.zeros(self.L_A + self.L_B, dtype=dt)
        _check(lib().bw_execute_host(self._p, _ptr(A), _ptr(B), _ptr(out)))
        return out

    def _spec_bytes(self):
        # residue words: 8 bytes under the 58-bit primes, 4 under the 31-bit ones
        word = 8 if self.prime_set in (SET_U64_REF, SET_HYBRID, SET_U64_PAIRS) else 4
        return self.k * self.N_exec * word

    def forward(self, raw, L, spec, stream=None):
        """Forward transforms of L raw limbs into spec (k planes of N_exec words)."""
        _check(lib().bw_forward(self._p, _dev(raw, L * (self.limb_bits // 8), "raw"), L,
                                _dev(spec, self._spec_bytes(), "spec"), _stream(stream)))

    def inverse(self, spec, stream=None):
        _check(lib().bw_inverse(self._p, _dev(spec, self._spec_bytes(), "spec"), _stream(stream)))

    def spectrum_order(self):
        order = np.zeros(self.N_exec, dtype=np.uint64)
        _check(lib().bw_spectrum_order(self._p, order.ctypes.data_as(C.POINTER(C.c_uint64))))
        return order

    def tune(self, a, b, stream=None, reps=3):
        """Autotune the pass split / tile cap on device operands a, b."""
        wb = self.limb_bits // 8
        _check(lib().bw_plan_tune(self._p, _dev(a, self.L_A * wb, "a"), _dev(b, self.L_B * wb, "b"),
                                  _stream(stream), reps))
        info = PlanInfo()
        _check(lib().bw_plan_get_info(self._p, C.byref(info)))
        self.info = info
        self.passes = [int(info.pass_log[i]) for i in range(info.npass)]

    def set_schedule(self, schedule):
        """SCHED_PASS_MAJOR: one launch per pass over all primes; SCHED_PRIME_MAJOR:
        per-prime chains; SCHED_HYBRID: forward passes pass-major, fused and
        inverse passes per prime."""
        _check(lib().bw_plan_set_schedule(self._p, schedule))
        info = PlanInfo()
        _check(lib().bw_plan_get_info(self._p, C.byref(info)))
        self.info = info

    def status(self):
        _check(lib().bw_plan_status(self._p))

    def set_timing(self, on):
        _check(lib().bw_plan_set_timing(self._p, 1 if on else 0))

    def timing(self):
        ms = (C.c_float * 64)()
        n = C.c_int()
        _check(lib().bw_plan_timing(self._p, ms, 64, C.byref(n)))
        return [float(ms[i]) for i in range(n.value)]


def multiply(A, B, limb_bits=None):
    """Host limb arrays (numpy, little-endian) -> product limbs (L_A + L_B)."""
    if limb_bits is None:
        limb_bits = 64 if np.asarray(A).dtype == np.uint64 else 32
    dt = _dt(limb_bits)
    A = np.ascontiguousarray(A, dtype=dt)
    B = np.ascontiguousarray(B, dtype=dt)
    out = np.zeros(len(A) + len(B), dtype=dt)
    ms = C.c_double()
    _check(lib().bw_multiply_host(limb_bits, _ptr(A), len(A), _ptr(B), len(B), _ptr(out), C.byref(ms)))
    return out


# ---- stage-level entry points (device tensors) ------------------------------
def zero_pad(src, dst, L, N, stream=None):
    wb = dst.element_size()
    _check(lib().bw_zero_pad(_dev(src, 4 * L, "src"), _dev(dst, wb * N, "dst"), wb, L, N, _stream(stream)))


def zero_pad_u64(src, dst, L, N, p, stream=None):
    _check(lib().bw_zero_pad_u64(_dev(src, 8 * L, "src"), _dev(dst, 8 * N, "dst"), L, N, p, _stream(stream)))


def pointwise(a, b, c, N, p, stream=None):
    wb = a.element_size()
    _check(lib().bw_pointwise(wb, _dev(a, wb * N, "a"), _dev(b, wb * N, "b"), _dev(c, wb * N, "c"), N, p,
                              _stream(stream)))


def crt_combine(prime_set, r0, r1, r2, N, o0, o1, o2=None, stream=None):
    wb = 8 if prime_set in (SET_U64_REF, SET_HYBRID, SET_U64_PAIRS) else 4
    _check(lib().bw_crt_combine(prime_set, _dev(r0, wb * N, "r0"), _dev(r1, wb * N, "r1"), _dev(r2, wb * N, "r2"),
                                N, _dev(o0, 8 * N, "o0"), _dev(o1, 8 * N, "o1"),
                                _dev(o2, 4 * N, "o2") if o2 is not None else None, _stream(stream)))


def carry128(hi, lo, N, out, n_out, out_bits=32, stream=None):
    _check(lib().bw_carry128(_dev(hi, 8 * N, "hi"), _dev(lo, 8 * N, "lo"), N,
                             _dev(out, n_out * (out_bits // 8), "out"), n_out, out_bits, _stream(stream)))


def carry160(lo, mid, hi, N, out, n_out, stream=None):
    _check(lib().bw_carry160(_dev(lo, 8 * N, "lo"), _dev(mid, 8 * N, "mid"), _dev(hi, 4 * N, "hi"), N,
                             _dev(out, 8 * n_out, "out"), n_out, _stream(stream)))


# ---- batched products ---------------------------------------------------------
class Batch:
    """Many products of uneven sizes (limb_bits[i] in {32, 64}) run together.

    Operands and results live in packed byte buffers: product i reads its A
    limbs at byte offset off_a[i] of the packed A buffer, and so on.  pack() /
    unpack() move between lists of numpy limb arrays and packed host buffers.
    """

    def __init__(self, limb_bits, L_A, L_B):
        n = len(L_A)
        if isinstance(limb_bits, int):
            limb_bits = [limb_bits] * n
        self.n = n
        self.bits = [int(b) for b in limb_bits]
        self.L_A = [int(x) for x in L_A]
        self.L_B = [int(x) for x in L_B]
        self._b = C.c_void_p()
        bits = (C.c_int * n)(*self.bits)
        la = (C.c_uint64 * n)(*self.L_A)
        lb = (C.c_uint64 * n)(*self.L_B)
        _check(lib().bw_batch_create(C.byref(self._b), n, bits, la, lb))
        self.info = BatchInfo()
        _check(lib().bw_batch_get_info(self._b, C.byref(self.info)))
        oa, ob, oc = (C.c_uint64 * n)(), (C.c_uint64 * n)(), (C.c_uint64 * n)()
        _check(lib().bw_batch_offsets(self._b, oa, ob, oc))
        self.off_a = [int(x) for x in oa]
        self.off_b = [int(x) for x in ob]
        self.off_c = [int(x) for x in oc]
        self.bytes_a, self.bytes_b, self.bytes_c = int(self.info.bytes_a), int(self.info.bytes_b), int(self.info.bytes_c)

    def close(self):
        if self._b:
            lib().bw_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def groups(self):
        out = []
        for gi in range(int(self.info.ngroups)):
            v = [C.c_int() for _ in range(5)]
            pl = (C.c_int * 4)()
            npass, gsh = C.c_int(), C.c_int()
            _check(lib().bw_batch_group_info(self._b, gi, C.byref(v[0]), C.byref(v[1]), C.byref(v[2]),
                                             C.byref(v[3]), pl, C.byref(npass), C.byref(gsh)))
            out.append({"limb_bits": v[0].value, "logN": v[1].value, "members": v[2].value, "slots": v[3].value,
                        "passes": [pl[j] for j in range(npass.value)], "gshift": gsh.value})
        return out

    def fused(self):
        """(products run by the fused whole-product launch, their largest log2 N)."""
        n, m = C.c_int(), C.c_int()
        _check(lib().bw_batch_fused_info(self._b, C.byref(n), C.byref(m)))
        return n.value, m.value

    def pack(self, As, Bs):
        """Lists of limb arrays -> packed host buffers (numpy uint8)."""
        pa = np.zeros(max(self.bytes_a, 1), dtype=np.uint8)
        pb = np.zeros(max(self.bytes_b, 1), dtype=np.uint8)
        for i in range(self.n):
            dt = _dt(self.bits[i])
            a = np.ascontiguousarray(As[i], dtype=dt).view(np.uint8)
            b = np.ascontiguousarray(Bs[i], dtype=dt).view(np.uint8)
            pa[self.off_a[i]:self.off_a[i] + a.size] = a
            pb[self.off_b[i]:self.off_b[i] + b.size] = b
        return pa, pb

    def unpack(self, pc):
        """Packed result bytes -> list of product limb arrays."""
        pc = np.asarray(pc).view(np.uint8)
        out = []
        for i in range(self.n):
            dt = _dt(self.bits[i])
            nb = (self.L_A[i] + self.L_B[i]) * (self.bits[i] // 8)
            out.append(pc[self.off_c[i]:self.off_c[i] + nb].view(dt).copy())
        return out

    def execute(self, a, b, c, stream=None):
        """Packed device buffers (torch uint8 tensors) -> packed results in c."""
        _check(lib().bw_batch_execute(self._b, _dev(a, self.bytes_a, "a"), _dev(b, self.bytes_b, "b"),
                                      _dev(c, self.bytes_c, "c"), _stream(stream)))

    def set_graph(self, on):
        """False/0: plain launches; True/1: replayed hipGraph; 2: hipGraph with
        one concurrent branch per size group."""
        _check(lib().bw_batch_set_graph(self._b, int(on)))

    def tune(self, a, b, stream=None, reps=3):
        _check(lib().bw_batch_tune(self._b, _dev(a, self.bytes_a, "a"), _dev(b, self.bytes_b, "b"),
                                   _stream(stream), reps))

    def status(self):
        _check(lib().bw_batch_status(self._b))


def set_fused_limit(max_log):
    """Largest log2 transform length (0 .. 11) that batches and servers created
    from now on run through the fused whole-product launch (0: none).  Returns
    the previous limit."""
    r = lib().bw_set_fused_limit(int(max_log))
    if r < 0:
        _check(r)
    return r


def batch_footprint(limb_bits, L_A, L_B):
    """Host-only: (workspace, packed A, packed B, packed C) bytes of a batch."""
    n = len(L_A)
    if isinstance(limb_bits, int):
        limb_bits = [limb_bits] * n
    out = [C.c_uint64() for _ in range(4)]
    _check(lib().bw_batch_footprint(n, (C.c_int * n)(*limb_bits), (C.c_uint64 * n)(*L_A), (C.c_uint64 * n)(*L_B),
                                    *[C.byref(o) for o in out]))
    return tuple(int(o.value) for o in out)


def tune_count():
    return int(lib().bw_tune_count())


def tune_clear():
    lib().bw_tune_clear()


# ---- serving -------------------------------------------------------------------
class Server:
    """Request loop with continuous batching on the current HIP device.

    submit() takes host limb arrays (numpy; pinned memory lets the copies
    overlap) and the result array (L_A + L_B limbs), which must stay alive
    until wait() returns for the ticket; the server keeps references to them.
    """

    def __init__(self, device_bytes=8 << 30, streams=3, max_batch=1024):
        cfg = ServerConfig(device_bytes, streams, max_batch)
        self._s = C.c_void_p()
        _check(lib().bw_server_create(C.byref(self._s), C.byref(cfg)))
        self._keep = {}

    def submit(self, A, B, out, limb_bits=None):
        if limb_bits is None:
            limb_bits = 64 if A.dtype == np.uint64 else 32
        dt = np.dtype(_dt(limb_bits))
        for name, x, n in (("A", A, len(A)), ("B", B, len(B)), ("out", out, len(A) + len(B))):
            if not isinstance(x, np.ndarray) or x.dtype != dt or not x.flags.c_contiguous or x.size < n:
                raise BWError(BW_ERR_ARG, "%s must be a contiguous numpy %s array of at least %d limbs"
                              % (name, dt.name, n))
        t = int(lib().bw_server_submit(self._s, limb_bits, _ptr(A), len(A), _ptr(B), len(B), _ptr(out)))
        if t < 0:
            _check(t)
        self._keep[t] = (A, B, out)
        return t

    def wait(self, ticket, timeout_ms=-1):
        """Blocks until the ticket completes; returns its latency in ms."""
        lat = C.c_double()
        rc = lib().bw_server_wait(self._s, ticket, timeout_ms, C.byref(lat))
        self._keep.pop(ticket, None)
        _check(rc)
        return lat.value

    def poll(self, ticket):
        rc = int(lib().bw_server_poll(self._s, ticket))
        if rc < 0:
            _check(rc)
        return rc == 1

    def stats(self):
        st = ServerStats()
        _check(lib().bw_server_get_stats(self._s, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self._s:
            lib().bw_server_destroy(self._s)
            self._s = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
