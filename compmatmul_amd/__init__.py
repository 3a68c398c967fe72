"""compmatmul_amd: MI355X-native compressed matrix multiplication.

Python host layer over the C-ABI library ``libcompmatmul_amd.so``
(declared in ``include/compmatmul_amd.h``).  The drop-in module with the
reference's interface is ``compmatmul_amd.compmatmul`` (importable as plain
``compmatmul`` with this directory on ``sys.path``); this module adds the
extended interface: float32 inputs, device-resident inputs and sketches,
sketch values, hash arrays, large-entry recovery, kernel timing, and
batches of many products (Batch).

There is no CPU fallback: if the native library is missing or no gfx950
device is present, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcompmatmul_amd.so")

FFT, FWHT = 0, 1
F64, F32 = 0, 1
HOST, DEVICE = 0, 1
FLAG_F64_REDUCE, FLAG_PLAIN, FLAG_LDS_KERNEL = 1, 2, 4
OK, EINVAL, ENOMEM, EDEVICE, ENOTSUP, EHASH = 0, -1, -2, -3, -4, -5
FAMILIES = {None: 0, "multiply add shift": 0, "multiply shift": 1,
            "tabulation8": 2, "tabulation16": 3}
VARIANTS = {"fft": FFT, "fwht": FWHT, FFT: FFT, FWHT: FWHT}

_LIB = None
_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


class CMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    """The loaded C-ABI library (raises ImportError when it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m compmatmul_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.cm_last_error.restype = ctypes.c_char_p
    L.cm_default_seed.restype = ctypes.c_uint64
    L.cm_init.argtypes = [ctypes.c_int]
    L.cm_hash_arrays.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int]
    L.cm_compress_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, _vp, ctypes.c_int64, _vp, ctypes.c_int64,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_uint, ctypes.POINTER(_vp)]
    L.cm_compress.argtypes = L.cm_compress_ex.argtypes[:14] + [ctypes.POINTER(_vp)]
    L.cm_decompress.argtypes = [_vp, _vp, ctypes.c_int64, ctypes.c_int]
    L.cm_recover.argtypes = [_vp, ctypes.c_double, ctypes.c_int64, _i32p, _i32p, _f64p, _i64p]
    L.cm_sketch_values.argtypes = [_vp, _f64p]
    L.cm_sketch_hashes.argtypes = [_vp, _u32p, _u32p]
    L.cm_sketch_free.argtypes = [_vp]
    L.cm_sketch_free.restype = None
    L.cm_dev_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.cm_dev_free.argtypes = [_vp]
    L.cm_memcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]
    for name in ("cm_matmul_fft", "cm_matmul_fwht"):
        getattr(L, name).argtypes = [ctypes.c_int] * 3 + [_vp] * 3 + [
            ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    L.cm_gemm.argtypes = [ctypes.c_int] * 3 + [_vp] * 3
    L.cm_timing_enable.argtypes = [ctypes.c_int]
    L.cm_timing_get.argtypes = [ctypes.c_char_p, _f64p, _i64p]
    L.cm_batch_plan.argtypes = [ctypes.c_int, _vp, ctypes.c_uint, ctypes.c_int,
                                ctypes.POINTER(_vp)]
    L.cm_batch_run.argtypes = [_vp]
    L.cm_batch_grain.argtypes = [_vp]
    L.cm_batch_recovered.argtypes = [_vp, ctypes.c_int, ctypes.c_int64, _i32p, _i32p, _vp, _i64p]
    L.cm_batch_values.argtypes = [_vp, ctypes.c_int, _vp]
    L.cm_batch_free.argtypes = [_vp]
    L.cm_batch_free.restype = None
    _LIB = L
    return L


def check(rc):
    if rc != OK:
        msg = lib().cm_last_error().decode()
        if rc == EINVAL:
            raise ValueError(msg)
        if rc == ENOMEM:
            raise MemoryError(msg)
        raise CMError(rc, msg)


def default_seed():
    return int(lib().cm_default_seed())


class DeviceArray:
    """A device buffer with a numpy-like shape and dtype (row-major)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = _vp()
        check(lib().cm_dev_alloc(max(self.nbytes, 1), ctypes.byref(self.ptr)))

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        out = cls(a.shape, a.dtype)
        if a.nbytes:
            check(lib().cm_memcpy(out.ptr, a.ctypes.data, a.nbytes, DEVICE, HOST))
        return out

    def upload(self, a):
        """Copy host array `a` (same shape and dtype) into this buffer."""
        a = np.ascontiguousarray(a, self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        if a.nbytes:
            check(lib().cm_memcpy(self.ptr, a.ctypes.data, a.nbytes, DEVICE, HOST))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().cm_memcpy(out.ctypes.data, self.ptr, self.nbytes, HOST, DEVICE))
        return out

    def free(self):
        if self.ptr:
            lib().cm_dev_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _operand(x, dtype):
    """(pointer, mem kind, leading dimension, keepalive) of a matrix operand."""
    if isinstance(x, DeviceArray):
        if x.dtype != dtype:
            raise ValueError(f"device operand has dtype {x.dtype}, expected {dtype}")
        return x.ptr, DEVICE, x.shape[1], x
    a = np.ascontiguousarray(x, dtype)
    return a.ctypes.data, HOST, a.shape[1], a


class Sketch:
    """Device-resident compressed product (d x b values plus hash arrays)."""

    def __init__(self, variant, dtype, d, b, m, p):
        self.handle = _vp()
        self.variant, self.dtype, self.d, self.b, self.m, self.p = variant, dtype, d, b, m, p

    def decompress(self, out=None):
        """Median-of-d estimate of C (m x p): numpy array or DeviceArray `out`."""
        if isinstance(out, DeviceArray):
            check(lib().cm_decompress(self.handle, out.ptr, self.p, DEVICE))
            return out
        C = np.empty((self.m, self.p), self.dtype) if out is None else out
        check(lib().cm_decompress(self.handle, C.ctypes.data, self.p, HOST))
        return C

    def recover(self, tau, max_out=1 << 16):
        """Entries with estimate > tau as (i, j, value) arrays sorted by (i, j)."""
        while True:
            oi = np.empty(max_out, np.int32)
            oj = np.empty(max_out, np.int32)
            ov = np.empty(max_out, np.float64)
            cnt = ctypes.c_int64()
            check(lib().cm_recover(self.handle, float(tau), max_out,
                                   oi.ctypes.data_as(_i32p), oj.ctypes.data_as(_i32p),
                                   ov.ctypes.data_as(_f64p), ctypes.byref(cnt)))
            if cnt.value <= max_out:
                k = cnt.value
                return oi[:k].copy(), oj[:k].copy(), ov[:k].copy()
            max_out = cnt.value

    def values(self):
        out = np.empty((self.d, self.b), np.float64)
        check(lib().cm_sketch_values(self.handle, out.ctypes.data_as(_f64p)))
        return out

    def hashes(self):
        hA = np.empty((self.d, self.m), np.uint32)
        hB = np.empty((self.d, self.p), np.uint32)
        check(lib().cm_sketch_hashes(self.handle, hA.ctypes.data_as(_u32p),
                                     hB.ctypes.data_as(_u32p)))
        return hA, hB

    def free(self):
        if self.handle:
            lib().cm_sketch_free(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def compress(variant, AT, B, d, b, seed=None, hash=None, dtype=None, flags=0, sketch=None):
    """Compressed product of AT (n x m) and B (n x p).

    `AT`/`B` are numpy arrays or DeviceArrays; dtype float64 or float32
    (default: the operands' dtype).  Pass `sketch` to reuse its device
    buffers (no allocation).  Returns a Sketch.
    """
    v = VARIANTS[variant]
    if dtype is None:
        dtype = AT.dtype if isinstance(AT, (DeviceArray, np.ndarray)) else np.float64
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
        raise ValueError("dtype must be float64 or float32")
    code = F64 if dtype == np.float64 else F32
    pa, ma, lda, ka = _operand(AT, dtype)
    pb, mb, ldb, kb = _operand(B, dtype)
    if ma != mb:
        raise ValueError("AT and B must both be host or both be device operands")
    n, m = AT.shape
    if B.shape[0] != n:
        raise ValueError("Shape mismatch! Note that AT is assumed to be transposed")
    p = B.shape[1]
    if hash not in FAMILIES:
        raise ValueError(f"Invalid hash `{hash}'")
    if seed is None:
        seed = default_seed()
    sk = sketch if sketch is not None else Sketch(v, dtype, d, b, m, p)
    check(lib().cm_compress_ex(v, code, m, n, p, pa, lda, pb, ldb, d, b, seed,
                               FAMILIES[hash], ma, flags, ctypes.byref(sk.handle)))
    sk.seed = seed
    del ka, kb
    return sk


def matmul(variant, AT, B, d, b, seed=None, hash=None, dtype=None):
    """Full compressed product C (m x p) as a numpy array."""
    sk = compress(variant, AT, B, d, b, seed, hash, dtype)
    try:
        return sk.decompress()
    finally:
        sk.free()


def hash_arrays(hash, d, b, seed, m, p):
    """Packed bucket|signbit arrays evaluated on the device (bit-exact)."""
    hA = np.empty((d, m), np.uint32)
    hB = np.empty((d, p), np.uint32)
    check(lib().cm_hash_arrays(FAMILIES[hash], d, b, seed, m, p, hA.ctypes.data,
                               hB.ctypes.data, HOST))
    return hA, hB


def synchronize():
    check(lib().cm_synchronize())


def timing_enable(on=True):
    check(lib().cm_timing_enable(1 if on else 0))


def timing_reset():
    check(lib().cm_timing_reset())


def timing_get(name):
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().cm_timing_get(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


class _Product(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int),
                ("p", ctypes.c_int), ("AT", _vp), ("ld_at", ctypes.c_int64), ("B", _vp),
                ("ld_b", ctypes.c_int64), ("C", _vp), ("ldc", ctypes.c_int64),
                ("d", ctypes.c_int), ("b", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("family", ctypes.c_int), ("tau", ctypes.c_double), ("max_out", ctypes.c_int64)]


BATCH_GRAPH = 1


class Product:
    """One product of a Batch: float32 DeviceArray operands AT (n x m) and
    B (n x p), optional DeviceArray output C (m x p), recovery threshold tau
    (None: no recovery) with room for max_out entries."""

    def __init__(self, variant, AT, B, d, b, seed, hash=None, C=None, tau=None, max_out=4096):
        for name, x in (("AT", AT), ("B", B), ("C", C)):
            if x is not None and (not isinstance(x, DeviceArray) or x.dtype != np.float32):
                raise ValueError(f"{name} must be a float32 DeviceArray")
        if AT.shape[0] != B.shape[0]:
            raise ValueError("Shape mismatch! Note that AT is assumed to be transposed")
        if C is not None and C.shape != (AT.shape[1], B.shape[1]):
            raise ValueError("C must be m x p")
        if hash not in FAMILIES:
            raise ValueError(f"Invalid hash `{hash}'")
        self.variant, self.AT, self.B, self.C = VARIANTS[variant], AT, B, C
        self.d, self.b, self.seed, self.hash = int(d), int(b), int(seed), hash
        self.tau, self.max_out = tau, int(max_out)
        self.n, self.m = AT.shape
        self.p = B.shape[1]

    def _c(self):
        q = _Product()
        q.variant, q.m, q.n, q.p = self.variant, self.m, self.n, self.p
        q.AT, q.ld_at = self.AT.ptr, self.m
        q.B, q.ld_b = self.B.ptr, self.p
        q.C, q.ldc = (self.C.ptr if self.C is not None else None), self.p
        q.d, q.b, q.seed, q.family = self.d, self.b, self.seed, FAMILIES[self.hash]
        q.tau = float("nan") if self.tau is None else float(self.tau)
        q.max_out = self.max_out if self.tau is not None else 0
        return q


class Batch:
    """Many float32 products planned once and run in one pass: one persistent
    compress launch per transform variant over XCD-local work queues, one
    launch per finishing step and per decompression class.  grain=0 uses the
    autotuned grain for these shapes; graph=True replays a HIP graph."""

    def __init__(self, products, grain=0, graph=False):
        self.products = list(products)
        arr = (_Product * max(1, len(self.products)))()
        for i, q in enumerate(self.products):
            arr[i] = q._c()
        self.handle = _vp()
        check(lib().cm_batch_plan(len(self.products), arr, BATCH_GRAPH if graph else 0,
                                  int(grain), ctypes.byref(self.handle)))

    @property
    def grain(self):
        return int(lib().cm_batch_grain(self.handle))

    def run(self):
        check(lib().cm_batch_run(self.handle))

    def recovered(self, i):
        """(i, j, value) arrays of product i's entries > tau, sorted by (i, j)."""
        cap = self.products[i].max_out
        oi = np.empty(max(1, cap), np.int32)
        oj = np.empty(max(1, cap), np.int32)
        ov = np.empty(max(1, cap), np.float32)
        cnt = ctypes.c_int64()
        check(lib().cm_batch_recovered(self.handle, i, cap, oi.ctypes.data_as(_i32p),
                                       oj.ctypes.data_as(_i32p), ov.ctypes.data, ctypes.byref(cnt)))
        if cnt.value > cap:
            raise RuntimeError(f"product {i}: {cnt.value} entries above tau, room for {cap}")
        k = cnt.value
        return oi[:k].copy(), oj[:k].copy(), ov[:k].copy()

    def values(self, i):
        q = self.products[i]
        out = np.empty((q.d, q.b), np.float32)
        check(lib().cm_batch_values(self.handle, i, out.ctypes.data))
        return out

    def free(self):
        if self.handle:
            lib().cm_batch_free(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dropin():
    """The reference-compatible `compmatmul` module (gemm/matmul_fft/matmul_fwht)."""
    from . import compmatmul as mod
    return mod
