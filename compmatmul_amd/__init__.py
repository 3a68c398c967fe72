"""compmatmul_amd: MI355X-native compressed matrix multiplication.

Python host layer over the C-ABI library ``libcompmatmul_amd.so``
(declared in ``include/compmatmul_amd.h``).  The drop-in module with the
reference's interface is ``compmatmul_amd.compmatmul`` (importable as plain
``compmatmul`` with this directory on ``sys.path``); this module adds the
extended interface: float32 inputs, device-resident inputs and sketches,
sketch values, hash arrays, large-entry recovery, kernel timing, batches
of many products (Batch) and a request server with continuous batching
(Server).

There is no CPU fallback: if the native library is missing or no gfx950
device is present, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# COMPMATMUL_AMD_LIB selects another build of the C-ABI library
LIB_PATH = os.environ.get("COMPMATMUL_AMD_LIB") or os.path.join(HERE, "libcompmatmul_amd.so")

FFT, FWHT = 0, 1
F64, F32 = 0, 1
HOST, DEVICE = 0, 1
FLAG_F64_REDUCE, FLAG_PLAIN, FLAG_LDS_KERNEL, FLAG_BF16_MFMA, FLAG_SINGLE_WAVE, FLAG_FP16_MFMA = 1, 2, 4, 8, 16, 32
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
    L.cm_matmul_noalloc.argtypes = [ctypes.c_int] * 5 + [_vp, ctypes.c_int64] * 3 + [
        ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    L.cm_timing_enable.argtypes = [ctypes.c_int]
    L.cm_timing_only.argtypes = [ctypes.c_char_p]
    L.cm_timing_get.argtypes = [ctypes.c_char_p, _f64p, _i64p]
    L.cm_probe_alu.argtypes = [ctypes.c_int] * 4 + [_f64p, _f64p]
    L.cm_probe_stream.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_size_t, ctypes.c_int, _f64p,
                                  _f64p]
    L.cm_batch_plan.argtypes = [ctypes.c_int, _vp, ctypes.c_uint, ctypes.c_int,
                                ctypes.POINTER(_vp)]
    L.cm_batch_run.argtypes = [_vp]
    L.cm_batch_grain.argtypes = [_vp]
    L.cm_batch_recovered.argtypes = [_vp, ctypes.c_int, ctypes.c_int64, _i32p, _i32p, _vp, _i64p]
    L.cm_batch_values.argtypes = [_vp, ctypes.c_int, _vp]
    L.cm_batch_free.argtypes = [_vp]
    L.cm_batch_free.restype = None
    L.cm_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(_vp)]
    L.cm_server_submit.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_uint64)]
    L.cm_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.c_double]
    L.cm_server_result.argtypes = [_vp, ctypes.c_uint64, ctypes.c_int64, _i32p, _i32p, _vp,
                                   _i64p, ctypes.POINTER(ctypes.c_double)]
    L.cm_server_decompress.argtypes = [_vp, ctypes.c_uint64, _vp]
    L.cm_server_release.argtypes = [_vp, ctypes.c_uint64]
    L.cm_server_values.argtypes = [_vp, ctypes.c_uint64, _vp]
    L.cm_server_stats_get.argtypes = [_vp, _vp]
    L.cm_server_destroy.argtypes = [_vp]
    L.cm_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.cm_host_free.argtypes = [_vp]
    L.cm_context_create.argtypes = [ctypes.c_int, _vp, ctypes.c_size_t, ctypes.POINTER(_vp)]
    L.cm_context_destroy.argtypes = [_vp]
    L.cm_ctx_compress.argtypes = [_vp] + L.cm_compress_ex.argtypes
    L.cm_ctx_decompress.argtypes = [_vp] + L.cm_decompress.argtypes
    L.cm_ctx_recover.argtypes = [_vp] + L.cm_recover.argtypes
    L.cm_ctx_synchronize.argtypes = [_vp]
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
    """Device-resident compressed product (d x b values plus hash arrays).
    A sketch made through a Context decompresses and recovers through it."""

    def __init__(self, variant, dtype, d, b, m, p):
        self.handle = _vp()
        self.context = None
        self.variant, self.dtype, self.d, self.b, self.m, self.p = variant, dtype, d, b, m, p

    def _decompress_call(self, ptr, mem):
        if self.context is not None:
            return lib().cm_ctx_decompress(self.context.handle, self.handle, ptr, self.p, mem)
        return lib().cm_decompress(self.handle, ptr, self.p, mem)

    def _check_out(self, out):
        """ValueError unless `out` can take the m x p estimate in this dtype."""
        if out.shape != (self.m, self.p):
            raise ValueError(f"out has shape {out.shape}, expected {(self.m, self.p)}")
        if np.dtype(out.dtype) != np.dtype(self.dtype):
            raise ValueError(f"out has dtype {out.dtype}, expected {np.dtype(self.dtype)}")
        if isinstance(out, np.ndarray):
            if not out.flags.c_contiguous:
                raise ValueError("out must be C-contiguous")
            if not out.flags.writeable:
                raise ValueError("out must be writeable")
        elif not isinstance(out, DeviceArray):
            raise ValueError("out must be a numpy array or a DeviceArray")

    def decompress(self, out=None):
        """Median-of-d estimate of C (m x p): numpy array or DeviceArray `out`."""
        if out is not None:
            self._check_out(out)
        if isinstance(out, DeviceArray):
            check(self._decompress_call(out.ptr, DEVICE))
            return out
        C = np.empty((self.m, self.p), self.dtype) if out is None else out
        check(self._decompress_call(C.ctypes.data, HOST))
        return C

    def recover(self, tau, max_out=1 << 16):
        """Entries with estimate > tau as (i, j, value) arrays sorted by (i, j)."""
        while True:
            oi = np.empty(max_out, np.int32)
            oj = np.empty(max_out, np.int32)
            ov = np.empty(max_out, np.float64)
            cnt = ctypes.c_int64()
            args = (self.handle, float(tau), max_out, oi.ctypes.data_as(_i32p),
                    oj.ctypes.data_as(_i32p), ov.ctypes.data_as(_f64p), ctypes.byref(cnt))
            if self.context is not None:
                check(lib().cm_ctx_recover(self.context.handle, *args))
            else:
                check(lib().cm_recover(*args))
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


def compress(variant, AT, B, d, b, seed=None, hash=None, dtype=None, flags=0, sketch=None,
             context=None):
    """Compressed product of AT (n x m) and B (n x p).

    `AT`/`B` are numpy arrays or DeviceArrays; dtype float64 or float32
    (default: the operands' dtype).  Pass `sketch` to reuse its device
    buffers (no allocation), `context` to run on a Context.  Returns a
    Sketch.
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
    args = (v, code, m, n, p, pa, lda, pb, ldb, d, b, seed, FAMILIES[hash], ma, flags,
            ctypes.byref(sk.handle))
    if context is not None:
        check(lib().cm_ctx_compress(context.handle, *args))
    else:
        check(lib().cm_compress_ex(*args))
    sk.seed = seed
    sk.context = context
    del ka, kb
    return sk


class Context:
    """A device, a stream and working memory of its own (cm_context_create).
    Sketches compressed through a context decompress and recover through it;
    calls on different contexts run concurrently from different threads.
    With workspace_bytes > 0 every per-call intermediate is carved from one
    device region of that size, made here: a product that needs more raises
    MemoryError, and the plain and matrix-core paths raise CMError."""

    def __init__(self, device=0, workspace_bytes=0):
        self.handle = _vp()
        check(lib().cm_context_create(int(device), None, int(workspace_bytes),
                                      ctypes.byref(self.handle)))

    def compress(self, variant, AT, B, d, b, seed=None, hash=None, dtype=None, flags=0,
                 sketch=None):
        return compress(variant, AT, B, d, b, seed, hash, dtype, flags, sketch, context=self)

    def synchronize(self):
        check(lib().cm_ctx_synchronize(self.handle))

    def destroy(self):
        if self.handle:
            check(lib().cm_context_destroy(self.handle))
            self.handle = _vp()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


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


def timing_enable(on=True, only=None):
    """Time kernel launches with HIP events; `only` restricts the timing to one
    kernel class (e.g. "compress"), so the other launches run without events."""
    check(lib().cm_timing_only(only.encode() if only else None))
    check(lib().cm_timing_enable(1 if on else 0))


def timing_reset():
    check(lib().cm_timing_reset())


def timing_get(name):
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().cm_timing_get(name.encode(), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


PROBE_ALU = {"fp32_add": 0, "fp32_fma": 1, "fp64_add": 2, "fp64_fma": 3}
PROBE_STREAM = {"copy": 0, "read": 1, "fill": 2}


def probe_alu(kind, blocks=4096, iters=4096, reps=5):
    """Calibration loop of known work (`kind` in PROBE_ALU): returns (ms per
    launch, flops per launch)."""
    ms, flops = ctypes.c_double(), ctypes.c_double()
    check(lib().cm_probe_alu(PROBE_ALU[kind], blocks, iters, reps, ctypes.byref(ms),
                             ctypes.byref(flops)))
    return ms.value, flops.value


def probe_stream(kind, dst, src, nbytes, reps=5):
    """Streaming copy, read or fill (`kind` in PROBE_STREAM) over `nbytes` of
    the DeviceArrays dst / src: returns (ms per launch, bytes moved per
    launch)."""
    ms, moved = ctypes.c_double(), ctypes.c_double()
    check(lib().cm_probe_stream(PROBE_STREAM[kind], dst.ptr if dst is not None else None,
                                src.ptr if src is not None else None, nbytes, reps,
                                ctypes.byref(ms), ctypes.byref(moved)))
    return ms.value, moved.value


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


class HostBuffer:
    """Pinned host memory viewed as a numpy array (`.array`), for operands
    that the server uploads asynchronously."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        shape = tuple(int(x) for x in np.atleast_1d(shape))
        nbytes = max(1, int(np.prod(shape)) * dtype.itemsize)
        self.ptr = _vp()
        check(lib().cm_host_alloc(nbytes, ctypes.byref(self.ptr)))
        raw = (ctypes.c_char * nbytes).from_address(self.ptr.value)
        self.array = np.frombuffer(raw, dtype=np.uint8)[:int(np.prod(shape)) * dtype.itemsize] \
            .view(dtype).reshape(shape)

    @classmethod
    def from_numpy(cls, a):
        out = cls(a.shape, a.dtype)
        out.array[...] = a
        return out

    def __del__(self):
        try:
            if self.ptr:
                lib().cm_host_free(self.ptr)
                self.ptr = _vp()
        except Exception:
            pass


class _Request(ctypes.Structure):
    _fields_ = [("variant", ctypes.c_int), ("m", ctypes.c_int), ("n", ctypes.c_int),
                ("p", ctypes.c_int), ("d", ctypes.c_int), ("b", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("family", ctypes.c_int), ("tau", ctypes.c_double),
                ("max_out", ctypes.c_int64), ("format", ctypes.c_int), ("AT", _vp), ("B", _vp),
                ("at_indptr", _vp), ("at_indices", _vp), ("at_values", _vp), ("b_indptr", _vp),
                ("b_indices", _vp), ("b_values", _vp), ("C", _vp), ("mem", ctypes.c_int)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int64), ("requests_done", ctypes.c_int64),
                ("requests_failed", ctypes.c_int64), ("max_waves_in_flight", ctypes.c_int),
                ("hbm_bytes", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t),
                ("launch_s", ctypes.c_double), ("retire_s", ctypes.c_double)]


DENSE, CSR = 0, 1


def _host_f32(x):
    if isinstance(x, HostBuffer):
        return x.array
    return np.ascontiguousarray(x, np.float32)


def _csr_parts(x):
    """(indptr, indices, values) of a CSR operand: a scipy.sparse matrix or a
    tuple of arrays."""
    if hasattr(x, "indptr"):
        ip, ix, v = x.indptr, x.indices, x.data
    else:
        ip, ix, v = x
    return (np.ascontiguousarray(ip, np.int32), np.ascontiguousarray(ix, np.int32),
            _host_f32(v))


class Server:
    """Request server with continuous batching on one GPU (cm_server_*).

    Requests queue in arrival order; a dispatcher thread admits them while
    their device footprint fits `hbm_bytes` and runs them as waves (batches)
    on up to `max_waves` streams at once.  Results: the entries above tau
    (and C when asked for), plus the request's latency.  The sketch stays on
    the device until release(ticket)."""

    def __init__(self, hbm_bytes=8 << 30, max_waves=3, max_wave_products=64, grain=64):
        self.handle = _vp()
        self._keep = {}
        check(lib().cm_server_create(int(hbm_bytes), int(max_waves), int(max_wave_products),
                                     int(grain), ctypes.byref(self.handle)))

    def submit(self, variant, AT, B, d, b, seed, hash=None, tau=None, max_out=4096,
               want_C=False, shape=None, C=None):
        """AT (n x m) and B (n x p): float32 arrays / HostBuffers, or CSR
        operands (scipy.sparse CSR or (indptr, indices, values); then pass
        shape=(m, n, p)).  DeviceArray operands (dense, or a CSR tuple of
        DeviceArrays) stay resident in HBM and are read in place; such a
        request writes C only into a float32 DeviceArray `C` (m x p).
        Returns the ticket."""
        if hash not in FAMILIES:
            raise ValueError(f"Invalid hash `{hash}'")
        r = _Request()
        keep = []
        sparse = hasattr(AT, "indptr") or isinstance(AT, tuple)
        dev = isinstance(AT, DeviceArray) or (isinstance(AT, tuple) and isinstance(AT[0], DeviceArray))
        if sparse:
            if shape is None:
                if not hasattr(AT, "shape"):
                    raise ValueError("pass shape=(m, n, p) with tuple CSR operands")
                shape = (AT.shape[1], AT.shape[0], B.shape[1])
            m, n, p = shape
            if dev:
                a, bb = tuple(AT), tuple(B)
                if not all(isinstance(x, DeviceArray) for x in a + bb):
                    raise ValueError("device CSR operands are tuples of DeviceArrays")
                ptrs_a, ptrs_b = [x.ptr for x in a], [x.ptr for x in bb]
            else:
                a, bb = _csr_parts(AT), _csr_parts(B)
                ptrs_a, ptrs_b = [x.ctypes.data for x in a], [x.ctypes.data for x in bb]
            keep += [a, bb]
            r.format = CSR
            r.at_indptr, r.at_indices, r.at_values = ptrs_a
            r.b_indptr, r.b_indices, r.b_values = ptrs_b
        else:
            if dev:
                if not isinstance(B, DeviceArray) or AT.dtype != np.float32 or B.dtype != np.float32:
                    raise ValueError("device operands are float32 DeviceArrays")
                a, bb = AT, B
            else:
                a, bb = _host_f32(AT), _host_f32(B)
            n, m = a.shape
            if bb.shape[0] != n:
                raise ValueError("Shape mismatch! Note that AT is assumed to be transposed")
            p = bb.shape[1]
            keep += [a, bb, AT, B]
            r.format = DENSE
            r.AT, r.B = (a.ptr, bb.ptr) if dev else (a.ctypes.data, bb.ctypes.data)
        r.variant, r.m, r.n, r.p, r.d, r.b = VARIANTS[variant], m, n, p, int(d), int(b)
        r.seed, r.family = int(seed), FAMILIES[hash]
        r.tau = float("nan") if tau is None else float(tau)
        r.max_out = int(max_out)
        r.mem = DEVICE if dev else HOST
        if dev:
            if want_C:
                raise ValueError("a device-resident request writes C into a DeviceArray: pass C=")
            if C is not None:
                if not isinstance(C, DeviceArray) or C.shape != (m, p) or C.dtype != np.float32:
                    raise ValueError("C must be a float32 m x p DeviceArray")
                keep.append(C)
                r.C = C.ptr
        elif want_C:
            C = np.empty((m, p), np.float32)
            keep.append(C)
            r.C = C.ctypes.data
        elif C is not None:
            raise ValueError("C= is for device-resident requests; use want_C=True")
        t = ctypes.c_uint64()
        check(lib().cm_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)))
        self._keep[t.value] = (keep, C, int(max_out), (m, p), (int(d), int(b)))
        return t.value

    def wait(self, ticket, timeout=None):
        """True when done (raises on a failed request), False on timeout."""
        rc = lib().cm_server_wait(self.handle, ticket, -1.0 if timeout is None else float(timeout))
        if rc == 1:
            return False
        check(rc)
        return True

    def result(self, ticket):
        """(rows, cols, values, latency_s, C or None) of a finished request."""
        keep, C, cap = self._keep[ticket][:3]
        oi = np.empty(max(1, cap), np.int32)
        oj = np.empty(max(1, cap), np.int32)
        ov = np.empty(max(1, cap), np.float32)
        cnt = ctypes.c_int64()
        lat = ctypes.c_double()
        check(lib().cm_server_result(self.handle, ticket, cap, oi.ctypes.data_as(_i32p),
                                     oj.ctypes.data_as(_i32p), ov.ctypes.data,
                                     ctypes.byref(cnt), ctypes.byref(lat)))
        k = min(cnt.value, cap)
        return oi[:k].copy(), oj[:k].copy(), ov[:k].copy(), lat.value, C

    def decompress(self, ticket):
        """All of C from the request's device-resident sketch."""
        m, p = self._keep[ticket][3]
        C = np.empty((m, p), np.float32)
        check(lib().cm_server_decompress(self.handle, ticket, C.ctypes.data))
        return C

    def values(self, ticket):
        """The request's device-resident sketch values (d x b, float32)."""
        out = np.empty(self._keep[ticket][4], np.float32)
        check(lib().cm_server_values(self.handle, ticket, out.ctypes.data))
        return out

    def release(self, ticket):
        check(lib().cm_server_release(self.handle, ticket))
        self._keep.pop(ticket, None)

    def stats(self):
        st = _ServerStats()
        check(lib().cm_server_stats_get(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def close(self):
        if self.handle:
            check(lib().cm_server_destroy(self.handle))
            self.handle = _vp()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dropin():
    """The reference-compatible `compmatmul` module (gemm/matmul_fft/matmul_fwht)."""
    from . import compmatmul as mod
    return mod
