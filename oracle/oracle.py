"""ctypes view of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a plain FP64 restatement of the reference algorithm
(mkarppa/compmatmul).  Only the tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only to check
results or as the timed CPU baseline; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FAMILIES = {None: 0, "multiply add shift": 0, "multiply shift": 1,
            "tabulation8": 2, "tabulation16": 3}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_rng_stream.argtypes = [ctypes.c_uint64, ctypes.c_int64, _u64p]
        L.orc_drng_stream.argtypes = [ctypes.c_uint64, ctypes.c_int64, _f64p]
        L.orc_default_seed.argtypes = [ctypes.c_uint64]
        L.orc_default_seed.restype = ctypes.c_uint64
        L.orc_param_words.argtypes = [ctypes.c_int]
        L.orc_param_words.restype = ctypes.c_int64
        L.orc_draw.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                               ctypes.c_uint64, _u64p]
        L.orc_hash.argtypes = [ctypes.c_int, _u64p, ctypes.c_uint64, ctypes.c_uint32]
        L.orc_hash.restype = ctypes.c_uint32
        L.orc_sign.argtypes = [ctypes.c_int, _u64p, ctypes.c_uint32]
        L.orc_hash_arrays.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                      _u32p, _u32p]
        L.orc_fft_r2c.argtypes = [ctypes.c_int, _f64p, _f64p]
        L.orc_fft_c2r.argtypes = [ctypes.c_int, _f64p, _f64p]
        L.orc_fwht.argtypes = [ctypes.c_int, _f64p]
        L.orc_compress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, _f64p, _f64p,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                   ctypes.c_int, _f64p]
        L.orc_decompress.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, _u32p, _u32p,
                                     _f64p, ctypes.c_int, _f64p]
        L.orc_median.argtypes = [ctypes.c_int, _f64p]
        L.orc_median.restype = ctypes.c_double
        L.orc_sketch_from_product.argtypes = [ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, _u32p, _u32p, _f64p,
                                              _f64p]
        L.orc_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p,
                               _f64p, _f64p, ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def rng_stream(seed, n):
    out = np.empty(n, np.uint64)
    lib().orc_rng_stream(seed, n, _p(out, _u64p))
    return out


def drng_stream(seed, n):
    out = np.empty(n, np.float64)
    lib().orc_drng_stream(seed, n, _p(out, _f64p))
    return out


def default_seed(millis):
    return int(lib().orc_default_seed(millis))


def draw(family, n, buckets, seed):
    """Parameters of n hash functions (rows of the returned uint64 array)."""
    w = lib().orc_param_words(family)
    out = np.empty(n * w, np.uint64)
    assert lib().orc_draw(family, n, buckets, seed, _p(out, _u64p)) == 0
    return out.reshape(n, w)


def hash_eval(family, params, buckets, key):
    params = np.ascontiguousarray(params, np.uint64)
    return int(lib().orc_hash(family, _p(params, _u64p), buckets, key))


def sign_eval(family, params, key):
    params = np.ascontiguousarray(params, np.uint64)
    return int(lib().orc_sign(family, _p(params, _u64p), key))


def hash_arrays(family, d, b, seed, m, p):
    """Packed bucket|signbit arrays (d x m for A's columns, d x p for B's rows)."""
    hA = np.empty((d, m), np.uint32)
    hB = np.empty((d, p), np.uint32)
    assert lib().orc_hash_arrays(family, d, b, seed, m, p,
                                 _p(hA, _u32p), _p(hB, _u32p)) == 0
    return hA, hB


def fft_r2c(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(len(x) // 2 + 1, np.complex128)
    assert lib().orc_fft_r2c(len(x), _p(x, _f64p), _p(out.view(np.float64), _f64p)) == 0
    return out


def fft_c2r(X, n):
    X = np.ascontiguousarray(X, np.complex128)
    out = np.empty(n, np.float64)
    assert lib().orc_fft_c2r(n, _p(X.view(np.float64), _f64p), _p(out, _f64p)) == 0
    return out


def fwht(x):
    x = np.array(x, np.float64)
    assert lib().orc_fwht(len(x), _p(x, _f64p)) == 0
    return x


def compress(variant, AT, B, d, b, seed, family=0, threads=0):
    """Sketch (d x b, FP64) of AT^T B; variant is 'fft' or 'fwht'."""
    AT = np.ascontiguousarray(AT, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    n, m = AT.shape
    p = B.shape[1]
    out = np.empty((d, b), np.float64)
    rc = lib().orc_compress(int(variant == "fwht"), family, m, n, p,
                            _p(AT, _f64p), _p(B, _f64p), d, b, seed, threads,
                            _p(out, _f64p))
    assert rc == 0
    return out


def decompress(variant, sketch, hA, hB, threads=0):
    d, b = sketch.shape
    m, p = hA.shape[1], hB.shape[1]
    sketch = np.ascontiguousarray(sketch, np.float64)
    hA = np.ascontiguousarray(hA, np.uint32)
    hB = np.ascontiguousarray(hB, np.uint32)
    C = np.empty((m, p), np.float64)
    lib().orc_decompress(int(variant == "fwht"), m, p, d, b, _p(hA, _u32p),
                         _p(hB, _u32p), _p(sketch, _f64p), threads, _p(C, _f64p))
    return C


def median(xs):
    xs = np.array(xs, np.float64)
    return float(lib().orc_median(len(xs), _p(xs, _f64p)))


def sketch_from_product(variant, C, hA, hB, b):
    C = np.ascontiguousarray(C, np.float64)
    d = hA.shape[0]
    m, p = C.shape
    out = np.empty((d, b), np.float64)
    lib().orc_sketch_from_product(int(variant == "fwht"), m, p, d, b,
                                  _p(np.ascontiguousarray(hA), _u32p),
                                  _p(np.ascontiguousarray(hB), _u32p),
                                  _p(C, _f64p), _p(out, _f64p))
    return out


def gemm(AT, B, threads=0):
    AT = np.ascontiguousarray(AT, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    n, m = AT.shape
    p = B.shape[1]
    C = np.empty((m, p), np.float64)
    lib().orc_gemm(m, n, p, _p(AT, _f64p), _p(B, _f64p), _p(C, _f64p), threads)
    return C


def matmul(variant, AT, B, d, b, seed, family=0, threads=0):
    """Full compressed product C = median-of-d decompression of the sketch."""
    n, m = AT.shape
    p = B.shape[1]
    sk = compress(variant, AT, B, d, b, seed, family, threads)
    hA, hB = hash_arrays(family, d, b, seed, m, p)
    return decompress(variant, sk, hA, hB, threads)
