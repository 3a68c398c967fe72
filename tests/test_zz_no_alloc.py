"""cm_matmul_noalloc: the compressed product with every intermediate in the
fixed device scratch, called through the C-ABI.

Its output must equal that of the allocating entry points on the same
operands and seed: cm_compress + cm_decompress on device operands (the
computation cm_matmul_fft / cm_matmul_fwht run), and the float64 drop-in
cm_matmul_fft itself.  The plans match (the same chunking of the inner
dimension fits the scratch here).  float32 runs the scheduled, atomic-free
scatter, so the bar is bit-for-bit; float64 runs the LDS kernel, whose
scatter adds with LDS atomics in an order that may change from run to run,
so its bar is 1e-12 of the largest |C| (the drop-in's own bar).  Two calls in
a row, the second on a smaller product, check that the shared scratch keeps
no state that leaks from one call into the next.
"""
import ctypes

import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads

pytestmark = pytest.mark.gpu

CASES = [  # variant, dtype, m, n, p, d, b
    ("fwht", np.float32, 1500, 700, 1200, 7, 1 << 14),
    ("fft", np.float32, 1500, 700, 1200, 5, 1 << 14),
    ("fft", np.float32, 3000, 96, 2500, 3, 1 << 16),
    ("fwht", np.float32, 300, 100, 250, 5, 4096),
    ("fft", np.float64, 200, 150, 180, 9, 1024),
    ("fwht", np.float64, 120, 64, 90, 5, 512),
]


def noalloc(variant, AT, B, d, b, seed):
    n, m = AT.shape
    p = B.shape[1]
    dt = AT.dtype
    dA, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    dC = cma.DeviceArray((m, p), dt)
    rc = cma.lib().cm_matmul_noalloc(cma.VARIANTS[variant], cma.F64 if dt == np.float64 else cma.F32,
                                     m, n, p, dA.ptr, ctypes.c_int64(m), dB.ptr, ctypes.c_int64(p),
                                     dC.ptr, ctypes.c_int64(p), d, b, ctypes.c_uint64(seed), 0)
    cma.check(rc)
    return dC.numpy()


def allocating(variant, AT, B, d, b, seed):
    dA, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    sk = cma.compress(variant, dA, dB, d, b, seed=seed, dtype=AT.dtype)
    dC = cma.DeviceArray((AT.shape[1], B.shape[1]), AT.dtype)
    sk.decompress(out=dC)
    sk.free()
    return dC.numpy()


def same(got, want):
    if got.dtype == np.float32:
        assert np.array_equal(got, want)
    else:
        diff = float(np.abs(got - want).max()) if got.size else 0.0
        print(f"  float64 max |diff| {diff:.3e}")
        assert diff <= 1e-12 * max(float(np.abs(want).max()), 1.0)


def operands(m, n, p, dtype, seed):
    AT, B, _ = workloads.planted(m, n, p, 16, 5.0, seed, dtype)
    return AT, B


@pytest.mark.parametrize("first,second", [(0, 3), (1, 4), (2, 5), (4, 0)])
def test_noalloc_matches_allocating_entry_twice(first, second):
    for idx, seed in ((first, 11), (second, 12)):  # the second product is smaller
        variant, dt, m, n, p, d, b = CASES[idx]
        print(f"noalloc case {idx}: {variant} {np.dtype(dt).name} {m}x{n}x{p} d={d} b={b}", flush=True)
        AT, B = operands(m, n, p, dt, seed)
        got = noalloc(variant, AT, B, d, b, seed)
        want = allocating(variant, AT, B, d, b, seed)
        assert got.dtype == dt and got.shape == (m, p)
        same(got, want)


def test_noalloc_matches_float64_dropin():
    m, n, p, d, b, seed = 200, 150, 180, 9, 1024, 77
    AT, B = operands(m, n, p, np.float64, 3)
    got = noalloc("fft", AT, B, d, b, seed)
    C = np.empty((m, p))
    cma.check(cma.lib().cm_matmul_fft(m, n, p, C.ctypes.data, AT.ctypes.data, B.ctypes.data, d, b,
                                      ctypes.c_uint64(seed), 1, 0))
    same(got, C)


def test_noalloc_rejects_a_product_beyond_the_scratch():
    # a 16 x 2^20 FP32 sketch alone fills the 64 MiB scratch: refused before
    # any launch, with the size in the message
    AT = cma.DeviceArray((4, 8), np.float32)
    B = cma.DeviceArray((4, 8), np.float32)
    C = cma.DeviceArray((8, 8), np.float32)
    rc = cma.lib().cm_matmul_noalloc(cma.FFT, cma.F32, 8, 4, 8, AT.ptr, ctypes.c_int64(8), B.ptr,
                                     ctypes.c_int64(8), C.ptr, ctypes.c_int64(8), 16, 1 << 20,
                                     ctypes.c_uint64(1), 0)
    assert rc == cma.ENOMEM
    assert "MiB" in cma.lib().cm_last_error().decode()
