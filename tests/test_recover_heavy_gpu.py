"""Recovery when most estimates pass tau (tau = 0: about half of all m p
entries), through every staged decompression form: two LDS buffers (4 KB
rows), the ring with 32 KB rows and with 64 KB rows, at 512 and 1024
threads (p >= 1024), FP32 and FP64, FFT and FWHT.  The staged kernels
claim a wave's recovered slots with one atomic, so these cases check that
the list is exactly {(i, j): C[i, j] > tau} of the same sketch's C (values
bit-equal), that the reported count is the full count when the list is
truncated, and that a truncated list holds distinct passing entries."""
import ctypes

import numpy as np
import pytest

import compmatmul_amd as cma

pytestmark = pytest.mark.gpu

# (dtype, b) per staged form: 4 KB rows, 32 KB rows, 64 KB rows
FORMS = [(np.float32, 1 << 10), (np.float32, 1 << 13), (np.float32, 1 << 14),
         (np.float64, 1 << 9), (np.float64, 1 << 12), (np.float64, 1 << 13)]


def operands(m, n, p, dtype, seed):
    rng = np.random.default_rng(seed)
    AT = rng.standard_normal((n, m)).astype(dtype)
    B = rng.standard_normal((n, p)).astype(dtype)
    return AT, B


@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("dtype,b", FORMS)
@pytest.mark.parametrize("d,p", [(5, 1100), (9, 700)])
def test_recover_most_entries(variant, dtype, b, d, p):
    m, n = 900, 48
    print(f"case {variant} {np.dtype(dtype).name} b={b} d={d} p={p}", flush=True)
    AT, B = operands(m, n, p, dtype, 1000 + d + b)
    dAT, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    sk = cma.compress(variant, dAT, dB, d, b, seed=31, dtype=dtype)
    dC = cma.DeviceArray((m, p), dtype)
    sk.decompress(dC)
    C = dC.numpy()
    want_i, want_j = np.nonzero(C > 0.0)
    assert len(want_i) > m * p // 4
    ii, jj, vv = sk.recover(0.0, max_out=m * p)
    assert np.array_equal(ii, want_i) and np.array_equal(jj, want_j)
    assert np.array_equal(vv, C[want_i, want_j].astype(np.float64))
    # truncated: the full count comes back, the stored entries pass and are distinct
    cap = 5000
    oi = np.empty(cap, np.int32)
    oj = np.empty(cap, np.int32)
    ov = np.empty(cap, np.float64)
    cnt = ctypes.c_int64()
    L = cma.lib()
    cma.check(L.cm_recover(sk.handle, 0.0, cap, oi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           oj.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                           ov.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(cnt)))
    assert cnt.value == len(want_i)
    keys = oi.astype(np.int64) * p + oj
    assert len(np.unique(keys)) == cap
    assert np.all(C[oi, oj] > 0.0)
    assert np.array_equal(ov, C[oi, oj].astype(np.float64))
    sk.free()
    for x in (dAT, dB, dC):
        x.free()


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_batch_recover_most_entries(variant):
    # the batch path's staged forms (FP32): every entry above tau = 0 is
    # listed, and each product's C holds the same values
    shapes = [(900, 48, 1100, 5, 1 << 10), (700, 40, 1300, 9, 1 << 13), (800, 56, 1024, 7, 1 << 14)]
    prods, Cs = [], []
    for k, (m, n, p, d, b) in enumerate(shapes):
        AT, B = operands(m, n, p, np.float32, 2000 + k)
        C = cma.DeviceArray((m, p), np.float32)
        prods.append(cma.Product(variant, cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B),
                                 d, b, 50 + k, C=C, tau=0.0, max_out=m * p))
        Cs.append(C)
    bt = cma.Batch(prods, grain=64)
    bt.run()
    cma.synchronize()
    for k, C in enumerate(Cs):
        print(f"product {k}: {shapes[k]}", flush=True)
        Ch = C.numpy()
        wi, wj = np.nonzero(Ch > 0.0)
        ii, jj, vv = bt.recovered(k)
        assert np.array_equal(ii, wi) and np.array_equal(jj, wj)
        assert np.array_equal(vv, Ch[wi, wj])
    bt.free()
