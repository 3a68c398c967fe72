"""The calibration kernels (cm_probe_alu, cm_probe_stream): the work they
report follows from their arguments, the copy and fill leave the expected
bytes, and buffers that are not device allocations of the stated size are
rejected on the host before any launch."""
import ctypes

import numpy as np
import pytest

import compmatmul_amd as cma

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,per_op", [("fp32_add", 1), ("fp32_fma", 2), ("fp64_add", 1),
                                         ("fp64_fma", 2)])
def test_alu_work(kind, per_op):
    ms, flops = cma.probe_alu(kind, blocks=512, iters=64, reps=2)
    assert ms > 0
    assert flops == 512 * 256 * 64 * 16 * per_op


def test_stream_copy_read_fill():
    n = (3 << 20) + 4  # floats; bytes a multiple of 16
    src_h = np.arange(n, dtype=np.float32)
    src = cma.DeviceArray.from_numpy(src_h)
    dst = cma.DeviceArray((n,), np.float32)
    ms, moved = cma.probe_stream("copy", dst, src, n * 4, reps=2)
    assert ms > 0 and moved == 2 * n * 4
    assert np.array_equal(dst.numpy(), src_h)
    ms, moved = cma.probe_stream("read", None, src, n * 4, reps=1)
    assert moved == n * 4
    ms, moved = cma.probe_stream("fill", dst, None, n * 4, reps=1)
    assert moved == n * 4
    assert np.all(dst.numpy() == 0.5)
    assert np.array_equal(src.numpy(), src_h)
    src.free()
    dst.free()


def test_stream_rejects_bad_buffers():
    a = cma.DeviceArray((1024,), np.float32)
    b = cma.DeviceArray((1024,), np.float32)
    host = np.zeros(1024, np.float32)
    L = cma.lib()
    ms, moved = ctypes.c_double(), ctypes.c_double()

    def call(kind, dst, src, nbytes):
        return L.cm_probe_stream(kind, dst, src, nbytes, 1, ctypes.byref(ms), ctypes.byref(moved))

    assert call(0, b.ptr, a.ptr, 4096) == cma.OK
    assert call(0, b.ptr, a.ptr, 4096 + 16) == cma.EINVAL       # beyond the allocation
    assert call(1, None, a.ptr, 4100) == cma.EINVAL             # not a multiple of 16
    assert call(2, host.ctypes.data, None, 4096) == cma.EINVAL  # host memory
    assert call(0, a.ptr, a.ptr, 4096) == cma.EINVAL            # overlapping copy
    assert call(2, None, None, 4096) == cma.EINVAL
    assert L.cm_probe_alu(9, 1, 1, 1, ctypes.byref(ms), ctypes.byref(moved)) == cma.EINVAL
    # work bounds: refused before any launch
    assert L.cm_probe_alu(0, 1 << 14, 1 << 13, 1, ctypes.byref(ms), ctypes.byref(moved)) == cma.EINVAL
    assert L.cm_probe_alu(0, 1, 1, 101, ctypes.byref(ms), ctypes.byref(moved)) == cma.EINVAL
    assert L.cm_probe_stream(0, b.ptr, a.ptr, 4096, 101, ctypes.byref(ms),
                             ctypes.byref(moved)) == cma.EINVAL
    a.free()
    b.free()
