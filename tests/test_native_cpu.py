"""CPU-side checks of the native library: it loads, exports every symbol the
C-ABI header declares, and fails loudly (no CPU fallback) without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import compmatmul_amd as cma

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "compmatmul_amd.h")


def declared_symbols():
    with open(HEADER) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for s in ("cm_matmul_fft", "cm_matmul_fwht", "cm_gemm", "cm_hash_arrays",
              "cm_compress", "cm_decompress", "cm_recover"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", cma.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    L = cma.lib()
    for s in declared_symbols():
        assert hasattr(L, s)


def test_dropin_module_imports():
    mod = cma.dropin()
    for name in ("gemm", "matmul_fft", "matmul_fwht"):
        assert callable(getattr(mod, name))


def _have_gpu():
    try:
        return ctypes.CDLL(cma.LIB_PATH).cm_init(0) == 0
    except OSError:
        return False


@pytest.mark.skipif(_have_gpu(), reason="checks the no-GPU behaviour")
def test_no_silent_cpu_fallback():
    AT = np.ones((4, 4))
    with pytest.raises(cma.CMError):
        cma.matmul("fwht", AT, AT, 3, 32, seed=1)
    mod = cma.dropin()
    from tests.fixtures import aligned_empty
    C = aligned_empty((4, 4))
    A = aligned_empty((4, 4))
    A[...] = 1
    with pytest.raises(RuntimeError):
        mod.matmul_fft(C, A, A, 3, 32, 5)
    with pytest.raises(cma.CMError):  # a context needs the device too
        cma.Context(0)


def test_context_entry_points_reject_null_handles():
    # argument checks that run before any device work
    L = cma.lib()
    out = ctypes.c_void_p()
    assert L.cm_context_create(0, None, 0, None) == cma.EINVAL
    assert L.cm_context_destroy(None) == cma.OK
    assert L.cm_ctx_synchronize(None) == cma.EINVAL
    assert L.cm_ctx_decompress(None, None, None, 0, cma.HOST) == cma.EINVAL
    cnt = ctypes.c_int64()
    assert L.cm_ctx_recover(None, None, 1.0, 0, None, None, None, ctypes.byref(cnt)) == cma.EINVAL
    assert L.cm_ctx_compress(None, 0, 1, 4, 4, 4, None, 4, None, 4, 3, 32, 1, 0, cma.HOST, 0,
                             ctypes.byref(out)) == cma.EINVAL


def test_dropin_validation_messages_without_gpu():
    # compmatmul.cpp:13-105: checks run before any device work
    from tests.fixtures import aligned_empty
    mod = cma.dropin()
    C = aligned_empty((4, 4))
    A = aligned_empty((4, 4))
    with pytest.raises(ValueError, match="only supports double precision"):
        mod.matmul_fft(C, A.astype(np.float32), A, 3, 32)
    with pytest.raises(ValueError, match=r"C must have ndim=2 \(got 1\)"):
        mod.matmul_fft(aligned_empty((16,)), A, A, 3, 32)
    with pytest.raises(ValueError, match="C-contiguous"):
        mod.matmul_fft(C, np.asfortranarray(np.ones((4, 3))), A, 3, 32)
    with pytest.raises(ValueError, match="Shape mismatch! Note that AT is assumed to be transposed"):
        mod.matmul_fft(C, aligned_empty((3, 4)), A, 3, 32)
    raw = np.empty(16 * 8 + 64, np.uint8)
    off = (-raw.ctypes.data) % 64 + 8
    mis = raw[off:off + 128].view(np.float64).reshape(4, 4)
    with pytest.raises(ValueError, match="AT is not aligned at a 64-byte boundary!"):
        mod.matmul_fft(C, mis, A, 3, 32)
    with pytest.raises(ValueError, match=r"b must be power of 2 greater than 1 \(got 12\)"):
        mod.matmul_fft(C, A, A, 3, 12)
    with pytest.raises(ValueError, match=r"b must be power of 2 greater than 1 \(got 1\)"):
        mod.matmul_fwht(C, A, A, 3, 1)
    with pytest.raises(ValueError, match=r"d must be positive \(got 0\)"):
        mod.matmul_fwht(C, A, A, 0, 32)
    with pytest.raises(ValueError, match="Invalid hash `foo' provided"):
        mod.matmul_fwht(C, A, A, 3, 32, 1, "foo")
    with pytest.raises(TypeError):
        mod.matmul_fwht(C, A, A, 3, 32, -1)


@pytest.mark.parametrize("bad", [
    np.empty((5, 4)),                        # wrong shape
    np.empty((4, 5), np.float32),            # wrong dtype for a float64 sketch
    np.empty((5, 4)).T,                      # right shape, not C-contiguous
])
def test_sketch_decompress_rejects_bad_out(bad):
    # the checks run before any library call, so no kernel is launched
    sk = cma.Sketch(cma.FFT, np.float64, 3, 64, 4, 5)
    with pytest.raises(ValueError):
        sk.decompress(out=bad)


def test_sketch_decompress_rejects_readonly_out():
    sk = cma.Sketch(cma.FWHT, np.float32, 3, 64, 4, 5)
    out = np.empty((4, 5), np.float32)
    out.flags.writeable = False
    with pytest.raises(ValueError):
        sk.decompress(out=out)
