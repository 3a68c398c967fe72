"""The three arithmetic modes of the FP32 FFT path against the oracle.

Modes: FP32 butterflies (the fused register kernels), FP64 accumulation of
the per-chunk spectra (CM_FLAG_F64_REDUCE), and BF16 matrix-core radix-16
DFT stages with hi/lo split correction (CM_FLAG_BF16_MFMA).  Each sketch is
compared with the oracle's FP64 sketch computed from the dense FP64 product
with the same hashes: relative L2 error <= 1e-5 in every mode (the BF16
mode's two halves carry ~16 significant bits; the dropped lo*lo term is
~2^-16 relative per stage).  The recovered planted set must equal the dense
product's {C > tau} in every mode.
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MODES = {"fp32": 0, "fp64_reduce": cma.FLAG_F64_REDUCE, "bf16_mfma": cma.FLAG_BF16_MFMA}
BARS = {"fp32": 1e-5, "fp64_reduce": 1e-5, "bf16_mfma": 1e-5}


@pytest.mark.parametrize("b", [1 << 12, 1 << 16])
def test_modes_against_oracle(b):
    m, n, p, d = 384, 1024, 320, 7
    AT, B, _ = workloads.planted(m, n, p, 24, 5.0, 2024, np.float32)
    C64 = AT.astype(np.float64).T @ B.astype(np.float64)
    truth = set(zip(*np.nonzero(C64 > 12.5)))
    hA, hB = orc.hash_arrays(0, d, b, 13794, m, p)
    want = orc.sketch_from_product("fft", C64, hA, hB, b)
    for name, flags in MODES.items():
        sk = cma.compress("fft", AT, B, d, b, seed=13794, dtype=np.float32, flags=flags)
        got = sk.values()
        err = got - want
        rel = float(np.linalg.norm(err) / np.linalg.norm(want))
        print(f"b={b} {name}: rel L2 {rel:.3e} max abs {np.abs(err).max():.3e} "
              f"rms {np.sqrt(np.mean(err ** 2)):.3e}")
        assert rel <= BARS[name], name
        ii, jj, _ = sk.recover(12.5)
        assert set(zip(ii.tolist(), jj.tolist())) == truth, name
        sk.free()


def test_bf16_mode_errors():
    AT, B, _ = workloads.planted(64, 32, 48, 4, 5.0, 1, np.float32)
    with pytest.raises(cma.CMError):  # b not a power of 16
        cma.compress("fft", AT, B, 3, 1 << 13, seed=1, dtype=np.float32, flags=cma.FLAG_BF16_MFMA)
    with pytest.raises(cma.CMError):  # FWHT has no BF16 mode
        cma.compress("fwht", AT, B, 3, 1 << 12, seed=1, dtype=np.float32, flags=cma.FLAG_BF16_MFMA)
    empty = cma.compress("fft", AT[:0], B[:0], 3, 1 << 12, seed=1, dtype=np.float32,
                         flags=cma.FLAG_BF16_MFMA)
    assert not empty.values().any()
