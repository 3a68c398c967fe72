"""The arithmetic modes of the FP32 FFT path against the oracle.

Modes: FP32 butterflies (the fused register kernels), FP64 accumulation of
the per-chunk spectra (CM_FLAG_F64_REDUCE), and matrix-core radix-16 DFT
stages with hi/lo split correction on BF16 (CM_FLAG_BF16_MFMA) or FP16
(CM_FLAG_FP16_MFMA) operands.  Each sketch is compared with the oracle's
FP64 sketch computed from the dense FP64 product with the same hashes:
relative L2 error <= 1e-5 in every mode (the BF16 mode's two halves carry
~16 significant bits, the dropped lo*lo term is ~2^-16 relative per stage;
FP16's carry ~22, with the input columns scaled by powers of two into its
range).  The recovered planted set must equal the dense product's {C > tau}
in every mode.
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MODES = {"fp32": 0, "fp64_reduce": cma.FLAG_F64_REDUCE, "bf16_mfma": cma.FLAG_BF16_MFMA,
         "fp16_mfma": cma.FLAG_FP16_MFMA}
BARS = {"fp32": 1e-5, "fp64_reduce": 1e-5, "bf16_mfma": 1e-5, "fp16_mfma": 1e-5}


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


@pytest.mark.parametrize("flag", [cma.FLAG_BF16_MFMA, cma.FLAG_FP16_MFMA], ids=["bf16", "fp16"])
def test_matrix_core_mode_errors(flag):
    AT, B, _ = workloads.planted(64, 32, 48, 4, 5.0, 1, np.float32)
    with pytest.raises(cma.CMError):  # b not a power of 16
        cma.compress("fft", AT, B, 3, 1 << 13, seed=1, dtype=np.float32, flags=flag)
    with pytest.raises(cma.CMError):  # FWHT has no matrix-core mode
        cma.compress("fwht", AT, B, 3, 1 << 12, seed=1, dtype=np.float32, flags=flag)
    with pytest.raises(ValueError):  # one operand type at a time (CM_EINVAL)
        cma.compress("fft", AT, B, 3, 1 << 12, seed=1, dtype=np.float32,
                     flags=cma.FLAG_BF16_MFMA | cma.FLAG_FP16_MFMA)
    empty = cma.compress("fft", AT[:0], B[:0], 3, 1 << 12, seed=1, dtype=np.float32, flags=flag)
    assert not empty.values().any()


@pytest.mark.parametrize("scale", [1e-12, 1e-4, 1e4, 1e12])
def test_fp16_mode_scaled_operands(scale):
    # FP16 holds magnitudes between about 6e-8 and 65504: the per-column
    # power-of-two scaling must keep operands of any magnitude exact to the
    # split's ~22 bits (both operands scaled, so the product scales by scale^2)
    m, n, p, d, b = 200, 96, 150, 5, 1 << 12
    AT, B, _ = workloads.planted(m, n, p, 8, 5.0, 77, np.float32)
    AT = (AT * np.float32(scale)).astype(np.float32)
    B = (B * np.float32(scale)).astype(np.float32)
    C64 = AT.astype(np.float64).T @ B.astype(np.float64)
    hA, hB = orc.hash_arrays(0, d, b, 13794, m, p)
    want = orc.sketch_from_product("fft", C64, hA, hB, b)
    sk = cma.compress("fft", AT, B, d, b, seed=13794, dtype=np.float32, flags=cma.FLAG_FP16_MFMA)
    got = sk.values()
    sk.free()
    rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    print(f"fp16 mode, operands x {scale:g}: rel L2 {rel:.3e}")
    assert rel <= 1e-5
