"""The pair-packed FFT keeps its accuracy when |A| and |B| differ.

The FFT kernels transform z_k = S a_k + i S' b_k and extract the product by
cancelling |A_k|^2 + |B_k|^2 (see the balance kernel), so without a balance
the FP32 error grows with the ratio |a_k| / |b_k|.  Each row pair is scaled by
(2^e_k, 2^-e_k) first.  These tests put ratios of 1e2 .. 1e4 between the
operands, for the whole product and varying per inner index k, through every
FFT path, and hold the sketch to the same bars as balanced data: within
1e-5 relative L2 of the oracle's FP64 closed form in FP32 (reference:
matmul_fft.hpp:95-105 transforms the two sketches separately), within 1e-10
in FP64.  An empty operand gives an exactly zero sketch.
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def rel_l2(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def imbalanced(m, n, p, ratio, mode, seed, dtype):
    """Planted operands with |a_k| / |b_k| = ratio ("global") or
    ratio^u, u uniform in [-1, 1] per inner index k ("per_k")."""
    AT, B, _ = workloads.planted(m, n, p, 8, 5.0, seed, np.float64)
    rng = np.random.default_rng(seed + 1)
    if mode == "global":
        f = np.full(n, float(ratio))
    else:
        f = float(ratio) ** rng.uniform(-1.0, 1.0, n)
    AT = AT * np.sqrt(f)[:, None]
    B = B / np.sqrt(f)[:, None]
    return AT.astype(dtype), B.astype(dtype)


def sketch_err(variant, AT, B, d, b, flags=0, dtype=np.float32, seed=13794):
    sk = cma.compress(variant, AT, B, d, b, seed=seed, dtype=dtype, flags=flags)
    hA, hB = sk.hashes()
    C = AT.astype(np.float64).T @ B.astype(np.float64)
    want = orc.sketch_from_product(variant, C, hA, hB, b)
    got = sk.values()
    sk.free()
    return rel_l2(got, want), got, want


PATHS = [
    ("slice-group", 1 << 14, 0),
    ("single-wave", 1 << 14, cma.FLAG_SINGLE_WAVE),
    ("register b=2^10", 1 << 10, 0),
    ("lds kernel", 1 << 12, cma.FLAG_LDS_KERNEL),
    ("plain", 1 << 12, cma.FLAG_PLAIN),
]


@pytest.mark.parametrize("mode", ["global", "per_k"])
@pytest.mark.parametrize("ratio", [1e2, 1e3, 1e4])
@pytest.mark.parametrize("path,b,flags", PATHS, ids=[p[0] for p in PATHS])
def test_fft_f32_imbalanced(path, b, flags, ratio, mode):
    print(f"fft fp32 {path} b={b} ratio={ratio:g} {mode}", flush=True)
    AT, B = imbalanced(300, 256, 200, ratio, mode, 31, np.float32)
    err, _, _ = sketch_err("fft", AT, B, 5, b, flags)
    print(f"  rel err {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("mode", ["global", "per_k"])
def test_fft_f32_imbalanced_large_rows(mode):
    # rows longer than one 2048-key staging segment, m not a multiple of it
    # (mixed A/B segments), b = 2^16
    AT, B = imbalanced(3000, 96, 2500, 1e4, mode, 41, np.float32)
    err, _, _ = sketch_err("fft", AT, B, 3, 1 << 16)
    print(f"fft fp32 3000x96x2500 b=2^16 {mode}: rel err {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("flag", [cma.FLAG_BF16_MFMA, cma.FLAG_FP16_MFMA], ids=["bf16", "fp16"])
def test_fft_matrix_core_modes_imbalanced(flag):
    AT, B = imbalanced(256, 128, 192, 1e3, "per_k", 51, np.float32)
    err, _, _ = sketch_err("fft", AT, B, 3, 1 << 12, flag)
    print(f"fft matrix-core mode {flag} ratio 1e3 per_k: rel err {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("mode", ["global", "per_k"])
def test_fft_f64_imbalanced(mode):
    AT, B = imbalanced(120, 90, 70, 1e4, mode, 61, np.float64)
    err, got, want = sketch_err("fft", AT, B, 5, 4096, dtype=np.float64)
    print(f"fft fp64 ratio 1e4 {mode}: rel err {err:.2e}")
    assert np.abs(got - want).max() <= 1e-10 * max(np.abs(want).max(), 1.0)


def test_fft_extreme_magnitudes_f64():
    # products of order 1 from operands of order 1e+150 and 1e-150: the
    # balance exponent reaches about +-500 and the squares of the rows would
    # overflow a plain double sum
    rng = np.random.default_rng(5)
    AT = rng.standard_normal((40, 30)) * 1e150
    B = rng.standard_normal((40, 20)) * 1e-150
    err, got, want = sketch_err("fft", AT, B, 3, 1024, dtype=np.float64)
    assert np.all(np.isfinite(got))
    assert np.abs(got - want).max() <= 1e-10 * max(np.abs(want).max(), 1.0)


@pytest.mark.parametrize("flags", [0, cma.FLAG_SINGLE_WAVE, cma.FLAG_LDS_KERNEL])
def test_fft_zero_operand_gives_zero_sketch(flags):
    # an all-zero B: every row pair has |b_k| = 0, so nothing is accumulated
    rng = np.random.default_rng(3)
    AT = rng.standard_normal((64, 500)).astype(np.float32)
    B = np.zeros((64, 300), np.float32)
    sk = cma.compress("fft", AT, B, 3, 1 << 14, seed=9, dtype=np.float32, flags=flags)
    assert np.array_equal(sk.values(), np.zeros((3, 1 << 14)))
    sk.free()


def test_fft_some_zero_rows():
    # rows k with a zero a_k or b_k add nothing; the rest are summed as usual
    AT, B = imbalanced(200, 64, 150, 10.0, "per_k", 71, np.float32)
    AT[::3] = 0.0
    B[1::5] = 0.0
    err, _, _ = sketch_err("fft", AT, B, 5, 1 << 13)
    assert err <= 1e-5


def test_batch_fft_imbalanced():
    specs = [(300, 256, 200, 1 << 14, 1e3, "global"), (700, 100, 900, 1 << 11, 1e4, "per_k"),
             (2500, 64, 1800, 1 << 16, 1e2, "per_k"), (100, 50, 80, 1 << 12, 1e4, "global")]
    prods, refs = [], []
    for i, (m, n, p, b, ratio, mode) in enumerate(specs):
        AT, B = imbalanced(m, n, p, ratio, mode, 80 + i, np.float32)
        dA, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
        prods.append(cma.Product("fft", dA, dB, 5, b, seed=100 + i))
        refs.append((AT, B, b, dA, dB))
    bt = cma.Batch(prods, grain=64)
    bt.run()
    cma.synchronize()
    for i, (AT, B, b, _, _) in enumerate(refs):
        hA, hB = orc.hash_arrays(0, 5, b, 100 + i, AT.shape[1], B.shape[1])
        want = orc.sketch_from_product("fft", AT.astype(np.float64).T @ B.astype(np.float64),
                                       hA, hB, b)
        err = rel_l2(bt.values(i), want)
        print(f"batch product {i} b={b}: rel err {err:.2e}")
        assert err <= 1e-5
    bt.free()


@pytest.mark.parametrize("dtype,b", [(np.float32, 1 << 14), (np.float64, 1 << 12)])
def test_fft_nan_row_spreads_as_without_balance(dtype, b):
    # a NaN in a_k keeps the row pair's scales at 1: the transform spreads it
    # over every frequency, so every sketch entry is NaN, as when the two
    # sketches are transformed separately (matmul_fft.hpp:95-105)
    AT, B = imbalanced(120, 48, 90, 1e3, "per_k", 17, dtype)
    AT[7, 11] = np.nan
    sk = cma.compress("fft", AT, B, 3, b, seed=9, dtype=dtype)
    got = sk.values()
    sk.free()
    assert np.isnan(got).all()
