"""The reference's Python boundary test (test.py:51-161), run against the
drop-in module on the GPU, plus the C++ end-to-end cases of tests.cpp."""
import numpy as np
import pytest

import compmatmul_amd as cma
from tests import fixtures as fx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def compmatmul():
    return cma.dropin()


def test_multiply(compmatmul):
    # test.py:51-88
    seed = 13794
    for AT, B, C_correct in fx.PY_PRODUCTS:
        for algo in ("gemm", "fft", "fwht"):
            for hash_fun in fx.HASHES:
                if algo == "gemm" and hash_fun is not None:
                    continue
                C = fx.aligned_empty(C_correct.shape)
                if algo == "gemm":
                    compmatmul.gemm(C, AT, B)
                    assert np.array_equal(C, C_correct)
                elif algo == "fft":
                    d, b = fx.TEST_PY_PARAMS[(algo, hash_fun)]
                    compmatmul.matmul_fft(C, AT, B, d, b, seed, hash_fun)
                    assert np.allclose(C, C_correct)
                else:
                    d, b = fx.TEST_PY_PARAMS[(algo, hash_fun)]
                    compmatmul.matmul_fwht(C, AT, B, d, b, seed, hash_fun)
                    assert np.array_equal(C, C_correct)


def test_seed_semantics(compmatmul):
    # test.py:90-119
    AT, B = fx.AT_4x2, fx.B_4x8_PY
    C1, C2, C3 = (fx.aligned_empty((2, 8)) for _ in range(3))
    d, b = 3, 8
    s1, s2 = 27548, 23340
    f = compmatmul.matmul_fft
    f(C1, AT, B, d, b, s1)
    f(C2, AT, B, d, b, s2)
    f(C3, AT, B, d, b, s1)
    assert np.allclose(C1, C3)
    assert not np.allclose(C1, C2)
    f(C2, AT, B, d, b, s1)
    assert np.allclose(C1, C2)
    f(C3, AT, B, d, b, s1, "multiply shift")
    f(C2, AT, B, d, b, s1, "multiply add shift")
    assert not np.allclose(C1, C3)
    assert np.allclose(C1, C2)
    g = compmatmul.matmul_fwht
    g(C1, AT, B, d, b, s1)
    g(C2, AT, B, d, b, s2)
    g(C3, AT, B, d, b, s1)
    assert np.array_equal(C1, C3)
    assert not np.allclose(C1, C2)
    g(C2, AT, B, d, b, s1)
    assert np.array_equal(C1, C2)
    g(C3, AT, B, d, b, s1, "multiply shift")
    g(C2, AT, B, d, b, s1, "multiply add shift")
    assert not np.allclose(C1, C3)
    assert np.array_equal(C1, C2)


def test_integer_16x16(compmatmul):
    # test.py:122-161
    AT, B, C_correct = fx.int_16x16()
    for algo in ("gemm", "fft", "fwht"):
        for hash_fun in fx.HASHES:
            if algo == "gemm" and hash_fun is not None:
                continue
            C = fx.aligned_empty(C_correct.shape)
            if algo == "gemm":
                compmatmul.gemm(C, AT, B)
                assert np.array_equal(C, C_correct)
            elif algo == "fft":
                d, b = fx.TEST_PY_PARAMS_16[(algo, hash_fun)]
                compmatmul.matmul_fft(C, AT, B, d, b, 13794, hash_fun)
                assert np.allclose(C, C_correct)
            else:
                d, b = fx.TEST_PY_PARAMS_16[(algo, hash_fun)]
                compmatmul.matmul_fwht(C, AT, B, d, b, 13794, hash_fun)
                assert np.array_equal(C, C_correct)


def test_default_seed_is_random(compmatmul):
    AT, B = fx.AT_4x2, fx.B_4x8_PY
    outs = []
    for _ in range(3):
        C = fx.aligned_empty((2, 8))
        compmatmul.matmul_fwht(C, AT, B, 3, 8)
        outs.append(C.copy())
    assert all(np.isfinite(o).all() for o in outs)


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_small_products_cpp_cases(compmatmul, variant):
    # tests.cpp:841-937 (FFT, 1e-6) and :1051-1145 (FWHT, exact), seed 25926
    fn = compmatmul.matmul_fft if variant == "fft" else compmatmul.matmul_fwht
    last = (5, 32) if variant == "fft" else (7, 32)
    for d, b in [(3, 8), last]:
        for AT, B, Cw in fx.SMALL_PRODUCTS:
            C = fx.aligned_empty(Cw.shape)
            fn(C, AT, B, d, b, 25926)
            if d > 1 and b >= 32:
                if variant == "fft":
                    assert np.abs(C - Cw).max() <= 1e-6
                else:
                    assert np.array_equal(C, Cw)


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_16x16_drng_cpp_cases(compmatmul, variant):
    # tests.cpp:939-986, :1147-1194: within 1e-12 of gemm for all 4 families
    names = {0: "multiply add shift", 1: "multiply shift", 2: "tabulation8",
             3: "tabulation16"}
    AT, B = fx.drng_16x16()
    C1 = fx.aligned_empty((16, 16))
    compmatmul.gemm(C1, AT, B)
    fn = compmatmul.matmul_fft if variant == "fft" else compmatmul.matmul_fwht
    for fam, d, b, seed in fx.CASES_16x16[variant]:
        C = fx.aligned_empty((16, 16))
        fn(C, AT, B, d, b, seed, names[fam])
        assert np.abs(C - C1).max() <= 1e-12
