"""Pins the CPU oracle before it is trusted to check results.

Each test cites the reference test or the mathematical fact it uses.
Known-answer values are the literals of the reference's tests.cpp/test.py.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from tests import fixtures as fx

MS, MAS, T8, T16 = 1, 0, 2, 3
M64 = (1 << 64) - 1
REF_RNG = os.path.join(os.path.dirname(orc.__file__), "_ref", "ref_rng")


# ---------------------------------------------------------------- rng
def test_rng_known_answer():
    # tests.cpp:90-96
    assert [int(x) for x in orc.rng_stream(1234, 4)] == [
        0x0bab45d9a0e3ae53, 0xd7c640660c19433e, 0xb0dedaa0d09a6691,
        0xdec9f41b58ec86eb]


@pytest.mark.skipif(not os.path.exists(REF_RNG), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed", [0, 1, 1234, 13794, 25926, 2**63 + 5])
def test_rng_matches_reference_rng_hpp(seed):
    out = subprocess.run([REF_RNG, str(seed), "300"], capture_output=True,
                         text=True, check=True).stdout.split()
    ref_u = [int(x, 16) for x in out[:300]]
    ref_d = [float(x) for x in out[300:]]
    assert [int(x) for x in orc.rng_stream(seed, 300)] == ref_u
    assert orc.drng_stream(seed, 300).tolist() == ref_d


def test_drng_grid_and_range():
    x = orc.drng_stream(99, 10000)
    assert (x >= 0).all() and (x < 1).all()
    assert np.all((x * 2.0**53) == np.floor(x * 2.0**53))  # 53-bit grid
    assert abs(x.mean() - 0.5) < 0.02


# ---------------------------------------------------------------- hashing
def test_multiply_shift_known_answers():
    # tests.cpp:146-223
    a = 2749059519329956733
    assert int(orc.draw(MS, 1, 64, 123456)[0, 0]) == a  # first (even) draw rejected
    assert int(orc.rng_stream(123456, 1)[0]) % 2 == 0
    prm = np.array([a], np.uint64)
    counts = np.zeros(64, int)
    for i in range(129):
        h = ((a * i) & M64) >> 58
        assert orc.hash_eval(MS, prm, 64, i) == h
        counts[h] += 1
        assert orc.sign_eval(MS, prm, i) == (1 if ((a * i) & M64) >> 63 else -1)
    assert (counts > 0).all()
    a1, a2 = 3688180003657674949, 11729574476469721337
    p2 = orc.draw(MS, 2, 32, 19850)
    assert [int(p2[0, 0]), int(p2[1, 0])] == [a1, a2]
    for i in range(1025):
        assert orc.hash_eval(MS, p2[0], 32, i) == ((a1 * i) & M64) >> 59
        assert orc.hash_eval(MS, p2[1], 32, i) == ((a2 * i) & M64) >> 59
    # key 0 always maps to bucket 0 with sign -1 (hashing.hpp:91-97)
    assert orc.hash_eval(MS, p2[0], 32, 0) == 0 and orc.sign_eval(MS, p2[0], 0) == -1


def test_multiply_add_shift_known_answers():
    # tests.cpp:228-308
    a, b = 2366053268901514180, 2749059519329956733
    assert [int(v) for v in orc.draw(MAS, 1, 128, 123456)[0]] == [a, b]
    prm = np.array([a, b], np.uint64)
    for i in range(513):
        v = (a * i + b) & M64
        assert orc.hash_eval(MAS, prm, 128, i) == v >> 57
        assert orc.sign_eval(MAS, prm, i) == (1 if v >> 63 else -1)
    p2 = orc.draw(MAS, 2, 128, 22201)
    assert [int(v) for v in p2.ravel()] == [
        2016826348251915964, 17696970887482165528,
        6666994550483350943, 1097363203091539945]
    # tests.cpp:796-807: draw_one(32, 8876), then 100 keys from rng(32149)
    a, b = 15506110424104283675, 3138147583905456810
    prm = orc.draw(MAS, 1, 32, 8876)[0]
    assert [int(v) for v in prm] == [a, b]
    for x in orc.rng_stream(32149, 100):
        x = int(x) & 0xffffffff
        h = ((a * x + b) & M64) >> 59
        assert orc.hash_eval(MAS, prm, 32, x) == h
        assert orc.sign_eval(MAS, prm, x) == (1 if h >> 4 else -1)


def test_sketch_parameters_known_answer():
    # tests.cpp:818-840: sketch(d=2, b=2, seed=3483) -> h1,h2,s1,s2 from seed+0..3
    want = {
        0: [1312110179027324775, 15465668779316680995,
            4555306672462241870, 12288159137770073612],
        1: [4329212686603688483, 10307990599335447505,
            12597734872139453857, 15119207240207437602],
        2: [11899074044546628492, 298345645907162146,
            8628211646415106536, 12647867801867038929],
        3: [7008367753992221008, 16948937881964684143,
            10319872211339334476, 13096518497071785733],
    }
    for off, vals in want.items():
        assert [int(v) for v in orc.draw(MAS, 2, 2, 3483 + off).ravel()] == vals


def _tab_expect(T, x, bits, shift):
    if bits == 8:
        h = T[0][x & 0xff] ^ T[1][(x >> 8) & 0xff] ^ T[2][(x >> 16) & 0xff] ^ T[3][(x >> 24) & 0xff]
    else:
        h = T[0][x & 0xffff] ^ T[1][(x >> 16) & 0xffff]
    return int(h) >> shift, (1 if int(h) >> 31 else -1)


@pytest.mark.parametrize("fam,bits,seed,n", [(T8, 8, 23711, 1), (T16, 16, 23712, 1),
                                             (T8, 8, 16321, 2), (T16, 16, 16322, 2)])
def test_tabulation_known_answers(fam, bits, seed, n):
    # tests.cpp:312-461: tables are consecutive rng words (low 32 bits),
    # row-major; the keys continue the same stream.
    m = 256 if n == 1 else 128
    words = (4 * 256) if bits == 8 else (2 * 65536)
    stream = orc.rng_stream(seed, n * words + 2048)
    prm = orc.draw(fam, n, m, seed)
    shift = 32 - int(np.log2(m))
    for f in range(n):
        T = (stream[f * words:(f + 1) * words] & 0xffffffff).astype(np.uint32)
        assert np.array_equal(prm[f].astype(np.uint32), T)
        T = T.reshape((4, 256) if bits == 8 else (2, 65536))
        for x in stream[n * words:n * words + 512]:
            x = int(x) & 0xffffffff
            h, s = _tab_expect(T, x, bits, shift)
            assert orc.hash_eval(fam, prm[f], m, x) == h
            assert orc.sign_eval(fam, prm[f], x) == s


@pytest.mark.parametrize("fam", [MAS, MS, T8, T16])
def test_hash_arrays_match_per_key_evaluation(fam):
    d, b, seed, m, p = 3, 64, 777, 40, 23
    hA, hB = orc.hash_arrays(fam, d, b, seed, m, p)
    h1, h2 = orc.draw(fam, d, b, seed), orc.draw(fam, d, b, seed + 1)
    s1, s2 = orc.draw(fam, d, b, seed + 2), orc.draw(fam, d, b, seed + 3)
    for t in range(d):
        for i in range(m):
            e = int(hA[t, i])
            assert e & 0x7fffffff == orc.hash_eval(fam, h1[t], b, i)
            assert (-1 if e >> 31 else 1) == orc.sign_eval(fam, s1[t], i)
        for j in range(p):
            e = int(hB[t, j])
            assert e & 0x7fffffff == orc.hash_eval(fam, h2[t], b, j)
            assert (-1 if e >> 31 else 1) == orc.sign_eval(fam, s2[t], j)


# ---------------------------------------------------------------- transforms
def test_fft_known_answer():
    # tests.cpp:576-615
    X = orc.fft_r2c(fx.FFT_X)
    assert np.abs(X - fx.FFT_X_CORRECT).max() < 1e-6
    assert np.abs(orc.fft_c2r(X, 32) - fx.FFT_X).max() < 1e-6


@pytest.mark.parametrize("n", [1, 2, 4, 8, 64, 256])
def test_fft_brute_force_dft(n):
    x = np.random.default_rng(n).standard_normal(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * np.outer(k, k) / n)
    want = (F @ x)[: n // 2 + 1]
    got = orc.fft_r2c(x)
    assert np.abs(got - want).max() < 1e-12 * max(1, n)
    assert np.abs(orc.fft_c2r(got, n) - x).max() < 1e-13 * max(1, n)


def test_fwht_known_answer():
    # tests.cpp:621-671
    X = orc.fwht(fx.FWHT_X)
    assert np.abs(X - fx.FWHT_X_CORRECT).max() < 1e-6
    assert np.abs(orc.fwht(X) / 32 - fx.FWHT_X).max() < 1e-6


@pytest.mark.parametrize("n", [1, 2, 8, 64, 128])
def test_fwht_brute_force_hadamard(n):
    f = np.arange(n)
    H = np.array([[(-1) ** bin(a & c).count("1") for c in f] for a in f], float)
    x = np.random.default_rng(n).standard_normal(n)
    assert np.abs(orc.fwht(x) - H @ x).max() < 1e-12 * max(1, n)
    assert np.array_equal(H @ H, n * np.eye(n))
    xi = np.random.default_rng(n).integers(-50, 50, n).astype(float)
    assert np.array_equal(orc.fwht(orc.fwht(xi)), n * xi)  # exact on integers


# ---------------------------------------------------------------- compress
@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("fam", [MAS, MS, T8, T16])
def test_compress_matches_closed_form(variant, fam):
    rng = np.random.default_rng(5 + fam)
    m, n, p, d, b, seed = 13, 9, 11, 3, 32, 4242
    AT = rng.standard_normal((n, m))
    B = rng.standard_normal((n, p))
    sk = orc.compress(variant, AT, B, d, b, seed, fam, threads=3)
    hA, hB = orc.hash_arrays(fam, d, b, seed, m, p)
    C = AT.T @ B
    want = orc.sketch_from_product(variant, C, hA, hB, b)
    # brute-force closed form, independent of the oracle's C code
    brute = np.zeros((d, b))
    for t in range(d):
        for i in range(m):
            for j in range(p):
                ea, eb = int(hA[t, i]), int(hB[t, j])
                ba, bb = ea & 0x7fffffff, eb & 0x7fffffff
                h = (ba ^ bb) if variant == "fwht" else (ba + bb) % b
                s = (-1 if ea >> 31 else 1) * (-1 if eb >> 31 else 1)
                brute[t, h] += s * C[i, j]
    assert np.abs(want - brute).max() < 1e-12
    assert np.abs(sk - want).max() < 1e-12


def test_compress_fwht_exact_on_integers():
    rng = np.random.default_rng(3)
    AT = rng.integers(-9, 10, (7, 10)).astype(float)
    B = rng.integers(-9, 10, (7, 6)).astype(float)
    hA, hB = orc.hash_arrays(0, 4, 16, 5, 10, 6)
    want = orc.sketch_from_product("fwht", AT.T @ B, hA, hB, 16)
    assert np.array_equal(orc.compress("fwht", AT, B, 4, 16, 5), want)


# ---------------------------------------------------------------- median
def test_median_semantics():
    # array.hpp:315-328: element d/2 of the sorted list (upper median if even)
    rng = np.random.default_rng(0)
    for d in range(1, 20):
        xs = rng.standard_normal(d)
        assert orc.median(xs) == np.sort(xs)[d // 2]
    assert orc.median([1.0, 2.0]) == 2.0
    assert orc.median([4.0, 1.0, 3.0, 2.0]) == 3.0


def test_decompress_definition():
    rng = np.random.default_rng(9)
    d, b, m, p = 5, 16, 6, 7
    sk = rng.standard_normal((d, b))
    hA, hB = orc.hash_arrays(2, d, b, 31, m, p)
    for variant in ("fft", "fwht"):
        C = orc.decompress(variant, sk, hA, hB)
        for i in range(m):
            for j in range(p):
                vals = []
                for t in range(d):
                    ea, eb = int(hA[t, i]), int(hB[t, j])
                    ba, bb = ea & 0x7fffffff, eb & 0x7fffffff
                    h = (ba ^ bb) if variant == "fwht" else (ba + bb) % b
                    s = (-1 if ea >> 31 else 1) * (-1 if eb >> 31 else 1)
                    vals.append(s * sk[t, h])
                assert C[i, j] == sorted(vals)[d // 2]


# ---------------------------------------------------------------- end to end
def test_gemm_exact():
    # tests.cpp:722-750
    for AT, B, C in fx.SMALL_PRODUCTS:
        assert np.array_equal(orc.gemm(AT, B), C)


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_small_products_reference_cases(variant):
    # tests.cpp:841-937 (FFT, 1e-6) and :1051-1145 (FWHT, exact), seed 25926
    last = (5, 32) if variant == "fft" else (7, 32)
    for d, b in [(1, 2), (1, 8), (1, 32), (3, 2), (3, 8), last]:
        if not (d > 1 and b >= 32):
            continue
        for AT, B, Cw in fx.SMALL_PRODUCTS:
            C = orc.matmul(variant, AT, B, d, b, 25926)
            if variant == "fft":
                assert np.abs(C - Cw).max() <= 1e-6
            else:
                assert np.array_equal(C, Cw)


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_16x16_drng_products(variant):
    # tests.cpp:939-986 and :1147-1194: drng(15954) inputs, 4 hash families
    AT, B = fx.drng_16x16()
    C1 = orc.gemm(AT, B)
    cases = fx.CASES_16x16[variant]
    for fam, d, b, seed in cases:
        C = orc.matmul(variant, AT, B, d, b, seed, fam)
        assert np.abs(C - C1).max() <= 1e-12


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_python_boundary_cases(variant):
    # test.py:51-88 (seed 13794) and :122-161 (16x16 integers)
    for AT, B, Cw in fx.PY_PRODUCTS:
        for h in fx.HASHES:
            d, b = fx.TEST_PY_PARAMS[(variant, h)]
            C = orc.matmul(variant, AT, B, d, b, 13794, orc.FAMILIES[h])
            if variant == "fft":
                assert np.allclose(C, Cw)
            else:
                assert np.array_equal(C, Cw)
    AT, B, Cw = fx.int_16x16()
    for h in fx.HASHES:
        d, b = fx.TEST_PY_PARAMS_16[(variant, h)]
        C = orc.matmul(variant, AT, B, d, b, 13794, orc.FAMILIES[h])
        if variant == "fft":
            assert np.allclose(C, Cw)
        else:
            assert np.array_equal(C, Cw)


def test_default_seed_is_first_draw_of_time_seeded_rng():
    # matmul_fft.hpp:30-32 with rng.hpp:52-57
    assert orc.default_seed(1234) == 0x0bab45d9a0e3ae53
