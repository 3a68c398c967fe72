"""Parity of the HIP path with the CPU oracle, through the C-ABI.

Bars: hash arrays and recovered sets bit-exact; FP64 sketches within 1e-10
relative (different summation order only) and exact on integer FWHT inputs;
FP32 sketches within 1e-5 relative L2 error of the oracle's FP64 sketch;
decompression of a given sketch bit-exact (selection of gathered values).
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FAMS = [None, "multiply shift", "tabulation8", "tabulation16"]


def rel_l2(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


# ---------------------------------------------------------------- hashing
@pytest.mark.parametrize("hash", FAMS)
@pytest.mark.parametrize("d,b,seed,m,p", [(3, 2, 5, 7, 9), (9, 4096, 13794, 512, 512),
                                          (15, 1 << 14, 13794, 2048, 2048),
                                          (5, 1 << 30, 99, 300, 70000)])
def test_hash_arrays_bit_exact(hash, d, b, seed, m, p):
    print(f"hash={hash} d={d} b={b} m={m} p={p}", flush=True)
    fam = orc.FAMILIES[hash]
    hA, hB = cma.hash_arrays(hash, d, b, seed, m, p)
    oA, oB = orc.hash_arrays(fam, d, b, seed, m, p)
    assert np.array_equal(hA, oA)
    assert np.array_equal(hB, oB)


# ---------------------------------------------------------------- sketches
@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("hash", FAMS)
@pytest.mark.parametrize("m,n,p,d,b", [(37, 19, 23, 3, 64), (64, 40, 50, 5, 1 << 13),
                                       (100, 33, 80, 2, 1 << 15), (5, 0, 4, 3, 32),
                                       (1, 300, 1, 1, 2)])
def test_sketch_f64_matches_oracle(variant, hash, m, n, p, d, b):
    print(f"{variant} {hash} m={m} n={n} p={p} d={d} b={b}", flush=True)
    rng = np.random.default_rng(m * 7 + n)
    AT = rng.standard_normal((n, m))
    B = rng.standard_normal((n, p))
    sk = cma.compress(variant, AT, B, d, b, seed=4242, hash=hash)
    got = sk.values()
    hA, hB = orc.hash_arrays(orc.FAMILIES[hash], d, b, 4242, m, p)
    assert np.array_equal(sk.hashes()[0], hA) and np.array_equal(sk.hashes()[1], hB)
    want = orc.sketch_from_product(variant, AT.T @ B, hA, hB, b)
    scale = max(np.abs(want).max(), 1.0)
    assert np.abs(got - want).max() <= 1e-10 * scale
    # decompression of this sketch is a selection: bit-exact vs the oracle's
    C = sk.decompress()
    assert np.array_equal(C, orc.decompress(variant, got, hA, hB))
    sk.free()


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_sketch_f64_exact_on_integers(variant):
    rng = np.random.default_rng(11)
    AT = rng.integers(-9, 10, (40, 30)).astype(float)
    B = rng.integers(-9, 10, (40, 20)).astype(float)
    sk = cma.compress(variant, AT, B, 4, 1 << 14, seed=7)
    hA, hB = sk.hashes()
    want = orc.sketch_from_product(variant, AT.T @ B, hA, hB, 1 << 14)
    got = sk.values()
    if variant == "fwht":
        assert np.array_equal(got, want)
    else:
        assert np.abs(got - np.round(got)).max() < 1e-9
        assert np.array_equal(np.round(got), want)


@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("b", [1 << 10, 1 << 14, 1 << 16])
def test_sketch_f32_relative_error(variant, b):
    m, n, p, d = 256, 512, 192, 5
    AT, B, _ = workloads.planted(m, n, p, 16, 5.0, 77, np.float32)
    sk = cma.compress(variant, AT, B, d, b, seed=13794, dtype=np.float32)
    hA, hB = sk.hashes()
    want = orc.sketch_from_product(variant, AT.astype(np.float64).T @ B.astype(np.float64),
                                   hA, hB, b)
    err = rel_l2(sk.values(), want)
    print(f"{variant} b={b} fp32 sketch rel err {err:.3e}")
    assert err <= 1e-5
    sk.free()


@pytest.mark.parametrize("variant", ["fft", "fwht"])
def test_plain_path_matches_fused(variant):
    AT, B, _ = workloads.planted(200, 300, 150, 8, 5.0, 5, np.float32)
    fused = cma.compress(variant, AT, B, 5, 4096, seed=1, dtype=np.float32)
    plain = cma.compress(variant, AT, B, 5, 4096, seed=1, dtype=np.float32,
                         flags=cma.FLAG_PLAIN)
    v1, v2 = fused.values(), plain.values()
    assert rel_l2(v2, v1) <= 2e-6


# ---------------------------------------------------------------- decompress
@pytest.mark.parametrize("d", list(range(1, 21)))
def test_median_rank_any_d(d):
    # array.hpp:323-328: nth_element at d/2 (upper median for even d)
    rng = np.random.default_rng(d)
    AT = rng.standard_normal((20, 33))
    B = rng.standard_normal((20, 29))
    for variant in ("fft", "fwht"):
        sk = cma.compress(variant, AT, B, d, 64, seed=d)
        vals = sk.values()
        hA, hB = sk.hashes()
        assert np.array_equal(sk.decompress(), orc.decompress(variant, vals, hA, hB))


@pytest.mark.parametrize("d", list(range(1, 17)))
@pytest.mark.parametrize("ties", [False, True])
def test_median_rank_f32(d, ties):
    # FP32 entries from d >= 10 on keep only the d - d/2 largest values seen
    # so far; small integer inputs make many equal values (and zeros of both
    # signs), which the selection must still place at rank d/2
    print(f"d={d} ties={ties}", flush=True)
    rng = np.random.default_rng(300 + d)
    if ties:
        AT = rng.integers(-2, 3, (6, 45)).astype(np.float32)
        B = rng.integers(-2, 3, (6, 38)).astype(np.float32)
    else:
        AT = rng.standard_normal((20, 45)).astype(np.float32)
        B = rng.standard_normal((20, 38)).astype(np.float32)
    for variant in ("fft", "fwht"):
        sk = cma.compress(variant, AT, B, d, 128, seed=d, dtype=np.float32)
        vals = sk.values()
        hA, hB = sk.hashes()
        want = orc.decompress(variant, vals.astype(np.float64), hA, hB)
        assert np.array_equal(sk.decompress().astype(np.float64), want), variant
        sk.free()


@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("d", [1, 2, 7, 15, 16])
def test_decompress_f32_64k_rows(variant, d):
    # FP32 sketch rows of 64 KB (b = 2^14) stream through the half-row LDS
    # ring; the median is a selection, so the result is bit-exact
    print(f"{variant} d={d} b=16384 fp32", flush=True)
    rng = np.random.default_rng(100 + d)
    AT = rng.standard_normal((24, 700)).astype(np.float32)
    B = rng.standard_normal((24, 530)).astype(np.float32)
    sk = cma.compress(variant, AT, B, d, 1 << 14, seed=31 + d, dtype=np.float32)
    vals = sk.values()
    hA, hB = sk.hashes()
    want = orc.decompress(variant, vals, hA, hB)
    assert np.array_equal(sk.decompress().astype(np.float64), want)
    sk.free()


@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("dtype,b", [(np.float64, 1 << 12), (np.float32, 1 << 13)])
@pytest.mark.parametrize("d", [1, 2, 3, 4, 5, 9, 13, 16])
@pytest.mark.parametrize("p", [530, 1300])
def test_decompress_32k_rows(variant, dtype, b, d, p):
    # 32 KB sketch rows (FP64 b = 2^12, FP32 b = 2^13): one ring slot per
    # row, up to four rows in flight, so the wait before each row counts the
    # copies of 0 .. 3 later rows; 512 and 1024 threads (p >= 1024; FP32
    # d = 13 keeps 512), ragged last tile and row block
    print(f"{variant} d={d} b={b} {np.dtype(dtype).name} p={p}", flush=True)
    rng = np.random.default_rng(300 + d)
    AT = rng.standard_normal((20, 45)).astype(dtype)
    B = rng.standard_normal((20, p)).astype(dtype)
    sk = cma.compress(variant, AT, B, d, b, seed=11 + d, dtype=dtype)
    vals = sk.values()
    hA, hB = sk.hashes()
    want = orc.decompress(variant, vals, hA, hB)
    assert np.array_equal(sk.decompress().astype(np.float64), want)
    sk.free()


@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("dtype,b,d", [(np.float32, 1 << 14, 1), (np.float32, 1 << 14, 10),
                                       (np.float32, 1 << 14, 15), (np.float32, 1 << 14, 16),
                                       (np.float64, 1 << 13, 9), (np.float64, 1 << 13, 16)])
def test_decompress_64k_rows_wide(variant, dtype, b, d):
    # 64 KB sketch rows with at least 1024 columns: the 1024-thread ring
    # kernel (d = 10 in FP32 keeps 512 threads); 1300 columns leave a ragged
    # last tile and 37 rows a ragged last row block
    print(f"{variant} d={d} b={b} {np.dtype(dtype).name} p=1300", flush=True)
    rng = np.random.default_rng(200 + d)
    AT = rng.standard_normal((16, 37)).astype(dtype)
    B = rng.standard_normal((16, 1300)).astype(dtype)
    sk = cma.compress(variant, AT, B, d, b, seed=7 + d, dtype=dtype)
    vals = sk.values()
    hA, hB = sk.hashes()
    want = orc.decompress(variant, vals, hA, hB)
    assert np.array_equal(sk.decompress().astype(np.float64), want)
    sk.free()


def test_decompress_into_device_buffer():
    rng = np.random.default_rng(1)
    AT = rng.standard_normal((30, 40)).astype(np.float32)
    B = rng.standard_normal((30, 50)).astype(np.float32)
    dAT, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    sk = cma.compress("fwht", dAT, dB, 5, 256, seed=3)
    hsk = cma.compress("fwht", AT, B, 5, 256, seed=3)
    out = cma.DeviceArray((40, 50), np.float32)
    sk.decompress(out)
    assert np.array_equal(out.numpy(), hsk.decompress())


# ---------------------------------------------------------------- end to end
@pytest.mark.parametrize("name", ["fft512_fp64", "fwht2048_fp32"])
def test_recovered_set_planted(name):
    c = workloads.CONFIGS[name]
    AT, B, plants = workloads.make(name)
    C_true = orc.gemm(AT.astype(np.float64), B.astype(np.float64))
    truth = set(zip(*np.nonzero(C_true > c["tau"])))
    sk = cma.compress(c["variant"], AT, B, c["d"], c["b"], seed=13794)
    ii, jj, vv = sk.recover(c["tau"])
    got = set(zip(ii.tolist(), jj.tolist()))
    assert got == truth
    assert len(truth) >= c["plants"]
    # the oracle's own median estimate from the GPU sketch gives the same set
    hA, hB = sk.hashes()
    C_orc = orc.decompress(c["variant"], sk.values(), hA, hB)
    assert set(zip(*np.nonzero(C_orc > c["tau"]))) == got
    # and the sketch agrees with the oracle's FP64 sketch
    want = orc.sketch_from_product(c["variant"], C_true, hA, hB, c["b"])
    err = rel_l2(sk.values(), want)
    print(f"{name}: |set|={len(got)} sketch rel err {err:.3e}")
    assert err <= (1e-10 if c["dtype"] == np.float64 else 1e-5)


def test_repeatable_and_padded_leading_dimension():
    rng = np.random.default_rng(2)
    AT = rng.integers(-5, 6, (64, 48)).astype(np.float64)
    B = rng.integers(-5, 6, (64, 40)).astype(np.float64)
    a = cma.compress("fwht", AT, B, 7, 512, seed=9).values()
    b = cma.compress("fwht", AT, B, 7, 512, seed=9).values()
    assert np.array_equal(a, b)
    ATp = np.zeros((64, 64))
    ATp[:, :48] = AT
    dp = cma.DeviceArray.from_numpy(ATp)
    dp.shape = (64, 64)
    # padded leading dimension through the C-ABI directly
    sk = cma.Sketch(cma.FWHT, np.dtype(np.float64), 7, 512, 48, 40)
    dB = cma.DeviceArray.from_numpy(B)
    import ctypes
    cma.check(cma.lib().cm_compress_ex(cma.FWHT, cma.F64, 48, 64, 40, dp.ptr, 64, dB.ptr, 40,
                                       7, 512, 9, 0, cma.DEVICE, 0, ctypes.byref(sk.handle)))
    assert np.array_equal(sk.values(), a)


def test_errors_raise():
    AT = np.ones((4, 4))
    with pytest.raises(ValueError, match="b must be power of 2"):
        cma.compress("fft", AT, AT, 3, 48, seed=1)
    with pytest.raises(ValueError, match="d must be positive"):
        cma.compress("fft", AT, AT, 0, 32, seed=1)


@pytest.mark.parametrize("m,n,p,b", [(2048, 96, 2048, 1 << 14), (100, 70, 37, 1 << 12),
                                     (1, 5, 3, 1 << 13), (300, 33, 0, 1 << 12),
                                     (60000, 3, 50, 1 << 12), (5000, 20, 4100, 1 << 15),
                                     (300, 40, 200, 1 << 10), (64, 9, 70, 1 << 10),
                                     (700, 50, 65536, 1 << 16), (4096, 30, 4096, 1 << 13),
                                     (2048, 96, 0, 1 << 14), (0, 7, 40, 1 << 14)])
def test_fwht_register_kernel_matches_lds_kernel(m, n, p, b):
    print(f"reg vs lds m={m} n={n} p={p} b={b}", flush=True)
    AT, B, _ = workloads.planted(m, n, max(p, 1), 4, 5.0, m + n, np.float32)
    B = np.ascontiguousarray(B[:, :p])
    d = 7
    reg = cma.compress("fwht", AT, B, d, b, seed=3, dtype=np.float32)
    lds = cma.compress("fwht", AT, B, d, b, seed=3, dtype=np.float32,
                       flags=cma.FLAG_LDS_KERNEL)
    again = cma.compress("fwht", AT, B, d, b, seed=3, dtype=np.float32)
    single = cma.compress("fwht", AT, B, d, b, seed=3, dtype=np.float32,
                          flags=cma.FLAG_SINGLE_WAVE)
    v1, v2, v3 = reg.values(), lds.values(), single.values()
    assert np.array_equal(v1, again.values())  # scheduled scatter: deterministic
    if p == 0 or m == 0:
        assert not v1.any() and not v2.any() and not v3.any()
        return
    assert rel_l2(v1, v2) <= 2e-6
    assert rel_l2(v3, v2) <= 2e-6  # slice-group and single-wave kernels
    hA, hB = reg.hashes()
    want = orc.sketch_from_product("fwht", AT.astype(np.float64).T @ B.astype(np.float64),
                                   hA, hB, b)
    assert rel_l2(v1, want) <= 1e-5


@pytest.mark.parametrize("m,n,p,b", [(2048, 40, 2048, 1 << 14), (100, 70, 37, 1 << 11),
                                     (1, 5, 3, 1 << 13), (300, 33, 0, 1 << 12),
                                     (60000, 3, 50, 1 << 11), (5000, 20, 4100, 1 << 16),
                                     (300, 40, 200, 1 << 10), (64, 9, 70, 1 << 10),
                                     (700, 50, 65000, 1 << 16), (4096, 30, 4096, 1 << 12),
                                     (2048, 96, 0, 1 << 14), (0, 7, 40, 1 << 14)])
def test_fft_register_kernel_matches_lds_kernel(m, n, p, b):
    print(f"fft reg vs lds m={m} n={n} p={p} b={b}", flush=True)
    AT, B, _ = workloads.planted(m, n, max(p, 1), 4, 5.0, m + n + 1, np.float32)
    B = np.ascontiguousarray(B[:, :p])
    d = 5
    reg = cma.compress("fft", AT, B, d, b, seed=11, dtype=np.float32)
    lds = cma.compress("fft", AT, B, d, b, seed=11, dtype=np.float32,
                       flags=cma.FLAG_LDS_KERNEL)
    again = cma.compress("fft", AT, B, d, b, seed=11, dtype=np.float32)
    single = cma.compress("fft", AT, B, d, b, seed=11, dtype=np.float32,
                          flags=cma.FLAG_SINGLE_WAVE)
    v1, v2, v3 = reg.values(), lds.values(), single.values()
    assert np.array_equal(v1, again.values())
    hA, hB = reg.hashes()
    want = orc.sketch_from_product("fft", AT.astype(np.float64).T @ B.astype(np.float64),
                                   hA, hB, b)
    if p == 0 or m == 0:
        # the pair packing extracts A*B by cancellation: with one operand
        # empty the residue is rounding of the other's |.|^2 (~1e-7
        # relative), not exactly zero
        scale = max(float((AT.astype(np.float64) ** 2).sum(axis=1).max(initial=0.0)),
                    float((B.astype(np.float64) ** 2).sum(axis=1).max(initial=0.0)), 1e-30)
        assert np.abs(v1).max() < 1e-6 * scale and np.abs(v2).max() < 1e-6 * scale
        assert np.abs(v3).max() < 1e-6 * scale
        return
    assert rel_l2(v1, want) <= 1e-5
    assert rel_l2(v2, want) <= 1e-5
    assert rel_l2(v3, want) <= 1e-5  # single-wave kernel


def test_timing_only_one_class():
    # events bracket only the named kernel class; None restores every class
    AT = np.random.default_rng(5).standard_normal((64, 96)).astype(np.float32)
    B = np.random.default_rng(6).standard_normal((64, 80)).astype(np.float32)
    try:
        for only, want in (("compress", {"compress"}), (None, {"compress", "hash", "decompress"})):
            cma.timing_enable(True, only=only)
            cma.timing_reset()
            sk = cma.compress("fwht", AT, B, 5, 1 << 14, seed=2, dtype=np.float32)
            sk.decompress()
            timed = {k for k in ("compress", "hash", "reduce", "combine", "decompress")
                     if cma.timing_get(k)[1] > 0}
            assert want <= timed and ("compress" in timed)
            if only:
                assert timed == {only}
            sk.free()
    finally:
        cma.timing_enable(False)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_decompress_host_ldc_keeps_padding(dtype):
    # a caller decompressing into a sub-block of a wider host matrix: only
    # the m x p window is written, columns p .. ldc-1 keep their contents
    m, p, pad = 37, 29, 11
    AT, B, _ = workloads.planted(m, 40, p, 4, 4.0, 17, dtype)
    sk = cma.compress("fft", AT, B, 5, 1 << 12, seed=3, dtype=dtype)
    dense = sk.decompress()
    wide = np.full((m, p + pad), -7.25, dtype)
    cma.check(cma.lib().cm_decompress(sk.handle, wide.ctypes.data, p + pad, cma.HOST))
    assert np.array_equal(wide[:, :p], dense)
    assert np.all(wide[:, p:] == dtype(-7.25))
    sk.free()


# ---------------------------------------------------------------- plain path
@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("dtype,b", [(np.float32, 2), (np.float32, 64), (np.float32, 4096),
                                     (np.float32, 1 << 14), (np.float32, 1 << 15),
                                     (np.float32, 1 << 16), (np.float64, 1024),
                                     (np.float64, 1 << 13), (np.float64, 1 << 15)])
def test_plain_path_matches_oracle(variant, dtype, b):
    # the unfused stages (single-pass rows up to 64 KB, column + row passes
    # beyond) against the oracle's closed form with the same hashes
    m, n, p, d = 230, 150, 170, 3
    print(f"plain {variant} {np.dtype(dtype).name} b={b}", flush=True)
    AT, B, _ = workloads.planted(m, n, p, 8, 5.0, 31 + b % 97, dtype)
    sk = cma.compress(variant, AT, B, d, b, seed=606, dtype=dtype, flags=cma.FLAG_PLAIN)
    hA, hB = sk.hashes()
    want = orc.sketch_from_product(variant, AT.astype(np.float64).T @ B.astype(np.float64), hA, hB, b)
    got = sk.values()
    if dtype == np.float64:
        assert np.abs(got - want).max() <= 1e-10 * max(np.abs(want).max(), 1.0)
    else:
        err = rel_l2(got, want)
        print(f"  rel err {err:.2e}")
        assert err <= 1e-5
    C = sk.decompress()
    assert np.array_equal(C.astype(np.float64), orc.decompress(variant, got, hA, hB))
    sk.free()


def test_plain_path_empty_and_multi_pass():
    # n = 0 gives a zero sketch; rows of 2^15 reals (FWHT) split in passes
    AT = np.zeros((0, 40), np.float32)
    B = np.zeros((0, 30), np.float32)
    sk = cma.compress("fft", AT, B, 3, 1 << 15, seed=1, dtype=np.float32, flags=cma.FLAG_PLAIN)
    assert not sk.values().any()
    sk.free()


# ------------------------------------------------- general kernel, every b
@pytest.mark.parametrize("variant", ["fft", "fwht"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("b", [2, 4, 8, 16, 128, 2048, 4096, 1 << 13, 1 << 14])
def test_general_kernel_every_last_pass(variant, dtype, b):
    # the general kernel runs its last transform pass in registers, one
    # instantiation per radix 2^((log2 b2 - 1) mod 4 + 1); with one slice the
    # inverse writes the sketch itself.  These b reach every radix, one and
    # several slices, in both precisions, against the oracle's closed form
    m, n, p, d = 90, 61, 75, 5
    print(f"general {variant} {np.dtype(dtype).name} b={b}", flush=True)
    AT, B, _ = workloads.planted(m, n, p, 4, 5.0, 7 + b % 89, dtype)
    sk = cma.compress(variant, AT, B, d, b, seed=2024, dtype=dtype, flags=cma.FLAG_LDS_KERNEL)
    hA, hB = sk.hashes()
    want = orc.sketch_from_product(variant, AT.astype(np.float64).T @ B.astype(np.float64), hA, hB, b)
    got = sk.values()
    if dtype == np.float64:
        assert np.abs(got - want).max() <= 1e-10 * max(np.abs(want).max(), 1.0)
    else:
        assert rel_l2(got, want) <= 1e-5
    assert np.array_equal(sk.decompress().astype(np.float64), orc.decompress(variant, got.astype(np.float64), hA, hB))
    sk.free()


# --------------------------------------------------- host transfers in chunks
@pytest.mark.parametrize("dtype,m,n,p", [(np.float32, 1500, 777, 900),   # 4.7 MB, 2 MiB chunks
                                         (np.float64, 700, 450, 530),    # 2.5 MB, 1 MiB chunks
                                         (np.float32, 300, 800, 200)])   # below 1 MiB: one copy
def test_host_operands_and_C_equal_device_path(dtype, m, n, p):
    # host operands go up, and a host C comes back, through two pinned slots
    # in chunks that do not divide the sizes; every value must arrive as on
    # the device-resident path
    AT, B, _ = workloads.planted(m, n, p, 4, 5.0, m + p, dtype)
    host = cma.compress("fft", AT, B, 5, 1 << 12, seed=99, dtype=dtype)
    dev = cma.compress("fft", cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B), 5,
                       1 << 12, seed=99, dtype=dtype)
    assert np.array_equal(host.values(), dev.values())
    C_host = host.decompress()
    dC = cma.DeviceArray((m, p), dtype)
    dev.decompress(dC)
    assert np.array_equal(C_host, dC.numpy())
    host.free()
    dev.free()
