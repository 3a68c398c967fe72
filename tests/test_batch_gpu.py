"""Batches of many FP32 products through the C-ABI (cm_batch_*).

Bars: every product's sketch within 1e-5 relative L2 error of the oracle's
FP64 sketch computed with the same hashes; its decompression bit-exact
against the oracle's decompression of the same sketch values; a product's
result bit-identical alone and inside a mixed batch (same grain); HIP graph
replays bit-identical to the eager run; recovered sets equal to the
oracle's thresholded estimate.
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

# variant, m, n, p, d, b, hash
SHAPES = [
    ("fwht", 300, 200, 150, 5, 1 << 10, None),
    ("fft", 300, 200, 150, 5, 1 << 10, None),
    ("fwht", 64, 129, 4096, 7, 1 << 11, "multiply shift"),
    ("fft", 1000, 77, 33, 9, 1 << 12, "tabulation8"),
    ("fwht", 2048, 300, 2048, 15, 1 << 14, None),
    ("fft", 512, 500, 700, 11, 1 << 16, None),
    ("fwht", 100, 1, 90, 3, 1 << 13, "tabulation16"),
    ("fft", 5, 0, 7, 2, 1 << 11, None),      # empty inner dimension
    ("fwht", 1, 3, 1, 1, 2, None),           # b = 2
    ("fft", 3, 5, 4, 16, 4, None),           # tiny b, d = 16
    ("fft", 90, 40, 0, 3, 1 << 12, None),    # empty C
    ("fwht", 3000, 150, 500, 9, 1 << 13, None),   # rows past one segment: staged path, G = 4
    ("fwht", 700, 64, 2500, 5, 1 << 12, None),    # rows past one segment: staged path, G = 2
    # empty operands on the slice-group FWHT kernels: no schedule is built
    # for the empty side, and none of its counts may be read
    ("fwht", 0, 50, 300, 5, 1 << 12, None),       # m = 0, paired rows, G = 2
    ("fwht", 2500, 30, 0, 3, 1 << 14, None),      # p = 0, staged rows, G = 4
    ("fwht", 0, 20, 2600, 3, 1 << 13, None),      # m = 0, staged rows, G = 4
]
SEED = 13794


def rel_l2(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def make_products(shapes, tau=None, seed_base=0):
    prods, host = [], []
    for k, (v, m, n, p, d, b, h) in enumerate(shapes):
        AT, B, _ = workloads.planted(m, n, max(p, 1), 4, 5.0, 1000 + k + seed_base, np.float32)
        B = np.ascontiguousarray(B[:, :p])
        C = cma.DeviceArray((m, p), np.float32)
        prods.append(cma.Product(v, cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B),
                                 d, b, SEED + k, hash=h, C=C, tau=tau))
        host.append((AT, B))
    return prods, host


def oracle_sketch(q, AT, B):
    fam = orc.FAMILIES[q.hash]
    hA, hB = orc.hash_arrays(fam, q.d, q.b, q.seed, q.m, q.p)
    C64 = AT.astype(np.float64).T @ B.astype(np.float64)
    name = "fwht" if q.variant == cma.FWHT else "fft"
    return orc.sketch_from_product(name, C64, hA, hB, q.b), hA, hB


@pytest.fixture(scope="module")
def mixed():
    prods, host = make_products(SHAPES)
    bt = cma.Batch(prods, grain=32)
    bt.run()
    cma.synchronize()
    return prods, host, bt


def test_batch_sketches_match_oracle(mixed):
    prods, host, bt = mixed
    for i, (q, (AT, B)) in enumerate(zip(prods, host)):
        print(f"product {i}: {SHAPES[i]}", flush=True)
        want, _, _ = oracle_sketch(q, AT, B)
        got = bt.values(i)
        if q.n == 0 or q.m == 0 or q.p == 0:
            # an empty operand gives an exactly zero sketch (for the FFT, the
            # balance scales of every row pair are 0)
            assert not got.any(), i
            continue
        assert rel_l2(got, want) <= 1e-5, i


def test_batch_decompress_bit_exact(mixed):
    prods, host, bt = mixed
    for i, q in enumerate(prods):
        if q.m == 0 or q.p == 0:
            continue
        fam = orc.FAMILIES[q.hash]
        hA, hB = orc.hash_arrays(fam, q.d, q.b, q.seed, q.m, q.p)
        vals = bt.values(i).astype(np.float64)
        want = orc.decompress("fwht" if q.variant == cma.FWHT else "fft", vals, hA, hB)
        got = q.C.numpy()
        assert np.array_equal(got.astype(np.float64), want), i


def test_batch_matches_single_product_path(mixed):
    prods, host, bt = mixed
    for i, (q, (AT, B)) in enumerate(zip(prods, host)):
        if q.n == 0 or q.m == 0 or q.p == 0:
            continue
        sk = cma.compress("fwht" if q.variant == cma.FWHT else "fft", AT, B, q.d, q.b,
                          seed=q.seed, hash=q.hash, dtype=np.float32)
        assert rel_l2(bt.values(i), sk.values()) <= 2e-6, i
        sk.free()


def test_alone_equals_in_mix(mixed):
    prods, host, bt = mixed
    for i in (0, 3, 4, 5):
        alone = cma.Batch([prods[i]], grain=32)
        alone.run()
        assert np.array_equal(alone.values(i=0), bt.values(i)), i
        alone.free()


def test_graph_replays_match_eager():
    shapes = SHAPES[:7]
    prods, host = make_products(shapes, seed_base=50)
    eager = cma.Batch(prods, grain=64)
    eager.run()
    want = [q.C.numpy() for q in prods]
    graph = cma.Batch(prods, grain=64, graph=True)
    AT, B = host[0]
    for r in range(3):
        for q in prods:  # each replay must rewrite the outputs
            q.C.upload(np.full(q.C.shape, np.nan, np.float32))
        graph.run()
        # eager work and a synchronisation between replays
        cma.compress("fwht", AT, B, 3, 1 << 12, seed=r, dtype=np.float32).free()
        cma.synchronize()
        for i, q in enumerate(prods):
            assert np.array_equal(q.C.numpy(), want[i]), (r, i)
    graph.free()
    eager.free()


def test_batch_recovers_planted_entries():
    shapes = [("fwht", 512, 1024, 700, 9, 1 << 14, None), ("fft", 700, 900, 512, 11, 1 << 13, None),
              ("fft", 256, 300, 256, 7, 1 << 10, "tabulation8")]
    prods = []
    host = []
    for k, (v, m, n, p, d, b, h) in enumerate(shapes):
        AT, B, _ = workloads.planted(m, n, p, 32, 5.0, 77 + k, np.float32)
        prods.append(cma.Product(v, cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B),
                                 d, b, SEED + k, hash=h, tau=12.5, max_out=1 << 14))
        host.append((AT, B))
    bt = cma.Batch(prods, grain=32)
    bt.run()
    for i, (q, (AT, B)) in enumerate(zip(prods, host)):
        fam = orc.FAMILIES[q.hash]
        hA, hB = orc.hash_arrays(fam, q.d, q.b, q.seed, q.m, q.p)
        est = orc.decompress(shapes[i][0], bt.values(i).astype(np.float64), hA, hB)
        want = set(zip(*np.nonzero(est > 12.5)))
        ii, jj, vv = bt.recovered(i)
        assert set(zip(ii.tolist(), jj.tolist())) == want, i
        truth = set(zip(*np.nonzero(AT.astype(np.float64).T @ B.astype(np.float64) > 12.5)))
        print(f"product {i}: recovered {len(want)} truth {len(truth)} "
              f"recall {len(want & truth) / max(1, len(truth)):.3f}")
        assert len(want & truth) >= 0.95 * len(truth)
    bt.free()


def test_autotuned_grain_is_cached():
    prods, _ = make_products(SHAPES[:4], seed_base=90)
    a = cma.Batch(prods)  # grain 0: tuned on first use
    assert a.grain in (16, 32, 64, 128, 256)
    b = cma.Batch(prods)
    assert b.grain == a.grain
    a.free()
    b.free()


def test_batch_errors():
    AT = cma.DeviceArray.from_numpy(np.zeros((4, 3), np.float32))
    B = cma.DeviceArray.from_numpy(np.zeros((4, 5), np.float32))
    with pytest.raises(cma.CMError):
        cma.Batch([cma.Product("fft", AT, B, 17, 64, 1)])
    with pytest.raises(ValueError):
        cma.Batch([cma.Product("fft", AT, B, 3, 48, 1)])
    with pytest.raises(ValueError):
        cma.Product("fft", cma.DeviceArray.from_numpy(np.zeros((4, 3))), B, 3, 64, 1)
