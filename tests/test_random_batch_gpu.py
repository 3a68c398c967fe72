"""A batch of seeded random products against the oracle.

64 products drawn from a fixed seed (variant, m, n, p, b = 2 .. 2^16,
d = 1 .. 16, hash family) run as one batch, replayed from its HIP graph
twice.  Every product's sketch is held to the oracle's FP64 closed form
(1e-5 relative L2), its C is bit-exact against the oracle's median of the
sketch the batch returned, and its recovered entries are exactly the
oracle's estimates above tau.  The draws spread the products over every
compress kernel kind, finishing step and decompression form of a batch.
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

HASHES = [None, "multiply shift", "tabulation8", "tabulation16"]
TAU = 4.0


def draw(count, seed=4242):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        out.append(dict(variant=str(rng.choice(["fft", "fwht"])),
                        m=int(rng.choice([1, 7, 64, 129, 300, 700])),
                        n=int(rng.choice([1, 9, 64, 250])),
                        p=int(rng.choice([1, 33, 128, 400, 700])),
                        d=int(rng.integers(1, 17)), b=1 << int(rng.integers(1, 17)),
                        hash=HASHES[int(rng.integers(0, 4))]))
    return out


def test_random_batch_against_oracle():
    specs = draw(64)
    prods, host = [], []
    for k, s in enumerate(specs):
        AT, B, _ = workloads.planted(s["m"], s["n"], s["p"], 4, 5.0, 500 + k, np.float32)
        prods.append(cma.Product(s["variant"], cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B),
                                 s["d"], s["b"], 77 + k, hash=s["hash"],
                                 C=cma.DeviceArray((s["m"], s["p"]), np.float32),
                                 tau=TAU, max_out=s["m"] * s["p"]))
        host.append((AT, B))
    bt = cma.Batch(prods, grain=0, graph=True)
    for _ in range(2):
        bt.run()
    cma.synchronize()
    for i, (s, q, (AT, B)) in enumerate(zip(specs, prods, host)):
        print(f"product {i}: {s}", flush=True)
        hA, hB = orc.hash_arrays(orc.FAMILIES[s["hash"]], s["d"], s["b"], q.seed, s["m"], s["p"])
        C64 = AT.astype(np.float64).T @ B.astype(np.float64)
        want = orc.sketch_from_product(s["variant"], C64, hA, hB, s["b"])
        got = bt.values(i).astype(np.float64)
        assert np.linalg.norm(got - want) <= 1e-5 * max(float(np.linalg.norm(want)), 1e-30), i
        est = orc.decompress(s["variant"], got, hA, hB)
        assert np.array_equal(q.C.numpy().astype(np.float64), est), i
        ii, jj, _ = bt.recovered(i)
        assert set(zip(ii.tolist(), jj.tolist())) == set(zip(*np.nonzero(est > TAU))), i
    bt.free()
