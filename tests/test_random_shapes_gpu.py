"""Seeded random products through every single-product path, against the oracle.

Shapes, b, d, variant, precision, hash family and path (default, single-wave,
general LDS kernel, plain) are drawn from a fixed seed, so the cases are the
same on every run.  Each product's sketch is held to the oracle's FP64 closed
form (1e-10 relative to its largest value in FP64, 1e-5 relative L2 in FP32),
its decompression is bit-exact against the oracle's median of the sketch it
returned, and the recovered entries are exactly the oracle's estimates above
tau.  The draws reach the cases the targeted tests cover one by one: empty
and one-element operands, b from 2 to 2^16, d from 1 to 16, sketch rows
through every decompression form.
"""
import numpy as np
import pytest

import compmatmul_amd as cma
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

HASHES = [None, "multiply shift", "tabulation8", "tabulation16"]
PATHS = [("default", 0), ("single-wave", cma.FLAG_SINGLE_WAVE),
         ("lds", cma.FLAG_LDS_KERNEL), ("plain", cma.FLAG_PLAIN)]


def draw_cases(count, seed=20261204):
    rng = np.random.default_rng(seed)
    cases = []
    for c in range(count):
        f64 = rng.random() < 0.3
        m = int(rng.choice([0, 1, 3, 17, 64, 130, 257, 600]))
        p = int(rng.choice([1, 2, 29, 64, 200, 513, 900]))
        n = int(rng.choice([0, 1, 5, 33, 120, 300]))
        lb = int(rng.integers(1, 17))
        d = int(rng.integers(1, 17))
        cases.append(dict(id=c, variant=str(rng.choice(["fft", "fwht"])),
                          dtype=np.float64 if f64 else np.float32, m=m, n=n, p=p, b=1 << lb, d=d,
                          hash=HASHES[int(rng.integers(0, 4))],
                          path=PATHS[int(rng.integers(0, 4))], seed=int(rng.integers(1, 1 << 30))))
    return cases


CASES = draw_cases(120)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['id']}-{c['variant']}-{np.dtype(c['dtype']).name}-"
                                                      f"{c['m']}x{c['n']}x{c['p']}-b{c['b']}-d{c['d']}-{c['path'][0]}")
def test_random_product_against_oracle(case):
    v, dt, m, n, p, b, d = (case[k] for k in ("variant", "dtype", "m", "n", "p", "b", "d"))
    name, flags = case["path"]
    print(f"case {case['id']}: {v} {np.dtype(dt).name} {m}x{n}x{p} b={b} d={d} "
          f"hash={case['hash']} path={name}", flush=True)
    rng = np.random.default_rng(case["seed"])
    AT = rng.standard_normal((n, m)).astype(dt)
    B = rng.standard_normal((n, p)).astype(dt)
    if n > 0 and m > 0 and p > 0:  # a few large entries to recover
        AT[0, 0] += dt(4.0)
        B[0, 0] += dt(4.0)
    sk = cma.compress(v, AT, B, d, b, seed=case["seed"], hash=case["hash"], dtype=dt, flags=flags)
    hA, hB = sk.hashes()
    famA, famB = orc.hash_arrays(orc.FAMILIES[case["hash"]], d, b, case["seed"], m, p)
    assert np.array_equal(hA, famA) and np.array_equal(hB, famB)
    C64 = AT.astype(np.float64).T @ B.astype(np.float64)
    want = orc.sketch_from_product(v, C64, hA, hB, b)
    got = sk.values()
    if dt == np.float64:
        assert np.abs(got - want).max() <= 1e-10 * max(np.abs(want).max(), 1.0)
    else:
        scale = max(float(np.linalg.norm(want)), 1e-30)
        # an exactly zero product (an empty operand) must give an exactly zero sketch
        assert (not want.any() and not got.any()) or np.linalg.norm(got - want) / scale <= 1e-5
    if m > 0 and p > 0:
        est = orc.decompress(v, got.astype(np.float64), hA, hB)
        C = sk.decompress()
        assert np.array_equal(C.astype(np.float64), est)
        tau = 3.0
        ii, jj, vv = sk.recover(tau)
        want_set = set(zip(*np.nonzero(est > tau)))
        assert set(zip(ii.tolist(), jj.tolist())) == want_set
    sk.free()
