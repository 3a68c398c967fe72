"""Synthetic planted-entry inputs for the benchmark and tests.

A (m x n) and B (n x p) hold iid N(0, 1/n) noise (variance 1/n).  Plants:
draw `plants` distinct (i, j) pairs and a random k for each, then set
A[i, k] = B[k, j] = magnitude, so C[i, j] gains magnitude^2.  Two plants
sharing k also create C[i1, j2] and C[i2, j1] of size magnitude^2, so the
ground truth of a recovery is {C_true > tau} from the dense product, not the
plant list.  Returned operands follow the reference's layout: AT (n x m,
A transposed) and B (n x p), both row-major.
"""
import numpy as np

DATA_SEED = 20260001


def planted(m, n, p, plants, magnitude, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    scale = 1.0 / np.sqrt(n)
    AT = (rng.standard_normal((n, m)) * scale).astype(dtype)
    B = (rng.standard_normal((n, p)) * scale).astype(dtype)
    plants = min(plants, m * p)
    flat = rng.choice(m * p, size=plants, replace=False)
    ii, jj = flat // p, flat % p
    kk = rng.integers(0, n, size=plants)
    AT[kk, ii] = magnitude
    B[kk, jj] = magnitude
    return AT, B, np.stack([ii, jj, kk], axis=1)


# workloads named by their shapes (m, n, p, plants, magnitude, variant, b, d, tau)
CONFIGS = {
    "fft512_fp64": dict(m=512, n=512, p=512, plants=32, magnitude=4.0, variant="fft",
                        b=4096, d=9, tau=8.0, dtype=np.float64),
    "fwht2048_fp32": dict(m=2048, n=2048, p=2048, plants=128, magnitude=6.0,
                          variant="fwht", b=1 << 14, d=15, tau=18.0, dtype=np.float32),
    "fft2048_fp32": dict(m=2048, n=2048, p=2048, plants=128, magnitude=6.0, variant="fft",
                         b=1 << 14, d=15, tau=18.0, dtype=np.float32),
    "fft4096x16384_fp32": dict(m=4096, n=16384, p=4096, plants=512, magnitude=5.0,
                               variant="fft", b=1 << 16, d=11, tau=12.5,
                               dtype=np.float32),
}


def make(name, seed_offset=0):
    c = CONFIGS[name]
    idx = list(CONFIGS).index(name) + 1
    AT, B, pl = planted(c["m"], c["n"], c["p"], c["plants"], c["magnitude"],
                        DATA_SEED + idx + seed_offset, c["dtype"])
    return AT, B, pl
