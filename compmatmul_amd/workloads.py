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


def planted(m, n, p, plants, magnitude, seed, dtype=np.float32, direct=False):
    """direct=True draws the noise in `dtype` itself (faster for float32;
    a different stream than the default float64 draw)."""
    rng = np.random.default_rng(seed)
    scale = 1.0 / np.sqrt(max(n, 1))
    if direct:
        AT = rng.standard_normal((n, m), dtype=dtype)
        AT *= dtype(scale)
        B = rng.standard_normal((n, p), dtype=dtype)
        B *= dtype(scale)
    else:
        AT = (rng.standard_normal((n, m)) * scale).astype(dtype)
        B = (rng.standard_normal((n, p)) * scale).astype(dtype)
    plants = min(plants, m * p) if n > 0 else 0
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


# many products of uneven shape: m, p log-uniform in [64, 4096], n in
# [128, 8192], b = 2^10 .. 2^16, d in {5, 7, 9, 11, 15}, FFT / FWHT 50/50,
# 16-256 plants of magnitude 5.0 each, recovery threshold 12.5
BATCH_SEED = 20260004


def batch_specs(count=512, seed=BATCH_SEED):
    rng = np.random.default_rng(seed)

    def logu(lo, hi, size):
        return np.exp(rng.uniform(np.log(lo), np.log(hi), size)).round().astype(int)

    m, p, n = logu(64, 4096, count), logu(64, 4096, count), logu(128, 8192, count)
    b = 1 << rng.integers(10, 17, count)
    d = rng.choice([5, 7, 9, 11, 15], count)
    variant = np.where(rng.random(count) < 0.5, "fft", "fwht")
    plants = rng.integers(16, 257, count)
    return [dict(variant=str(variant[i]), m=int(m[i]), n=int(n[i]), p=int(p[i]), b=int(b[i]),
                 d=int(d[i]), plants=int(plants[i]), magnitude=5.0, tau=12.5)
            for i in range(count)]


def make(name, seed_offset=0):
    c = CONFIGS[name]
    idx = list(CONFIGS).index(name) + 1
    AT, B, pl = planted(c["m"], c["n"], c["p"], c["plants"], c["magnitude"],
                        DATA_SEED + idx + seed_offset, c["dtype"])
    return AT, B, pl
