"""Benchmark of the compressed matrix product on one MI355X.

Workload (the metric's single-GPU configuration): plain method, FWHT
(XOR-convolution) variant, A 2048x2048 and B 2048x2048 fp32 with iid
N(0, 1/2048) noise plus 128 planted entries of magnitude 6.0, sketch length
b = 2^14, d = 15 repetitions, multiply-add-shift hashes with seed 13794.

One step = one full product from inputs already resident in HBM: hash
functions drawn and evaluated, fused sketch -> transform -> spectral MAC,
reduction and inverse, then the median-of-d decompression of all of C into
HBM (the reference's matmul_fwht).  The recovered set (> tau = 18) is checked
against the FP64 dense product after the timed region.

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: this
build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import compmatmul_amd as cma  # noqa: E402
from compmatmul_amd import workloads  # noqa: E402

WORKLOAD = "fwht2048_fp32"
HASH_SEED = 13794
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def fwht_flops(m, n, p, d, b):
    """Algorithmic FLOPs of the compress kernel (sketch + 2 forward WHTs per
    (t, k) + spectral MAC); the d inverse transforms run in the reduce kernel."""
    lb = int(np.log2(b))
    return 2 * d * n * (m + p) + 2 * d * n * b * lb + 2 * d * n * b


def cpu_baseline(c, AT, B):
    """The FP64 oracle on the host: compress over a bounded slice of the inner
    dimension, scaled to the full product, plus the full decompression."""
    sys.path.insert(0, HERE)
    from oracle import oracle as orc
    threads = min(16, os.cpu_count() or 1)
    n = AT.shape[0]
    ns = 128
    AT64 = np.ascontiguousarray(AT[:ns], np.float64)
    B64 = np.ascontiguousarray(B[:ns], np.float64)
    orc.compress(c["variant"], AT64[:8], B64[:8], c["d"], c["b"], HASH_SEED, 0, threads)
    t0 = time.perf_counter()
    sk = orc.compress(c["variant"], AT64, B64, c["d"], c["b"], HASH_SEED, 0, threads)
    t1 = time.perf_counter()
    hA, hB = orc.hash_arrays(0, c["d"], c["b"], HASH_SEED, c["m"], c["p"])
    t2 = time.perf_counter()
    orc.decompress(c["variant"], sk, hA, hB, threads)
    t3 = time.perf_counter()
    per_product = (t1 - t0) * (n / ns) + (t3 - t2)
    return {"value": 1.0 / per_product, "unit": "products/s", "cores": threads,
            "kind": "port",
            "sample": f"FP64 oracle (C++/OpenMP restatement, not FFTW/FXT): compress over "
                      f"{ns} of {n} inner rows scaled x{n // ns}, plus the full "
                      f"{c['m']}x{c['p']} median-of-{c['d']} decompression"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain", action="store_true", help="time the unfused path instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the product shards by inner dimension or by product; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    c = workloads.CONFIGS[WORKLOAD]
    m, n, p, d, b = c["m"], c["n"], c["p"], c["d"], c["b"]
    AT, B, _ = workloads.make(WORKLOAD)
    print(f"[bench] workload {WORKLOAD} m={m} n={n} p={p} d={d} b={b}", file=sys.stderr,
          flush=True)
    dAT = cma.DeviceArray.from_numpy(AT)
    dB = cma.DeviceArray.from_numpy(B)
    dC = cma.DeviceArray((m, p), np.float32)
    flags = cma.FLAG_PLAIN if args.plain else 0
    sk = None

    def step():
        nonlocal sk
        sk = cma.compress(c["variant"], dAT, dB, d, b, seed=HASH_SEED, dtype=np.float32,
                          flags=flags, sketch=sk)
        sk.decompress(dC)

    for _ in range(args.warmup):
        step()
    cma.synchronize()
    cma.timing_enable(True)
    cma.timing_reset()
    cma.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    cma.synchronize()
    elapsed = time.perf_counter() - t0
    kname = "plain_transform" if args.plain else "compress"
    k_ms, k_n = cma.timing_get(kname)
    totals = {k: cma.timing_get(k) for k in ("hash", "compress", "reduce", "combine",
                                            "decompress", "plain_sketch", "plain_transform",
                                            "plain_mac")}
    cma.timing_enable(False)
    print("[bench] per-kernel ms/launch: " + ", ".join(
        f"{k}={v[0] / v[1]:.4f}" for k, v in totals.items() if v[1]), file=sys.stderr, flush=True)

    # parity of the last product: recovered set vs the dense FP64 product
    C_est = dC.numpy()
    C_true = AT.astype(np.float64).T @ B.astype(np.float64)
    truth = set(zip(*np.nonzero(C_true > c["tau"])))
    ii, jj, _ = sk.recover(c["tau"])
    got = set(zip(ii.tolist(), jj.tolist()))
    est_set = set(zip(*np.nonzero(C_est > c["tau"])))
    recall = len(got & truth) / max(1, len(truth))
    precision = len(got & truth) / max(1, len(got))

    ms_per_step = 1e3 * elapsed / args.steps
    flops = fwht_flops(m, n, p, d, b)
    k_avg_s = (k_ms / max(1, k_n)) * 1e-3
    achieved = flops / k_avg_s / 1e12 if k_avg_s > 0 else 0.0
    out = {
        "metric": "products/s",
        "value": args.steps / elapsed,
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: N(0,1/2048) noise + 128 planted entries of 6.0 (seeded)",
        "config": {"workload": "fwht 2048x2048x2048 fp32, b=2^14, d=15, 128 plants, tau=18",
                   "m": m, "n": n, "p": p, "b": b, "d": d, "variant": "fwht",
                   "hash": "multiply add shift", "path": "plain" if args.plain else "fused",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": None, "kernel": kname,
                     "kernel_ms": k_avg_s * 1e3, "flops_per_launch": flops},
        "detail": {"recall": recall, "precision": precision, "recovered": len(got),
                   "truth": len(truth), "C_threshold_set_equal": est_set == got,
                   "kernel_ms": {k: (v[0] / v[1] if v[1] else None) for k, v in totals.items()}},
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(c, AT, B)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
