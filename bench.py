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

DEFAULT_WORKLOAD = "fwht2048_fp32"
HASH_SEED = 13794
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# largest sketch length of the unfused (plain) FP32 path: one LDS-resident
# transform of at most 64 KB (FFT: complex)
PLAIN_MAX_B = {"fft": 1 << 13, "fwht": 1 << 14}


def compress_flops(variant, m, n, p, d, b):
    """Algorithmic FLOPs of the compress kernel: sketch + forward transforms
    per (t, k) + spectral MAC (the d inverse transforms run in the reduce
    kernel).  FWHT: 2dn(m+p) + 2dn*b*log2(b) + 2dn*b.  FFT (two real
    transforms of length b per (t, k), 2.5 b log2 b flops each, complex MAC
    over b/2+1 bins): 2dn(m+p) + 2dn*2.5*b*log2(b) + 8dn(b/2+1)."""
    lb = int(np.log2(b))
    if variant == "fwht":
        return 2 * d * n * (m + p) + 2 * d * n * b * lb + 2 * d * n * b
    return 2 * d * n * (m + p) + 2 * d * n * 2.5 * b * lb + 8 * d * n * (b // 2 + 1)


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


def copy_probe(nbytes):
    src = cma.DeviceArray((nbytes // 4,), np.float32)
    dst = cma.DeviceArray((nbytes // 4,), np.float32)
    for _ in range(5):
        cma.check(cma.lib().cm_memcpy(dst.ptr, src.ptr, nbytes, cma.DEVICE, cma.DEVICE))
    cma.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 5}), flush=True)


def traffic_from_profile(name):
    """HBM bytes per compress launch for `name`, from the committed counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the same counters
    read on a device-to-device copy of known size), or None."""
    path = os.path.join(HERE, "profiles", f"{name}_compress_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("traffic_bytes")


def bench_batch(args):
    """512 products of uneven shape in one batch (HIP graph replay of one
    plan): products/s over the whole batch, recall on every 16th product."""
    specs = workloads.batch_specs(512)
    t0 = time.perf_counter()
    prods, sample = [], {}
    for i, s in enumerate(specs):
        AT, B, _ = workloads.planted(s["m"], s["n"], s["p"], s["plants"], s["magnitude"],
                                     workloads.BATCH_SEED + 1 + i, np.float32, direct=True)
        prods.append(cma.Product(s["variant"], cma.DeviceArray.from_numpy(AT),
                                 cma.DeviceArray.from_numpy(B), s["d"], s["b"], HASH_SEED + i,
                                 C=cma.DeviceArray((s["m"], s["p"]), np.float32),
                                 tau=s["tau"], max_out=1 << 16))
        if i % 16 == 0:
            sample[i] = (AT, B)
    print(f"[bench] batch inputs generated and uploaded in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    bt = cma.Batch(prods, grain=args.grain, graph=True)
    print(f"[bench] plan (grain {bt.grain}) in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        bt.run()
    cma.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bt.run()
    cma.synchronize()
    elapsed = time.perf_counter() - t0
    # one eager run of the same plan shapes with kernel timing on
    eager = cma.Batch(prods, grain=bt.grain)
    cma.timing_enable(True)
    cma.timing_reset()
    eager.run()
    cma.synchronize()
    stages = {k: cma.timing_get(k)[0] for k in ("hash", "zero", "lists", "compress_fwht",
                                               "compress_fft", "finish", "decompress")}
    cma.timing_enable(False)
    eager.free()
    flops = sum(compress_flops(s["variant"], s["m"], s["n"], s["p"], s["d"], s["b"]) for s in specs)
    comp_s = (stages["compress_fwht"] + stages["compress_fft"]) * 1e-3
    achieved = flops / comp_s / 1e12 if comp_s > 0 else 0.0
    hits = truth_n = got_n = 0
    for i, (AT, B) in sample.items():
        C_true = AT.astype(np.float64).T @ B.astype(np.float64)
        truth = set(zip(*np.nonzero(C_true > specs[i]["tau"])))
        ii, jj, _ = bt.recovered(i)
        got = set(zip(ii.tolist(), jj.tolist()))
        hits += len(got & truth)
        truth_n += len(truth)
        got_n += len(got)
    nbytes = sum(4 * (s["m"] * s["n"] + s["n"] * s["p"] + s["m"] * s["p"]) for s in specs)
    batch_s = elapsed / args.steps
    return {
        "metric": "products/s",
        "value": len(specs) / batch_s,
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * batch_s,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 512 products, N(0,1/n) noise + 16-256 planted entries of 5.0 each (seeded)",
        "config": {"workload": "batch of 512 products: m, p log-uniform [64, 4096], n [128, 8192], "
                               "b 2^10..2^16, d {5,7,9,11,15}, FFT/FWHT 50/50, tau 12.5",
                   "grain": bt.grain, "graph": True, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "compress_fwht + compress_fft (persistent)",
                     "kernel_ms": comp_s * 1e3, "flops_per_launch": flops},
        "detail": {"stage_ms": stages, "input_bytes": nbytes,
                   "hbm_GBps_whole_batch": nbytes / batch_s / 1e9,
                   "sample_products": len(sample), "recall": hits / max(1, truth_n),
                   "precision": hits / max(1, got_n)},
    }


def bench_precision(args):
    """The precision form (FFT, 4096 x 16384 x 4096, b = 2^16, d = 11) in three
    arithmetic modes: FP32 butterflies, FP64 cross-chunk spectral
    accumulation, BF16 matrix-core DFT stages with hi/lo split correction.
    Sketch errors against the device FP64 path with the same hashes (itself
    within 1e-10 of the oracle in the tests); estimate errors against the
    dense FP64 product; recovered set against {C > tau}."""
    name = "fft4096x16384_fp32"
    c = workloads.CONFIGS[name]
    m, n, p, d, b, tau = c["m"], c["n"], c["p"], c["d"], c["b"], c["tau"]
    AT, B, _ = workloads.make(name)
    C_true = AT.astype(np.float64).T @ B.astype(np.float64)
    truth = set(zip(*np.nonzero(C_true > tau)))
    ref = cma.compress("fft", cma.DeviceArray.from_numpy(AT.astype(np.float64)),
                       cma.DeviceArray.from_numpy(B.astype(np.float64)), d, b, seed=HASH_SEED,
                       dtype=np.float64)
    ref_vals = ref.values()
    C_ref = ref.decompress()
    ref.free()
    dAT, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    modes = {"fp32": 0, "fp64_cross_chunk": cma.FLAG_F64_REDUCE, "bf16_mfma_split": cma.FLAG_BF16_MFMA}
    out = {}
    for mode, flags in modes.items():
        sk = None
        for _ in range(max(1, args.warmup)):
            sk = cma.compress("fft", dAT, dB, d, b, seed=HASH_SEED, dtype=np.float32, flags=flags,
                              sketch=sk)
        cma.synchronize()
        cma.timing_enable(True)
        cma.timing_reset()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sk = cma.compress("fft", dAT, dB, d, b, seed=HASH_SEED, dtype=np.float32, flags=flags,
                              sketch=sk)
        cma.synchronize()
        step_s = (time.perf_counter() - t0) / args.steps
        k_ms = cma.timing_get("compress")
        cma.timing_enable(False)
        v = sk.values()
        e = v - ref_vals
        C_est = sk.decompress().astype(np.float64)
        ii, jj, _ = sk.recover(tau)
        got = set(zip(ii.tolist(), jj.tolist()))
        out[mode] = {
            "sketch": {"max_abs": float(np.abs(e).max()), "rms": float(np.sqrt(np.mean(e * e))),
                       "rel_l2": float(np.linalg.norm(e) / np.linalg.norm(ref_vals))},
            "C_vs_fp64_sketch_estimate": {"max_abs": float(np.abs(C_est - C_ref).max()),
                                          "rms": float(np.sqrt(np.mean((C_est - C_ref) ** 2)))},
            "C_vs_dense_fp64": {"max_abs": float(np.abs(C_est - C_true).max()),
                                "rms": float(np.sqrt(np.mean((C_est - C_true) ** 2)))},
            "recall": len(got & truth) / max(1, len(truth)),
            "precision": len(got & truth) / max(1, len(got)),
            "recovered": len(got), "truth": len(truth),
            "compress_ms": k_ms[0] / max(1, k_ms[1]), "compress_products_per_s": 1.0 / step_s,
        }
        print(f"[bench] {mode}: {json.dumps(out[mode])}", file=sys.stderr, flush=True)
        sk.free()
    flops = compress_flops("fft", m, n, p, d, b)
    k_s = out["fp32"]["compress_ms"] * 1e-3
    return {
        "metric": "products/s", "value": out["fp32"]["compress_products_per_s"],
        "unit": "products/s (compress)", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 / out["fp32"]["compress_products_per_s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: N(0,1/{n}) noise + {c['plants']} planted entries of {c['magnitude']}",
        "config": {"workload": f"precision form: fft {m}x{n}x{p}, b={b}, d={d}, tau={tau}, "
                               "three arithmetic modes", "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": flops / k_s / 1e12, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / k_s / 1e12 / PEAK_FP32_TFLOPS,
                     "traffic": None, "kernel": "compress (fp32 mode)",
                     "kernel_ms": k_s * 1e3, "flops_per_launch": flops},
        "detail": {"modes": out},
    }


def serve_datasets(count, seed):
    """`count` request templates from the batch shape distribution; 30% hold
    CSR operands at 1% density (N(0,1) values plus the planted entries),
    the rest dense operands in pinned host memory."""
    import scipy.sparse as sparse
    specs = workloads.batch_specs(count, seed)
    rng = np.random.default_rng(seed + 1)
    out = []
    for i, s in enumerate(specs):
        m, n, p = s["m"], s["n"], s["p"]
        if rng.random() < 0.3:
            A = sparse.random(n, m, density=0.01, format="lil", random_state=seed + 2 * i,
                              dtype=np.float32, data_rvs=lambda k: rng.standard_normal(k))
            Bs = sparse.random(n, p, density=0.01, format="lil", random_state=seed + 2 * i + 1,
                               dtype=np.float32, data_rvs=lambda k: rng.standard_normal(k))
            for _ in range(s["plants"]):
                k = rng.integers(n)
                A[k, rng.integers(m)] = s["magnitude"]
                Bs[k, rng.integers(p)] = s["magnitude"]
            out.append(dict(s, csr=True, AT=A.tocsr(), B=Bs.tocsr()))
        else:
            AT, B, _ = workloads.planted(m, n, p, s["plants"], s["magnitude"], seed + 2 * i,
                                         np.float32, direct=True)
            out.append(dict(s, csr=False, AT=cma.HostBuffer.from_numpy(AT),
                            B=cma.HostBuffer.from_numpy(B)))
    return out


def bench_serve(args):
    """Poisson arrivals at a fixed QPS into the continuous-batching server;
    the same arrivals replayed through one product per launch sequence."""
    templates = serve_datasets(64, workloads.BATCH_SEED + 100)
    rng = np.random.default_rng(7)
    nreq = args.requests
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, nreq))
    picks = rng.integers(0, len(templates), nreq)
    seeds = HASH_SEED + np.arange(nreq)

    def operands(t):
        return t["AT"], t["B"]

    sv = cma.Server(hbm_bytes=args.hbm_gb << 30, max_waves=args.waves,
                    max_wave_products=args.wave_products, grain=args.grain or 64)
    # warm-up: one request of every template
    for k, t in enumerate(templates):
        a, b = operands(t)
        tk = sv.submit(t["variant"], a, b, t["d"], t["b"], 1 + k, tau=t["tau"], max_out=1 << 16)
        sv.wait(tk)
        sv.release(tk)
    import threading
    tickets = [None] * nreq
    t0 = time.perf_counter() + 0.05

    def feeder():
        for r in range(nreq):
            dt = t0 + arrivals[r] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t = templates[picks[r]]
            a, b = operands(t)
            tickets[r] = sv.submit(t["variant"], a, b, t["d"], t["b"], int(seeds[r]),
                                   tau=t["tau"], max_out=1 << 16)

    th = threading.Thread(target=feeder)
    th.start()
    lat = np.zeros(nreq)
    results = {}
    for r in range(nreq):
        while tickets[r] is None:
            time.sleep(1e-4)
        sv.wait(tickets[r])
        ii, jj, _, lat[r], _ = sv.result(tickets[r])
        if picks[r] < 16 and picks[r] not in results:
            results[picks[r]] = set(zip(ii.tolist(), jj.tolist()))
        sv.release(tickets[r])
    th.join()
    t_end = time.perf_counter()
    st = sv.stats()
    sv.close()
    served_s = t_end - t0

    # one product per launch sequence: the same arrivals, served in order,
    # by the fused single-product path and by the unfused (plain) path where
    # its LDS limit allows the sketch length (else fused)
    dev = {}
    for k, t in enumerate(templates):  # operands resident on the device, as for the server
        if t["csr"]:
            a, b = t["AT"].toarray(), t["B"].toarray()
        else:
            a, b = t["AT"].array, t["B"].array
        dev[k] = (cma.DeviceArray.from_numpy(a), cma.DeviceArray.from_numpy(b))

    def one_at_a_time(plain):
        lat_ = np.zeros(nreq)
        t1 = time.perf_counter() + 0.05
        for r in range(nreq):
            dt = t1 + arrivals[r] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t = templates[picks[r]]
            a, b = dev[picks[r]]
            fl = cma.FLAG_PLAIN if plain and t["b"] <= PLAIN_MAX_B[t["variant"]] else 0
            sk = cma.compress(t["variant"], a, b, t["d"], t["b"], seed=int(seeds[r]),
                              dtype=np.float32, flags=fl)
            sk.recover(t["tau"])
            sk.free()
            lat_[r] = time.perf_counter() - (t1 + arrivals[r])
        return lat_, time.perf_counter() - t1

    base_lat, base_s = one_at_a_time(False)
    plain_lat, plain_s = one_at_a_time(True) if args.plain_baseline else (None, None)

    # recall of the served results on the first templates
    hits = truth_n = got_n = 0
    for k, got in results.items():
        t = templates[k]
        a = t["AT"].toarray() if t["csr"] else t["AT"].array
        b = t["B"].toarray() if t["csr"] else t["B"].array
        truth = set(zip(*np.nonzero(a.astype(np.float64).T @ b.astype(np.float64) > t["tau"])))
        hits += len(got & truth)
        truth_n += len(truth)
        got_n += len(got)
    q = lambda x, f: float(np.quantile(x, f))
    return {
        "metric": "products/s",
        "value": nreq / served_s,
        "unit": "products/s",
        "n_gpus": 1,
        "steps": nreq,
        "warmup": len(templates),
        "ms_per_step": 1e3 * served_s / nreq,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: batch shape distribution, 30% CSR at 1% density, planted entries (seeded)",
        "config": {"workload": f"serving: Poisson arrivals at {args.qps} QPS, {nreq} requests over "
                               f"{len(templates)} request templates",
                   "qps": args.qps, "max_waves": args.waves,
                   "max_wave_products": args.wave_products, "grain": args.grain or 64,
                   "parallelism": "single GPU"},
        "detail": {"latency_ms": {"p50": 1e3 * q(lat, 0.5), "p99": 1e3 * q(lat, 0.99),
                                  "max": 1e3 * float(lat.max())},
                   "one_product_per_launch": {"products_per_s": nreq / base_s,
                                              "latency_ms": {"p50": 1e3 * q(base_lat, 0.5),
                                                             "p99": 1e3 * q(base_lat, 0.99)}},
                   "plain_one_product_per_launch": None if plain_lat is None else {
                       "products_per_s": nreq / plain_s,
                       "latency_ms": {"p50": 1e3 * q(plain_lat, 0.5),
                                      "p99": 1e3 * q(plain_lat, 0.99)},
                       "note": "unfused path for b within its LDS limit, fused otherwise"},
                   "server": st, "recall": hits / max(1, truth_n),
                   "precision": hits / max(1, got_n)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain", action="store_true", help="time the unfused path instead")
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD,
                    choices=sorted(workloads.CONFIGS) + ["batch512", "serve", "precision"],
                    help="named workload (default: the metric's configuration)")
    ap.add_argument("--grain", type=int, default=0,
                    help="batch512: work-item grain (0 = autotuned); serve: 0 = 64")
    ap.add_argument("--qps", type=float, default=300.0, help="serve: arrival rate")
    ap.add_argument("--requests", type=int, default=600, help="serve: requests")
    ap.add_argument("--waves", type=int, default=3, help="serve: waves on the device at once")
    ap.add_argument("--wave-products", type=int, default=64, help="serve: requests per wave")
    ap.add_argument("--hbm-gb", type=int, default=64, help="serve: server HBM budget (GiB)")
    ap.add_argument("--plain-baseline", action="store_true",
                    help="serve: also replay the arrivals through the unfused path")
    ap.add_argument("--kernel-times", action="store_true",
                    help="time every kernel class with HIP events (default: only the "
                         "dominant one, since each event pair slows the stream)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counter readings)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the product shards by inner dimension or by product; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    name = args.workload
    if name == "batch512":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if name == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if name == "precision":
        print(json.dumps(bench_precision(args)), flush=True)
        return
    c = workloads.CONFIGS[name]
    m, n, p, d, b = c["m"], c["n"], c["p"], c["d"], c["b"]
    dtype = np.dtype(c["dtype"])
    AT, B, _ = workloads.make(name)
    print(f"[bench] workload {name} m={m} n={n} p={p} d={d} b={b}", file=sys.stderr,
          flush=True)
    dAT = cma.DeviceArray.from_numpy(AT)
    dB = cma.DeviceArray.from_numpy(B)
    dC = cma.DeviceArray((m, p), dtype)
    flags = cma.FLAG_PLAIN if args.plain else 0
    sk = None

    def step():
        nonlocal sk
        sk = cma.compress(c["variant"], dAT, dB, d, b, seed=HASH_SEED, dtype=dtype,
                          flags=flags, sketch=sk)
        sk.decompress(dC)

    for _ in range(args.warmup):
        step()
    cma.synchronize()
    kname = "plain_transform" if args.plain else "compress"
    cma.timing_enable(True, only=None if args.kernel_times else kname)
    cma.timing_reset()
    cma.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    cma.synchronize()
    elapsed = time.perf_counter() - t0
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
    flops = compress_flops(c["variant"], m, n, p, d, b)
    peak = PEAK_FP32_TFLOPS if dtype == np.float32 else PEAK_FP64_TFLOPS
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
        "dtype": "f32" if dtype == np.float32 else "f64",
        "data": f"synthetic: N(0,1/{n}) noise + {c['plants']} planted entries of "
                f"{c['magnitude']} (seeded)",
        "config": {"workload": f"{c['variant']} {m}x{n}x{p} {dtype.name}, b={b}, d={d}, "
                               f"{c['plants']} plants, tau={c['tau']}",
                   "m": m, "n": n, "p": p, "b": b, "d": d, "variant": c["variant"],
                   "hash": "multiply add shift", "path": "plain" if args.plain else "fused",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None if args.plain else traffic_from_profile(name),
                     "kernel": kname,
                     "kernel_ms": k_avg_s * 1e3, "flops_per_launch": flops},
        "detail": {"recall": recall, "precision": precision, "recovered": len(got),
                   "truth": len(truth), "C_threshold_set_equal": est_set == got,
                   "kernel_ms": {k: (v[0] / v[1] if v[1] else None) for k, v in totals.items()}},
    }
    if not args.no_cpu_baseline and name == DEFAULT_WORKLOAD:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(c, AT, B)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
