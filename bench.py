"""Benchmark of the compressed matrix product on one MI355X.

Default workload (the metric's configuration: products/s and p50/p99 latency
at a fixed QPS, against the plain GPU path at equal recall): the serving
form.  Requests arrive as a Poisson stream at a fixed QPS; each request's
shape is drawn from the batch distribution (m, p log-uniform in [64, 4096],
n in [128, 8192], b = 2^10 .. 2^16, d in {5, 7, 9, 11, 15}, FFT / FWHT
50/50), 30% of them with CSR operands at 1% density (N(0, 1) values plus
planted entries of 5.0), the rest dense N(0, 1/n) plus planted entries; the
recovered entries (> tau = 12.5) are each request's result.  The operands
of the 64 request templates are resident in HBM before the timed region
(dense arrays, or CSR arrays the server expands on the device).

One step = one request, from its arrival to its recovered list on the host.
value = requests served / wall time of the timed requests.  The same
arrivals are then replayed through one product per launch sequence on the
fused single-product path and on the plain (unfused) path; their recall is
compared on the first templates.  The roofline block is the dominant
kernel class of the serving run (the persistent compress kernels), timed
with HIP events on the wave streams it runs on.  The FP64 oracle on the
host cores is the CPU baseline, on a bounded sample of the templates.

Other workloads (--workload): the named single-product configurations, the
512-product batch, the precision form.  --gpus N with N > 1 exits: this
build measures one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import compmatmul_amd as cma  # noqa: E402
from compmatmul_amd import workloads  # noqa: E402

DEFAULT_WORKLOAD = "serve"
HASH_SEED = 13794
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# The compress kernels run FP32 adds and multiplies on the vector ALUs (no
# matrix cores); 157.3 TFLOP/s is MI355X's published FP32 peak (packed FMA).
COMPUTE_BOUND = "valu"
PEAK_NOTE = ("FP32 vector ALU: 157.3 TFLOP/s is the packed-FMA peak; with packed FP32 "
             "disabled, add-only work peaks at 39.3 TFLOP/s")
PEAK_NOTE_FP64 = ("FP64 vector ALU: 78.6 TFLOP/s is the FMA peak; add-only FP64 work "
                  "peaks at 39.3 TFLOP/s")


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


def copy_probe(nbytes):
    """Kernels of known traffic for the FETCH_SIZE / WRITE_SIZE scales: 5
    device-to-device copies of `nbytes` (read and write nbytes each), then 5
    fills of `nbytes` by the HIP runtime's hipMemset (write only).  Sizes
    above the 256 MB MALL keep the counters on HBM traffic."""
    src = cma.DeviceArray((nbytes // 4,), np.float32)
    dst = cma.DeviceArray((nbytes // 4,), np.float32)
    for _ in range(5):
        cma.check(cma.lib().cm_memcpy(dst.ptr, src.ptr, nbytes, cma.DEVICE, cma.DEVICE))
    cma.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")  # already loaded by the library
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    fills = 0
    for _ in range(5):
        fills += hip.hipMemset(dst.ptr, 0, nbytes) == 0
    ok = hip.hipDeviceSynchronize() == 0
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 5, "fills": fills, "fills_ok": ok}),
          flush=True)


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


def achievable_peaks(nbytes=1 << 30):
    """Rates this device sustains on calibration kernels of known work: FP32
    and FP64 add and FMA loops (16 independent chains per thread, 16 waves
    per CU, no packed FP32 instructions), and a streaming copy, read and
    fill of 1 GiB (four times the 256 MB MALL, so HBM serves them).  They
    sit beside the published peaks as the denominators of `achievable`."""
    out = {}
    for kind, iters in (("fp32_add", 4096), ("fp32_fma", 4096), ("fp64_add", 2048),
                        ("fp64_fma", 2048)):
        ms, flops = cma.probe_alu(kind, iters=iters)
        out[f"{kind}_tflops"] = flops / (ms * 1e-3) / 1e12
    src = cma.DeviceArray((nbytes // 4,), np.float32)
    dst = cma.DeviceArray((nbytes // 4,), np.float32)
    try:
        cma.probe_stream("fill", src, None, nbytes, reps=1)  # finite data for the reads
        for kind in ("copy", "read", "fill"):
            ms, moved = cma.probe_stream(kind, dst, src, nbytes)
            out[f"hbm_{kind}_gbs"] = moved / (ms * 1e-3) / 1e9
    finally:
        src.free()
        dst.free()
    return out


def add_achievable(roof, peaks, f64=False):
    """Puts the sustained rate of the calibration kernel that bounds `roof`
    beside its published peak: the streaming copy for an HBM-bound kernel;
    for a compute-bound one the add loop (the transforms' butterflies are
    adds and subtractions) and the FMA loop (twiddle products can fuse)."""
    if roof is None or not peaks:
        return
    if roof.get("unit") == "GB/s":
        bw = peaks["hbm_copy_gbs"]
        roof["achievable"] = {"copy_gbs": bw, "frac_copy": roof["achieved"] / bw}
    else:
        pre = "fp64" if f64 else "fp32"
        add, fma = peaks[f"{pre}_add_tflops"], peaks[f"{pre}_fma_tflops"]
        roof["achievable"] = {f"{pre}_add_loop_tflops": add, f"{pre}_fma_loop_tflops": fma,
                              "frac_add_loop": roof["achieved"] / add,
                              "frac_fma_loop": roof["achieved"] / fma}


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
        "roofline": {"bound": COMPUTE_BOUND, "peak_note": PEAK_NOTE, "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "compress_fwht + compress_fft (persistent)",
                     "kernel_ms": comp_s * 1e3, "flops_per_launch": flops},
        "detail": {"stage_ms": stages, "input_bytes": nbytes,
                   "hbm_GBps_whole_batch": nbytes / batch_s / 1e9,
                   "sample_products": len(sample), "recall": hits / max(1, truth_n),
                   "precision": hits / max(1, got_n)},
    }


def bench_precision(args):
    """The precision form (FFT, 4096 x 16384 x 4096, b = 2^16, d = 11) in its
    arithmetic modes: FP32 butterflies, FP64 cross-chunk spectral
    accumulation, and matrix-core DFT stages with hi/lo split correction on
    BF16 or (column-scaled) FP16 operands.
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
    modes = {"fp32": 0, "fp64_cross_chunk": cma.FLAG_F64_REDUCE, "bf16_mfma_split": cma.FLAG_BF16_MFMA,
             "fp16_mfma_split": cma.FLAG_FP16_MFMA}
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
                               "four arithmetic modes (FP32, FP64 cross-chunk, BF16 and FP16 matrix-core splits)", "parallelism": "single GPU"},
        "roofline": {"bound": COMPUTE_BOUND, "peak_note": PEAK_NOTE, "achieved": flops / k_s / 1e12, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / k_s / 1e12 / PEAK_FP32_TFLOPS,
                     "traffic": None, "kernel": "compress (fp32 mode)",
                     "kernel_ms": k_s * 1e3, "flops_per_launch": flops},
        "detail": {"modes": out},
    }


SERVE_TEMPLATES = 64
SERVE_SEED = workloads.BATCH_SEED + 100
SERVE_REQ_PER_STEP = 64  # one serving step: a window of 64 requests
SERVE_QPS = 1000.0       # the fixed arrival rate of the serving workload


def serve_templates(count, seed):
    """`count` request templates from the batch shape distribution, operands
    resident in HBM: 30% CSR at 1% density (N(0,1) values plus the planted
    entries; device CSR arrays, and a dense device copy for the
    single-product paths, which take dense operands), the rest dense."""
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
            A, Bs = A.tocsr(), Bs.tocsr()

            def dev_csr(x):
                return tuple(cma.DeviceArray.from_numpy(np.ascontiguousarray(v, dt))
                             for v, dt in ((x.indptr, np.int32), (x.indices, np.int32),
                                           (x.data, np.float32)))
            AT, B = A.toarray(), Bs.toarray()
            out.append(dict(s, csr=True, host=(AT, B), serve=(dev_csr(A), dev_csr(Bs)),
                            dense=(cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B))))
        else:
            AT, B, _ = workloads.planted(m, n, p, s["plants"], s["magnitude"], seed + 2 * i,
                                         np.float32, direct=True)
            dev = (cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B))
            out.append(dict(s, csr=False, host=(AT, B), serve=dev, dense=dev))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads the oracle uses: OMP_NUM_THREADS when set, else the core count
    (at most 16, the host share of one GPU)."""
    env = os.environ.get("OMP_NUM_THREADS")
    n = int(env) if env and env.isdigit() else (os.cpu_count() or 1)
    return max(1, min(16, n))


def time_oracle(products, threads, runs=5):
    """1 warm-up and `runs` timed passes of the FP64 oracle (compress plus the
    full median-of-d decompression) over `products`; the p50 pass time."""
    from oracle import oracle as orc
    def one():
        for q in products:
            sk = orc.compress(q["variant"], q["AT"], q["B"], q["d"], q["b"], HASH_SEED, 0, threads)
            hA, hB = orc.hash_arrays(0, q["d"], q["b"], HASH_SEED, q["AT"].shape[1], q["B"].shape[1])
            orc.decompress(q["variant"], sk, hA, hB, threads)
    t0 = time.perf_counter()
    one()
    ts = [time.perf_counter() - t0] if runs == 0 else []  # runs = 0: the warm-up's time
    for _ in range(runs):
        t0 = time.perf_counter()
        one()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline_serving(templates, budget_gflop=60.0):
    """The oracle on a bounded sample of the serving templates: templates in
    order while their compress work stays within `budget_gflop`."""
    threads = cpu_threads()
    sample, flops = [], 0.0
    for t in templates:
        f = compress_flops(t["variant"], t["m"], t["n"], t["p"], t["d"], t["b"])
        if sample and flops + f > budget_gflop * 1e9:
            continue
        sample.append(dict(variant=t["variant"], d=t["d"], b=t["b"],
                           AT=np.ascontiguousarray(t["host"][0], np.float64),
                           B=np.ascontiguousarray(t["host"][1], np.float64)))
        flops += f
        if flops > budget_gflop * 1e9 * 0.8:
            break
    p50, ts = time_oracle(sample, threads)
    shapes = ", ".join(f"{q['variant']} {q['AT'].shape[1]}x{q['AT'].shape[0]}x{q['B'].shape[1]} "
                       f"b={q['b']} d={q['d']}" for q in sample)
    return {"value": len(sample) / p50, "unit": "products/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"FP64 oracle (C++/OpenMP restatement, not FFTW/MKL), compress plus full "
                      f"decompression of {len(sample)} serving templates ({shapes}); 1 warm-up, "
                      f"p50 of {len(ts)} runs",
            "runs_s": ts}


def cpu_full_configs(threads, names=("fft512_fp64", "fwht2048_fp32")):
    """The CPU timing of the two smaller configurations in full: the FP64
    oracle, 1 warm-up and 5 timed runs (one, when a run exceeds 15 s)."""
    out = {}
    for name in names:
        c = workloads.CONFIGS[name]
        AT, B, _ = workloads.make(name)
        q = dict(variant=c["variant"], d=c["d"], b=c["b"], AT=np.ascontiguousarray(AT, np.float64),
                 B=np.ascontiguousarray(B, np.float64))
        first, _ = time_oracle([q], threads, runs=0)  # the warm-up alone
        runs = 5 if first < 15.0 else 1
        p50, ts = time_oracle([q], threads, runs=runs)
        out[name] = {"products_per_s": 1.0 / p50, "p50_s": p50, "runs_s": ts, "cores": threads}
        print(f"[bench] cpu {name}: p50 {p50:.3f} s over {len(ts)} runs", file=sys.stderr, flush=True)
    return out


def plain_hbm_bytes(variant, m, n, p, d, b, esz=4):
    """HBM bytes the plain path's stages move: the sketches written, the
    transform's read and write (twice for rows past one 128 KB tile), the
    spectra read by the MAC."""
    rows_bytes = d * n * b * esz * 2  # FFT: one complex row; FWHT: an A and a B row
    tile = (1 << 15) if variant == "fwht" else (1 << 14)
    passes = 1 if b <= tile else 2
    return rows_bytes * (1 + 2 * passes + 1)


def serve_arrivals(nreq, qps, templates, seed=7):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, nreq))
    picks = rng.integers(0, len(templates), nreq)
    return arrivals, picks


def bench_serve(args):
    """The serving form: Poisson arrivals at a fixed QPS into the
    continuous-batching server, operands resident in HBM; then the same
    arrivals replayed one product per launch sequence (fused, plain)."""
    t_setup = time.perf_counter()
    templates = serve_templates(SERVE_TEMPLATES, SERVE_SEED)
    print(f"[bench] serving: {len(templates)} templates resident in HBM "
          f"({sum(t['csr'] for t in templates)} CSR) in {time.perf_counter() - t_setup:.1f} s",
          file=sys.stderr, flush=True)
    nreq, qps = args.steps * SERVE_REQ_PER_STEP, args.qps
    arrivals, picks = serve_arrivals(nreq, qps, templates)
    seeds = HASH_SEED + np.arange(nreq)
    grain = args.grain or 64
    max_out = 1 << 16
    sv = cma.Server(hbm_bytes=args.hbm_gb << 30, max_waves=args.waves,
                    max_wave_products=args.wave_products, grain=grain)

    def submit(t, seed):
        a, b = t["serve"]
        shape = (t["m"], t["n"], t["p"]) if t["csr"] else None
        return sv.submit(t["variant"], a, b, t["d"], t["b"], int(seed), tau=t["tau"],
                         max_out=max_out, shape=shape)

    # warm-up: `warmup` windows of requests over every template, then drained
    warm = [submit(templates[k % len(templates)], 1 + k)
            for k in range(args.warmup * SERVE_REQ_PER_STEP)]
    for tk in warm:
        sv.wait(tk)
        sv.release(tk)
    import threading
    cma.timing_enable(True, only="compress*")
    cma.timing_reset()
    tickets = [None] * nreq
    t0 = time.perf_counter() + 0.05

    def feeder():
        for r in range(nreq):
            dt = t0 + arrivals[r] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[r] = submit(templates[picks[r]], seeds[r])

    th = threading.Thread(target=feeder)
    th.start()
    lat = np.zeros(nreq)
    served = {}
    for r in range(nreq):
        while tickets[r] is None:
            time.sleep(1e-4)
        sv.wait(tickets[r])
        ii, jj, _, lat[r], _ = sv.result(tickets[r])
        if picks[r] < 16 and picks[r] not in served:
            served[int(picks[r])] = set(zip(ii.tolist(), jj.tolist()))
        sv.release(tickets[r])
    th.join()
    t_end = time.perf_counter()
    kern = {k: cma.timing_get(k) for k in ("compress_fft", "compress_fwht")}
    cma.timing_enable(False)
    st = sv.stats()
    sv.close()
    served_s = t_end - t0
    print(f"[bench] served {nreq} requests in {served_s:.2f} s", file=sys.stderr, flush=True)

    # roofline of the dominant kernel class over the timed region: the
    # algorithmic FLOPs of the requests' compress work / the summed launch
    # durations (events on each wave's stream)
    flops = {"compress_fft": 0.0, "compress_fwht": 0.0}
    for r in range(nreq):
        t = templates[picks[r]]
        flops["compress_" + t["variant"]] += compress_flops(t["variant"], t["m"], t["n"], t["p"],
                                                           t["d"], t["b"])
    dom = max(kern, key=lambda k: kern[k][0])
    dom_s = kern[dom][0] * 1e-3
    achieved = flops[dom] / dom_s / 1e12 if dom_s > 0 else 0.0

    def one_at_a_time(plain, count):
        """The same arrivals, one product per launch sequence, in order."""
        lat_ = np.zeros(count)
        rec = {}
        t1 = time.perf_counter() + 0.05
        for r in range(count):
            dt = t1 + arrivals[r] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t = templates[picks[r]]
            a, b = t["dense"]
            sk = cma.compress(t["variant"], a, b, t["d"], t["b"], seed=int(seeds[r]),
                              dtype=np.float32, flags=cma.FLAG_PLAIN if plain else 0)
            ii, jj, _ = sk.recover(t["tau"])
            sk.free()
            lat_[r] = time.perf_counter() - (t1 + arrivals[r])
            if picks[r] < 16 and picks[r] not in rec:
                rec[int(picks[r])] = set(zip(ii.tolist(), jj.tolist()))
        return lat_, time.perf_counter() - t1, rec

    q = lambda x, f: float(np.quantile(x, f))
    fused_lat, fused_s, fused_rec = one_at_a_time(False, nreq)
    print(f"[bench] fused one at a time: {nreq / fused_s:.0f} products/s", file=sys.stderr, flush=True)
    nplain = min(nreq, args.plain_requests)
    plain_lat, plain_s, plain_rec = one_at_a_time(True, nplain)
    print(f"[bench] plain one at a time: {nplain / plain_s:.0f} products/s", file=sys.stderr, flush=True)

    # recall of each path against {C_true > tau} on the first templates seen
    def recall(rec):
        hits = truth_n = got_n = 0
        for k, got in rec.items():
            t = templates[k]
            AT, B = t["host"]
            truth = set(zip(*np.nonzero(AT.astype(np.float64).T @ B.astype(np.float64) > t["tau"])))
            hits += len(got & truth)
            truth_n += len(truth)
            got_n += len(got)
        return {"recall": hits / max(1, truth_n), "precision": hits / max(1, got_n),
                "templates": len(rec)}
    common = sorted(set(served) & set(fused_rec) & set(plain_rec))
    rc = {name: recall({k: d[k] for k in common})
          for name, d in (("server", served), ("fused_one_at_a_time", fused_rec),
                          ("plain_one_at_a_time", plain_rec))}
    value = nreq / served_s
    plain_rate = nplain / plain_s
    return {
        "metric": "products/s",
        "value": value,
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * served_s / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: batch shape distribution, 30% CSR at 1% density, planted entries of "
                "5.0 (seeded); operands resident in HBM",
        "config": {"workload": f"serving: Poisson arrivals at {qps:g} QPS, {nreq} requests "
                               f"({args.steps} steps of {SERVE_REQ_PER_STEP}) over "
                               f"{len(templates)} templates (m, p in [64, 4096], n in [128, 8192], "
                               "b 2^10..2^16, d {5,7,9,11,15}, FFT/FWHT 50/50, 30% CSR, tau 12.5)",
                   "qps": qps, "max_waves": args.waves, "max_wave_products": args.wave_products,
                   "grain": grain, "parallelism": "single GPU"},
        "roofline": {"bound": COMPUTE_BOUND, "peak_note": PEAK_NOTE, "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": None, "kernel": f"{dom} (persistent batch compress kernels)",
                     "kernel_ms": 1e3 * dom_s / max(1, kern[dom][1]),
                     "launches": kern[dom][1], "flops_timed_region": flops[dom]},
        "detail": {
            "latency_ms": {"p50": 1e3 * q(lat, 0.5), "p99": 1e3 * q(lat, 0.99),
                           "max": 1e3 * float(lat.max())},
            "vs_plain_one_product_per_launch": value / plain_rate,
            "vs_fused_one_product_per_launch": value / (nreq / fused_s),
            "fused_one_product_per_launch": {"products_per_s": nreq / fused_s,
                                             "latency_ms": {"p50": 1e3 * q(fused_lat, 0.5),
                                                            "p99": 1e3 * q(fused_lat, 0.99)}},
            "plain_one_product_per_launch": {"products_per_s": plain_rate, "requests": nplain,
                                             "latency_ms": {"p50": 1e3 * q(plain_lat, 0.5),
                                                            "p99": 1e3 * q(plain_lat, 0.99)}},
            "recall_equal": rc, "server": st,
            "compress_ms_total": {k: v[0] for k, v in kern.items()},
            "compress_launches": {k: v[1] for k, v in kern.items()},
            "device_tflops_timed_region": (flops["compress_fft"] + flops["compress_fwht"]) / served_s / 1e12,
        },
    }, templates


def bench_single(args, name, plain=False, quiet=False):
    """One named configuration: `steps` products after `warmup`, the compress
    class timed with HIP events; recovered set checked against the dense
    FP64 product after the timed region."""
    c = workloads.CONFIGS[name]
    m, n, p, d, b = c["m"], c["n"], c["p"], c["d"], c["b"]
    dtype = np.dtype(c["dtype"])
    AT, B, _ = workloads.make(name)
    print(f"[bench] workload {name} m={m} n={n} p={p} d={d} b={b}{' plain' if plain else ''}",
          file=sys.stderr, flush=True)
    # `rotate` copies of the operands, used in turn: with more bytes than the
    # 256 MB MALL, each product reads its inputs from HBM (traffic counters)
    rot = max(1, getattr(args, "rotate", 1))
    # --host: operands handed over in host memory and C returned there, as
    # the drop-in entry points take them (the timed products include the
    # transfers over PCIe)
    host = bool(getattr(args, "host", False))
    dATs = [AT] if host else [cma.DeviceArray.from_numpy(AT) for _ in range(rot)]
    dBs = [B] if host else [cma.DeviceArray.from_numpy(B) for _ in range(rot)]
    rot = 1 if host else rot
    dC = np.empty((m, p), dtype) if host else cma.DeviceArray((m, p), dtype)
    flags = cma.FLAG_PLAIN if plain else 0
    sk = None
    calls = 0

    def step():
        nonlocal sk, calls
        dAT, dB = dATs[calls % rot], dBs[calls % rot]
        calls += 1
        sk = cma.compress(c["variant"], dAT, dB, d, b, seed=HASH_SEED, dtype=dtype,
                          flags=flags, sketch=sk)
        sk.decompress(dC)

    for _ in range(args.warmup):
        step()
    cma.synchronize()
    kname = "plain*" if plain else "compress"
    cma.timing_enable(True, only=None if args.kernel_times else kname)
    cma.timing_reset()
    cma.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    cma.synchronize()
    elapsed = time.perf_counter() - t0
    totals = {k: cma.timing_get(k) for k in ("hash", "compress", "reduce", "combine",
                                            "decompress", "plain_sketch", "plain_transform",
                                            "plain_mac")}
    cma.timing_enable(False)
    print("[bench] per-kernel ms/launch: " + ", ".join(
        f"{k}={v[0] / v[1]:.4f}" for k, v in totals.items() if v[1]), file=sys.stderr, flush=True)
    C_est = dC if host else dC.numpy()
    C_true = AT.astype(np.float64).T @ B.astype(np.float64)
    truth = set(zip(*np.nonzero(C_true > c["tau"])))
    ii, jj, _ = sk.recover(c["tau"])
    got = set(zip(ii.tolist(), jj.tolist()))
    est_set = set(zip(*np.nonzero(C_est > c["tau"])))
    recall = len(got & truth) / max(1, len(truth))
    precision = len(got & truth) / max(1, len(got))
    flops = compress_flops(c["variant"], m, n, p, d, b)
    peak = PEAK_FP32_TFLOPS if dtype == np.float32 else PEAK_FP64_TFLOPS
    if plain:  # the stages' time per product
        k_avg_s = sum(totals[k][0] for k in ("plain_sketch", "plain_transform", "plain_mac")) * 1e-3 / args.steps
        nbytes = plain_hbm_bytes(c["variant"], m, n, p, d, b, dtype.itemsize)
        roof = {"bound": "hbm", "achieved": nbytes / k_avg_s / 1e9, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": nbytes / k_avg_s / 1e9 / PEAK_HBM_GBS, "traffic": None,
                "kernel": "plain stages (sketch + transform + MAC)", "kernel_ms": k_avg_s * 1e3,
                "bytes_per_product": nbytes}
    else:
        k_ms, k_n = totals["compress"]
        k_avg_s = (k_ms / max(1, k_n)) * 1e-3
        achieved = flops / k_avg_s / 1e12 if k_avg_s > 0 else 0.0
        roof = {"bound": COMPUTE_BOUND,
                "peak_note": PEAK_NOTE if dtype == np.float32 else PEAK_NOTE_FP64,
                "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": traffic_from_profile(name),
                "traffic_source": "profile constant: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                                  f"(profiles/{name}_compress_traffic.md)",
                "kernel": kname, "kernel_ms": k_avg_s * 1e3, "flops_per_launch": flops}
    return {
        "metric": "products/s",
        "value": args.steps / elapsed,
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if dtype == np.float32 else "f64",
        "data": f"synthetic: N(0,1/{n}) noise + {c['plants']} planted entries of "
                f"{c['magnitude']} (seeded)",
        "config": {"workload": f"{c['variant']} {m}x{n}x{p} {dtype.name}, b={b}, d={d}, "
                               f"{c['plants']} plants, tau={c['tau']}",
                   "m": m, "n": n, "p": p, "b": b, "d": d, "variant": c["variant"],
                   "hash": "multiply add shift", "path": "plain" if plain else "fused",
                   "operands": "host (PCIe-inclusive)" if host else "device-resident",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "detail": {"recall": recall, "precision": precision, "recovered": len(got),
                   "truth": len(truth), "C_threshold_set_equal": est_set == got,
                   "kernel_ms": {k: (v[0] / v[1] if v[1] else None) for k, v in totals.items()}},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps (serve: windows of 64 requests; else products)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain", action="store_true", help="time the unfused path instead")
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD,
                    choices=sorted(workloads.CONFIGS) + ["batch512", "serve", "precision"],
                    help="named workload (default: the metric's configuration, serving)")
    ap.add_argument("--grain", type=int, default=0,
                    help="batch512: work-item grain (0 = autotuned); serve: 0 = 64")
    ap.add_argument("--qps", type=float, default=SERVE_QPS, help="serve: arrival rate")
    ap.add_argument("--waves", type=int, default=3, help="serve: waves on the device at once")
    ap.add_argument("--wave-products", type=int, default=64, help="serve: requests per wave")
    ap.add_argument("--hbm-gb", type=int, default=64, help="serve: server HBM budget (GiB)")
    ap.add_argument("--plain-requests", type=int, default=512,
                    help="serve: arrivals replayed through the plain path")
    ap.add_argument("--no-detail-legs", action="store_true",
                    help="serve: skip the single-product fwht 2048^3 legs in the detail")
    ap.add_argument("--kernel-times", action="store_true",
                    help="time every kernel class with HIP events (default: only the "
                         "dominant one, since each event pair slows the stream)")
    ap.add_argument("--host", action="store_true",
                    help="single-product workloads: operands in host memory and C returned "
                         "there (the drop-in entry points' hand-over; transfers timed)")
    ap.add_argument("--rotate", type=int, default=1, metavar="N",
                    help="single-product workloads: cycle through N device copies of the "
                         "operands (N x their bytes above the 256 MB MALL: inputs come from HBM)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the kernels of known traffic that scale FETCH_SIZE / "
                         "WRITE_SIZE counter readings: 5 device-to-device copies, and with "
                         "torch 5 sums and 5 fills, of MIB MiB each")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the calibration kernels (sustained ALU and HBM rates)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe << 20)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(a product would shard by ranges of the inner dimension, a batch or "
              "a request stream by product)",
              file=sys.stderr)
        sys.exit(2)

    peaks = None
    if not args.no_peaks:
        peaks = achievable_peaks()
        print("[bench] sustained rates: " + ", ".join(f"{k}={v:.1f}" for k, v in peaks.items()),
              file=sys.stderr, flush=True)

    def finish(out, f64=False):
        if peaks:
            add_achievable(out.get("roofline"), peaks, f64)
            out["detail"]["achievable_peaks"] = peaks
        print(json.dumps(out), flush=True)

    name = args.workload
    if name == "batch512":
        finish(bench_batch(args))
        return
    if name == "precision":
        finish(bench_precision(args))
        return
    if name != "serve":
        out = bench_single(args, name, plain=args.plain)
        if not args.no_cpu_baseline and name in ("fft512_fp64", "fwht2048_fp32"):
            full = cpu_full_configs(cpu_threads(), [name])[name]
            out["cpu_baseline"] = {"value": full["products_per_s"], "unit": "products/s",
                                   "cores": full["cores"], "kind": "port", "cpu_model": cpu_model(),
                                   "sample": f"FP64 oracle (C++/OpenMP restatement, not FFTW/MKL): "
                                             f"the full product, 1 warm-up, p50 of "
                                             f"{len(full['runs_s'])} runs"}
        finish(out, f64=out["dtype"] == "f64")
        return

    out, templates = bench_serve(args)
    if not args.no_detail_legs:
        # the fwht 2048^3 single-product workload, fused and plain, in the detail
        import types
        leg = types.SimpleNamespace(steps=100, warmup=10, kernel_times=False)
        f = bench_single(leg, "fwht2048_fp32")
        pl = bench_single(types.SimpleNamespace(steps=5, warmup=1, kernel_times=False),
                          "fwht2048_fp32", plain=True)
        out["detail"]["fwht2048_fp32"] = {
            "products_per_s": f["value"], "roofline": f["roofline"], "recall": f["detail"]["recall"],
            "plain_products_per_s": pl["value"], "plain_roofline": pl["roofline"],
            "fused_vs_plain": f["value"] / pl["value"]}
        if peaks:
            add_achievable(f["roofline"], peaks)
            add_achievable(pl["roofline"], peaks)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_serving(templates)
        out["detail"]["cpu_full_configs"] = cpu_full_configs(cpu_threads())
    finish(out)


if __name__ == "__main__":
    main()
