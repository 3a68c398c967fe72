"""The three larger workloads, checked against the oracle at their real sizes.

* Precision form: A 4096 x 16384, B 16384 x 4096 (AT 16384 x 4096), FFT,
  b = 2^16, d = 11, 512 planted entries of 5.0, in its arithmetic modes
  (FP32 butterflies; FP64 accumulation of the per-chunk spectra; BF16 and
  FP16 matrix-core radix-16 stages with hi/lo split).  Each sketch within 1e-5
  relative L2 of the oracle's FP64 closed form with the same hashes
  (p_t[h] = sum over h1(i) + h2(j) = h of s1 s2 C[i, j]); the recovered set
  equal to {C_true > 12.5} of the dense FP64 product.
* Batched form: the 512 products of the batch distribution in one plan
  (autotuned grain) replayed as a HIP graph; every 16th product's sketch
  within 1e-5 of the closed form, its C bit-exact against the oracle's
  median-of-d decompression of the sketch (array.hpp:323-328), its recovered
  list equal to the oracle estimate's entries above tau, and the replays
  bit-identical.
* Serving form: a Poisson stream at a fixed QPS into the continuous-batching
  server, 30% of the requests CSR at 1% density; every served request's
  sketch, C and recovered list checked as above, and bit-identical to the
  same request served alone.

Reference semantics: matmul_fft.hpp:135-147 and matmul_fwht.hpp:122-134
(decompress), matmul_fft.hpp:25-129 / matmul_fwht.hpp:25-114 (sketch).
"""
import threading
import time

import numpy as np
import pytest
import scipy.sparse as sparse

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SEED = 13794


def rel_l2(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def oracle_check(name, variant, AT64, B64, d, b, seed, values, C, rec, tau, bar=1e-5):
    """Sketch within `bar` of the closed form; C bit-exact against the oracle's
    decompression of `values`; `rec` (rows, cols) equal to its entries > tau."""
    m, p = AT64.shape[1], B64.shape[1]
    hA, hB = orc.hash_arrays(0, d, b, seed, m, p)
    C64 = AT64.T @ B64
    want = orc.sketch_from_product(variant, C64, hA, hB, b)
    err = rel_l2(values.astype(np.float64), want)
    assert err <= bar, f"{name}: sketch rel err {err:.3e}"
    est = orc.decompress(variant, values.astype(np.float64), hA, hB)
    if C is not None:
        assert np.array_equal(C.astype(np.float64), est), f"{name}: C differs from the oracle's"
    if rec is not None:
        got = set(zip(rec[0].tolist(), rec[1].tolist()))
        assert got == set(zip(*np.nonzero(est > tau))), f"{name}: recovered set"
    return err, C64


# ------------------------------------------------------------ precision form
def test_precision_form_modes():
    c = workloads.CONFIGS["fft4096x16384_fp32"]
    m, n, p, d, b, tau = c["m"], c["n"], c["p"], c["d"], c["b"], c["tau"]
    print(f"precision form {m}x{n}x{p} b={b} d={d}", flush=True)
    AT, B, _ = workloads.make("fft4096x16384_fp32")
    AT64, B64 = AT.astype(np.float64), B.astype(np.float64)
    C64 = AT64.T @ B64
    truth = set(zip(*np.nonzero(C64 > tau)))
    hA, hB = orc.hash_arrays(0, d, b, SEED, m, p)
    want = orc.sketch_from_product("fft", C64, hA, hB, b)
    dA, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    modes = {"fp32": 0, "fp64_reduce": cma.FLAG_F64_REDUCE, "bf16_mfma": cma.FLAG_BF16_MFMA,
             "fp16_mfma": cma.FLAG_FP16_MFMA}
    for name, flags in modes.items():
        print(f"  mode {name}", flush=True)
        sk = cma.compress("fft", dA, dB, d, b, seed=SEED, dtype=np.float32, flags=flags)
        got = sk.values()
        e = got - want
        rel = rel_l2(got, want)
        ii, jj, _ = sk.recover(tau)
        rec = set(zip(ii.tolist(), jj.tolist()))
        est = sk.decompress().astype(np.float64)
        print(f"  {name}: sketch rel L2 {rel:.3e}, max abs {np.abs(e).max():.3e}, "
              f"rms {np.sqrt(np.mean(e ** 2)):.3e}; C-hat vs dense max {np.abs(est - C64).max():.3f} "
              f"rms {np.sqrt(np.mean((est - C64) ** 2)):.4f}; recall "
              f"{len(rec & truth) / len(truth):.3f} precision {len(rec & truth) / max(1, len(rec)):.3f}",
              flush=True)
        assert rel <= 1e-5, name
        assert rec == truth, name
        sk.free()


# ------------------------------------------------------------ batched form
def test_batched_form_512_products_graph():
    specs = workloads.batch_specs(512)
    prods, sample = [], {}
    t0 = time.perf_counter()
    for i, s in enumerate(specs):
        AT, B, _ = workloads.planted(s["m"], s["n"], s["p"], s["plants"], s["magnitude"],
                                     workloads.BATCH_SEED + 1 + i, np.float32, direct=True)
        prods.append(cma.Product(s["variant"], cma.DeviceArray.from_numpy(AT),
                                 cma.DeviceArray.from_numpy(B), s["d"], s["b"], SEED + i,
                                 C=cma.DeviceArray((s["m"], s["p"]), np.float32),
                                 tau=s["tau"], max_out=1 << 16))
        if i % 16 == 0:
            sample[i] = (AT.astype(np.float64), B.astype(np.float64))
    print(f"512 products generated in {time.perf_counter() - t0:.1f} s", flush=True)
    bt = cma.Batch(prods, grain=0, graph=True)
    print(f"autotuned grain {bt.grain}", flush=True)
    replays = []
    for r in range(2):
        for i in sample:  # each replay must rewrite the outputs
            prods[i].C.upload(np.full(prods[i].C.shape, np.nan, np.float32))
        bt.run()
        cma.synchronize()
        replays.append({i: (bt.values(i), prods[i].C.numpy(), bt.recovered(i)) for i in sample})
    for i, (AT64, B64) in sample.items():
        s = specs[i]
        vals, C, rec = replays[0][i]
        v2, C2, rec2 = replays[1][i]
        assert np.array_equal(vals, v2) and np.array_equal(C, C2), i
        assert all(np.array_equal(x, y) for x, y in zip(rec, rec2)), i
        err, _ = oracle_check(f"product {i}", s["variant"], AT64, B64, s["d"], s["b"], SEED + i,
                              vals, C, rec, s["tau"])
        print(f"product {i} {s['variant']} {s['m']}x{s['n']}x{s['p']} b={s['b']} d={s['d']}: "
              f"rel err {err:.2e}, recovered {len(rec[0])}", flush=True)
    bt.free()


# ------------------------------------------------------------ serving form
def serving_requests(count, seed):
    """Shapes of the batch distribution; 30% CSR at density 1% with N(0, 1)
    values plus planted entries, the rest dense planted operands."""
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
            out.append(dict(s, csr=True, AT=A, B=Bs, AT64=A.toarray().astype(np.float64),
                            B64=Bs.toarray().astype(np.float64)))
        else:
            AT, B, _ = workloads.planted(m, n, p, s["plants"], s["magnitude"], seed + 2 * i,
                                         np.float32, direct=True)
            out.append(dict(s, csr=False, AT=AT, B=B, AT64=AT.astype(np.float64),
                            B64=B.astype(np.float64)))
    return out


MAX_OUT = 1 << 20  # room for every entry above tau (noise included) of any request


def test_serving_form_poisson_with_csr():
    count, qps = 40, 250.0
    reqs = serving_requests(count, workloads.BATCH_SEED + 300)
    ncsr = sum(r["csr"] for r in reqs)
    print(f"{count} requests, {ncsr} CSR, Poisson at {qps} QPS", flush=True)
    arrivals = np.cumsum(np.random.default_rng(3).exponential(1.0 / qps, count))
    sv = cma.Server(hbm_bytes=32 << 30, max_waves=3, max_wave_products=16, grain=64)
    tickets = [None] * count
    t0 = time.perf_counter() + 0.02

    def feeder():
        for k, r in enumerate(reqs):
            dt = t0 + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets[k] = sv.submit(r["variant"], r["AT"], r["B"], r["d"], r["b"], SEED + k,
                                   tau=r["tau"], max_out=MAX_OUT, want_C=True)

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    mixed = []
    for k, t in enumerate(tickets):
        assert sv.wait(t, timeout=120), k
        ii, jj, vv, lat, C = sv.result(t)
        assert len(ii) < MAX_OUT, k
        mixed.append((sv.values(t), C, ii, jj, vv))
        sv.release(t)
    st = sv.stats()
    print(st, flush=True)
    assert st["requests_done"] == count and st["requests_failed"] == 0
    # each request alone, one at a time: bit-identical to its run in the mix
    for k, r in enumerate(reqs):
        t = sv.submit(r["variant"], r["AT"], r["B"], r["d"], r["b"], SEED + k, tau=r["tau"],
                      max_out=MAX_OUT, want_C=True)
        assert sv.wait(t, timeout=120), k
        ii, jj, vv, _, C = sv.result(t)
        vals = sv.values(t)
        sv.release(t)
        mv, mC, mi, mj, mvv = mixed[k]
        assert np.array_equal(vals, mv) and np.array_equal(C, mC), k
        assert np.array_equal(ii, mi) and np.array_equal(jj, mj) and np.array_equal(vv, mvv), k
    sv.close()
    for k, r in enumerate(reqs):
        mv, mC, mi, mj, _ = mixed[k]
        err, C64 = oracle_check(f"request {k}", r["variant"], r["AT64"], r["B64"], r["d"], r["b"],
                                SEED + k, mv, mC, (mi, mj), r["tau"])
        truth = set(zip(*np.nonzero(C64 > r["tau"])))
        got = set(zip(mi.tolist(), mj.tolist()))
        print(f"request {k} {'csr' if r['csr'] else 'dense'} {r['variant']} "
              f"{r['m']}x{r['n']}x{r['p']} b={r['b']} d={r['d']}: rel err {err:.2e}, "
              f"recall {len(got & truth) / max(1, len(truth)):.3f}", flush=True)
