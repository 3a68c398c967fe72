"""The request server (cm_server_*): continuous batching on one GPU.

Bars: every request's result (recovered entries and C) bit-identical
whether it runs alone or mixed with other requests in overlapping waves;
identical to the batch path with the same grain (itself pinned to the
oracle in test_batch_gpu.py); CSR operands give the dense result bit for
bit, and their sketch, C and recovered set match the oracle; device-resident
requests give the host requests' results; the retained sketch decompresses
to the same C; admission respects the HBM budget.  The Poisson serving form
is checked against the oracle request by request in test_configs_gpu.py.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRAIN = 64
# variant, m, n, p, d, b
SHAPES = [("fwht", 300, 400, 250, 5, 1 << 10), ("fft", 260, 300, 310, 7, 1 << 12),
          ("fwht", 700, 520, 640, 9, 1 << 14), ("fft", 128, 900, 96, 11, 1 << 11),
          ("fft", 400, 200, 380, 5, 1 << 15), ("fwht", 90, 1500, 1200, 15, 1 << 13)]


def data(k, m, n, p):
    AT, B, _ = workloads.planted(m, n, p, 24, 5.0, 500 + k, np.float32)
    return AT, B


def run_alone(sv, reqs):
    out = []
    for (v, m, n, p, d, b), (AT, B), seed in reqs:
        t = sv.submit(v, AT, B, d, b, seed, tau=12.5, max_out=1 << 15, want_C=True)
        assert sv.wait(t, timeout=120)
        out.append(sv.result(t))
        sv.release(t)
    return out


def test_alone_equals_in_mix_and_batch_path():
    reqs = [(s, data(k, *s[1:4]), 4000 + k) for k, s in enumerate(SHAPES)]
    sv = cma.Server(hbm_bytes=2 << 30, max_waves=3, max_wave_products=2, grain=GRAIN)
    alone = run_alone(sv, reqs)
    tickets = [sv.submit(v, AT, B, d, b, seed, tau=12.5, max_out=1 << 15, want_C=True)
               for (v, m, n, p, d, b), (AT, B), seed in reqs]
    for k, t in enumerate(tickets):
        print(f"request {k}: {SHAPES[k]}", flush=True)
        assert sv.wait(t, timeout=120)
        ii, jj, vv, lat, C = sv.result(t)
        a = alone[k]
        assert np.array_equal(ii, a[0]) and np.array_equal(jj, a[1]) and np.array_equal(vv, a[2])
        assert np.array_equal(C, a[4])
        assert lat > 0
        # the retained sketch decompresses to the same C
        assert np.array_equal(sv.decompress(t), C)
        sv.release(t)
    st = sv.stats()
    print(st)
    assert st["requests_done"] == 2 * len(SHAPES) and st["requests_failed"] == 0
    assert st["max_waves_in_flight"] >= 2  # waves overlapped on the device
    sv.close()
    # same grain through the batch path: bit-identical C
    for k, ((v, m, n, p, d, b), (AT, B), seed) in enumerate(reqs):
        C = cma.DeviceArray((m, p), np.float32)
        bt = cma.Batch([cma.Product(v, cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B),
                                    d, b, seed, C=C)], grain=GRAIN)
        bt.run()
        assert np.array_equal(C.numpy(), alone[k][4]), k
        bt.free()


def test_csr_operands_match_dense():
    rng = np.random.default_rng(9)
    m, n, p = 500, 800, 450
    A = sp.random(n, m, density=0.01, format="csr", random_state=1, dtype=np.float32,
                  data_rvs=lambda k: rng.standard_normal(k).astype(np.float32))
    Bs = sp.random(n, p, density=0.01, format="csr", random_state=2, dtype=np.float32,
                   data_rvs=lambda k: rng.standard_normal(k).astype(np.float32))
    A, Bs = A.tolil(), Bs.tolil()
    for t in range(20):  # planted entries
        i, j, k = rng.integers(m), rng.integers(p), rng.integers(n)
        A[k, i] = 5.0
        Bs[k, j] = 5.0
    A, Bs = A.tocsr(), Bs.tocsr()
    sv = cma.Server(hbm_bytes=1 << 30, max_waves=2, grain=GRAIN)
    for v in ("fft", "fwht"):
        t1 = sv.submit(v, A, Bs, 7, 1 << 12, 77, tau=12.5, want_C=True)
        t2 = sv.submit(v, A.toarray(), Bs.toarray(), 7, 1 << 12, 77, tau=12.5, want_C=True)
        assert sv.wait(t1, timeout=120) and sv.wait(t2, timeout=120)
        r1, r2 = sv.result(t1), sv.result(t2)
        assert np.array_equal(r1[4], r2[4])
        assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])
        # against the oracle: the sketch within 1e-5 of the closed form, the
        # recovered set exactly the oracle estimate's entries above tau
        hA, hB = orc.hash_arrays(0, 7, 1 << 12, 77, m, p)
        C64 = A.T.toarray().astype(np.float64) @ Bs.toarray().astype(np.float64)
        want = orc.sketch_from_product(v, C64, hA, hB, 1 << 12)
        vals = sv.values(t1).astype(np.float64)
        assert np.linalg.norm(vals - want) <= 1e-5 * np.linalg.norm(want)
        est = orc.decompress(v, vals, hA, hB)
        assert np.array_equal(r1[4].astype(np.float64), est)
        assert set(zip(r1[0].tolist(), r1[1].tolist())) == set(zip(*np.nonzero(est > 12.5)))
        sv.release(t1)
        sv.release(t2)
    sv.close()


def test_admission_by_footprint():
    # the footprint of one request, measured alone; then a budget of two and
    # a half footprints admits at most two requests at a time
    ops = [data(40 + k, 256, 512, 256) for k in range(6)]
    sv = cma.Server(hbm_bytes=1 << 30, max_waves=3, max_wave_products=8, grain=GRAIN)
    t = sv.submit("fwht", *ops[0], 5, 1 << 12, 0, tau=12.5)
    assert sv.wait(t, timeout=120)
    sv.release(t)
    one = sv.stats()["peak_bytes"]
    sv.close()
    assert 0 < one < 64 << 20
    budget = one * 5 // 2
    sv = cma.Server(hbm_bytes=budget, max_waves=3, max_wave_products=8, grain=GRAIN)
    tickets = [sv.submit("fwht", AT, B, 5, 1 << 12, k, tau=12.5) for k, (AT, B) in enumerate(ops)]
    for t in tickets:
        assert sv.wait(t, timeout=120)
        sv.release(t)
    st = sv.stats()
    assert st["requests_done"] == 6 and st["peak_bytes"] <= budget
    # a request larger than the whole budget fails instead of waiting forever
    AT, B = data(99, 2048, 2048, 2048)
    t = sv.submit("fwht", AT, B, 15, 1 << 14, 1, tau=18.0)
    with pytest.raises(MemoryError):
        sv.wait(t, timeout=60)
    sv.close()


def test_bad_csr_rejected():
    sv = cma.Server(hbm_bytes=64 << 20, grain=GRAIN)
    ip = np.array([0, 1, 2], np.int32)
    ix = np.array([0, 9], np.int32)  # column 9 out of range for m = 4
    vals = np.ones(2, np.float32)
    with pytest.raises(ValueError):
        sv.submit("fft", (ip, ix, vals), (ip, np.array([0, 1], np.int32), vals), 3, 64, 1,
                  shape=(4, 2, 4))
    sv.close()


def test_device_resident_requests_match_host_requests():
    # dense and CSR operands resident in HBM, read in place, C written into
    # a device array: the same sketch, C and recovered list as the request
    # with host operands
    rng = np.random.default_rng(21)
    m, n, p = 420, 650, 380
    A = sp.random(n, m, density=0.01, format="csr", random_state=5, dtype=np.float32,
                  data_rvs=lambda k: rng.standard_normal(k).astype(np.float32))
    Bs = sp.random(n, p, density=0.01, format="csr", random_state=6, dtype=np.float32,
                   data_rvs=lambda k: rng.standard_normal(k).astype(np.float32))
    AT, B = data(77, m, n, p)
    sv = cma.Server(hbm_bytes=1 << 30, max_waves=2, grain=GRAIN)
    dev_csr = lambda x: tuple(cma.DeviceArray.from_numpy(np.ascontiguousarray(v, dt))
                              for v, dt in ((x.indptr, np.int32), (x.indices, np.int32),
                                            (x.data, np.float32)))
    cases = [("dense", (AT, B), (cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)), None),
             ("csr", (A, Bs), (dev_csr(A), dev_csr(Bs)), (m, n, p))]
    for v in ("fft", "fwht"):
        for name, host, dev, shape in cases:
            print(f"{v} {name}", flush=True)
            th = sv.submit(v, host[0], host[1], 7, 1 << 13, 31, tau=12.5, want_C=True)
            dC = cma.DeviceArray((m, p), np.float32)
            td = sv.submit(v, dev[0], dev[1], 7, 1 << 13, 31, tau=12.5, C=dC, shape=shape)
            assert sv.wait(th, timeout=120) and sv.wait(td, timeout=120)
            rh, rd = sv.result(th), sv.result(td)
            assert np.array_equal(sv.values(th), sv.values(td))
            assert np.array_equal(rh[4], dC.numpy())
            assert all(np.array_equal(x, y) for x, y in zip(rh[:3], rd[:3]))
            sv.release(th)
            sv.release(td)
    sv.close()
