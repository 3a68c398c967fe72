"""Contexts (cm_context_create and the cm_ctx_* entry points) on the GPU.

A context runs the same kernels as the default context on its own stream and
buffers, so its sketches equal the default path's bit for bit (the same
chunking of k whenever its workspace leaves room for the default's partial
spectra), and both equal the oracle's closed form within the FP32 bar.  Two
contexts driven from two threads at once give the results each gives alone.
A workspace budget carves every per-call intermediate from one region: a
product that needs more raises MemoryError (CM_ENOMEM) naming the size, the
context stays usable, and the paths that allocate their own buffers are
refused.
"""
import threading

import numpy as np
import pytest

import compmatmul_amd as cma
from compmatmul_amd import workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SEED = 13794


def planted(m, n, p, seed, dtype=np.float32):
    AT, B, _ = workloads.planted(m, n, p, 8, 5.0, seed, dtype)
    return AT, B


@pytest.mark.parametrize("variant,b", [("fwht", 1 << 12), ("fft", 1 << 14), ("fft", 1 << 10)])
@pytest.mark.parametrize("budget", [0, 256 << 20])
def test_context_equals_default_path(variant, b, budget):
    m, n, p, d = 300, 200, 260, 7
    AT, B = planted(m, n, p, 3)
    ref = cma.compress(variant, AT, B, d, b, seed=SEED, dtype=np.float32)
    ctx = cma.Context(0, budget)
    sk = ctx.compress(variant, AT, B, d, b, seed=SEED, dtype=np.float32)
    got, want = sk.values(), ref.values()
    print(f"{variant} b={b} budget={budget >> 20} MiB: max |diff| {np.abs(got - want).max():.3e}")
    assert np.array_equal(got, want)
    hA, hB = sk.hashes()
    C64 = AT.astype(np.float64).T @ B.astype(np.float64)
    closed = orc.sketch_from_product(variant, C64, hA, hB, b)
    assert np.linalg.norm(got - closed) / np.linalg.norm(closed) <= 1e-5
    assert np.array_equal(sk.decompress(), ref.decompress())
    ri, rj, rv = ref.recover(12.5)
    ci, cj, cv = sk.recover(12.5)
    assert np.array_equal(ri, ci) and np.array_equal(rj, cj) and np.array_equal(rv, cv)
    # device operands and device C through the context
    dAT, dB = cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B)
    dC = cma.DeviceArray((m, p), np.float32)
    sk2 = ctx.compress(variant, dAT, dB, d, b, seed=SEED, dtype=np.float32)
    sk2.decompress(dC)
    ctx.synchronize()
    assert np.array_equal(dC.numpy(), ref.decompress())
    for s in (sk, sk2, ref):
        s.free()
    ctx.destroy()


def test_two_contexts_from_two_threads():
    shapes = [("fwht", 700, 300, 500, 9, 1 << 13), ("fft", 500, 260, 640, 5, 1 << 14),
              ("fwht", 129, 77, 2100, 15, 1 << 12), ("fft", 1024, 512, 96, 11, 1 << 12)]
    inputs = [planted(m, n, p, 40 + i) for i, (_, m, n, p, _, _) in enumerate(shapes)]
    alone = []
    for (v, m, n, p, d, b), (AT, B) in zip(shapes, inputs):
        sk = cma.compress(v, AT, B, d, b, seed=SEED, dtype=np.float32)
        alone.append((sk.values(), sk.decompress()))
        sk.free()
    ctxs = [cma.Context(0), cma.Context(0, 512 << 20)]
    results = [[None] * len(shapes) for _ in ctxs]
    errors = []

    def run(ci):
        try:
            for rep in range(3):
                for k, ((v, m, n, p, d, b), (AT, B)) in enumerate(zip(shapes, inputs)):
                    sk = ctxs[ci].compress(v, AT, B, d, b, seed=SEED, dtype=np.float32)
                    results[ci][k] = (sk.values(), sk.decompress())
                    sk.free()
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(ci,)) for ci in range(len(ctxs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for ci in range(len(ctxs)):
        for k in range(len(shapes)):
            vals, C = results[ci][k]
            assert np.array_equal(vals, alone[k][0]), (ci, k)
            assert np.array_equal(C, alone[k][1]), (ci, k)
    for c in ctxs:
        c.destroy()


def test_workspace_budget():
    ctx = cma.Context(0, 8 << 20)
    small = planted(200, 64, 180, 5)
    sk = ctx.compress("fwht", *small, 5, 1 << 12, seed=SEED, dtype=np.float32)
    ref = cma.compress("fwht", *small, 5, 1 << 12, seed=SEED, dtype=np.float32)
    assert np.array_equal(sk.values(), ref.values())
    # host operands of 2 x 16 MB cannot be staged in 8 MiB
    big = planted(2048, 2048, 2048, 6)
    with pytest.raises(MemoryError, match="MiB context workspace"):
        ctx.compress("fft", *big, 5, 1 << 14, seed=SEED, dtype=np.float32)
    # the context is still usable after the failure
    sk2 = ctx.compress("fwht", *small, 5, 1 << 12, seed=SEED, dtype=np.float32, sketch=sk)
    assert np.array_equal(sk2.values(), ref.values())
    # recovery buffers of a huge max_out do not fit either
    with pytest.raises(MemoryError):
        sk2.recover(12.5, max_out=1 << 22)
    ii, jj, _ = sk2.recover(12.5)
    ri, rj, _ = ref.recover(12.5)
    assert np.array_equal(ii, ri) and np.array_equal(jj, rj)
    # the paths with buffers of their own are refused on a budget
    for flags in (cma.FLAG_PLAIN, cma.FLAG_BF16_MFMA, cma.FLAG_FP16_MFMA):
        with pytest.raises(cma.CMError):
            ctx.compress("fft", *small, 5, 1 << 12, seed=SEED, dtype=np.float32, flags=flags)
    for s in (sk2, ref):
        s.free()
    ctx.destroy()


@pytest.mark.parametrize("hash", ["tabulation16", "tabulation8", "multiply shift"])
def test_budget_context_hash_families(hash):
    # tables and parameters are uploaded into the workspace
    AT, B = planted(3000, 64, 700, 8)
    ctx = cma.Context(0, 128 << 20)
    sk = ctx.compress("fwht", AT, B, 16, 1 << 12, seed=SEED, hash=hash, dtype=np.float32)
    ref = cma.compress("fwht", AT, B, 16, 1 << 12, seed=SEED, hash=hash, dtype=np.float32)
    for a, r in zip(sk.hashes(), ref.hashes()):
        assert np.array_equal(a, r)
    assert np.array_equal(sk.values(), ref.values())
    sk.free()
    ref.free()
    ctx.destroy()


def test_float64_context_exact_on_integers():
    # FWHT of integer inputs is exact in FP64 on every path
    rng = np.random.default_rng(9)
    AT = rng.integers(-3, 4, (40, 33)).astype(np.float64)
    B = rng.integers(-3, 4, (40, 29)).astype(np.float64)
    ctx = cma.Context(0, 64 << 20)
    sk = ctx.compress("fwht", AT, B, 5, 256, seed=SEED)
    hA, hB = sk.hashes()
    want = orc.sketch_from_product("fwht", AT.T @ B, hA, hB, 256)
    assert np.array_equal(sk.values(), want)
    sk.free()
    ctx.destroy()


def test_sketch_outlives_its_context():
    AT, B = planted(150, 90, 120, 11)
    ctx = cma.Context(0)
    sk = ctx.compress("fft", cma.DeviceArray.from_numpy(AT), cma.DeviceArray.from_numpy(B), 5,
                      1 << 12, seed=SEED, dtype=np.float32)
    ref = cma.compress("fft", AT, B, 5, 1 << 12, seed=SEED, dtype=np.float32)
    ctx.destroy()
    sk.context = None  # the default context reads and decompresses it now
    assert np.array_equal(sk.values(), ref.values())
    assert np.array_equal(sk.decompress(), ref.decompress())
    sk.free()
    ref.free()
