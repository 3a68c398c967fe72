"""The multiply tool and the C++ known-answer runner on the GPU."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MULTIPLY = os.path.join(ROOT, "compmatmul_amd", "multiply")
RUNNER = os.path.join(ROOT, "tests", "cpp", "tests_runner")


def test_runner_all_cases():
    r = subprocess.run([RUNNER], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout


@pytest.mark.parametrize("algo,extra,tol", [("gemm", [], 0.0),
                                            ("fft", ["-b", "65536", "-d", "9"], 1e-9),
                                            ("fwht", ["-b", "65536", "-d", "9"], 1e-9)])
def test_multiply_end_to_end(tmp_path, algo, extra, tol):
    rng = np.random.default_rng(3)
    A = rng.integers(-5, 6, (40, 30)).astype(float)  # n x m (A transposed)
    B = rng.integers(-5, 6, (40, 20)).astype(float)
    np.save(tmp_path / "A.npy", np.asfortranarray(A))  # Fortran order is converted
    np.save(tmp_path / "B.npy", B)
    out = tmp_path / "C.npy"
    r = subprocess.run([MULTIPLY, str(tmp_path / "A.npy"), str(tmp_path / "B.npy"), str(out),
                        "-a", algo, "--seed", "13794", *extra], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    C = np.load(out)
    assert C.shape == (30, 20) and C.dtype == np.float64
    want = A.T @ B
    # 600 integer entries in 65536 buckets: an entry collides in a repetition
    # with probability ~1%, so the median of 9 is the entry itself up to
    # FP64 transform rounding
    assert np.abs(C - want).max() <= tol * max(1.0, np.abs(want).max())
    # -f overwrites
    r = subprocess.run([MULTIPLY, str(tmp_path / "A.npy"), str(tmp_path / "B.npy"), str(out),
                        "-a", "gemm", "-f"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and np.array_equal(np.load(out), want)
