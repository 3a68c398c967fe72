"""The multiply tool and the C++ known-answer runner, on the CPU.

The tool's argument and input checks run before any device work, so their
messages and exit codes are checked here (multiply.cpp:91-166 of the
reference); the runner's [cpu] cases cover the random streams, drawn hash
parameters, tabulation tables and the .npy reader/writer.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MULTIPLY = os.path.join(ROOT, "compmatmul_amd", "multiply")
RUNNER = os.path.join(ROOT, "tests", "cpp", "tests_runner")


def run(*args):
    r = subprocess.run([MULTIPLY, *map(str, args)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stderr


def test_runner_cpu_cases():
    r = subprocess.run([RUNNER, "[cpu]"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout


@pytest.fixture
def files(tmp_path):
    A = np.arange(1, 17, dtype=float).reshape(4, 4).T.copy()
    B = np.arange(17, 33, dtype=float).reshape(4, 4)
    np.save(tmp_path / "A.npy", A)
    np.save(tmp_path / "B.npy", B)
    np.save(tmp_path / "A32.npy", A.astype(np.float32))
    np.save(tmp_path / "A3.npy", np.ones((2, 2, 2)))
    np.save(tmp_path / "B5.npy", np.ones((5, 3)))
    np.save(tmp_path / "I8.npy", np.ones((4, 4), np.int64))
    return tmp_path


@pytest.mark.parametrize("extra,msg", [
    (["-a", "fft"], "Got algorithm `fft' but the parameter b is not set!"),
    (["-a", "fwht", "-b", "12", "-d", "3"], "The parameter b must be a positive power of 2 (got 12)"),
    (["-a", "fft", "-b", "0", "-d", "3"], "The parameter b must be a positive power of 2 (got 0)"),
    (["-a", "fft", "-b", "16"], "Got algorithm `fft' but the parameter d is not set!"),
    (["-a", "fwht", "-b", "16", "-d", "0"], "The parameter d must be positive (got 0)"),
    (["-a", "qr"], "does not meet constraint"),
])
def test_parameter_checks(files, extra, msg):
    rc, err = run(files / "A.npy", files / "B.npy", files / "C.npy", *extra)
    assert rc == 1 and msg in err
    assert not (files / "C.npy").exists()


def test_input_checks(files):
    rc, err = run(files / "A32.npy", files / "B.npy", files / "C.npy")
    assert rc == 1 and "Word size mismatch: A is expected to have word size of 8, but got 4" in err
    rc, err = run(files / "A3.npy", files / "B.npy", files / "C.npy")
    assert rc == 1 and "Dimension mismatch: A has 3 dims, but 2 expected!" in err
    rc, err = run(files / "A.npy", files / "B5.npy", files / "C.npy")
    assert rc == 1 and "Dimension mismatch: B is expected to have 4 rows, but got 5 instead" in err
    rc, err = run(files / "I8.npy", files / "B.npy", files / "C.npy")
    assert rc == 1 and "little-endian float64" in err
    rc, err = run(files / "A.npy", files / "missing.npy", files / "C.npy")
    assert rc == 1 and "cannot open" in err


def test_existing_output_and_arguments(files):
    (files / "C.npy").write_bytes(b"x")
    rc, err = run(files / "A.npy", files / "B.npy", files / "C.npy")
    assert rc == 1 and "already exists!" in err
    rc, err = run(files / "A.npy", files / "B.npy")
    assert rc == 1 and "Required argument missing" in err
    rc, err = run(files / "A.npy", files / "B.npy", files / "C.npy", "-b", "x")
    assert rc == 1 and "for arg -b" in err
