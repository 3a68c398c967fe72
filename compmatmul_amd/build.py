"""Builds the native parts of compmatmul_amd in-tree.

* ``libcompmatmul_amd.so``: HIP kernels for gfx950 plus the C-ABI
  (hipcc, one object per source, compiled in parallel).
* ``compmatmul<ext>``: the pybind11 drop-in module ``compmatmul``, linked
  against the C-ABI library with an ``$ORIGIN`` rpath.
* ``multiply``: the command-line tool (cli/), and ``tests/cpp/tests_runner``,
  the C++ known-answer test executable, both linked against the library.

Usage: ``python -m compmatmul_amd.build`` (or ``build()`` from Python);
``--changed`` recompiles only the sources changed since their objects.
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
LIB = os.path.join(HERE, "libcompmatmul_amd.so")
EXT = os.path.join(HERE, "compmatmul" + sysconfig.get_config_var("EXT_SUFFIX"))
MULTIPLY = os.path.join(HERE, "multiply")
RUNNER = os.path.join(ROOT, "tests", "cpp", "tests_runner")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 VALU instructions stay disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]

SOURCES = ["hash.hip", "balance.hip", "tables.hip", "noalloc.hip", "compress.hip", "compress_reg.hip", "compress_grp.hip", "compress_fft_reg.hip", "compress_fft_grp.hip", "decompress.hip",
           "plain.hip", "sparse.hip", "fft_mfma.hip", "api.cpp", "batch.cpp", "serve.cpp", "host_hash.cpp", "host_copy.cpp", "probe.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src, changed_only=False):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if changed_only and os.path.exists(out):
        # skip an object newer than its source and every shared header
        deps = [path] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
        deps.append(os.path.join(ROOT, "include", "compmatmul_amd.h"))
        if os.path.getmtime(out) > max(os.path.getmtime(d) for d in deps):
            return out
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC] + HIP_FLAGS + lang + ["-c", path, "-o", out])
    return out


def build(verbose=False, changed_only=False):
    """Compile and link everything; changed_only=True reuses up-to-date objects."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, changed_only), srcs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs)
    os.replace(LIB + ".tmp", LIB)
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
          "-I", pybind11.get_include(), "-I", py_inc,
          os.path.join(CSRC, "pybind_compmatmul.cpp"),
          "-L", HERE, "-lcompmatmul_amd", "-Wl,-rpath,$ORIGIN",
          "-o", EXT + ".tmp"])
    os.replace(EXT + ".tmp", EXT)
    cli = os.path.join(HERE, "cli")
    _run(["g++", "-O2", "-std=c++17", "-I", ROOT, os.path.join(cli, "multiply.cpp"),
          os.path.join(cli, "npy.cpp"), "-L", HERE, "-lcompmatmul_amd", "-Wl,-rpath,$ORIGIN",
          "-o", MULTIPLY + ".tmp"])
    os.replace(MULTIPLY + ".tmp", MULTIPLY)
    _run(["g++", "-O2", "-std=c++17", "-I", ROOT, os.path.join(ROOT, "tests", "cpp", "runner.cpp"),
          os.path.join(cli, "npy.cpp"), "-L", HERE, "-lcompmatmul_amd",
          "-Wl,-rpath,$ORIGIN/../../compmatmul_amd", "-o", RUNNER + ".tmp"])
    os.replace(RUNNER + ".tmp", RUNNER)
    if verbose:
        print("built", LIB, EXT, MULTIPLY, "and", RUNNER)
    return LIB


if __name__ == "__main__":
    build(verbose=True, changed_only="--changed" in sys.argv)
    sys.exit(0)
