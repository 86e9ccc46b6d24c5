"""Build the gfx950 library in-tree with hipcc: ``libfagra_gpu.so`` and ``libfagra_gpu.a``.

Usage: ``python -m fagra_amd.build`` (or ``build()`` from Python). Each source is compiled once
to a position-independent object under ``_build/``; the shared library (loaded by the Python
binding) and the static archive (for hosts that link the C ABI statically, such as a Rust crate
built with ``cargo:rustc-link-lib=static=fagra_gpu``) are made from the same objects. Both
export the C ABI declared in ``include/fagra_gpu.h``. The C driver ``tests/c/lm_bal.c`` is then
linked against the static archive (``tests/c/lm_bal``), which the tests run on the GPU.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ_DIR = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libfagra_gpu.so")
ARCHIVE = os.path.join(HERE, "libfagra_gpu.a")
C_DRIVER = os.path.join(ROOT, "tests", "c", "lm_bal.c")
C_DRIVER_BIN = os.path.join(ROOT, "tests", "c", "lm_bal")
SOURCES = ["fg_device.hip", "fg_mlp.hip", "fg_generic.hip", "fg_covariance.hip", "fg_lm.cpp", "fg_bal.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FAGRA_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # Packed FP32 VALU forms stay off: ordinary float code is compiled to scalar VALU ops.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in line and line.strip())
    if res.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({res.returncode}):\n{out}")
    if verbose and out:
        print(out)


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OBJ_DIR, os.path.basename(s) + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 1)) as pool:
        list(pool.map(lambda so: _run([HIPCC, *FLAGS, "-c", so[0], "-o", so[1]], verbose), zip(srcs, objs)))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    if os.path.exists(ARCHIVE + ".tmp"):
        os.remove(ARCHIVE + ".tmp")
    _run(["ar", "rcs", ARCHIVE + ".tmp", *objs], verbose)
    os.replace(ARCHIVE + ".tmp", ARCHIVE)
    if os.path.exists(C_DRIVER):
        obj = os.path.join(OBJ_DIR, "lm_bal.o")
        _run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", f"-I{os.path.join(ROOT, 'include')}", "-c", C_DRIVER,
              "-o", obj], verbose)
        _run([HIPCC, f"--offload-arch={ARCH}", obj, ARCHIVE, "-o", C_DRIVER_BIN + ".tmp"], verbose)
        os.replace(C_DRIVER_BIN + ".tmp", C_DRIVER_BIN)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
