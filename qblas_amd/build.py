"""Build recipe for libqblas_hip.so (gfx950) -- the only native artefact of the package.

Each ``csrc/*.hip`` file is compiled to an object with hipcc (in parallel), then linked into
``qblas_amd/libqblas_hip.so``.  Objects go to ``qblas_amd/csrc/_build/`` (git-ignored).
Run as ``python -m qblas_amd.build`` or through ``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libqblas_hip.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("QBLAS_ARCH", "gfx950")

# -packed-fp32-ops off: the compiler emits no packed FP32 instructions (v_pk_{add,mul,fma}_f32).
# -ffp-contract=off: HIP's default (fast) fuses a*b + c across statements, which breaks the
# error-free transforms (TwoProd/TwoSum) of the double-double form.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function", "-Wno-bitwise-instead-of-logical",
    "-I", os.path.join(HERE, "..", "include"),
]

SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    # keep warnings visible but short
    warn = "\n".join(l for l in r.stderr.splitlines() if "warning:" in l and "packed-fp32-ops" not in l)
    return obj, warn


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    for _, w in results:
        if w and verbose:
            print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
