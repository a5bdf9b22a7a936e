"""The C++ drop-in: a program written against the reference's C++ API and C ABI
(include/quadblas/quadblas.hpp over libqblas_hip.so) compiles with g++ -Wall -Wextra -Werror (CPU),
and on the GPU passes the reference's own known-answer cases bit for bit
(tests/cpu/cpp_dropin_check.cpp); the reference's public names and values are all there
(tests/cpu/cpp_names_check.cpp, CPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "qblas_amd")


def _build(tmp_path, name="cpp_dropin_check"):
    if not os.path.exists(os.path.join(LIBDIR, "libqblas_hip.so")):
        pytest.skip("libqblas_hip.so not built")
    exe = tmp_path / name
    src = os.path.join(ROOT, "tests", "cpu", name + ".cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), src, "-L", LIBDIR, "-lqblas_hip", "-Wl,-rpath," + LIBDIR], check=True)
    return exe


def test_cpp_dropin_compiles(tmp_path):
    assert _build(tmp_path).exists()


def test_cpp_public_names(tmp_path):
    """Every public name of cpp_classes.hpp, types.hpp, constants.hpp and quadblas.hpp's version
    block, used as a caller of the reference would (copy-initialisation of Vector included), with
    the reference's values (tests/cpu/cpp_names_check.cpp).  Runs without a GPU."""
    exe = _build(tmp_path, "cpp_names_check")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_cpp_dropin_known_answers_on_gpu(tmp_path):
    exe = _build(tmp_path)
    print("cpp drop-in check", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
