"""CPU checks of the shipped library that need no GPU:
  * the binary128 core (qblas_amd/csrc/q128.hpp, the same source the kernels use) compiled for the
    host and compared bit-for-bit with libquadmath on random, structured (ties, all-ones, carries,
    boundary exponent offsets), subnormal/overflow and special operands, plus conversions;
  * libqblas_hip.so loads and exports every function include/qblas_hip.h declares.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_q128_core_matches_libquadmath(tmp_path):
    exe = tmp_path / "q128_host_check"
    src = os.path.join(ROOT, "tests", "cpu", "q128_host_check.cpp")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-o", str(exe), src, "-lquadmath"], check=True)
    for seed in (1, 2):
        r = subprocess.run([str(exe), str(seed), "400000"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-3000:]
        assert "mismatches 0 " in r.stdout


def _declared_functions():
    hdr = open(os.path.join(ROOT, "include", "qblas_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b((?:quadblas|qblas)_\w+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol(Q):
    names = _declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", Q.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = Q.lib()
    for n in names:
        assert hasattr(lib, n)


def test_host_side_state(Q):
    assert Q.get_version().startswith("QuadBLAS 1.0.0")
    Q.set_num_threads(5)
    assert Q.get_num_threads() == 5
    Q.set_num_threads(0)
    assert Q.get_num_threads() >= 1
    assert Q.get_kc() == 126
    Q.set_kc(256)
    assert Q.get_kc() == 256
    Q.set_kc(126)
    assert Q.is_aligned(64) == 1 and Q.is_aligned(48) == 0
    with pytest.raises(RuntimeError):
        Q.set_mode(9)
    m = Q.get_mode()
    Q.set_mode("exact")
    assert Q.get_mode() == Q.MODE_EXACT
    Q.set_mode(m)


def test_exact_pipeline_matches_oracle(O, tmp_path):
    """Host run of the EXACT form's arithmetic (qblas_amd/csrc/exact.hpp) with the kernels' wave and
    partial structure, bit-for-bit against the oracle's exact accumulator, incl. crafted ties that
    only a far-away beta*y breaks and the signs of exact zeros."""
    exe = tmp_path / "exact_host_check"
    src = os.path.join(ROOT, "tests", "cpu", "exact_host_check.cpp")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-o", str(exe), src, "-L" + odir, "-loracle",
                    "-Wl,-rpath," + odir, "-lquadmath"], check=True)
    for seed in (1, 2):
        r = subprocess.run([str(exe), str(seed), "200"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-3000:]
        assert " mismatches 0" in r.stdout
