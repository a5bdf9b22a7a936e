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


def _decode(hi, lo):
    """(sign, m, e) with value = (-1)^sign m 2^e for a finite binary128."""
    be = (hi >> 48) & 0x7FFF
    frac = ((hi & ((1 << 48) - 1)) << 64) | lo
    if be == 0:
        return hi >> 63, frac, 1 - 16383 - 112
    return hi >> 63, frac | (1 << 112), be - 16383 - 112


def test_sqrt_is_correctly_rounded(tmp_path):
    # qb::sqrt_packed (the square root of Vector::norm / quadblas_qnrm2) against the definition of
    # round-to-nearest: r = m 2^e with (2m-1)^2 2^(2e) <= 4a <= (2m+1)^2 2^(2e), in exact integers
    exe = tmp_path / "q128_host_check"
    src = os.path.join(ROOT, "tests", "cpu", "q128_host_check.cpp")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-o", str(exe), src, "-lquadmath"], check=True)
    out = subprocess.run([str(exe), "sqrt", "7", "20000"], capture_output=True, text=True, check=True).stdout
    n = 0
    for line in out.splitlines():
        ahi, alo, rhi, rlo = (int(v, 16) for v in line.split())
        sa, ma, ea = _decode(ahi, alo)
        sr, mr, er = _decode(rhi, rlo)
        assert sa == 0 and sr == 0 and mr >> 112 == 1, line        # positive, normal result
        # 4a vs (2m -+ 1)^2 2^(2e): scale both sides to a common power of two
        lo_side, hi_side = (2 * mr - 1) ** 2, (2 * mr + 1) ** 2
        sh = ea + 2 - 2 * er                                          # 4a = ma 2^(ea+2)
        lhs = ma << sh if sh >= 0 else ma
        lo_s, hi_s = (lo_side, hi_side) if sh >= 0 else (lo_side << -sh, hi_side << -sh)
        assert lo_s <= lhs <= hi_s, line
        assert lhs != lo_s and lhs != hi_s, line                        # never a tie
        n += 1
    assert n > 15000


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


def test_ozaki_epilogue_top_bits(tmp_path):
    """The OZAKI finish's epilogue with general alpha, beta (exact.hpp finish_reg_top: the sum cut to
    its top 128 bits before the product with alpha) against the full-width exact epilogue
    (finish_reg_t), on random, alpha = +-1, subnormal-alpha, beta*y = 0 and cancelling cases:
    bit-identical when nothing was cut, else within one ulp plus the reported bound 2^terr."""
    exe = tmp_path / "epilogue_host_check"
    src = os.path.join(ROOT, "tests", "cpu", "epilogue_host_check.cpp")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-o", str(exe), src, "-lquadmath"], check=True)
    for seed in (1, 2):
        r = subprocess.run([str(exe), str(seed), "100000"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-3000:]
        assert " mismatches 0 " in r.stdout

