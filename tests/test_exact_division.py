"""The division-free slab test divides with div_cr1, one correction step
after a reciprocal (rta_amd/csrc/hip/exact.h); div_cr is the two-step form
Markstein's theorem covers.  These tests check both bit for bit against
IEEE-754 division: on the host with hardware FMA (random operands over the
admitted ranges, slab-like operands, quotients next to representable values,
and one significand pair at every admitted exponent pair, the case the
scaling argument extends the enumeration to), and on the GPU against the
device's own division, div_cr1 over all 2^46 significand pairs."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import pytest

import rta_amd

HERE = os.path.dirname(os.path.abspath(__file__))


def _build_checker(tmp_path):
    exe = str(tmp_path / "exact_division_check")
    subprocess.run(["g++", "-O2", "-mfma", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "exact_division_check.cpp")], check=True)
    return exe


def test_div_cr_and_div_cr1_match_ieee_on_host(tmp_path):
    exe = _build_checker(tmp_path)
    out = subprocess.run([exe, "30000000", "11"], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr + out.stdout
    checked, bad, bad1 = [int(x) for x in out.stdout.split()[1::2]]
    assert checked > 50_000_000 and bad == 0 and bad1 == 0


def test_checker_detects_outside_range(tmp_path):
    """Sanity of the checker itself: dropping the operand-range guard must
    produce mismatches (overflowing intermediates)."""
    src = open(os.path.join(HERE, "exact_division_check.cpp")).read()
    src = src.replace("if (b == 0.0f || !divisor_fast_ok(b) || !dividend_fast_ok(a)) return;",
                      "if (b == 0.0f) return;")
    src = src.replace("% 121) - 60", "% 250) - 125").replace("% 157) - 96", "% 250) - 125")
    src = src.replace('"../rta_amd/csrc/hip/exact.h"',
                      '"' + os.path.join(os.path.dirname(HERE), "rta_amd/csrc/hip/exact.h") + '"')
    path = tmp_path / "unguarded.cpp"
    path.write_text(src)
    exe = str(tmp_path / "unguarded")
    subprocess.run(["g++", "-O2", "-mfma", "-ffp-contract=off", "-o", exe, str(path)], check=True)
    out = subprocess.run([exe, "2000000", "3"], capture_output=True, text=True)
    assert out.returncode == 1


@pytest.mark.gpu
def test_div_cr_matches_ieee_on_gpu(gpu):
    bad = C.c_uint64(0)
    ex = (C.c_float * 2)()
    rta_amd.call("rta_hip_check_division", C.c_uint64(1 << 31), C.c_uint64(12345), C.byref(bad),
                 ex)
    assert bad.value == 0, f"mismatch example a={ex[0]!r} b={ex[1]!r}"


def _pairs(mb_begin, mb_count):
    bad = C.c_uint64(0)
    ex = (C.c_uint32 * 2)()
    rta_amd.call("rta_hip_check_division_pairs", mb_begin, mb_count, C.byref(bad), ex)
    return bad.value, (ex[0], ex[1])


@pytest.mark.gpu
def test_div_cr1_significand_pairs_on_gpu(gpu):
    """div_cr1 (one correction step) against IEEE division for every pair of
    significands A, B in [1, 2): all 2^23 divisor significands in 64 calls of
    2^17 (2^46 pairs, about 40 s), which is what exact.h relies on."""
    total = 0
    for k in range(64):
        mb0, cnt = k << 17, 1 << 17
        bad, ex = _pairs(mb0, cnt)
        print(f"divisor significands [{mb0}, {mb0 + cnt}): {bad} mismatches", flush=True)
        assert bad == 0, f"mismatch A=0x{ex[0]:08x} B=0x{ex[1]:08x}"
        total += cnt
    assert total == 1 << 23


@pytest.mark.gpu
def test_division_pairs_rejects_bad_range(gpu):
    bad = C.c_uint64(0)
    with pytest.raises(rta_amd.RtaError):
        rta_amd.call("rta_hip_check_division_pairs", (1 << 23) - 4, 8, C.byref(bad), None)
