"""The step's own float transcendentals (beamlab_amd/csrc/bl_devmath.h) against the host C library, on the CPU:
expf, powf(x, 2), sinf, cosf and atanf on every one of the 2^32 float inputs, atan2f on 1e8 drawn pairs (uniform bits,
the quaternion logarithm's (|v|, w) arguments, arguments within a factor of 4). The header is compiled for the host by
tests/devmath_check.cpp; the device build compiles the same source. Also recomputes the 2/pi window table."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _two_over_pi_bits(nbits):
    """floor(2/pi * 2^nbits) by Machin's formula in integers."""
    prec = nbits + 64
    one = 1 << prec

    def arctan_inv(x):
        s, t, k, sign = 0, one // x, 1, 1
        while t:
            s += sign * (t // k)
            t //= x * x
            k += 2
            sign = -sign
        return s

    pi = 4 * (4 * arctan_inv(5) - arctan_inv(239))
    return ((2 * one * one) // pi) >> (prec - nbits)


def test_inv_pio4_table():
    src = open(os.path.join(ROOT, "beamlab_amd", "csrc", "bl_devmath.h")).read()
    body = src[src.index("#define BL_INV_PIO4"):]
    body = body[body.index("{") + 1:body.index("}")]
    table = [int(t.strip().rstrip("u"), 16) for t in body.replace("\\", "").split(",")]
    bits = _two_over_pi_bits(8 * 27)  # 27 bytes of 2/pi after the point
    want = [(bits >> (8 * 27 - 8 * (i + 1))) & 0xffffffff for i in range(24)]
    assert table == want


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    out = str(tmp_path_factory.mktemp("devmath") / "devmath_check")
    cmd = [gxx, "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-Wno-unknown-pragmas",
           "-I", os.path.join(ROOT, "beamlab_amd", "csrc"), os.path.join(HERE, "devmath_check.cpp"), "-o", out, "-lm"]
    subprocess.run(cmd, check=True)
    return out


def _run(checker, *args):
    res = subprocess.run([checker, *args], check=True, capture_output=True, text=True, timeout=600)
    rows = {}
    for line in res.stdout.splitlines():
        name, checked, bad, first = line.split()
        rows[name] = (int(checked), int(bad), first)
    return rows


@pytest.mark.parametrize("fn", ["table", "expf", "pow2", "sinf", "cosf", "atanf", "acosf"])
def test_exhaustive(checker, fn):
    rows = _run(checker, fn)
    assert len(rows) == 1
    (name, (checked, bad, first)), = rows.items()
    assert bad == 0, f"{name}: {bad} of {checked} inputs differ from the C library (first {first})"


def test_atan2f_pairs(checker):
    rows = _run(checker, "atan2f", "100000000")
    checked, bad, first = rows["atan2f"]
    assert checked >= 99_000_000 and bad == 0, f"atan2f: {bad} pairs differ (first y,x bits {first})"
