"""The FP64 erfc/exp of the fast pair term (graspa_amd/csrc/device_math.hpp,
coefficients in fast_math_coeffs.hpp), re-evaluated step by step in numpy and
checked against scipy's erfc/exp -- libraries independent of the fit -- and
against the regeneration of the coefficient header."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest
from scipy.special import erfc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "graspa_amd", "csrc")


def _coeffs():
    txt = open(os.path.join(CSRC, "fast_math_coeffs.hpp")).read()
    scal = {m.group(1): float(m.group(2)) for m in re.finditer(r"constexpr (?:double|int) (\w+) = ([-0-9.e+]+);", txt)}
    arrs = {}
    for m in re.finditer(r"constexpr double (\w+)\[\d+\] = \{([^}]*)\};", txt, re.S):
        arrs[m.group(1)] = [float(v) for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
    return scal, arrs


def _exp_fast(y, s, a):
    n = np.rint(y * s["kLog2e"])
    r = y - n * s["kLn2Hi"]
    r = r - n * s["kLn2Lo"]
    p = np.full_like(y, a["kExpPoly"][-1])
    for c in reversed(a["kExpPoly"][:-1]):
        p = p * r + c
    return np.ldexp(p, n.astype(np.int64))


def _erfc_fast(x, s, a):
    d = s["kErfcC"] + x
    y = (1.0 / d.astype(np.float32)).astype(np.float32).astype(np.float64)   # FP32 reciprocal seed
    y = y + y * (1.0 - d * y)                                                # one Newton step
    t = s["kErfcC"] * y
    u = t * s["kErfcUScale"] + s["kErfcUOff"]
    F = np.full_like(x, a["kErfcPoly"][-1])
    for c in reversed(a["kErfcPoly"][:-1]):
        F = F * u + c
    r = _exp_fast(-x * x, s, a) * t * F
    return np.where(x < s["kErfcXMax"], r, 0.0)


def test_exp_fast_accuracy():
    s, a = _coeffs()
    y = np.linspace(-40.0, 0.0, 400001)
    rel = np.abs(_exp_fast(y, s, a) / np.exp(y) - 1.0)
    assert rel.max() < 2e-14


def test_erfc_fast_accuracy():
    s, a = _coeffs()
    x = np.concatenate([np.linspace(0.0, 6.0, 600001)[:-1], [1e-12, 0.5, 1.0, 2.0, 3.18, 5.999999]])
    ref = erfc(x)
    got = _erfc_fast(x, s, a)
    rel = np.abs(got / ref - 1.0)
    assert rel.max() < 5e-13
    # beyond the fitted range the term is dropped: erfc(6) = 2.15e-17
    assert _erfc_fast(np.array([6.0, 7.5]), s, a).tolist() == [0.0, 0.0]
    assert erfc(6.0) < 2.2e-17


def test_erfc_fast_pins_a_dropped_term():
    """A coefficient dropped from the fit breaks the accuracy bound."""
    s, a = _coeffs()
    x = np.linspace(0.0, 6.0, 60001)
    a2 = dict(a)
    a2["kErfcPoly"] = a["kErfcPoly"][:-1]
    rel = np.abs(_erfc_fast(x, s, a2) / erfc(x) - 1.0)
    assert rel.max() > 1e-10


def _coulomb_table(alpha, cut2):
    import ctypes as C
    from graspa_amd import engine
    L = engine.lib()
    n = L.graspa_coulomb_table(alpha, cut2, None, 0)
    assert n > 0 and n % (DEG + 1) == 0
    buf = (C.c_double * n)()
    assert L.graspa_coulomb_table(alpha, cut2, buf, n) == n
    return np.frombuffer(buf, dtype=np.float64).copy()


MBITS, DEG = 4, 7          # intervals per octave = 2**MBITS; polynomial degree


def _coul_table_eval(tab, rr):
    """erfc(alpha r)/r the way the fast kernels read the table: interval from
    the exponent and top MBITS mantissa bits of r^2 as FP32, degree-DEG
    polynomial in r^2 - s_i."""
    shift = 23 - MBITS
    bits = rr.astype(np.float32).view(np.uint32)
    idx = (bits >> shift).astype(np.int64) - ((127 - 7) << MBITS)
    # the kernels clamp below only: inside the cutoff the index never passes the last row
    idx = np.maximum(idx, 0)
    assert idx.max() < tab.size // (DEG + 1)
    si = (bits & np.uint32((0xFFFFFFFF << shift) & 0xFFFFFFFF)).view(np.float32).astype(np.float64)
    d = rr - si
    c = tab.reshape(-1, DEG + 1)[idx]
    g = c[:, DEG]
    for k in range(DEG - 1, -1, -1):
        g = g * d + c[:, k]
    return g


@pytest.mark.parametrize("alpha,cut", [(0.265058, 12.0), (0.285, 10.0), (0.2, 16.0)])
def test_coulomb_table_accuracy(alpha, cut):
    """The fast form's real-space table against scipy's erfc: absolute error
    of erfc(alpha r)/r below 1e-13 (1e-9 internal units for the largest charge
    products) for every r^2 the overlap rule lets through (0.01 .. cutoff^2)."""
    tab = _coulomb_table(alpha, cut * cut)
    rng = np.random.default_rng(3)
    rr = np.concatenate([np.geomspace(0.01, cut * cut * (1 - 1e-12), 400001),
                         rng.uniform(0.01, cut * cut, 200000),
                         2.0 ** np.arange(-6, int(np.log2(cut * cut)) + 1)])   # octave edges
    rr = rr[rr < cut * cut]
    r = np.sqrt(rr)
    ref = erfc(alpha * r) / r
    got = _coul_table_eval(tab, rr)
    err = np.abs(got - ref)
    assert err.max() < 1e-13, err.max()
    assert np.max(err / ref) < 1e-9


@pytest.mark.parametrize("cut2", [144.0, 256.0 * (1.0 - 1e-9), 64.0 * (1.0 - 2.0 ** -26), 100.0, 256.0])
def test_coulomb_table_covers_fp32_rounded_cutoff(cut2):
    """The lookups clamp the row index below only: r^2 just under cut^2, whose
    FP32 rounding can reach the next power of two, still falls in the last
    octave of the table."""
    tab = _coulomb_table(0.265058, cut2)
    nrow = tab.size // (DEG + 1)
    rr = np.array([np.nextafter(cut2, 0.0), cut2 * (1.0 - 2.0 ** -25), cut2 * (1.0 - 1e-12)])
    bits = rr.astype(np.float32).view(np.uint32)
    idx = (bits >> (23 - MBITS)).astype(np.int64) - ((127 - 7) << MBITS)
    assert idx.max() < nrow, (idx, nrow)
    _coul_table_eval(tab, rr)


def test_coulomb_table_pins_a_wrong_interval():
    """Reading the neighbouring interval's polynomial breaks the bound."""
    tab = _coulomb_table(0.265058, 144.0)
    rr = np.geomspace(0.5, 100.0, 1001)
    shifted = np.roll(tab.reshape(-1, DEG + 1), 1, axis=0).ravel()
    r = np.sqrt(rr)
    assert np.abs(_coul_table_eval(shifted, rr) - erfc(0.265058 * r) / r).max() > 1e-6


def test_coefficient_header_is_regenerated_identically():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_fast_math.py")], capture_output=True, text=True,
                         check=True).stdout
    assert out.strip() == open(os.path.join(CSRC, "fast_math_coeffs.hpp")).read().strip()
