"""Host logic of the drift check and the statistics writer (no GPU).

Checked against what the mathematics fixes: the total tail correction
against a brute-force sum over atom pairs, block averages against numpy on a
per-cycle series split into blocks, the heat of adsorption against the
closed form of an exactly linear energy-loading relation, the Widom block
against the Henry-coefficient definition; and the written text against the
reference's drift parser rules (Examples/test_examples.py:32-91: the first
number after "Total Energy:" following "ENERGY DRIFT" / "GPU DRIFT", and
"Work took X seconds")."""
import math
import re

import numpy as np
import pytest

from graspa_amd import report, synth
from graspa_amd.deck import ENERGY_TO_KELVIN, GAS_CONSTANT


def _tail_system():
    s = synth.build(n_fw=40, L=20.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=5, P=1e5,
                    probs=dict(translation=1.0, swap=1.0))
    rng = np.random.default_rng(1)
    T = s.ntypes
    t = rng.uniform(-50.0, -5.0, (T, T))
    s.tail = (t + t.T) / 2
    s.use_tail = (rng.uniform(size=(T, T)) < 0.7).astype(np.int32)
    s.use_tail = np.maximum(s.use_tail, s.use_tail.T)
    return s


def test_tail_total_equals_pair_sum_over_atoms():
    s = _tail_system()
    n_mol = [3, 2]
    types = list(s.fw_type)
    for c, nm in zip(s.components, n_mol):
        types += list(c.types) * nm
    types = np.array(types)
    brute = 0.0
    for a in types:                     # every ordered pair of atoms, self pairs included
        for b in types:
            if s.use_tail[a, b]:
                brute += s.tail[a, b]
    assert report.tail_total(s, n_mol) == pytest.approx(brute / s.volume, rel=1e-12)
    s.use_tail[:] = 0
    assert report.tail_total(s, n_mol) == 0.0


def _blocks_from_series(N, E, ns=1, block=None):
    """Block sums as the device gathers them, from per-cycle series
    N[cycle][species] and E[cycle][term]."""
    n_cyc = len(N)
    bs = block or max(1, n_cyc // 5)
    blk = {"n_block": 5, "block_size": bs, "cycles": np.zeros(5, np.int64), "n": np.zeros((5, 8)), "n2": np.zeros((5, 8)),
           "nn": np.zeros((5, 8, 8)), "en": np.zeros((5, 8)), "e": np.zeros((5, 10)),
           "widom_w": np.zeros((5, 8)), "widom_n": np.zeros((5, 8))}
    for i in range(n_cyc):
        b = min(i // bs, 4)
        blk["cycles"][b] += 1
        e = np.asarray(E[i])
        ads = e.sum() - e[0] - e[3] - e[6]
        blk["e"][b] += e
        for c in range(ns):
            blk["n"][b][c + 1] += N[i][c]
            blk["n2"][b][c + 1] += N[i][c] ** 2
            blk["en"][b][c + 1] += ads * N[i][c]
            for d in range(ns):
                blk["nn"][b][c + 1][d + 1] += N[i][c] * N[i][d]
    return blk


def _record(s, blk, drift=None, counters=None):
    z = {k: 0.0 for k in report.TERMS}
    d = drift or report.DriftReport(dict(z), dict(z), dict(z), dict(z))
    ns = len(s.components)
    return report.RunRecord(system=s, names=[c.name for c in s.components], framework_name="synthetic",
                            precision=0, counters=counters if counters is not None else np.zeros((ns, 9), np.int64),
                            n_mol=[0] * ns, drift=d, blocks=blk, seconds=1.25, cycles=int(blk["cycles"].sum()))


def test_block_averages_match_numpy():
    s = synth.config0()
    rng = np.random.default_rng(2)
    N = rng.integers(0, 30, size=(53, 1)).astype(float)
    E = np.zeros((53, 10))
    blk = _blocks_from_series(N, E)
    text = report.write_output(_record(s, blk), __import__("io").StringIO())
    sec = text.split("BLOCK AVERAGES (LOADING: # MOLECULES)")[1].split("COMPONENT [1]")[1]
    got = [float(x) for x in re.findall(r"Average: ([-0-9.]+), ErrorBar: ([-0-9.]+)", sec)[0]]
    bs = 53 // 5
    b0 = N[:bs, 0]
    assert got[0] == pytest.approx(b0.mean(), abs=1e-5)
    assert got[1] == pytest.approx(2 * b0.std(), abs=1e-5)
    # last block takes the remainder
    last = N[4 * bs:, 0]
    lines = re.findall(r"BLOCK \[4\], Blocksize: (\d+), Average: ([-0-9.]+)", sec)
    assert int(lines[0][0]) == len(last)
    assert float(lines[0][1]) == pytest.approx(last.mean(), abs=1e-5)
    means = [N[:bs].mean(), N[bs:2 * bs].mean(), N[2 * bs:3 * bs].mean(), N[3 * bs:4 * bs].mean(), last.mean()]
    ov = re.search(r"Overall: Average: ([-0-9.]+), ErrorBar: ([-0-9.]+)", sec)
    assert float(ov.group(1)) == pytest.approx(np.mean(means), abs=1e-5)
    assert float(ov.group(2)) == pytest.approx(2 * np.std(means), abs=1e-5)


def test_heat_of_adsorption_closed_form():
    """E = a N exactly: cov(E, N) / var(N) = a, so Q = (a k - T) 0.01 / k
    (k = energy_to_kelvin) in every block."""
    s = synth.config0()
    rng = np.random.default_rng(4)
    a = -2500.0
    N = rng.integers(0, 12, size=(100, 1)).astype(float)
    E = np.zeros((100, 10))
    E[:, 1] = a * N[:, 0]                   # host-guest VDW carries the whole energy
    E[:, 0] = 777.0                         # host-host terms are excluded
    blk = _blocks_from_series(N, E)
    q = report.heat_of_adsorption(_record(s, blk))
    expect = (a * ENERGY_TO_KELVIN - s.temperature) * 0.01 / ENERGY_TO_KELVIN
    assert np.allclose(q[0], expect, rtol=1e-9)


def test_heat_of_adsorption_two_species_uncorrelated():
    """Two species whose loadings are uncorrelated and E = a1 N1 + a2 N2:
    the inverse covariance is diagonal and Q_i = (a_i k - T) 0.01 / k."""
    s = synth.build(n_fw=40, L=20.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=5)
    n1 = np.array([0, 1, 0, 1] * 25, float)
    n2 = np.array([0, 0, 1, 1] * 25, float)    # independent of n1 within every block of 20
    a1, a2 = -1800.0, -900.0
    E = np.zeros((100, 10))
    E[:, 4] = a1 * n1 + a2 * n2
    blk = _blocks_from_series(np.stack([n1, n2], 1), E, ns=2)
    q = report.heat_of_adsorption(_record(s, blk))
    for i, a in enumerate((a1, a2)):
        assert np.allclose(q[i], (a * ENERGY_TO_KELVIN - s.temperature) * 0.01 / ENERGY_TO_KELVIN, rtol=1e-9)


def test_widom_block_definitions():
    s = synth.config1()
    rec = _record(s, _blocks_from_series(np.zeros((5, 1)), np.zeros((5, 10))))
    rho = s.framework_density()
    wr, mu, henry, fug = report.widom_block(rec, 12.5, 5.0, rho)
    assert wr == pytest.approx(2.5)
    assert henry == pytest.approx(2.5 / (GAS_CONSTANT * s.temperature * rho))
    assert mu == pytest.approx(-ENERGY_TO_KELVIN / s.beta * math.log(2.5))


def _parse_like_reference(text):
    drift_total = re.compile(r"Total Energy:\s+([-+]?[0-9]*\.?[0-9]+)")
    out = {}
    for key in ("ENERGY DRIFT", "GPU DRIFT"):
        m = re.search(key, text)
        out[key] = float(drift_total.search(text, m.end()).group(1)) if m else None
    w = re.search(r"Work took\s+([-+]?[0-9]*\.?[0-9]+)\s+seconds", text)
    out["work"] = float(w.group(1)) if w else None
    return out


def test_written_drift_sections_parse_like_the_reference():
    s = synth.config0()
    init = {k: 0.0 for k in report.TERMS}
    init.update(hg_vdw=-5000.0, hg_real=-300.0, hg_ewald=-40.0, gg_vdw=-60.0)
    run = {k: 0.0 for k in report.TERMS}
    run.update(hg_vdw=-1200.0, gg_ewald=12.0)
    final = {k: init[k] + run[k] for k in report.TERMS}
    final["hg_vdw"] -= 2.0e-6                 # a drift of 2e-6 in one term
    f64 = dict(final)
    f64["hg_real"] += 3.0e-7
    d = report.DriftReport(init, run, final, f64)
    assert d.drift["hg_vdw"] == pytest.approx(2.0e-6, abs=1e-9)
    assert d.max_abs_drift() == pytest.approx(2.0e-6, abs=1e-9)
    blk = _blocks_from_series(np.ones((10, 1)), np.zeros((10, 10)))
    text = report.write_output(_record(s, blk, drift=d), __import__("io").StringIO())
    p = _parse_like_reference(text)
    assert p["ENERGY DRIFT"] == pytest.approx(0.0, abs=1e-5)       # %.5f of 2e-6
    assert p["ENERGY DRIFT"] < 1e-3                                 # the reference's pass rule
    assert p["GPU DRIFT"] == pytest.approx(0.0, abs=1e-5)
    assert p["work"] == pytest.approx(1.25)
    # the energy lines keep the reference's layout: 28-column label, value, [K], [kJ/mol]
    line = [l for l in text.splitlines() if l.startswith("VDW [Host-Guest]:")][0]
    assert re.match(r"VDW \[Host-Guest\]: {11}-5000\.00000 \(-6013\.62142 \[K\], -50\.00000 \[kJ/mol\]\)$", line)


def test_identity_swap_statistics_lines():
    """Print_IdentitySwap_Statistics (print_statistics.cuh:214-224): one line
    per other species with the attempts and acceptances from this one."""
    from graspa_amd import engine
    s = synth.build(n_fw=40, L=20.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=5)
    cnt = np.zeros((2, engine.N_COUNTERS), np.int64)
    cnt[0][engine.CNT_IDSW_TOT + 1], cnt[0][engine.CNT_IDSW_ACC + 1] = 40, 7
    cnt[1][engine.CNT_IDSW_TOT + 0], cnt[1][engine.CNT_IDSW_ACC + 0] = 33, 5
    cnt[1][engine.CNT_IDSW_TOT + 1] = 9        # to its own species: counted, not printed
    blk = _blocks_from_series(np.ones((10, 2)), np.zeros((10, 10)), ns=2)
    text = report.write_output(_record(s, blk, counters=cnt), __import__("io").StringIO())
    lines = [l for l in text.splitlines() if l.startswith("Identity Swap Performed")]
    assert lines == ["Identity Swap Performed, FROM [CO2 (1)] TO [N2 (2)]: 40 (7 Accepted)",
                     "Identity Swap Performed, FROM [N2 (2)] TO [CO2 (1)]: 33 (5 Accepted)"]
