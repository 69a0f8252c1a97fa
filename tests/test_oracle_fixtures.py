"""The oracle against gRASPA's own printed energies for its example decks.

Expected values come from the reference's committed outputs (see
tests/golden/golden_values.json, "source" fields); inputs are the same decks,
stored as .npz by tests/golden/make_fixtures.py.  gRASPA prints 5 decimals, so
the comparison is to half a unit in the last printed digit.
"""
import numpy as np
import pytest

from conftest import load_golden_system

NIST = ["nist_box1", "nist_box2", "nist_box3", "nist_box4"]
NIST_KVECTORS = {"nist_box1": 831, "nist_box2": 1068, "nist_box3": 838, "nist_box4": 1028}  # NIST readme table


def close5(a, b):
    return abs(a - b) <= 0.5e-5 + 1e-12 * abs(b)


@pytest.mark.parametrize("name", NIST + ["tip4p_restart"])
def test_total_vdw_real(name, golden_values, oracle_lib):
    from oracle.binding import OracleSystem
    g = golden_values[name]
    o = OracleSystem(load_golden_system(name))
    e = o.total_vdw_real()
    assert close5(e.gg_vdw, g["gg_vdw"]), (e.gg_vdw, g["gg_vdw"])
    assert close5(e.gg_real, g["gg_real"]), (e.gg_real, g["gg_real"])
    assert e.hg_vdw == 0.0 and e.hh_vdw == 0.0


@pytest.mark.parametrize("name", NIST + ["tip4p_restart"])
def test_ewald_total(name, golden_values, oracle_lib):
    from oracle.binding import OracleSystem
    g = golden_values[name]
    o = OracleSystem(load_golden_system(name))
    f, sk_fw, sk_ads, selfe, intra, nact = o.ewald_total()
    assert close5(f.gg_ewald, g["gg_fourier"])
    assert close5(selfe[1], g["self_guest"])
    assert close5(-intra[1], g["intra_exclusion_guest"])
    assert close5(f.gg_ewald - selfe[1] - intra[1], g["gg_ewald"])
    assert np.all(sk_fw == 0.0)  # no framework atoms in these decks
    if name in NIST_KVECTORS:
        assert nact == NIST_KVECTORS[name]


@pytest.mark.parametrize("name", NIST + ["tip4p_restart", "co2_mfi"])
def test_template_exclusion(name, golden_values, oracle_lib):
    from oracle.binding import OracleSystem
    g = golden_values[name]
    o = OracleSystem(load_golden_system(name))
    intra, selfe = o.exclusion(1)
    assert close5(intra, g["template_intra"])
    assert close5(selfe, g["template_self"])


@pytest.mark.parametrize("name", NIST)
def test_tail(name, golden_values, oracle_lib):
    from oracle.binding import OracleSystem, lib, _i64p
    g = golden_values[name]
    s = load_golden_system(name)
    o = OracleSystem(s)
    assert close5(s.tail[0, 0], g["tail_pair_00"])
    n = s.pseudo_counts([len(s.components[0].init_pos) // s.components[0].molsize])
    assert close5(lib().or_tail_total(o.box, n.ctypes.data_as(_i64p)), g["tail_total"])


def test_charged_framework(golden_values, oracle_lib):
    """Framework-only S(k), self and intra-exclusion of MFI-2x2x2 (2304 atoms,
    CIF charges, RASPA-mode alpha/kmax on a cuboid box)."""
    from oracle.binding import OracleSystem
    g = golden_values["co2_mfi"]
    s = load_golden_system("co2_mfi")
    assert len(s.fw_type) == g["n_framework_atoms"]
    assert s.nvec == g["n_structure_factors"]
    assert abs(s.volume - g["volume"]) < 1e-5
    assert abs(s.beta - g["beta"]) < 0.5e-5
    o = OracleSystem(s)
    f, sk_fw, sk_ads, selfe, intra, nact = o.ewald_total()
    assert close5(f.hh_ewald, g["hh_fourier"])
    assert close5(selfe[0], g["self_framework"])
    assert close5(-intra[0], g["intra_exclusion_framework"])
    assert np.allclose(sk_ads, 0.0, atol=1e-9)  # stored guest S(k) = total - framework, empty box
