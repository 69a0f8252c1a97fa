"""Deck reader: synthetic decks written in gRASPA's formats read back to the
same system; the reference's unit conversions."""
import os

import numpy as np
import pytest

from graspa_amd import deck, synth


def test_units():
    # fxn_main.h:115-127: beta at 300 K and 1e5 Pa in internal units
    assert abs(deck.beta_of(300.0) - 1.0 / (0.8314462618 * 300.0 * (1.380649e-23 / 1.380649e-23))) < 1e-3 * 0.004
    assert abs(deck.pressure_internal(1.6605402e7) - 1.0) < 1e-15


def test_ewald_parameters_raspa_mode():
    s = synth.config1()
    # alpha from EwaldPrecision 1e-6 and rc 12 (read_data.cpp:693-702)
    assert abs(s.alpha - 0.265058) < 1e-6
    assert s.kmax == (7, 7, 7)
    assert abs(s.recip_cut - (1.05 * 7) ** 2) < 1e-12


def test_write_read_round_trip(tmp_path):
    s = synth.build(n_fw=300, L=20.0, species=("CO2", "N2"), mol_fraction=[0.3, 0.7], seed=9)
    synth.write_deck(s, str(tmp_path), cycles=10)
    r = deck.read_deck(str(tmp_path))
    assert r.pseudo_names == s.pseudo_names
    assert np.allclose(r.cell, s.cell)
    assert np.allclose(r.fw_pos, s.fw_pos, atol=1e-9)
    assert np.allclose(r.fw_charge, s.fw_charge, atol=1e-11)
    assert np.array_equal(r.fw_type, s.fw_type)
    assert np.allclose(r.eps, s.eps, rtol=1e-9)
    assert np.allclose(r.sig, s.sig, rtol=1e-9)
    assert [c.name for c in r.components] == ["CO2", "N2"]
    assert abs(r.components[1].mol_fraction - 0.7) < 1e-12
    assert np.allclose(r.components[0].template_pos, s.components[0].template_pos)
    assert r.kmax == s.kmax and abs(r.alpha - s.alpha) < 1e-12


def test_npz_round_trip(tmp_path):
    s = synth.config0()
    p = str(tmp_path / "s.npz")
    deck.save_system(p, s)
    r = deck.load_system(p)
    assert np.array_equal(r.fw_pos, s.fw_pos) and r.kmax == s.kmax and r.alpha == s.alpha


def _with_lines(tmp_path, globals_=(), component=()):
    from graspa_amd import synth
    s = synth.config0()
    synth.write_deck(s, str(tmp_path), cycles=10)
    p = tmp_path / "simulation.input"
    lines = p.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.lstrip().startswith("Component ") and "MoleculeName" in ln)
    lines = lines[:k] + list(globals_) + lines[k:k + 1] + list(component) + lines[k + 1:]
    p.write_text("\n".join(lines) + "\n")


def test_output_keywords(tmp_path):
    _with_lines(tmp_path, ["NumberOfBlocks 8", "PrintEvery 50", "MoviesEvery 100", "SaveOutputToFile yes"])
    r = deck.read_deck(str(tmp_path))
    assert (r.n_blocks, r.print_every, r.movies_every, r.save_output) == (8, 50, 100, True)


@pytest.mark.parametrize("glob,comp", [(["GibbsVolumeChangeProbability 0.1"], []),
                                       (["NPTVolumeChangeProbability 1.0"], []),
                                       (["RunTMMC yes"], []),
                                       (["NumberOfSimulations 2"], []),
                                       ([], ["CBCFProbability 1.0"]),
                                       ([], ["RotationSpecialProbability 0.5"])])
def test_unsupported_methods_are_refused(tmp_path, glob, comp):
    _with_lines(tmp_path, glob, comp)
    with pytest.raises(NotImplementedError):
        deck.read_deck(str(tmp_path))


def test_switched_off_methods_are_accepted(tmp_path):
    _with_lines(tmp_path, ["GibbsVolumeChangeProbability 0.0", "RunTMMC no", "NumberOfSimulations 1"],
                ["CBCFProbability 0.0"])
    deck.read_deck(str(tmp_path))


@pytest.mark.parametrize("value", ["yes", "true", "YES"])
def test_lj1264_is_refused(tmp_path, value):
    """UseLJ1264 yes|true (read_data.cpp:506-514) switches the pair potential to
    the 12-6-4 polynomial (maths.cuh:454-476), which the engine does not run."""
    _with_lines(tmp_path, ["UseLJ1264 " + value])
    with pytest.raises(NotImplementedError, match="UseLJ1264"):
        deck.read_deck(str(tmp_path))


def test_lj1264_off_is_accepted(tmp_path):
    _with_lines(tmp_path, ["UseLJ1264 no"])
    deck.read_deck(str(tmp_path))


def test_non_lj_mixing_rule_rows_are_refused(tmp_path):
    """A lennard-jones-1264 row holds C12, C6, C4, C10 (read_data.cpp:820-834);
    reading its columns as epsilon and sigma would give wrong energies."""
    _with_lines(tmp_path)
    p = tmp_path / "force_field_mixing_rules.def"
    lines = p.read_text().splitlines()
    k = next(i for i, ln in enumerate(lines) if "lennard-jones" in ln)
    t = lines[k].split()
    lines[k] = "%s lennard-jones-1264 1.0e6 1.0e3 10.0 0.0" % t[0]
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(NotImplementedError, match="lennard-jones-1264"):
        deck.read_deck(str(tmp_path))


REF_LJ1264 = "/root/reference/Examples/TIP4PEW-MgMOF-LJ1264"


@pytest.mark.skipif(not os.path.isdir(REF_LJ1264), reason="the reference's example decks are not mounted here")
def test_reference_lj1264_deck_is_refused():
    with pytest.raises(NotImplementedError):
        deck.read_deck(REF_LJ1264)
