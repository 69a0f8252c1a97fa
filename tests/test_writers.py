"""Restart, movie and AllData writers (host only).

The restart round trip goes through the deck reader, the restatement of the
reference's restart parser (read_data.cpp:3000-3221) that the NIST and
Restart-Examples fixtures pin: the 478 TIP4P molecules read from the
reference's own restart example are written and read back unchanged, and so
are molecules of two species with gaps in their molecule order."""
import copy
import os

import numpy as np
import pytest

from graspa_amd import deck, synth, writers

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _molecules_of(system):
    out = []
    for c in system.components:
        n = 0 if c.init_pos is None else len(c.init_pos)
        q = c.init_charge if c.init_charge is not None else np.tile(c.charges, n // c.molsize)
        out.append((np.asarray(c.init_pos).reshape(-1, 3), np.asarray(q), np.tile(c.types, n // c.molsize),
                    np.asarray(c.init_molid)))
    return out


def test_restart_round_trip_reference_example(tmp_path):
    s = deck.load_system(os.path.join(GOLDEN, "tip4p_restart.npz"))
    mols = _molecules_of(s)
    p = tmp_path / "Restart" / "System_0" / "restartfile"
    writers.write_restart(str(p), s, mols, max_translation=[(0.3, 0.3, 0.3)], max_rotation=[(0.2, 0.2, 0.2)])
    t = copy.deepcopy(s)
    for c in t.components:
        c.init_pos = c.init_molid = c.init_charge = None
    deck.read_restart(str(p), t)
    c0, c1 = s.components[0], t.components[0]
    assert len(c1.init_pos) == len(c0.init_pos) == 1912
    np.testing.assert_allclose(c1.init_pos, c0.init_pos, rtol=1e-14, atol=1e-12)
    np.testing.assert_array_equal(c1.init_molid, c0.init_molid)
    np.testing.assert_allclose(c1.init_charge, c0.init_charge, rtol=1e-5)   # charges keep 6 digits, as written


def test_restart_round_trip_two_species(tmp_path):
    s = synth.build(n_fw=50, L=22.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=3)
    rng = np.random.default_rng(0)
    mols = []
    for c, n in zip(s.components, (4, 3)):
        centres = rng.uniform(0, 22.0, (n, 1, 3))
        pos = (centres + np.asarray(c.template_pos)[None]).reshape(-1, 3)
        mols.append((pos, np.tile(c.charges, n), np.tile(c.types, n), np.repeat(np.arange(n), c.molsize)))
    p = tmp_path / "restartfile"
    writers.write_restart(str(p), s, mols)
    t = copy.deepcopy(s)
    deck.read_restart(str(p), t)
    for (pos, q, _ty, mid), c in zip(mols, t.components):
        # the reader unwraps each molecule about its first atom; intact molecules come back as written
        np.testing.assert_allclose(c.init_pos, pos, rtol=1e-14, atol=1e-12)
        np.testing.assert_array_equal(c.init_molid, mid)
        np.testing.assert_allclose(c.init_charge, q, rtol=1e-5)


def test_lammps_movie_wraps_molecules_and_keeps_geometry(tmp_path):
    s = synth.build(n_fw=30, L=20.0, species=("CO2",), seed=4)
    c = s.components[0]
    pos = (np.array([[[-3.0, 25.0, 7.0]], [[10.0, 10.0, 19.9]]]) + np.asarray(c.template_pos)[None]).reshape(-1, 3)
    mols = [(pos, np.tile(c.charges, 2), np.tile(c.types, 2), np.repeat(np.arange(2), 3))]
    p = tmp_path / "Movies" / "System_0" / "result_0.data"
    writers.write_lammps_movie(str(p), s, mols)
    lines = p.read_text().splitlines()
    assert lines[2] == "%d atoms" % (30 + 6)
    atoms = lines[lines.index("Atoms") + 2:]
    assert len(atoms) == 36
    xyz = np.array([[float(v) for v in a.split()[4:7]] for a in atoms])
    fw = xyz[:30]
    assert np.all((fw >= 0) & (fw < 20.0))
    ads = xyz[30:].reshape(2, 3, 3)
    assert np.all((ads[:, 0] >= 0) & (ads[:, 0] < 20.0))              # first atom of each molecule in the box
    np.testing.assert_allclose(ads - ads[:, :1], pos.reshape(2, 3, 3) - pos.reshape(2, 3, 3)[:, :1], atol=1e-8)
    mol_col = [int(a.split()[1]) for a in atoms]
    assert mol_col[:30] == [1] * 30 and mol_col[30:] == [2, 2, 2, 3, 3, 3]


def test_all_data_dump(tmp_path):
    s = synth.build(n_fw=10, L=20.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=4)
    mols = []
    for c, n in zip(s.components, (2, 1)):
        pos = np.tile(np.asarray(c.template_pos), (n, 1)) + 5.0
        mols.append((pos, np.tile(c.charges, n), np.tile(c.types, n), np.repeat(np.arange(n), c.molsize)))
    writers.write_all_data(str(tmp_path / "AllData" / "System_0"), s, mols)
    a = (tmp_path / "AllData" / "System_0" / "Component_1.data").read_text().splitlines()
    b = (tmp_path / "AllData" / "System_0" / "Component_2.data").read_text().splitlines()
    assert len(a) == 1 + 6 and len(b) == 1 + s.components[1].molsize
    assert a[4].split()[:2] == ["1", "0"]
