"""A deck run end to end on the GPU: deck files in, output.txt, restart file
and movie snapshots out; the drift check of the written output passes the
reference's rule (Examples/test_examples.py: ENERGY DRIFT total below 1e-3)
and the restart file reads back into a box with the same energy."""
import copy
import os
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_run_deck(tmp_path):
    from graspa_amd import deck, engine, report, run, synth
    s = synth.build(n_fw=300, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=41, P=1e6, cycles=10,
                    probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5))
    d = tmp_path / "deck"
    synth.write_deck(s, str(d), cycles=10)
    sim = d / "simulation.input"
    txt = sim.read_text().replace("NumberOfInitializationCycles 0", "NumberOfInitializationCycles 3")
    k = txt.index("Component 0")
    sim.write_text(txt[:k] + "MoviesEvery 5\nNumberOfBlocks 2\n" + txt[k:])   # global keywords precede the components
    print("run_deck test", flush=True)
    rec = run.run_deck(str(d), precision=engine.GRASPA_MIXED, moves_per_launch=64, capacity=256)
    out = (d / "Output" / "System_0" / "output.txt").read_text()
    m = re.search("ENERGY DRIFT", out)
    val = float(re.compile(r"Total Energy:\s+([-+]?[0-9]*\.?[0-9]+)").search(out, m.end()).group(1))
    assert abs(val) < 1e-3
    assert re.search(r"Work took\s+[0-9.]+\s+seconds", out)
    assert rec.blocks["n_block"] == 2 and int(rec.blocks["cycles"].sum()) == 10
    movies = sorted(os.listdir(d / "Movies" / "System_0"))
    assert len(movies) >= 1
    t = copy.deepcopy(deck.read_deck(str(d)))
    deck.read_restart(str(d / "Restart" / "System_0" / "restartfile"), t)
    assert [len(c.init_pos) // c.molsize for c in t.components] == rec.n_mol
    b = engine.Box(t, precision=0, capacity=256)
    e = report.recompute_totals(b, t, rec.n_mol)
    b.close()
    for k in report.TERMS:
        assert abs(e[k] - rec.drift.final_fp64[k]) <= 1e-9 * max(1.0, abs(e[k])), k


def test_run_deck_creates_molecules(tmp_path):
    """CreateNumberOfMolecules: the deck's molecules are created before the
    cycles (CreateMolecule_InOneBox) and the written drift check passes."""
    from graspa_amd import engine, run, synth
    s = synth.build(n_fw=300, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=43, P=1e5, cycles=6,
                    probs=dict(translation=1.0, rotation=1.0, identity=1.0))
    d = tmp_path / "deck"
    synth.write_deck(s, str(d), cycles=6)
    sim = d / "simulation.input"
    txt = sim.read_text()
    parts = txt.split("CreateNumberOfMolecules 0")
    assert len(parts) == 3
    sim.write_text(parts[0] + "CreateNumberOfMolecules 9" + parts[1] + "CreateNumberOfMolecules 5" + parts[2])
    print("run_deck create test", flush=True)
    logs = []
    rec = run.run_deck(str(d), precision=engine.GRASPA_MIXED, moves_per_launch=64, capacity=256, log=logs.append)
    assert any("Need to create 9" in l for l in logs)
    out = (d / "Output" / "System_0" / "output.txt").read_text()
    m = re.search("ENERGY DRIFT", out)
    val = float(re.compile(r"Total Energy:\s+([-+]?[0-9]*\.?[0-9]+)").search(out, m.end()).group(1))
    assert abs(val) < 1e-3
    # the initial stage is the box after creation; without insertions and
    # deletions the 14 molecules stay (identity swaps may change the species)
    assert rec.drift.initial["hg_vdw"] != 0.0 and rec.drift.initial["gg_vdw"] != 0.0
    assert sum(rec.n_mol) == 14
