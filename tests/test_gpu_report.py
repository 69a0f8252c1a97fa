"""Block statistics of the device-resident driver and the drift check, on
the GPU.

The block sums are checked against what the driver reports by other routes:
the production loading sums and Widom sums of graspa_batch_box_stats, and the
running energy at each cycle end rebuilt from the move records (running
energy at the end minus the accepted energy changes after each cycle).  The
drift check must find the recomputed final energy equal to the initial one
plus the running change (the reference's final check, fxn_main.h:454-542,
passes below 1e-3 internal units)."""
import io
import re

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(seed=21):
    from graspa_amd import synth
    return synth.build(n_fw=300, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=seed, P=1e6,
                       cycles=12, probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5))


@pytest.mark.parametrize("prec", [0, 1])
def test_block_stats_and_drift(engine_mod, prec):
    from graspa_amd import report
    s = _system()
    print("report test: precision", prec, flush=True)
    box = engine_mod.Box(s, precision=prec, capacity=256)
    box.init_ewald()
    initial = report.recompute_totals(box, s, [0, 0])
    bt = engine_mod.Batch([box], [7], [0], trace_cap=20000, cycle_cap=64)
    bt.run(4, production=False)
    n_prod = 12
    bt.run(n_prod, production=True)
    blk = bt.block_stats(0)
    st = bt.stats(0)
    assert blk["n_block"] == 5 and blk["block_size"] == n_prod // 5
    assert list(blk["cycles"]) == [2, 2, 2, 2, 4]        # the last block takes the remainder
    # loadings and Widom sums agree with the production sums of the box
    np.testing.assert_allclose(blk["n"].sum(0)[1:3], st.load_sum, rtol=0, atol=0)
    np.testing.assert_allclose(blk["n2"].sum(0)[1:3], st.load_sum2, rtol=0, atol=0)
    np.testing.assert_allclose(blk["widom_w"].sum(0)[1:3], st.widom_w, rtol=1e-12, atol=0)
    np.testing.assert_allclose(blk["widom_n"].sum(0)[1:3], st.widom_n, rtol=0, atol=0)
    assert st.widom_n.sum() > 0
    # per-cycle running energy rebuilt from the move records
    tr = bt.trace(0)
    assert len(tr) > 0
    run_end = np.array([getattr(st.running, k) for k in report.TERMS])
    de = np.array([[getattr(t.de, k) for k in report.TERMS] for t in tr])
    moved = np.array([t.accepted != 0 and t.move != 4 for t in tr])    # Widom moves change nothing
    de[~moved] = 0.0
    cyc = np.array([t.cycle for t in tr])
    start = run_end - de.sum(0)
    first = cyc.min()
    e_blocks = np.zeros((blk["n_block"], 10))
    for c in range(first, first + n_prod):
        e_c = start + de[cyc <= c].sum(0)
        e_blocks[min((c - first) // blk["block_size"], 4)] += e_c
    np.testing.assert_allclose(blk["e"], e_blocks, rtol=1e-9, atol=1e-6)
    # drift check and the written report
    rec = report.record_from_batch(bt, 0, box, s, initial, seconds=1.0, capacity=256)
    assert rec.drift.max_abs_drift() < (1e-5 if prec == 0 else 1e-3), rec.drift.drift
    assert max(abs(v) for v in rec.drift.gpu_drift.values()) < 1e-3
    text = report.write_output(rec, io.StringIO())
    m = re.search("ENERGY DRIFT", text)
    val = float(re.compile(r"Total Energy:\s+([-+]?[0-9]*\.?[0-9]+)").search(text, m.end()).group(1))
    assert abs(val) < 1e-3
    assert "Rosenbluth Summary For Component [1]" in text
    # resume: the final configuration written as a restart file and read into
    # a fresh box has the same total energy
    import copy
    import os
    import tempfile
    from graspa_amd import deck, writers
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "Restart", "System_0", "restartfile")
        writers.write_restart(path, s, writers.box_molecules(box))
        t = copy.deepcopy(s)
        deck.read_restart(path, t)
    fresh = engine_mod.Box(t, precision=0, capacity=256)
    e = report.recompute_totals(fresh, t, rec.n_mol)
    fresh.close()
    for k in report.TERMS:
        assert abs(e[k] - rec.drift.final_fp64[k]) <= 1e-9 * max(1.0, abs(e[k])), (k, e[k], rec.drift.final_fp64[k])
    bt.close()
    box.close()
