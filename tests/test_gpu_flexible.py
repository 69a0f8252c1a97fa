"""Flexible n-butane (TraPPE bend and dihedral, grown bead by bead, 16 first
bead trials x 8 trial positions per bead) on the device-resident driver:
move for move the same as the oracle's batched driver in FP64, within the
fast form's tolerance in the mixed form, on the stated 3,000-atom, 30 A
uncharged framework and on a 400-atom variant at 1e7 Pa where chains go in
and move.  Entry points without flexible growth refuse such a component."""
import pytest

from test_gpu_configs import _check_sequence, _device_batch, _oracle_batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", ["3000_atoms", "400_atoms_1e7Pa"])
def test_flexible_butane_device_driver(engine_mod, variant):
    from graspa_amd import synth
    s = synth.config2(flexible=True) if variant == "3000_atoms" else synth.config2(n_fw=400, P=1e7, flexible=True)
    assert s.components[0].chain is not None and s.n_trial_pos == 16 and s.n_trial_orient == 8
    cycles = 25 if variant == "3000_atoms" else 60
    o, mc, otr = _oracle_batch(s, 13, 0, cycles)
    boxes, bt = _device_batch(engine_mod, [s], [13], [0], cycles, engine_mod.GRASPA_FP64)
    n_acc = _check_sequence(otr, bt.trace(0))
    assert int(bt.stats(0).n_mol[0]) == o.comps[1].n_mol
    if variant != "3000_atoms":
        # chains were inserted, deleted and moved
        assert n_acc > 50 and o.comps[1].n_mol > 5
        assert any(x.accepted and x.move == 3 for x in otr)
        assert any(x.accepted and x.move in (0, 1) for x in otr)
    bt.close()
    boxes[0].close()
    boxes, bt = _device_batch(engine_mod, [s], [13], [0], cycles, engine_mod.GRASPA_MIXED)
    _check_sequence(otr, bt.trace(0), fp64=False)
    bt.close()
    boxes[0].close()


def test_flexible_chain_refused_where_not_grown(engine_mod):
    from graspa_amd import synth
    s = synth.config2(n_fw=400, flexible=True)
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_MIXED)
    box.init_ewald()
    with pytest.raises(engine_mod.EngineError):
        engine_mod.MonteCarlo(box)
    with pytest.raises(engine_mod.EngineError):
        box.widom_batch(1, 64, seed=1, results=False)
    box.close()
