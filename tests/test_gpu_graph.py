"""A queued Widom batch captured as a HIP graph: every replay equals the
enqueued batch of the same first index, bit for bit, with eager batches and
synchronisations between the replays, in both precision forms; a replay after
the box changed is refused until the batch is captured again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("sum_w", "sum_w2", "sum_we", "n", "n_first_bead", "n_ok", "screened_pairs", "evaluated_pairs")


def _box(engine_mod, prec):
    from graspa_amd import synth
    s = synth.config1(n_fw=750)
    b = engine_mod.Box(s, precision=prec)
    b.init_ewald()
    intra, selfe = engine_mod.exclusion_energy(s, 1)
    return s, b, intra + selfe


@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_graph_replays_equal_enqueued_batches(engine_mod, form):
    prec = engine_mod.GRASPA_MIXED if form == "mixed" else engine_mod.GRASPA_FP64
    n, seed, f0, reps = 3000, 77, 500, 4
    _, ref_box, excl = _box(engine_mod, prec)
    ref = []
    for r in range(reps):
        ref_box.widom_totals(reset=True)
        ref_box.widom_enqueue(1, n, seed, stream=2, first_index=f0 + r * n, excl=excl)
        ref.append(ref_box.widom_totals(reset=True))
    ref_box.close()
    _, box, excl = _box(engine_mod, prec)
    print("graph capture: %s form" % form, flush=True)
    box.widom_graph(1, n, seed, stream=2, first_index=f0, excl=excl)
    for r in range(reps):
        # eager work and a synchronisation between the replays
        box.widom_batch(1, 500, seed + 100 + r, first_index=0, excl=excl, results=False)
        box.sync()
        box.widom_totals(reset=True)
        box.widom_graph_launch(1)
        got = box.widom_totals(reset=True)
        assert got.n == n and got.n_ok > 0
        for k in KEYS:
            assert getattr(got, k) == getattr(ref[r], k), (r, k, getattr(got, k), getattr(ref[r], k))
    # several replays queued at once continue the index sequence
    box.widom_totals(reset=True)
    box.widom_graph_launch(2)
    got = box.widom_totals(reset=True)
    assert got.n == 2 * n
    box.close()


def test_graph_refused_after_the_box_changes(engine_mod):
    s, box, excl = _box(engine_mod, engine_mod.GRASPA_MIXED)
    box.widom_graph(1, 2000, 5, excl=excl)
    box.widom_graph_launch(1)
    # a molecule committed: the batch's arguments (molecule counts) changed
    c = s.components[0]
    pos = np.asarray(c.template_pos) + np.array([3.0, 4.0, 5.0])
    box.set_molecules(1, pos, np.asarray(c.charges), np.asarray(c.types), np.zeros(c.molsize, dtype=np.int64))
    box.init_ewald()
    with pytest.raises(engine_mod.EngineError):
        box.widom_graph_launch(1)
    box.widom_graph(1, 2000, 5, excl=excl)
    box.widom_totals(reset=True)
    box.widom_graph_launch(1)
    assert box.widom_totals(reset=True).n == 2000
    box.close()
