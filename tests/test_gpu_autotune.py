"""Launch shapes of the Widom pair-energy kernel and their autotuner.

A shape only decides which wave evaluates which trial site, so every shape
must give the same Widom results bit for bit, in both precision forms; the
autotuner must keep one of the shapes it tried, and the results after tuning
must equal those before.  The FP64 form is also checked against the oracle in
a non-default shape, so a shape-specific indexing slip cannot hide behind the
shape-to-shape comparison."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("rosen", "energy", "vdw", "real", "ewald", "first_bead", "orientation", "ok")


def _box(engine_mod, prec, n_fw=500):
    from graspa_amd import synth
    s = synth.config1(n_fw=n_fw)
    box = engine_mod.Box(s, precision=prec)
    box.init_ewald()
    intra, selfe = engine_mod.exclusion_energy(s, 1)
    return s, box, intra + selfe


def _run(engine_mod, box, excl, n=400, seed=77):
    res, summ = box.widom_batch(1, n, seed=seed, excl=excl)
    return engine_mod.widom_results_to_numpy(res), summ


def _same(a, b):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("prec_name", ["GRASPA_FP64", "GRASPA_MIXED"])
def test_every_shape_gives_identical_results(engine_mod, prec_name):
    prec = getattr(engine_mod, prec_name)
    s, box, excl = _box(engine_mod, prec)
    assert box.eval_shape() == {"waves": 12, "per_cu": 0, "framework_lds": 1}
    ref, ref_summ = _run(engine_mod, box, excl)
    assert ref["ok"].sum() > 0
    for waves in (4, 8, 12, 16):
        for fw_lds in (1, 0):
            for per_cu in (1, 0):
                box.set_eval_shape(waves, per_cu, fw_lds)
                got, summ = _run(engine_mod, box, excl)
                _same(ref, got)
                assert summ.sum_w == ref_summ.sum_w and summ.evaluated_pairs == ref_summ.evaluated_pairs
    box.close()


def test_shape_fp64_matches_oracle(engine_mod):
    """A non-default shape (4 waves, one workgroup per CU, framework from L2)
    against the FP64 oracle on the same draws."""
    from oracle import binding as ob
    s, box, excl = _box(engine_mod, engine_mod.GRASPA_FP64, n_fw=400)
    o = ob.OracleSystem(s)
    _, skf, ska, _, _, _ = o.ewald_total()
    ref = ob.widom_records_to_numpy(o.widom_batch(1, ska, skf, seed=31, n_insert=200))
    box.set_eval_shape(4, 1, 0)
    res, _ = box.widom_batch(1, 200, seed=31, excl=excl)
    got = engine_mod.widom_results_to_numpy(res)
    box.close()
    assert np.array_equal(got["ok"], ref["ok"])
    ok = ref["ok"] == 1
    assert ok.sum() > 100
    assert np.array_equal(got["first_bead"][ok], ref["first_bead"][ok])
    assert np.array_equal(got["orientation"][ok], ref["orientation"][ok])
    err = np.abs(got["energy"][ok] - ref["energy"][ok])
    assert err.max() <= 1e-10 * max(1.0, np.abs(ref["energy"][ok]).max())


def test_autotune_keeps_a_tried_shape_and_results(engine_mod):
    s, box, excl = _box(engine_mod, engine_mod.GRASPA_MIXED)
    before, _ = _run(engine_mod, box, excl)
    chosen, tried = box.autotune_widom(1, 2000, seed=5, reps=2, excl=excl)
    assert len(tried) >= 4
    shapes = [t for t, _ in tried]
    assert chosen in shapes
    assert all(ms > 0.0 for _, ms in tried)
    best = min(ms for _, ms in tried)
    assert [ms for t, ms in tried if t == chosen][0] == best
    after, _ = _run(engine_mod, box, excl)
    _same(before, after)
    box.close()


def test_bad_shapes_rejected(engine_mod):
    s, box, excl = _box(engine_mod, engine_mod.GRASPA_MIXED, n_fw=200)
    for bad in ((5, 0, 1), (12, -1, 1), (12, 0, 2), (12, 65, 1)):
        with pytest.raises(engine_mod.EngineError):
            box.set_eval_shape(*bad)
    assert box.eval_shape() == {"waves": 12, "per_cu": 0, "framework_lds": 1}
    box.set_eval_shape(8, 2, 0)
    box.set_eval_shape(default=True)
    assert box.eval_shape() == {"waves": 12, "per_cu": 0, "framework_lds": 1}
    box.close()


def test_queued_batches_totals_match_synchronous(engine_mod):
    """Batches queued with graspa_widom_enqueue sum on the device to exactly
    the per-batch summaries of synchronous graspa_widom_batch calls over the
    same draws (same additions in the same order), and a reset clears them."""
    s, box, excl = _box(engine_mod, engine_mod.GRASPA_MIXED)
    box.widom_totals(reset=True)
    keys = ("sum_w", "sum_w2", "sum_we", "n", "n_first_bead", "n_ok", "screened_pairs", "evaluated_pairs")
    host = dict.fromkeys(keys, 0.0)
    for k in range(4):
        _, summ = box.widom_batch(1, 300, seed=9, first_index=300 * k, excl=excl, results=False)
        for key in keys:
            host[key] += getattr(summ, key)
    sync_tot = box.widom_totals(reset=True)
    for k in range(4):
        box.widom_enqueue(1, 300, seed=9, first_index=300 * k, excl=excl)
    queued = box.widom_totals(reset=True)
    empty = box.widom_totals()
    box.close()
    assert host["n"] == 1200 and host["n_ok"] > 0
    for key in keys:
        assert getattr(sync_tot, key) == host[key], key
        assert getattr(queued, key) == host[key], key
        assert getattr(empty, key) == 0.0, key


def test_autotune_leaves_the_running_totals_unchanged(engine_mod):
    """A tuning run between queued production batches leaves the device-side
    running totals as they were: the timed batches are not counted."""
    s, box, excl = _box(engine_mod, engine_mod.GRASPA_MIXED)
    box.widom_totals(reset=True)
    for k in range(2):
        box.widom_enqueue(1, 300, seed=9, first_index=300 * k, excl=excl)
    box.sync()
    box.autotune_widom(1, 500, seed=5, reps=1, excl=excl)
    box.widom_enqueue(1, 300, seed=9, first_index=600, excl=excl)
    tuned = box.widom_totals(reset=True)
    for k in range(3):
        box.widom_enqueue(1, 300, seed=9, first_index=300 * k, excl=excl)
    plain = box.widom_totals(reset=True)
    box.close()
    assert plain.n == 900.0
    for key in ("sum_w", "sum_w2", "sum_we", "n", "n_first_bead", "n_ok", "screened_pairs", "evaluated_pairs"):
        assert getattr(tuned, key) == getattr(plain, key), key
