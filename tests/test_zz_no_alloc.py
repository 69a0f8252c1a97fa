"""The allocation-free Widom entry point (graspa_widom_enqueue_noalloc)
against graspa_widom_enqueue: the same batches through the C-ABI give the same
running totals, bit for bit, on two calls in a row (the second smaller, so
the fixed scratch block holds the first call's leftovers), in both precision
forms, with and without the voxel maps that only the allocating form builds
(without them the work counters differ: trials a map would have dropped are
screened or evaluated instead, with the same outcome); a batch larger than
the block is refused before any GPU work."""
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("sum_w", "sum_w2", "sum_we", "n", "n_first_bead", "n_ok")
WORK = ("screened_pairs", "evaluated_pairs")


def _box(engine_mod, precision):
    from graspa_amd import synth
    s = synth.config1(n_fw=750)
    box = engine_mod.Box(s, precision=precision)
    box.init_ewald()
    intra, selfe = engine_mod.exclusion_energy(s, 1)
    return box, intra + selfe


def _one(box, fn, n, seed, excl):
    box.widom_totals(reset=True)
    fn(1, n, seed, stream=3, first_index=17, excl=excl)
    return box.widom_totals(reset=True)


def _same(a, b, with_work):
    for k in FIELDS + (WORK if with_work else ()):
        assert getattr(a, k) == getattr(b, k), (k, getattr(a, k), getattr(b, k))


@pytest.mark.parametrize("form", ["mixed", "fp64"])
def test_noalloc_equals_allocating_form(engine_mod, form):
    prec = engine_mod.GRASPA_MIXED if form == "mixed" else engine_mod.GRASPA_FP64
    sizes = ((6000, 41), (1500, 42))
    # fresh boxes: the allocation-free form runs before any voxel map exists
    print("no-alloc widom: %s form, fresh box" % form, flush=True)
    fresh, excl = _box(engine_mod, prec)
    got_fresh = [_one(fresh, fresh.widom_enqueue_noalloc, n, seed, excl) for n, seed in sizes]
    fresh.close()
    ref_box, excl = _box(engine_mod, prec)
    ref = [_one(ref_box, ref_box.widom_enqueue, n, seed, excl) for n, seed in sizes]
    for a, b in zip(got_fresh, ref):
        assert a.n_ok > 0 and a.n_first_bead > a.n_ok
        # without the voxel maps, trials a map would drop go through the
        # screen or the exact evaluation instead: same outcome, more work
        _same(a, b, with_work=False)
        assert a.evaluated_pairs >= b.evaluated_pairs
    # the same box once its maps exist: every counter agrees
    print("no-alloc widom: %s form, maps built" % form, flush=True)
    got = [_one(ref_box, ref_box.widom_enqueue_noalloc, n, seed, excl) for n, seed in sizes]
    for a, b in zip(got, ref):
        _same(a, b, with_work=True)
    ref_box.close()


def test_noalloc_refuses_batches_beyond_its_block(engine_mod):
    box, excl = _box(engine_mod, engine_mod.GRASPA_MIXED)
    with pytest.raises(RuntimeError) as err:
        box.widom_enqueue_noalloc(1, 1_000_000, 5, excl=excl)
    assert "scratch" in str(err.value)
    # nothing was queued, and the entry point still works
    t = box.widom_totals(reset=True)
    assert t.n == 0
    t = _one(box, box.widom_enqueue_noalloc, 2000, 7, excl)
    assert t.n == 2000
    box.close()
