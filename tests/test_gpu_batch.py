"""Device-resident batched GCMC (graspa_batch_*) against the oracle.

The plain FP64 form draws the same counter-based uniforms as the oracle's
batched driver (or_batch_cycles), so it must reproduce the oracle's move
sequence, accept/reject decisions and molecule counts exactly and its energy
changes to 1e-10 relative.  The fast form must keep every per-move energy
change within 1e-5 kJ/mol (1e-3 internal units) of the oracle.  A box's
trajectory must be bit-identical alone and inside a batch of other boxes, and
the running energy must equal the change of the recomputed totals."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENERGY_KEYS = ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald", "tail")
MIXED_DE_TOL = 1e-3      # internal units (10 J/mol): 1e-5 kJ/mol


def _mixture(seed=11, n_fw=200, L=24.0, P=5e6, tail=False):
    from graspa_amd import synth
    s = synth.build(n_fw=n_fw, L=L, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=seed, T=300.0, P=P,
                    cycles=10, probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5, reinsertion=1.0,
                                          identity=1.0))
    if tail:
        # tail corrections on a random subset of type pairs
        rng = np.random.default_rng(3)
        T = s.ntypes
        t = rng.uniform(-80.0, -5.0, (T, T))
        s.tail = (t + t.T) / 2
        u = (rng.uniform(size=(T, T)) < 0.6).astype(np.int32)
        s.use_tail = np.maximum(u, u.T)
    return s


def _oracle_run(system, seed, stream, cycles, n_max=20000):
    from oracle import binding as ob
    o = ob.OracleSystem(system, capacity=128)
    mc = ob.make_mc(system, o, production=1)
    _, skf, ska, _, _, _ = o.ewald_total()
    tr = (ob.OrTrace * n_max)()
    mi = C.c_int64(0)
    n = ob.lib().or_batch_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), seed, stream,
                                 C.byref(mi), 0, cycles, tr, n_max)
    assert 0 < n <= n_max
    return o, mc, tr[:n]


def _engine_batch(engine_mod, systems, seeds, streams, cycles, prec, trace_cap=20000, moves_per_launch=0):
    boxes = []
    for s in systems:
        b = engine_mod.Box(s, precision=prec, capacity=128)
        b.init_ewald()
        boxes.append(b)
    bt = engine_mod.Batch(boxes, seeds, streams, trace_cap=trace_cap, cycle_cap=4096)
    print("batch case n_boxes=%d cycles=%d prec=%d" % (len(boxes), cycles, prec), flush=True)
    moves = bt.run(cycles, production=True, moves_per_launch=moves_per_launch)
    return boxes, bt, moves


def _same_sequence(otr, gtr):
    assert len(otr) == len(gtr)
    for i, (a, b) in enumerate(zip(otr, gtr)):
        assert (a.cycle, a.step, a.move, a.comp, a.mol) == (b.cycle, b.step, b.move, b.comp, b.mol), i
        assert a.accepted == b.accepted and a.overlap == b.overlap, (i, a.move, a.pacc, b.pacc)


@pytest.mark.parametrize("moves_per_launch", [0, 7])
def test_batch_fp64_reproduces_oracle(engine_mod, moves_per_launch):
    s = _mixture(tail=True)
    assert s.has_tail
    cycles = 20
    o, mc, otr = _oracle_run(s, seed=77, stream=3, cycles=cycles)
    boxes, bt, moves = _engine_batch(engine_mod, [s], [77], [3], cycles, engine_mod.GRASPA_FP64,
                                     moves_per_launch=moves_per_launch)
    gtr = bt.trace(0)
    assert moves == len(otr)
    _same_sequence(otr, gtr)
    n_acc = {m: 0 for m in range(5)}
    for i, (a, b) in enumerate(zip(otr, gtr)):
        n_acc[a.move] = n_acc.get(a.move, 0) + a.accepted
        for k in ENERGY_KEYS:
            x, y = getattr(a.de, k), getattr(b.de, k)
            assert abs(x - y) <= 1e-9 + 1e-10 * abs(x), (i, k, x, y)
        assert abs(a.pacc - b.pacc) <= 1e-10 * abs(a.pacc) + 1e-300, (i, a.pacc, b.pacc)
        assert abs(a.rosen - b.rosen) <= 1e-10 * abs(a.rosen) + 1e-300, (i, a.rosen, b.rosen)
    assert n_acc[2] > 3 and n_acc[0] + n_acc[1] > 3
    assert n_acc.get(6, 0) >= 1              # reinsertions accepted
    assert n_acc.get(7, 0) >= 4              # identity swaps accepted
    st = bt.stats(0)
    # identity-swap counters: attempts and acceptances per (old, new) species pair
    from graspa_amd import engine
    for ci in (1, 2):
        for cj in (1, 2):
            assert st.counters[ci - 1][engine.CNT_IDSW_TOT + cj - 1] == mc.idsw_total[ci][cj]
            assert st.counters[ci - 1][engine.CNT_IDSW_ACC + cj - 1] == mc.idsw_acc[ci][cj]
    assert st.moves == len(otr) and st.cycles == cycles
    for ci in (1, 2):
        assert int(st.n_mol[ci - 1]) == o.comps[ci].n_mol
    tot = lambda e: sum(getattr(e, k) for k in ENERGY_KEYS)
    assert abs(tot(st.running) - tot(mc.running)) <= 1e-9 * max(1.0, abs(tot(mc.running)))
    # the device store after the run holds the oracle's molecules
    bt.sync()
    for ci in (1, 2):
        pos, q, t, m = boxes[0].get_molecules(ci)
        opos, oq, ot, om = o.comp_arrays(ci)
        na = len(q)
        assert np.allclose(pos, opos[:na], rtol=0, atol=1e-10)
        assert np.array_equal(t, ot[:na])
    bt.close()


def test_batch_mixed_within_tolerance(engine_mod):
    s = _mixture()
    cycles = 12
    o, mc, otr = _oracle_run(s, seed=78, stream=0, cycles=cycles)
    boxes, bt, moves = _engine_batch(engine_mod, [s], [78], [0], cycles, engine_mod.GRASPA_MIXED)
    gtr = bt.trace(0)
    _same_sequence(otr, gtr)
    errs = []
    for a, b in zip(otr, gtr):
        if a.accepted or a.move == 4:
            for k in ENERGY_KEYS:
                errs.append(abs(getattr(a.de, k) - getattr(b.de, k)))
    assert len(errs) > 20
    assert max(errs) <= MIXED_DE_TOL, max(errs)
    print("mixed batch max |dE| = %.3e internal units" % max(errs))
    bt.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_alone_equals_in_mix(engine_mod, prec):
    """A box's trajectory does not depend on the other boxes of the batch."""
    from graspa_amd import synth
    a = _mixture(seed=21)
    others = [synth.config0(seed=5, n_fw=400, L=26.0, cycles=5), _mixture(seed=22, n_fw=350, L=25.0, P=2e5)]
    others[0].pressure_pa = 3e6
    _, alone, _ = _engine_batch(engine_mod, [a], [9], [4], 8, prec)
    _, mix, _ = _engine_batch(engine_mod, [others[0], _mixture(seed=21), others[1]], [1, 9, 2], [0, 4, 1], 8, prec,
                              moves_per_launch=13)
    ta, tm = alone.trace(0), mix.trace(1)
    assert len(ta) == len(tm) > 100
    for i, (x, y) in enumerate(zip(ta, tm)):
        assert (x.move, x.comp, x.mol, x.accepted, x.overlap) == (y.move, y.comp, y.mol, y.accepted, y.overlap), i
        assert x.pacc == y.pacc and x.rosen == y.rosen, i
        for k in ENERGY_KEYS:
            assert getattr(x.de, k) == getattr(y.de, k), (i, k)
    sa, sm = alone.stats(0), mix.stats(1)
    assert np.array_equal(sa.n_mol, sm.n_mol)
    assert np.array_equal(sa.counters, sm.counters)
    # every box's busy time is inside the launches that ran it
    busy = mix.busy_seconds()
    kern_ms, launches = mix.kernel_time()
    assert launches > 1 and np.all(busy > 0), busy
    assert busy.max() <= kern_ms * 1e-3 * 1.01 + 1e-4, (busy, kern_ms)
    alone.close()
    mix.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_running_energy_equals_recomputed_totals(engine_mod, prec):
    """Running sum of accepted dE = change of the recomputed totals; the
    running structure factors equal the rebuilt ones."""
    from graspa_amd import engine, synth
    systems = [_mixture(seed=31, n_fw=600, L=26.0, P=2e6), synth.config3_box(1)]
    boxes = []
    e0, f0 = [], []
    for s in systems:
        b = engine_mod.Box(s, precision=prec, capacity=256)
        f0.append(b.init_ewald())
        e0.append(b.total_energy())
        boxes.append(b)
    bt = engine_mod.Batch(boxes, [3, 4], [0, 1], trace_cap=0, cycle_cap=64)
    bt.run(6, production=True)
    bt.sync()
    for i, (s, b) in enumerate(zip(systems, boxes)):
        st = bt.stats(i)
        assert st.cycles == 6
        _, sk_run = b.structure_factors()
        e1 = b.total_energy()
        f1 = b.init_ewald()
        _, sk_new = b.structure_factors()
        assert np.abs(sk_run - sk_new).max() <= 1e-9 * max(1.0, np.abs(sk_new).max())
        r = st.running
        for k in ("hg_vdw", "gg_vdw", "hg_real", "gg_real"):
            d = getattr(e1, k) - getattr(e0[i], k)
            assert abs(getattr(r, k) - d) <= 1e-7 * max(1.0, abs(d)), (i, k, getattr(r, k), d)
        excl = 0.0
        for ci in range(1, len(s.components) + 1):
            intra, selfe = engine.exclusion_energy(s, ci)
            excl += (st.n_mol[ci - 1] - 0) * (intra + selfe)
        df = (f1.gg_ewald + f1.hg_ewald) - (f0[i].gg_ewald + f0[i].hg_ewald) - excl
        assert abs(r.gg_ewald + r.hg_ewald - df) <= 1e-7 * max(1.0, abs(df)), (i, r.gg_ewald + r.hg_ewald, df)
    times = bt.cycle_times(0, 64)
    assert len(times) == 6 and np.all(times > 0)
    bt.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_large_replicated_box_reproduces_oracle(engine_mod, prec):
    """A 48 A box (2x2x2 copies of a 24 A framework cell, 2,400 atoms, so
    minimum images and cutoffs span several framework cells): the FP64 form
    reproduces the oracle's batched driver move for move, the fast form stays
    within the tolerance on the same sequence."""
    from graspa_amd import synth
    s = synth.build(n_fw=300, L=24.0, ucells=2, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=12, T=300.0,
                    P=5e6, cycles=10, probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5, reinsertion=1.0,
                                                 identity=1.0))
    assert len(s.fw_type) == 2400 and s.cell[0] == 48.0
    cycles = 10
    o, mc, otr = _oracle_run(s, seed=79, stream=2, cycles=cycles)
    boxes, bt, moves = _engine_batch(engine_mod, [s], [79], [2], cycles, prec)
    gtr = bt.trace(0)
    _same_sequence(otr, gtr)
    n_acc = 0
    for i, (a, b) in enumerate(zip(otr, gtr)):
        n_acc += a.accepted
        if not (a.accepted or a.move == 4):
            continue
        for k in ENERGY_KEYS:
            x, y = getattr(a.de, k), getattr(b.de, k)
            tol = 1e-9 + 1e-10 * abs(x) if prec == 0 else MIXED_DE_TOL
            assert abs(x - y) <= tol, (i, k, x, y)
    assert n_acc > 10
    bt.sync()
    for ci in (1, 2):
        pos, q, t, m = boxes[0].get_molecules(ci)
        opos, oq, ot, om = o.comp_arrays(ci)
        assert len(q) == o.comps[ci].n_mol * o.comps[ci].molsize
        assert np.allclose(pos, opos[:len(q)], rtol=0, atol=1e-10 if prec == 0 else 1e-6)
    bt.close()
