"""The plain FP64 GCMC driver against the oracle's restatement of RunMoves.

Both draw from the same generator in the reference's order (the oracle calls
the C library's rand(), the engine its restatement of it), so the plain FP64
form must reproduce the oracle's move sequence, accept/reject decisions,
molecule counts and energy changes exactly (energies to 1e-10 relative).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_pair(system, cycles, prec=0):
    from graspa_amd import engine
    from oracle import binding as ob
    import ctypes as C

    o = ob.OracleSystem(system, capacity=256)
    mc = ob.make_mc(system, o, production=1)
    assert ob.lib().or_mc_setup(C.byref(mc)) == 0
    of, skf, ska, _, _, _ = o.ewald_total()
    n_max = 200000
    otr = (ob.OrTrace * n_max)()
    n_o = ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ska.ctypes.data_as(ob._f64p),
                                skf.ctypes.data_as(ob._f64p), 0, cycles, otr, n_max)
    assert n_o > 0

    box = engine.Box(system, precision=prec, capacity=256)
    box.init_ewald()
    gm = engine.MonteCarlo(box)
    n_g, gtr = gm.cycles(0, cycles, production=True, max_trace=n_max)
    e, nm, cnt, ww, wn = gm.stats()
    res = dict(n_o=n_o, n_g=n_g, otr=otr, gtr=gtr, omc=mc, e=e, nm=nm, cnt=cnt, o=o, box=box)
    return res


def test_gcmc_fp64_reproduces_oracle_sequence():
    from graspa_amd import synth
    s = synth.config0(seed=1, n_fw=300, L=24.0, cycles=40)
    s.pressure_pa = 1e6
    r = _run_pair(s, 40)
    assert r["n_o"] == r["n_g"]
    n = r["n_o"]
    otr, gtr = r["otr"], r["gtr"]
    n_acc = 0
    for i in range(n):
        a, b = otr[i], gtr[i]
        assert (a.cycle, a.step, a.move, a.comp, a.mol) == (b.cycle, b.step, b.move, b.comp, b.mol), i
        assert a.accepted == b.accepted and a.overlap == b.overlap, (i, a.move)
        n_acc += a.accepted
        for k in ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald"):
            x, y = getattr(a.de, k), getattr(b.de, k)
            assert abs(x - y) <= 1e-9 + 1e-10 * abs(x), (i, k, x, y)
        if a.pacc > 0:
            assert abs(a.pacc - b.pacc) <= 1e-9 * a.pacc, (i, a.pacc, b.pacc)
    assert n_acc > 10
    # molecule counts and the running energy agree
    assert int(r["nm"][0]) == r["o"].comps[1].n_mol
    assert abs(r["e"].total() - r["omc"].running.total()) <= 1e-8 * max(1.0, abs(r["omc"].running.total()))
    # GPU state after the run equals the oracle's state
    pos, q, t, m = r["box"].get_molecules(1)
    opos, oq, ot, om = r["o"].comp_arrays(1)
    na = len(q)
    assert np.allclose(pos, opos[:na], atol=1e-12)
    assert np.array_equal(m, om[:na])


def test_reinsertion_fp64_reproduces_oracle_sequence():
    """CO2/N2 mixture with translation, rotation, Widom, reinsertion and swap
    moves in the reference's random-number order."""
    from graspa_amd import synth
    s = synth.build(n_fw=200, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=11, T=300.0, P=5e6,
                    cycles=12, probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5, reinsertion=1.0))
    r = _run_pair(s, 12)
    assert r["n_o"] == r["n_g"]
    otr, gtr = r["otr"], r["gtr"]
    n_reins = 0
    for i in range(r["n_o"]):
        a, b = otr[i], gtr[i]
        assert (a.cycle, a.step, a.move, a.comp, a.mol) == (b.cycle, b.step, b.move, b.comp, b.mol), i
        assert a.accepted == b.accepted and a.overlap == b.overlap, (i, a.move)
        if a.move == 6:
            n_reins += a.accepted
            assert abs(a.rosen - b.rosen) <= 1e-10 * abs(a.rosen) + 1e-300, (i, a.rosen, b.rosen)
        for k in ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald"):
            x, y = getattr(a.de, k), getattr(b.de, k)
            assert abs(x - y) <= 1e-9 + 1e-10 * abs(x), (i, k, x, y)
        if a.pacc > 0 and a.pacc < 1e300:
            assert abs(a.pacc - b.pacc) <= 1e-9 * a.pacc, (i, a.pacc, b.pacc)
    assert n_reins >= 2
    assert abs(r["e"].total() - r["omc"].running.total()) <= 1e-8 * max(1.0, abs(r["omc"].running.total()))
    for ci in (1, 2):
        pos, q, t, m = r["box"].get_molecules(ci)
        opos, oq, ot, om = r["o"].comp_arrays(ci)
        na = len(q)
        assert np.allclose(pos, opos[:na], atol=1e-12)


def test_identity_swap_fp64_reproduces_oracle_sequence():
    """CO2/N2 mixture with identity swaps (and the other moves), tail
    corrections on, in the reference's random-number order: the species pair,
    molecule, growth on the old first bead, retrace, Ewald, tail and
    acceptance must match the oracle move for move."""
    from graspa_amd import engine, synth
    s = synth.build(n_fw=200, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=11, T=300.0, P=5e6,
                    cycles=20, probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5, reinsertion=1.0,
                                          identity=2.0))
    rng = np.random.default_rng(3)
    t = rng.uniform(-80.0, -5.0, (s.ntypes, s.ntypes))
    s.tail = (t + t.T) / 2
    u = (rng.uniform(size=(s.ntypes, s.ntypes)) < 0.6).astype(np.int32)
    s.use_tail = np.maximum(u, u.T)
    r = _run_pair(s, 20)
    assert r["n_o"] == r["n_g"]
    otr, gtr = r["otr"], r["gtr"]
    n_id = 0
    for i in range(r["n_o"]):
        a, b = otr[i], gtr[i]
        assert (a.cycle, a.step, a.move, a.comp, a.mol) == (b.cycle, b.step, b.move, b.comp, b.mol), i
        assert a.accepted == b.accepted and a.overlap == b.overlap, (i, a.move)
        if a.move == 7:
            n_id += a.accepted
            assert abs(a.rosen - b.rosen) <= 1e-10 * abs(a.rosen) + 1e-300, (i, a.rosen, b.rosen)
        for k in ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald", "tail"):
            x, y = getattr(a.de, k), getattr(b.de, k)
            assert abs(x - y) <= 1e-9 + 1e-10 * abs(x), (i, k, x, y)
        if a.pacc > 0 and a.pacc < 1e300:
            assert abs(a.pacc - b.pacc) <= 1e-9 * a.pacc, (i, a.pacc, b.pacc)
    assert n_id >= 4
    for ci in (1, 2):
        for cj in (1, 2):
            assert r["cnt"][ci - 1][engine.CNT_IDSW_TOT + cj - 1] == r["omc"].idsw_total[ci][cj]
            assert r["cnt"][ci - 1][engine.CNT_IDSW_ACC + cj - 1] == r["omc"].idsw_acc[ci][cj]
    assert abs(r["e"].total() - r["omc"].running.total()) <= 1e-8 * max(1.0, abs(r["omc"].running.total()))
    for ci in (1, 2):
        pos, q, t, m = r["box"].get_molecules(ci)
        opos, oq, ot, om = r["o"].comp_arrays(ci)
        na = len(q)
        assert na == r["o"].comps[ci].n_mol * r["o"].comps[ci].molsize
        assert np.allclose(pos, opos[:na], atol=1e-12)
        assert np.array_equal(m, om[:na])


def test_create_molecules_fp64_reproduces_oracle():
    """CreateMolecule_InOneBox in the reference's draw order: the plain FP64
    driver creates the same molecules as the oracle, with the same energy."""
    import ctypes as C
    from graspa_amd import engine, synth
    from oracle import binding as ob
    s = synth.build(n_fw=300, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=17, T=300.0, P=1e5,
                    cycles=4, probs=dict(translation=1.0, rotation=1.0, swap=2.0))
    o = ob.OracleSystem(s, capacity=64)
    mc = ob.make_mc(s, o, production=1)
    assert ob.lib().or_mc_setup(C.byref(mc)) == 0
    _, skf, ska, _, _, _ = o.ewald_total()
    ode = ob.OrEnergy()
    for comp, n in ((1, 10), (2, 6)):
        assert ob.lib().or_mc_create_molecules(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), comp, n,
                                               100000, C.byref(ode)) == n
    box = engine.Box(s, precision=engine.GRASPA_FP64, capacity=64)
    box.init_ewald()
    gm = engine.MonteCarlo(box)
    gde = {}
    for comp, n in ((1, 10), (2, 6)):
        e = gm.create_molecules(comp, n)
        for k in ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald"):
            gde[k] = gde.get(k, 0.0) + getattr(e, k)
    for k, v in gde.items():
        x = getattr(ode, k)
        assert abs(x - v) <= 1e-9 + 1e-10 * abs(x), (k, x, v)
    for ci in (1, 2):
        pos, q, t, m = box.get_molecules(ci)
        opos, oq, ot, om = o.comp_arrays(ci)
        na = len(q)
        assert na == o.comps[ci].n_mol * o.comps[ci].molsize
        assert np.allclose(pos, opos[:na], atol=1e-12)
    # the cycles that follow stay in step with the oracle
    n_max = 20000
    otr = (ob.OrTrace * n_max)()
    n_o = ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 0, 4, otr, n_max)
    n_g, gtr = gm.cycles(0, 4, production=True, max_trace=n_max)
    assert n_o == n_g
    for i in range(n_o):
        a, b = otr[i], gtr[i]
        assert (a.move, a.comp, a.mol, a.accepted) == (b.move, b.comp, b.mol, b.accepted), i
    gm.close()
    box.close()
