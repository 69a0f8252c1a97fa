"""The stated workloads on the GPU, each against the oracle or a physical law.

- plain FP64 GCMC of CO2 in the 1,536-atom, 24 A framework (translate : rotate
  : insert : delete 1:1:1:1): the reference-order driver and the
  device-resident driver reproduce the oracle move for move over 100 cycles;
- rigid n-butane CBMC (3,000-atom, 30 A uncharged framework, 16 x 8 trials)
  and a 400-atom variant at 1e7 Pa where molecules go in: the device driver
  reproduces the oracle's batched driver move for move in FP64, and the fast
  form stays within 1e-5 kJ/mol per move on the same sequence;
- 256 independent boxes (frameworks of 500-6,000 atoms per cell, up to 47,520
  atoms after replication, 18-40 A, CO2/N2) in one persistent launch: the
  largest box and a sample of others reproduce the oracle in FP64, and every
  box's running energy equals the change of its recomputed totals;
- a served Poisson stream of jobs from the 256-box distribution: every job's
  result alone equals its result in the mix, in both forms;
- the ideal gas through the device driver: Poisson loading with the SI
  ideal-gas mean.
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENERGY_KEYS = ("hg_vdw", "hg_real", "gg_vdw", "gg_real", "hg_ewald", "gg_ewald", "tail")
MIXED_DE_TOL = 1e-3      # internal units (10 J/mol): 1e-5 kJ/mol
BOLTZMANN = 1.380649e-23


def _oracle_batch(system, seed, stream, cycles, n_max=40000, capacity=256):
    from oracle import binding as ob
    o = ob.OracleSystem(system, capacity=capacity)
    mc = ob.make_mc(system, o, production=1)
    _, skf, ska, _, _, _ = o.ewald_total()
    tr = (ob.OrTrace * n_max)()
    mi = C.c_int64(0)
    n = ob.lib().or_batch_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), seed, stream,
                                 C.byref(mi), 0, cycles, tr, n_max)
    assert 0 < n <= n_max
    return o, mc, tr[:n]


def _check_sequence(otr, gtr, fp64=True):
    """Same moves, decisions and overlaps; energies to 1e-10 relative (FP64)
    or within the fast form's tolerance.  Returns the number of accepted moves."""
    assert len(otr) == len(gtr), (len(otr), len(gtr))
    n_acc = 0
    for i, (a, b) in enumerate(zip(otr, gtr)):
        assert (a.cycle, a.step, a.move, a.comp, a.mol) == (b.cycle, b.step, b.move, b.comp, b.mol), i
        assert a.accepted == b.accepted and a.overlap == b.overlap, (i, a.move, a.pacc, b.pacc)
        n_acc += a.accepted
        if fp64:
            for k in ENERGY_KEYS:
                x, y = getattr(a.de, k), getattr(b.de, k)
                assert abs(x - y) <= 1e-9 + 1e-10 * abs(x), (i, k, x, y)
            assert abs(a.pacc - b.pacc) <= 1e-10 * abs(a.pacc) + 1e-300, (i, a.pacc, b.pacc)
        elif a.accepted or a.move == 4:
            for k in ENERGY_KEYS:
                assert abs(getattr(a.de, k) - getattr(b.de, k)) <= MIXED_DE_TOL, (i, k)
    return n_acc


def _device_batch(engine_mod, systems, seeds, streams, cycles, prec, trace_cap=40000, capacity=256):
    boxes = []
    for s in systems:
        b = engine_mod.Box(s, precision=prec, capacity=capacity)
        b.init_ewald()
        boxes.append(b)
    bt = engine_mod.Batch(boxes, seeds, streams, trace_cap=trace_cap, cycle_cap=max(cycles, 1))
    print("device batch boxes=%d cycles=%d prec=%d" % (len(boxes), cycles, prec), flush=True)
    bt.run(cycles, production=True)
    return boxes, bt


# --------------------------------------------------------------------------
# plain FP64 GCMC, one box: 1,536 atoms, 24 A
# --------------------------------------------------------------------------
def test_config0_reference_order_driver_reproduces_oracle(engine_mod):
    from graspa_amd import synth
    from oracle import binding as ob
    s = synth.config0()
    assert len(s.fw_type) == 1536 and s.cell[0] == 24.0 and s.kmax == (7, 7, 7)
    cycles = 100
    o = ob.OracleSystem(s, capacity=256)
    mc = ob.make_mc(s, o, production=1)
    assert ob.lib().or_mc_setup(C.byref(mc)) == 0
    _, skf, ska, _, _, _ = o.ewald_total()
    n_max = 40000
    otr = (ob.OrTrace * n_max)()
    n_o = ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 0, cycles, otr, n_max)
    print("config0 parity driver cycles=%d" % cycles, flush=True)
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_FP64, capacity=256)
    box.init_ewald()
    gm = engine_mod.MonteCarlo(box)
    n_g, gtr = gm.cycles(0, cycles, production=True, max_trace=n_max)
    assert n_g == n_o >= 20 * cycles
    _check_sequence(otr[:n_o], gtr[:n_g])
    e, nm, cnt, _, _ = gm.stats()
    assert int(nm[0]) == o.comps[1].n_mol
    # every move type was attempted (translations and rotations only with a
    # molecule in the box)
    moves = {otr[i].move for i in range(n_o)}
    assert {2, 3} <= moves
    gm.close()
    box.close()


def test_config0_device_driver_reproduces_oracle(engine_mod):
    from graspa_amd import synth
    s = synth.config0()
    cycles = 100
    o, mc, otr = _oracle_batch(s, 31, 0, cycles)
    boxes, bt = _device_batch(engine_mod, [s], [31], [0], cycles, engine_mod.GRASPA_FP64)
    _check_sequence(otr, bt.trace(0))
    st = bt.stats(0)
    assert st.cycles == cycles and int(st.n_mol[0]) == o.comps[1].n_mol
    bt.close()
    boxes[0].close()


# --------------------------------------------------------------------------
# rigid n-butane CBMC, uncharged, 16 x 8 trials
# --------------------------------------------------------------------------
@pytest.mark.parametrize("variant", ["3000_atoms", "400_atoms_1e7Pa"])
def test_config2_butane_device_driver(engine_mod, variant):
    from graspa_amd import synth
    s = synth.config2() if variant == "3000_atoms" else synth.config2(n_fw=400, P=1e7)
    assert s.no_charges and s.n_trial_pos == 16 and s.n_trial_orient == 8 and s.components[0].molsize == 4
    cycles = 25 if variant == "3000_atoms" else 60
    o, mc, otr = _oracle_batch(s, 11, 0, cycles)
    boxes, bt = _device_batch(engine_mod, [s], [11], [0], cycles, engine_mod.GRASPA_FP64)
    n_acc = _check_sequence(otr, bt.trace(0))
    assert int(bt.stats(0).n_mol[0]) == o.comps[1].n_mol
    bt.close()
    boxes[0].close()
    if variant != "3000_atoms":
        # molecules were inserted and moved
        assert n_acc > 50 and o.comps[1].n_mol > 5
        assert any(x.accepted and x.move in (0, 1) for x in otr)
    boxes, bt = _device_batch(engine_mod, [s], [11], [0], cycles, engine_mod.GRASPA_MIXED)
    _check_sequence(otr, bt.trace(0), fp64=False)
    bt.close()
    boxes[0].close()


# --------------------------------------------------------------------------
# 256 boxes in one persistent launch
# --------------------------------------------------------------------------
CONFIG3_SAMPLE = (10, 160, 0, 77, 200)    # 10: the largest box (47,520 atoms); 160: 572 atoms, 36 A


def test_config3_all_boxes_one_launch(engine_mod):
    from graspa_amd import engine, synth
    systems = [synth.config3_box(b) for b in range(256)]
    assert len(systems[10].fw_type) == 47520
    cycles = 10
    boxes, e0, f0 = [], [], []
    for s in systems:
        b = engine_mod.Box(s, precision=engine_mod.GRASPA_FP64, capacity=256)
        f0.append(b.init_ewald())
        e0.append(b.total_energy())
        boxes.append(b)
    bt = engine_mod.Batch(boxes, [3000 + b for b in range(256)], list(range(256)), trace_cap=4000, cycle_cap=cycles)
    print("config3 256 boxes, one launch, %d cycles" % cycles, flush=True)
    l0 = bt.kernel_time()[1]
    moves = bt.run(cycles, production=True)
    assert bt.kernel_time()[1] - l0 == 1          # one persistent launch for all boxes and cycles
    assert moves > 256 * 20 * cycles - 1
    bt.sync()
    # every box: running energy = change of the recomputed totals, running S(k) = rebuilt S(k)
    for i, (s, b) in enumerate(zip(systems, boxes)):
        st = bt.stats(i)
        assert st.cycles == cycles
        _, sk_run = b.structure_factors()
        e1 = b.total_energy()
        f1 = b.init_ewald()
        _, sk_new = b.structure_factors()
        assert np.abs(sk_run - sk_new).max() <= 1e-9 * max(1.0, np.abs(sk_new).max()), i
        r = st.running
        for k in ("hg_vdw", "gg_vdw", "hg_real", "gg_real"):
            d = getattr(e1, k) - getattr(e0[i], k)
            assert abs(getattr(r, k) - d) <= 1e-7 * max(1.0, abs(d)), (i, k, getattr(r, k), d)
        excl = 0.0
        for ci in range(1, len(s.components) + 1):
            intra, selfe = engine.exclusion_energy(s, ci)
            excl += st.n_mol[ci - 1] * (intra + selfe)
        df = (f1.gg_ewald + f1.hg_ewald) - (f0[i].gg_ewald + f0[i].hg_ewald) - excl
        assert abs(r.gg_ewald + r.hg_ewald - df) <= 1e-7 * max(1.0, abs(df)), (i, r.gg_ewald + r.hg_ewald, df)
    # the sample against the oracle's batched driver, move for move
    for i in CONFIG3_SAMPLE:
        print("config3 box %d (%d atoms) vs oracle" % (i, len(systems[i].fw_type)), flush=True)
        o, mc, otr = _oracle_batch(systems[i], 3000 + i, i, cycles)
        _check_sequence(otr, bt.trace(i))
        for ci in (1, 2):
            assert int(bt.stats(i).n_mol[ci - 1]) == o.comps[ci].n_mol
    bt.close()
    for b in boxes:
        b.close()


def test_config3_fast_form_alone_equals_in_batch(engine_mod):
    """Fast form: a box gives bit-identical results in the 256-box launch and
    alone."""
    from graspa_amd import synth
    systems = [synth.config3_box(b) for b in range(256)]
    cycles = 8
    boxes, bt = _device_batch(engine_mod, systems, [3000 + b for b in range(256)], list(range(256)), cycles,
                              engine_mod.GRASPA_MIXED, trace_cap=2000)
    for i in (10, 254, 33):
        ab, at = _device_batch(engine_mod, [systems[i]], [3000 + i], [i], cycles, engine_mod.GRASPA_MIXED,
                               trace_cap=2000)
        ta, tm = at.trace(0), bt.trace(i)
        assert len(ta) == len(tm) > 0
        for x, y in zip(ta, tm):
            assert (x.move, x.comp, x.mol, x.accepted, x.overlap, x.pacc) == (y.move, y.comp, y.mol, y.accepted,
                                                                              y.overlap, y.pacc)
            for k in ENERGY_KEYS:
                assert getattr(x.de, k) == getattr(y.de, k)
        at.close()
        ab[0].close()
    bt.close()
    for b in boxes:
        b.close()


# --------------------------------------------------------------------------
# served stream of config-3 jobs
# --------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [0, 1])
def test_config4_served_jobs_equal_alone(engine_mod, prec):
    from graspa_amd import serve
    jobs = serve.poisson_jobs(10, 200.0, seed=4000, cycles_lo=50, cycles_hi=150)
    for j in jobs:
        j.capacity = 256
    print("config4 served jobs=10 slots=4 prec=%d" % prec, flush=True)
    srv = serve.Server(n_slots=4, precision=prec, moves_per_launch=400, cycle_cap=200)
    done = srv.run(jobs)
    srv.close()
    assert sorted(j.job_id for j in done) == list(range(10))
    assert srv.launches < sum(j.stats.moves for j in done) / 400
    for j in done:
        box = engine_mod.Box(j.system, precision=prec, capacity=j.capacity)
        box.init_ewald()
        bt = engine_mod.Batch([box], [j.seed], [j.stream], cycle_cap=200)
        bt.run(j.cycles, production=True)
        ref = bt.stats(0)
        bt.close()
        box.close()
        st = j.stats
        assert st.cycles == j.cycles == ref.cycles and st.moves == ref.moves
        assert np.array_equal(st.n_mol, ref.n_mol) and np.array_equal(st.counters, ref.counters)
        for k in ENERGY_KEYS:
            assert getattr(st.running, k) == getattr(ref.running, k), (j.job_id, k)


# --------------------------------------------------------------------------
# ideal gas through the device driver
# --------------------------------------------------------------------------
def test_device_driver_ideal_gas_loading(engine_mod):
    """Non-interacting rigid CO2-shaped molecules (epsilon 0, no charges; one
    framework atom that interacts with nothing): 16 independent boxes, swap
    moves only.  The loading must be Poisson with mean y phi P V / (k_B T)."""
    from graspa_amd import synth
    z, L, T = 3.0, 24.0, 300.0
    P = z * BOLTZMANN * T / (L ** 3 * 1e-30)
    systems = []
    for r in range(16):
        s = synth.build(n_fw=1, L=L, species=("CO2",), seed=1, T=T, P=P, charges=False, probs=dict(swap=1.0))
        s.eps = np.zeros_like(np.asarray(s.eps))
        s.shift = np.zeros_like(np.asarray(s.shift))
        systems.append(s)
    cycles = 1500
    boxes, bt = _device_batch(engine_mod, systems, [70 + r for r in range(16)], list(range(16)), 100,
                              engine_mod.GRASPA_FP64, trace_cap=0)
    samples = []
    for blk in range(cycles // 100):
        bt.run(100, production=True)
        samples.append([bt.stats(i).n_mol[0] for i in range(16)])
    n = np.array(samples, dtype=float)          # [block][box], one snapshot every 100 cycles
    mean, var = n.mean(), n.var()
    err = n.mean(0).std(ddof=1) / math.sqrt(16)
    assert abs(mean - z) < 4.0 * max(err, math.sqrt(z / n.size)), (mean, z, err)
    assert abs(var / mean - 1.0) < 0.25, (var, mean)
    bt.close()
    for b in boxes:
        b.close()
