"""The oracle's GCMC drivers against their own energy bookkeeping.

After any number of steps the running sum of accepted energy changes must
equal the change of the total energy recomputed from scratch (pair sums over
all atoms, Fourier sums from structure factors rebuilt from the atoms, minus
the rigid exclusion of every molecule), and the running adsorbate structure
factor must equal the rebuilt one.  A dropped or mis-signed term in any move's
energy change (translation, rotation, insertion, deletion, reinsertion,
identity swap, the Ewald same/cross split, the exclusion correction, the tail
change) or a commit that stores the wrong molecule breaks one of these
equalities.  Both drivers are checked: the
reference's random-number order (or_mc_cycles) and the batched form with
counter-based streams (or_batch_cycles)."""
import ctypes as C

import numpy as np
import pytest


def _system(P=5e6, n_fw=200, reinsertion=1.0, identity=1.0, tail=True):
    from graspa_amd import synth
    s = synth.build(n_fw=n_fw, L=24.0, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=11, T=300.0, P=P,
                    cycles=10, probs=dict(translation=1.0, rotation=1.0, swap=2.0, widom=0.5,
                                          reinsertion=reinsertion, identity=identity))
    if tail:
        # tail corrections on a random subset of type pairs, so the tail
        # change of insertions, deletions and identity swaps is booked too
        rng = np.random.default_rng(3)
        T = s.ntypes
        t = rng.uniform(-80.0, -5.0, (T, T))
        s.tail = (t + t.T) / 2
        u = (rng.uniform(size=(T, T)) < 0.6).astype(np.int32)
        s.use_tail = np.maximum(u, u.T)
    return s


def _totals(o):
    e = o.total_vdw_real()
    f, _, ska, _, _, _ = o.ewald_total()
    excl = 0.0
    for ci in range(1, o.ncomp):
        i, sf = o.exclusion(ci)
        excl += o.comps[ci].n_mol * (i + sf)
    return dict(hg_vdw=e.hg_vdw, gg_vdw=e.gg_vdw, hg_real=e.hg_real, gg_real=e.gg_real,
                ewald=f.gg_ewald + f.hg_ewald - excl), ska


def _n_mol(o):
    return [o.comps[c].n_mol for c in range(1, o.ncomp)]


def _check_drift(o, mc, before, after, ska_running, ska_rebuilt):
    r = mc.running
    for k in ("hg_vdw", "gg_vdw", "hg_real", "gg_real"):
        d = after[k] - before[k]
        assert abs(getattr(r, k) - d) <= 1e-9 * max(1.0, abs(d)), (k, getattr(r, k), d)
    de = after["ewald"] - before["ewald"]
    assert abs(r.gg_ewald + r.hg_ewald - de) <= 1e-9 * max(1.0, abs(de)), (r.gg_ewald + r.hg_ewald, de)
    scale = max(1.0, float(np.abs(ska_rebuilt).max()))
    assert np.abs(ska_running - ska_rebuilt).max() <= 1e-11 * scale


@pytest.mark.parametrize("form", ["reference_order", "batched"])
def test_running_energy_equals_recomputed_totals(oracle_lib, form):
    from oracle import binding as ob
    from graspa_amd import report
    s = _system()
    assert s.has_tail
    o = ob.OracleSystem(s, capacity=128)
    mc = ob.make_mc(s, o, production=1)
    before, ska = _totals(o)
    tail0 = report.tail_total(s, _n_mol(o))
    f0, skf, ska, _, _, _ = o.ewald_total()
    n_max = 5000
    tr = (ob.OrTrace * n_max)()
    if form == "reference_order":
        assert ob.lib().or_mc_setup(C.byref(mc)) == 0
        n = ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 0, 20, tr, n_max)
    else:
        mi = C.c_int64(0)
        n = ob.lib().or_batch_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 77, 3,
                                     C.byref(mi), 0, 20, tr, n_max)
        assert mi.value == n
    assert n > 0
    moves = [tr[i].move for i in range(min(n, n_max))]
    acc = [tr[i].accepted for i in range(min(n, n_max))]
    # every move type ran and some were accepted (6: reinsertion, 7: identity swap)
    for m in (0, 1, 2, 3, 4, 6, 7):
        assert m in moves, m
    assert sum(a for a, m in zip(acc, moves) if m == 2) > 3
    assert sum(a for a, m in zip(acc, moves) if m in (0, 1)) > 3
    assert sum(a for a, m in zip(acc, moves) if m == 6) >= 1
    # identity swaps between the two species were accepted in both directions
    idsw = np.ctypeslib.as_array(mc.idsw_acc)
    assert idsw[1][2] >= 1 and idsw[2][1] >= 1, idsw[:3, :3]
    # (an identity swap drawn while the box holds no adsorbate is not attempted)
    assert 0 < int(np.ctypeslib.as_array(mc.idsw_total).sum()) <= sum(1 for m in moves if m == 7)
    after, ska_rebuilt = _totals(o)
    _check_drift(o, mc, before, after, ska, ska_rebuilt)
    dt = report.tail_total(s, _n_mol(o)) - tail0
    assert abs(mc.running.tail - dt) <= 1e-9 * max(1.0, abs(dt)), (mc.running.tail, dt)
    # pseudo-atom counts follow the molecules
    n = s.pseudo_counts(_n_mol(o))
    assert [mc.n_pseudo[t] for t in range(len(n))] == [int(v) for v in n]


def test_batched_form_is_deterministic_and_stream_dependent(oracle_lib):
    """Same (seed, stream) gives the same trajectory; another stream another."""
    from oracle import binding as ob

    def run(seed, stream):
        s = _system()
        o = ob.OracleSystem(s, capacity=128)
        mc = ob.make_mc(s, o, production=1)
        _, skf, ska, _, _, _ = o.ewald_total()
        tr = (ob.OrTrace * 2000)()
        mi = C.c_int64(0)
        n = ob.lib().or_batch_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), seed, stream,
                                     C.byref(mi), 0, 5, tr, 2000)
        return [(tr[i].move, tr[i].comp, tr[i].mol, tr[i].accepted, tr[i].de.hg_vdw) for i in range(n)]

    a, b, c = run(5, 0), run(5, 0), run(5, 1)
    assert a == b
    assert a != c


def test_reinsertion_samples_the_boltzmann_distribution(oracle_lib):
    """One methane molecule among 150 framework atoms, reinsertion moves only:
    the oracle's chain must sample exp(-beta U(r)), so its mean energy equals
    <U> = int U e^{-beta U} / int e^{-beta U} over the box, estimated here by
    plain Monte Carlo integration over 400,000 uniform points.  Accepting the
    grown configuration without the retrace weight (or with the retrace
    missing the other first-bead weights) samples the CBMC growth
    distribution instead, which is shifted towards low energies."""
    from graspa_amd import synth
    from oracle import binding as ob
    s = synth.build(n_fw=150, L=25.0, species=("methane",), seed=5, T=200.0, P=1e3, charges=False, K=10,
                    probs=dict(reinsertion=1.0))
    T = s.ntypes
    fw = np.asarray(s.fw_pos)
    t_fw = np.asarray(s.fw_type)
    site_t = int(s.components[0].types[0])
    eps = np.asarray(s.eps)[t_fw, site_t]
    sig2 = np.asarray(s.sig)[t_fw, site_t] ** 2
    L = s.cell[0]

    def energy(p):
        d = fw[None, :, :] - p[:, None, :]
        d -= L * np.round(d / L)
        rr = (d * d).sum(-1)
        x = (sig2[None, :] / rr) ** 3
        e = 4.0 * eps[None, :] * (x * x - x)
        return np.where(rr < s.cut_vdw ** 2, e, 0.0).sum(1)

    rng = np.random.default_rng(0)
    pts = rng.uniform(0, L, (400000, 3))
    U = np.concatenate([energy(pts[i:i + 20000]) for i in range(0, len(pts), 20000)])
    w = np.exp(-s.beta * (U - U.min()))
    u_ref = float((U * w).sum() / w.sum())
    start = pts[np.argmin(U)]
    c = s.components[0]
    c.init_pos = start[None, :].copy()
    c.init_molid = np.zeros(1, dtype=np.int64)
    c.init_charge = np.zeros(1)
    o = ob.OracleSystem(s, capacity=4)
    mc = ob.make_mc(s, o, production=1)
    _, skf, ska, _, _, _ = o.ewald_total()
    e0 = o.total_vdw_real().hg_vdw
    assert abs(e0 - energy(start[None, :])[0]) < 1e-9 * max(1.0, abs(e0))
    assert ob.lib().or_mc_setup(C.byref(mc)) == 0
    n_cyc = 1500
    tr = (ob.OrTrace * (20 * n_cyc))()
    n = ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 0, n_cyc, tr, 20 * n_cyc)
    assert n == 20 * n_cyc
    de = np.array([tr[i].de.hg_vdw for i in range(n)])
    e = e0 + np.cumsum(de)
    acc = np.mean([tr[i].accepted for i in range(n)])
    assert 0.02 < acc < 0.98
    blocks = e.reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(e.mean() - u_ref) < 4.0 * err + 0.01 * abs(u_ref), (e.mean(), u_ref, err)


def test_identity_swap_samples_the_semigrand_distribution(oracle_lib):
    """One molecule that is either methane or argon among 150 framework atoms,
    reinsertion and identity-swap moves only.  Identity swaps with symmetric
    proposals and Pacc = (f_new / f_old) W_new / W_old leave
    pi(s, r) ~ f_s exp(-beta U_s(r)) invariant, so the fraction of cycles
    spent as methane is x_m Z_m / (x_m Z_m + x_a Z_a) with
    Z_s = int exp(-beta U_s) over the box (plain Monte Carlo integration over
    300,000 uniform points).  An inverted prefactor, a dropped retrace weight
    or a mis-signed energy change moves the fraction far outside the bound."""
    from graspa_amd import synth
    from oracle import binding as ob
    x_m = 0.3
    s = synth.build(n_fw=150, L=25.0, species=("methane", "argon"), mol_fraction=[x_m, 1.0 - x_m], seed=5,
                    T=200.0, P=1e3, charges=False, K=10, probs=dict(reinsertion=1.0, identity=1.0))
    fw = np.asarray(s.fw_pos)
    t_fw = np.asarray(s.fw_type)
    L = s.cell[0]

    def energy(p, site_t):
        eps = np.asarray(s.eps)[t_fw, site_t]
        sig2 = np.asarray(s.sig)[t_fw, site_t] ** 2
        d = fw[None, :, :] - p[:, None, :]
        d -= L * np.round(d / L)
        rr = (d * d).sum(-1)
        x = (sig2[None, :] / rr) ** 3
        e = 4.0 * eps[None, :] * (x * x - x)
        return np.where(rr < s.cut_vdw ** 2, e, 0.0).sum(1)

    rng = np.random.default_rng(1)
    pts = rng.uniform(0, L, (300000, 3))
    Um = np.concatenate([energy(pts[i:i + 20000], int(s.components[0].types[0])) for i in range(0, len(pts), 20000)])
    Ua = np.concatenate([energy(pts[i:i + 20000], int(s.components[1].types[0])) for i in range(0, len(pts), 20000)])
    u0 = min(Um.min(), Ua.min())
    Zm = np.exp(-s.beta * (Um - u0)).mean()
    Za = np.exp(-s.beta * (Ua - u0)).mean()
    p_ref = x_m * Zm / (x_m * Zm + (1.0 - x_m) * Za)
    # start as methane at the methane energy minimum of the sample
    c = s.components[0]
    c.init_pos = pts[np.argmin(Um)][None, :].copy()
    c.init_molid = np.zeros(1, dtype=np.int64)
    c.init_charge = np.zeros(1)
    o = ob.OracleSystem(s, capacity=4)
    mc = ob.make_mc(s, o, production=1)
    _, skf, ska, _, _, _ = o.ewald_total()
    assert ob.lib().or_mc_setup(C.byref(mc)) == 0
    tr = (ob.OrTrace * 64)()
    n_cyc = 3000
    is_m = np.zeros(n_cyc)
    for cyc in range(n_cyc):
        ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), cyc, 1, tr, 64)
        assert o.comps[1].n_mol + o.comps[2].n_mol == 1
        is_m[cyc] = o.comps[1].n_mol
    acc = np.ctypeslib.as_array(mc.idsw_acc)
    assert acc[1][2] > 20 and acc[2][1] > 20, acc[:3, :3]
    blocks = is_m.reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / np.sqrt(len(blocks))
    assert abs(is_m.mean() - p_ref) < 4.0 * err + 0.02, (is_m.mean(), p_ref, err)


def test_create_molecules_then_cycles_keep_the_bookkeeping(oracle_lib):
    """CreateMolecule_InOneBox: the requested molecules of each species are
    created from an empty box (insertions accepted whenever the construction
    succeeds, no move counters), their energy change equals the change of the
    recomputed totals, and the cycles that follow keep the bookkeeping."""
    from oracle import binding as ob
    s = _system(P=1e5)
    o = ob.OracleSystem(s, capacity=128)
    mc = ob.make_mc(s, o, production=1)
    before, _ = _totals(o)
    _, skf, ska, _, _, _ = o.ewald_total()
    assert ob.lib().or_mc_setup(C.byref(mc)) == 0
    de = ob.OrEnergy()
    for comp, n in ((1, 12), (2, 7)):
        got = ob.lib().or_mc_create_molecules(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), comp, n,
                                              100000, C.byref(de))
        assert got == n
        assert o.comps[comp].n_mol == n
    assert mc.ins_total[1] == 0 and mc.ins_acc[1] == 0
    mid, ska_rebuilt = _totals(o)
    for k in ("hg_vdw", "gg_vdw", "hg_real", "gg_real"):
        d = mid[k] - before[k]
        assert abs(getattr(de, k) - d) <= 1e-9 * max(1.0, abs(d)), (k, getattr(de, k), d)
    # the running energy holds the creation; the cycles that follow add to it
    assert mc.total_molecules == 1 + 12 + 7
    tr = (ob.OrTrace * 5000)()
    ob.lib().or_mc_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 0, 5, tr, 5000)
    after, ska_rebuilt = _totals(o)
    _check_drift(o, mc, before, after, ska, ska_rebuilt)
