"""Parity of the HIP engine with the FP64 oracle, through the C-ABI.

Plain FP64 form: per-move energies to 1e-10 relative (the reference's GPU FP64
kernels differ from its CPU FP64 code in summation order only).  Fast
("mixed") form -- FP64 pair arithmetic with an FP32-seeded reciprocal refined
by a Newton step and a tabulated erfc(ar)/r, an FP32 overlap screen, FP64
accumulation and FP64 Ewald: the same constructions as the oracle and
per-move |dE| within the tolerances stated in each test.  Inputs are the reference's own decks (NIST
SPC/E, TIP4P restart, charged MFI framework) and seeded synthetic boxes.
"""
import numpy as np
import pytest

from conftest import load_golden_system

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-10


def _close(a, b, rtol=FP64_RTOL, atol=1e-9):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.all(np.abs(a - b) <= atol + rtol * np.abs(b))


def _oracle(s):
    from oracle.binding import OracleSystem
    return OracleSystem(s)


# --------------------------------------------------------------------------
# totals and structure factors on the reference's decks
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["nist_box1", "nist_box4", "tip4p_restart"])
def test_totals_match_oracle_and_reference(name, engine_mod, golden_values):
    s = load_golden_system(name)
    g = golden_values[name]
    print("case", name, flush=True)
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_FP64)
    f = box.init_ewald()
    e = box.total_energy()
    o = _oracle(s)
    oe = o.total_vdw_real()
    of, skf, ska, se, intra, _ = o.ewald_total()
    assert _close(e.gg_vdw, oe.gg_vdw) and _close(e.gg_real, oe.gg_real)
    assert _close(f.gg_ewald, of.gg_ewald)
    assert abs(e.gg_vdw - g["gg_vdw"]) < 0.5e-5 and abs(e.gg_real - g["gg_real"]) < 0.5e-5
    assert abs(f.gg_ewald - g["gg_fourier"]) < 0.5e-5
    gf, ga = box.structure_factors()
    assert np.allclose(ga, ska, rtol=1e-10, atol=1e-9)
    box.close()


def test_charged_framework_structure_factor(engine_mod, golden_values):
    s = load_golden_system("co2_mfi")
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_FP64)
    f = box.init_ewald()
    assert abs(f.hh_ewald - golden_values["co2_mfi"]["hh_fourier"]) < 0.5e-5
    o = _oracle(s)
    of, skf, ska, _, _, _ = o.ewald_total()
    gf, ga = box.structure_factors()
    assert np.allclose(gf, skf, rtol=1e-10, atol=1e-9)
    box.close()


# --------------------------------------------------------------------------
# per-move primitives
# --------------------------------------------------------------------------
def _synth(n_fw=600, L=24.0, seed=4, n_mol=12):
    """Synthetic CO2 box with a few guests placed on a coarse grid."""
    from graspa_amd import synth
    s = synth.config0(seed=seed, n_fw=n_fw, L=L, cycles=10)
    c = s.components[0]
    rng = np.random.default_rng(seed)
    pos = []
    for m in range(n_mol):
        cen = rng.uniform(0, L, 3)
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        pos += [cen - 1.16 * v, cen, cen + 1.16 * v]
    c.init_pos = np.array(pos)
    c.init_molid = np.repeat(np.arange(n_mol), 3).astype(np.int64)
    c.init_charge = np.tile(c.charges, n_mol)
    return s


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_site_energies(prec, engine_mod):
    s = _synth()
    o = _oracle(s)
    P = engine_mod.GRASPA_FP64 if prec == "fp64" else engine_mod.GRASPA_MIXED
    box = engine_mod.Box(s, precision=P)
    rng = np.random.default_rng(1)
    sites = rng.uniform(0, 24.0, (40, 3))
    types = np.array([3, 4] * 20, dtype=np.int32)
    q = np.where(types == 3, -0.35, 0.70)
    for self_comp, self_molid in ((-1, -1), (1, 5)):
        got, gov = box.site_energies(sites, types, q, self_comp, self_molid)
        ref, rov = o.site_energies(sites, types, q, self_comp, self_molid)
        assert np.array_equal(gov, rov)
        keep = rov == 0
        if prec == "fp64":
            assert _close(got[keep], ref[keep])
        else:
            err = np.abs(got[keep] - ref[keep])
            # FP64 arithmetic with FP32-seeded reciprocals and the fitted erfc (~1e-13 relative)
            print("mixed site energies: max |err| %.3e" % err.max())
            assert np.all(err <= 1e-6 + 1e-9 * np.abs(ref[keep])), err.max()
    box.close()


def test_single_body_and_ewald_delta(engine_mod):
    s = _synth()
    o = _oracle(s)
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_FP64)
    box.init_ewald()
    of, skf, ska, _, _, _ = o.ewald_total()
    c = s.components[0]
    old = c.init_pos[6:9]
    new = old + np.array([0.4, -0.2, 0.3])
    e, ov = box.single_body_delta(1, 2, old, new, c.types, c.charges)
    en, ovn = o.site_energies(new, c.types, c.charges, 1, 2)
    eo, _ = o.site_energies(old, c.types, c.charges, 1, 2)
    d = en.sum(0) - eo.sum(0)
    assert ov == int(ovn.any())
    assert _close([e.hg_vdw, e.hg_real, e.gg_vdw, e.gg_real], d, atol=1e-8)
    same, cross = box.ewald_delta(old, c.charges, new, c.charges)
    osame, ocross, sk_new = o.ewald_delta(ska, skf, old, c.charges, new, c.charges)
    assert _close(same, osame, atol=1e-8) and _close(cross, ocross, atol=1e-8)
    box.ewald_accept()
    _, ga = box.structure_factors()
    assert np.allclose(ga, sk_new, rtol=1e-10, atol=1e-9)
    box.close()


def test_commit_and_total_energy_consistency(engine_mod):
    """Running sums of accepted dE equal the change of the recomputed totals."""
    s = _synth()
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_FP64)
    f0 = box.init_ewald()
    e0 = box.total_energy()
    c = s.components[0]
    old = c.init_pos[3:6]
    new = old + np.array([-0.5, 0.25, 0.1])
    de, ov = box.single_body_delta(1, 1, old, new, c.types, c.charges)
    same, cross = box.ewald_delta(old, c.charges, new, c.charges)
    box.commit_move(1, 1, new)
    box.ewald_accept()
    e1 = box.total_energy()
    f1 = box.init_ewald()
    tot = lambda e: e.hg_vdw + e.hg_real + e.gg_vdw + e.gg_real
    run = de.hg_vdw + de.hg_real + de.gg_vdw + de.gg_real
    assert abs((tot(e1) - tot(e0)) - run) < 1e-7 * max(1, abs(tot(e0)))
    fo = lambda f: f.gg_ewald + f.hg_ewald
    assert abs((fo(f1) - fo(f0)) - (same + 2 * cross)) < 1e-7 * max(1, abs(fo(f0)))
    # insertion then deletion of the same molecule returns the totals
    m = np.array([[5.0, 5.0, 4.0], [5.0, 5.0, 5.16], [5.0, 5.0, 6.32]])
    box.commit_insert(1, m, c.charges, c.types, 12)
    box.commit_delete(1, 12)
    e2 = box.total_energy()
    assert abs(tot(e2) - tot(e1)) < 1e-8 * max(1, abs(tot(e1)))
    box.close()


# --------------------------------------------------------------------------
# batched Widom
# --------------------------------------------------------------------------
def _widom_pair(n_fw, n_insert, seed, prec, engine_mod, first_index=0, stream=0, system=None):
    from graspa_amd import synth
    from oracle.binding import widom_records_to_numpy
    s = synth.config1(n_fw=n_fw) if system is None else system
    print("widom case n_fw=%d n=%d seed=%d prec=%d" % (len(s.fw_type), n_insert, seed, prec), flush=True)
    o = _oracle(s)
    _, skf, ska, _, _, _ = o.ewald_total()
    intra, selfe = o.exclusion(1)
    ref = widom_records_to_numpy(o.widom_batch(1, ska, skf, seed=seed, n_insert=n_insert, first_index=first_index,
                                               stream=stream))
    box = engine_mod.Box(s, precision=prec)
    box.init_ewald()
    res, summ = box.widom_batch(1, n_insert, seed=seed, stream=stream, first_index=first_index, excl=intra + selfe)
    got = engine_mod.widom_results_to_numpy(res)
    box.close()
    return s, ref, got, summ


def test_widom_fp64_matches_oracle(engine_mod):
    s, ref, got, summ = _widom_pair(400, 300, 21, engine_mod.GRASPA_FP64, engine_mod)
    assert ref["ok"].sum() > 250
    assert np.array_equal(got["ok"], ref["ok"])
    assert np.array_equal(got["first_bead"][ref["ok"] == 1], ref["first_bead"][ref["ok"] == 1])
    assert np.array_equal(got["orientation"][ref["ok"] == 1], ref["orientation"][ref["ok"] == 1])
    assert _close(got["energy"], ref["energy"], atol=1e-8)
    assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=0.0)
    assert abs(summ.sum_w - ref["rosen"].sum()) <= 1e-9 * abs(ref["rosen"].sum())
    assert summ.n == 300 and summ.n_ok == ref["ok"].sum()


@pytest.mark.parametrize("K", [16, 20])
def test_widom_fp64_matches_oracle_trial_counts(engine_mod, K):
    """The first-bead selection takes four insertions per wave (one per 16-lane
    row, ordered sums by a scan along the row) for K <= 16 trials and one
    insertion per wave beyond: both must make the oracle's selections with its
    weights, K = 16 filling a row and K = 20 taking the other path."""
    from graspa_amd import synth
    s = synth.config1(n_fw=400)
    s.n_trial_pos = K
    _, ref, got, summ = _widom_pair(400, 200, 31 + K, engine_mod.GRASPA_FP64, engine_mod, system=s)
    assert ref["ok"].sum() > 150
    assert np.array_equal(got["ok"], ref["ok"])
    ok = ref["ok"] == 1
    assert np.array_equal(got["first_bead"][ok], ref["first_bead"][ok])
    assert np.array_equal(got["orientation"][ok], ref["orientation"][ok])
    assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=0.0)
    assert summ.n == 200 and summ.n_ok == ref["ok"].sum()


MIXED_DE_TOL = 1e-3    # internal units (10 J/mol): 1e-5 kJ/mol per insertion


def _check_mixed(ref, got):
    """The fast form makes the same constructions as the FP64 oracle (same
    outcome, first bead and orientation for every insertion) and its energies
    and weights agree within the stated tolerances."""
    assert np.array_equal(got["ok"], ref["ok"])
    ok = ref["ok"] == 1
    assert np.array_equal(got["first_bead"][ok], ref["first_bead"][ok])
    assert np.array_equal(got["orientation"][ok], ref["orientation"][ok])
    same_sel = ok
    err = np.abs(got["energy"][same_sel] - ref["energy"][same_sel])
    print("mixed |dE| max %.3e internal (%.3e kJ/mol), median %.3e over %d" %
          (err.max(), err.max() * 0.01, np.median(err), err.size))
    assert np.all(err <= MIXED_DE_TOL)
    w_rel = np.abs(got["rosen"][same_sel] - ref["rosen"][same_sel]) / np.maximum(ref["rosen"][same_sel], 1e-300)
    assert np.max(w_rel) < 1e-4
    return err


def test_widom_mixed_within_tolerance(engine_mod):
    """Mixed form vs FP64 oracle: the same constructions and |dE| per
    insertion within 1e-5 kJ/mol."""
    s, ref, got, summ = _widom_pair(400, 300, 23, engine_mod.GRASPA_MIXED, engine_mod)
    err = _check_mixed(ref, got)
    assert np.median(err) < 1e-6


@pytest.mark.parametrize("cut_vdw,cut_coul,shifted", [(11.0, 12.0, True), (12.0, 10.0, False)])
@pytest.mark.parametrize("framework_lds", [1, 0])
def test_widom_mixed_unequal_cutoffs_and_shift(engine_mod, cut_vdw, cut_coul, shifted, framework_lds):
    """The fast pair kernel's variants beside the default one (equal cutoffs,
    no LJ shift): separate LJ and Coulomb cutoff branches, and force-field
    rows whose shift is read, with the framework in LDS and read from memory;
    the same constructions as the FP64 oracle and energies within tolerance."""
    from graspa_amd import synth
    from graspa_amd.deck import _shift_value
    from oracle.binding import widom_records_to_numpy
    s = synth.config1(n_fw=400)
    s.cut_vdw, s.cut_coul = cut_vdw, cut_coul
    nt = len(s.pseudo_names)
    # shifted LJ (Get_Shifted_Value, read_data.cpp:737-748) at the VDW cutoff, or truncated
    sh = np.zeros((nt, nt))
    if shifted:
        for i in range(nt):
            for j in range(nt):
                sh[i, j] = _shift_value(s.eps[i, j], s.sig[i, j], cut_vdw * cut_vdw)
        assert np.any(sh != 0.0)
    s.shift = sh
    print("widom unequal cutoffs vdw %.1f coul %.1f shifted %d framework_lds %d" %
          (cut_vdw, cut_coul, shifted, framework_lds), flush=True)
    o = _oracle(s)
    _, skf, ska, _, _, _ = o.ewald_total()
    intra, selfe = o.exclusion(1)
    ref = widom_records_to_numpy(o.widom_batch(1, ska, skf, seed=29, n_insert=300))
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_MIXED)
    box.init_ewald()
    box.set_eval_shape(12, 0, framework_lds)
    res, _ = box.widom_batch(1, 300, seed=29, excl=intra + selfe)
    got = engine_mod.widom_results_to_numpy(res)
    box.close()
    assert ref["ok"].sum() > 200
    _check_mixed(ref, got)


def test_widom_mixed_n2_massless_site(engine_mod):
    """N2 has a massless centre with no LJ (epsilon 0): its pairs are overlaps
    only by r^2 < 0.01, so the fast pair loop must take its general form (the
    explicit r^2 test); same constructions as the FP64 oracle."""
    from graspa_amd import synth
    s = synth.build(n_fw=400, L=24.0, species=("N2",), seed=11, T=300.0, P=1e5, probs=dict(widom=1.0))
    s.use_max_step, s.max_step_per_cycle = True, 1
    _, ref, got, _ = _widom_pair(400, 300, 37, engine_mod.GRASPA_MIXED, engine_mod, system=s)
    assert ref["ok"].sum() > 200
    _check_mixed(ref, got)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_widom_single_site_species(engine_mod, form):
    """A one-site molecule (united-atom methane, uncharged): no orientations,
    so the chain kernel's placed list goes straight to the finish kernel
    (no Fourier term); the same insertions, weights and energies as the
    FP64 oracle."""
    from graspa_amd import synth
    s = synth.build(n_fw=400, L=24.0, species=("methane",), seed=13, T=300.0, P=1e5, probs=dict(widom=1.0))
    prec = engine_mod.GRASPA_FP64 if form == "fp64" else engine_mod.GRASPA_MIXED
    _, ref, got, summ = _widom_pair(400, 300, 41, prec, engine_mod, system=s)
    assert ref["ok"].sum() > 200
    if form == "fp64":
        assert np.array_equal(got["ok"], ref["ok"])
        ok = ref["ok"] == 1
        assert np.array_equal(got["first_bead"][ok], ref["first_bead"][ok])
        assert _close(got["energy"], ref["energy"], atol=1e-8)
        assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=0.0)
    else:
        _check_mixed(ref, got)
    assert summ.n == 300 and summ.n_ok == ref["ok"].sum()


def test_widom_result_independent_of_batching(engine_mod):
    """Insertion i gives the same result alone, inside a larger batch, or
    offset by first_index: the counter-based streams make each result a pure
    function of (seed, stream, index)."""
    from graspa_amd import synth
    s = synth.config1(n_fw=400)
    box = engine_mod.Box(s, precision=engine_mod.GRASPA_MIXED)
    box.init_ewald()
    res_all, _ = box.widom_batch(1, 512, seed=5)
    a = engine_mod.widom_results_to_numpy(res_all)
    res_sub, _ = box.widom_batch(1, 37, seed=5, first_index=100)
    b = engine_mod.widom_results_to_numpy(res_sub)
    for k in ("rosen", "energy", "first_bead", "orientation", "ok"):
        assert np.array_equal(a[k][100:137], b[k])
    box.close()


def test_widom_dense_config_sample(engine_mod):
    """The stated config (1,536 atoms, 24 A) in FP64 against the oracle on a sample."""
    s, ref, got, summ = _widom_pair(1536, 128, 99, engine_mod.GRASPA_FP64, engine_mod)
    assert np.array_equal(got["ok"], ref["ok"])
    assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=1e-300)
    # the overlap screen did the rejecting: most first-bead trials never reach
    # the full pair evaluation (the reference evaluates 128 x 10 x 1536 pairs
    # for the first beads alone)
    assert summ.screened_pairs > 0 and summ.evaluated_pairs < 0.25 * 128 * 10 * 1536
    assert summ.evaluated_pairs % 1536 == 0


def test_widom_dense_config_mixed(engine_mod):
    s, ref, got, summ = _widom_pair(1536, 2000, 98, engine_mod.GRASPA_MIXED, engine_mod)
    assert np.array_equal(got["ok"] == 1, ref["ok"] == 1)
    k = ref["ok"] == 1
    assert np.array_equal(got["first_bead"][k], ref["first_bead"][k])
    if (ref["ok"] == 1).any():
        _check_mixed(ref, got)
    assert summ.n == 2000 and summ.n_first_bead >= summ.n_ok == (ref["ok"] == 1).sum()


def _triclinic_system(n_fw=400):
    from graspa_amd import synth
    s = synth.config1(n_fw=n_fw)
    s.cell = s.cell.copy()
    s.cell[3] = 3.5     # b tilted along x
    s.cell[6] = -2.0    # c tilted along x
    s.cell[7] = 1.5     # c tilted along y
    s.setup_ewald()
    assert not s.cubic
    return s


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_widom_triclinic_cell(prec, engine_mod):
    """Triclinic cell: minimum image through fractional coordinates in both
    forms; trial positions use diag(cell) as the reference does (mc_widom.h:122-213)."""
    P = engine_mod.GRASPA_FP64 if prec == "fp64" else engine_mod.GRASPA_MIXED
    s, ref, got, summ = _widom_pair(0, 300, 31, P, engine_mod, system=_triclinic_system())
    assert (ref["ok"] == 1).sum() > 100
    if prec == "fp64":
        assert np.array_equal(got["ok"], ref["ok"])
        k = ref["ok"] == 1
        assert np.array_equal(got["first_bead"][k], ref["first_bead"][k])
        assert _close(got["energy"], ref["energy"], atol=1e-8)
        assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=0.0)
    else:
        _check_mixed(ref, got)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_widom_large_framework_streams_from_cache(prec, engine_mod):
    """12,000 framework atoms (192 KB of coordinates) do not fit in LDS: the
    kernel reads them through the caches instead; same results."""
    from graspa_amd import synth
    s = synth.config1(n_fw=12000, L=48.0)
    P = engine_mod.GRASPA_FP64 if prec == "fp64" else engine_mod.GRASPA_MIXED
    s, ref, got, summ = _widom_pair(0, 64, 41, P, engine_mod, system=s)
    if prec == "fp64":
        assert np.array_equal(got["ok"], ref["ok"])
        assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=1e-300)
        assert _close(got["energy"], ref["energy"], atol=1e-7)
    else:
        k = ref["ok"] == 1
        assert np.array_equal(got["first_bead"][k], ref["first_bead"][k])
        if k.any():
            _check_mixed(ref, got)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_widom_with_guests(prec, engine_mod):
    """Widom insertions into a loaded box: guest-guest pairs, the guests in the
    overlap screen and in the energies."""
    s = _synth(n_fw=400, n_mol=20)
    s.n_trial_pos, s.n_trial_orient = 10, 10
    P = engine_mod.GRASPA_FP64 if prec == "fp64" else engine_mod.GRASPA_MIXED
    s, ref, got, summ = _widom_pair(0, 300, 51, P, engine_mod, system=s)
    assert (ref["ok"] == 1).sum() > 50
    if prec == "fp64":
        assert np.array_equal(got["ok"], ref["ok"])
        assert _close(got["energy"], ref["energy"], atol=1e-8)
        assert _close(got["rosen"], ref["rosen"], rtol=1e-9, atol=0.0)
    else:
        _check_mixed(ref, got)
    assert summ.evaluated_pairs % (400 + 60) == 0
