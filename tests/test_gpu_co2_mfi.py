"""The reference's CO2-in-MFI example, run by the device-resident driver and
compared with the statistics the reference printed for it.

The deck (Examples/CO2-MFI: MFI 2x2x2 with CIF charges, rigid 3-site CO2,
298 K, 1e4 Pa, VDW cutoff 12.8 A, Coulomb 12 A, Ewald precision 1e-6,
translation : rotation : reinsertion : swap 1:1:1:1, 5,000 initialization
cycles) is the committed fixture tests/golden/co2_mfi.npz.  The reference ran
it once (Examples/CO2-MFI/output.txt:171-188) and printed its acceptance
counts, its final maximum steps and (through the adsorbate self energy,
output.txt:232-233, 343615.73682 / 13215.98988 per molecule) its final
loading of 26 molecules.

Each form runs 16 independent seeds for the deck's 5,000 cycles.  Per
quantity, the spread of the 16 runs estimates the run-to-run standard
deviation s of one run, and the mean must lie within 4 s sqrt(1 + 1/16) of the
reference's single run.  This pins the parts of the acceptance rules that
parity with the oracle cannot (the oracle is a restatement): GetPrefactor's
units (mc_utilities.h:315-351), the insertion / deletion Pacc
(move_struct.h:34-59), the translation / rotation Metropolis test
(mc_single_particle.h:236-263) and the reinsertion weight ratio
(move_struct.h:170-407).

The maximum translation is deterministic here: the acceptance stays far
below 0.5, so every update (every 500 cycles once a translation was
attempted, mc_utilities.h:612-638) multiplies it by 0.95, and the reference's
printed values are 0.1 x cell x 0.95^9 to all ten printed digits.
"""
import numpy as np
import pytest

from conftest import load_golden_system

pytestmark = pytest.mark.gpu

CYCLES = 5000
N_SEEDS = 16
# Examples/CO2-MFI/output.txt:171-188 (accepted, performed) and the final loading
REFERENCE = {
    "translation": (819, 24719),
    "rotation": (12037, 25130),
    "insertion": (4218, 13735),
    "deletion": (4192, 13766),
    "reinsertion": (4565, 27518),
}
REFERENCE_FINAL_MOLECULES = 26
REFERENCE_MAX_TRANSLATION = (2.5237707363, 2.5082666008, 1.6869255701)
REFERENCE_MAX_ROTATION = 0.4455488660
# counter indices (graspa_engine.h): (accepted, performed)
COUNTERS = {"translation": (1, 0), "rotation": (3, 2), "insertion": (5, 4), "deletion": (7, 6),
            "reinsertion": (10, 9)}


def _run(engine_mod, precision):
    s = load_golden_system("co2_mfi")
    assert s.n_init == CYCLES and s.n_prod == 0 and len(s.components) == 1
    boxes = []
    for _ in range(N_SEEDS):
        b = engine_mod.Box(s, precision=precision, capacity=256)
        b.init_ewald()
        boxes.append(b)
    print("co2-mfi: %d boxes x %d cycles, precision %d" % (N_SEEDS, CYCLES, precision), flush=True)
    bt = engine_mod.Batch(boxes, [101 + i for i in range(N_SEEDS)], list(range(N_SEEDS)))
    bt.run(CYCLES, production=False)
    out = {"cell": np.array([s.cell[0], s.cell[4], s.cell[8]])}
    for k in COUNTERS:
        out[k] = []
    out["n_final"], out["max_trans"], out["max_rot"] = [], [], []
    for i in range(N_SEEDS):
        st = bt.stats(i)
        assert st.cycles == CYCLES
        c = st.counters[0]
        for k, (a, t) in COUNTERS.items():
            out[k].append((int(c[a]), int(c[t])))
        out["n_final"].append(int(st.n_mol[0]))
        mt, mr = bt.max_steps(i)
        out["max_trans"].append(mt[0].copy())
        out["max_rot"].append(mr[0].copy())
    bt.close()
    for b in boxes:
        b.close()
    return out


@pytest.fixture(scope="module", params=["fp64", "mixed"])
def co2_mfi_runs(request, engine_mod):
    prec = engine_mod.GRASPA_FP64 if request.param == "fp64" else engine_mod.GRASPA_MIXED
    return _run(engine_mod, prec)


def _within(samples, ref, what):
    x = np.asarray(samples, dtype=np.float64)
    m, s = x.mean(), x.std(ddof=1)
    bound = 4.0 * s * np.sqrt(1.0 + 1.0 / len(x))
    print("%s: reference %.4f, runs %.4f +- %.4f (bound %.4f)" % (what, ref, m, s, bound), flush=True)
    assert abs(m - ref) <= bound, (what, ref, m, s)


def test_acceptance_ratios_match_reference_run(co2_mfi_runs):
    for k, (ra, rt) in REFERENCE.items():
        ratios = [a / t for a, t in co2_mfi_runs[k]]
        assert all(t > 1000 for _, t in co2_mfi_runs[k]), k
        _within(ratios, ra / rt, k + " acceptance")


def test_final_loading_matches_reference_run(co2_mfi_runs):
    _within(co2_mfi_runs["n_final"], REFERENCE_FINAL_MOLECULES, "final molecules")


def test_attempt_mix_matches_reference_run(co2_mfi_runs):
    # the four move kinds are drawn 1:1:1:1 and swaps split evenly between
    # insertion and deletion; the reference's counts (translations up to the
    # last update only, output.txt:43) are one sample of the same mix
    tot = {k: np.array([t for _, t in co2_mfi_runs[k]], dtype=np.float64) for k in COUNTERS}
    frac_ins = tot["insertion"] / (tot["insertion"] + tot["deletion"])
    _within(frac_ins, REFERENCE["insertion"][1] / (REFERENCE["insertion"][1] + REFERENCE["deletion"][1]),
            "insertion share of swaps")
    moves = sum(tot.values())
    ref_moves = sum(t for _, t in REFERENCE.values())
    # every run does 5,000 cycles of max(20, N + 1) steps; the reference's total
    # lacks the translations and rotations of its last 499 cycles
    assert np.all(moves > 0.9 * ref_moves) and np.all(moves < 1.3 * ref_moves), (moves, ref_moves)


def test_max_steps_follow_reference_schedule(co2_mfi_runs):
    cell = co2_mfi_runs["cell"]
    # the reference's printed values are 0.1 cell 0.95^9 to every printed digit
    assert np.allclose(np.array(REFERENCE_MAX_TRANSLATION), 0.1 * cell * 0.95 ** 9, rtol=0, atol=6e-11)
    theta0 = 30.0 / (180.0 / 3.1415)
    assert abs(REFERENCE_MAX_ROTATION - theta0 * 1.05 ** 3 * 0.95 ** 6) < 6e-11
    for mt in co2_mfi_runs["max_trans"]:
        # nine updates, or ten when a translation was attempted in cycle 0
        ok = [np.allclose(mt, 0.1 * cell * 0.95 ** k, rtol=1e-12, atol=0) for k in (9, 10)]
        assert any(ok), mt
    for mr in co2_mfi_runs["max_rot"]:
        assert mr[0] == mr[1] == mr[2]
        ok = [abs(mr[0] - theta0 * 1.05 ** a * 0.95 ** (u - a)) < 1e-12 * theta0 for u in (9, 10) for a in range(u + 1)]
        assert any(ok), mr
