"""Energy drift check and output.txt-style statistics of a GCMC run.

The drift check recomputes the total energy of a box's final configuration
(pair terms by the total-energy kernel, Fourier terms from structure factors
rebuilt from the positions, exclusion energies of the rigid molecules and the
tail correction) and compares it with the initial totals plus the running sum
of accepted energy changes, term by term, as the reference's final check does
(fxn_main.h:454-542: CHECK DELTA_E, ENERGY DRIFT, GPU DRIFT).  A box in the
fast form is also recomputed in the plain FP64 form, whose difference is
reported as the GPU drift.

The writer prints the sections of the reference's output.txt that describe a
run: move statistics (print_statistics.cuh:39-214), the energy summary
(fxn_main.h:454-542), production averages of the energy, the Rosenbluth
summary of Widom moves (print_statistics.cuh:100-200), heat of adsorption and
loading block averages (print_statistics.cuh:456-645), with the reference's
line formats, so that parsers of those files (the reference's
Examples/test_examples.py drift check among them) read them unchanged.
Block statistics come from the device-resident driver
(graspa_batch_block_stats).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import engine
from .deck import AVOGADRO, BOLTZMANN, ENERGY_TO_KELVIN, GAS_CONSTANT

TERMS = ("hh_vdw", "hg_vdw", "gg_vdw", "hh_real", "hg_real", "gg_real", "hh_ewald", "hg_ewald", "gg_ewald", "tail")
ENERGY_TO_KJ = 0.01          # internal energy unit (10 J/mol) in kJ/mol


def _zero():
    return {k: 0.0 for k in TERMS}


def _total(e):
    return sum(e[k] for k in TERMS)


def _sub(a, b):
    return {k: a[k] - b[k] for k in TERMS}


def _add(a, b):
    return {k: a[k] + b[k] for k in TERMS}


def tail_total(system, n_mol) -> float:
    """TotalTailCorrection (TailCorrection_Energy_Functions.h:3-21):
    sum over pseudo-atom type pairs of tail_ij N_i N_j / V, framework atoms
    and adsorbate molecules (n_mol per species) counted by type."""
    if not system.has_tail:
        return 0.0
    n = system.pseudo_counts(list(n_mol)).astype(np.float64)
    t = np.where(np.asarray(system.use_tail) != 0, np.asarray(system.tail, dtype=np.float64), 0.0)
    return float(n @ t @ n) / system.volume


def exclusion_total(system, n_mol) -> float:
    """Self plus intra-molecular exclusion energy of every rigid molecule
    present (ewald_preparation.h:351-366, one term per molecule)."""
    ex = 0.0
    for ci, nm in enumerate(n_mol, start=1):
        intra, selfe = engine.exclusion_energy(system, ci)
        ex += nm * (intra + selfe)
    return ex


def recompute_totals(box, system, n_mol) -> Dict[str, float]:
    """Total energy of the box's current configuration by term (internal
    units): VDW and real-space Coulomb from graspa_total_energy, Fourier
    terms from structure factors rebuilt from the positions (this resets the
    box's running S(k) to the rebuilt one), minus the exclusion energies of
    the molecules, and the tail correction."""
    pair = box.total_energy().as_dict()
    fourier = box.init_ewald().as_dict()
    e = _zero()
    for k in ("hh_vdw", "hg_vdw", "gg_vdw", "hh_real", "hg_real", "gg_real"):
        e[k] = pair[k]
    if not system.no_charges:
        e["hg_ewald"] = fourier["hg_ewald"]
        e["gg_ewald"] = fourier["gg_ewald"] - exclusion_total(system, n_mol)
    e["tail"] = tail_total(system, n_mol)
    return e


def recompute_totals_fp64(box, system, n_mol, capacity=None) -> Dict[str, float]:
    """The same totals in the plain FP64 form: a temporary FP64 box holding
    the molecules of `box`."""
    ref = engine.Box(system, precision=engine.GRASPA_FP64, capacity=capacity)
    try:
        for ci in range(1, len(system.components) + 1):
            pos, q, t, m = box.get_molecules(ci)
            ref.set_molecules(ci, pos, q, t, m)
        return recompute_totals(ref, system, n_mol)
    finally:
        ref.close()


@dataclass
class DriftReport:
    initial: Dict[str, float]
    running: Dict[str, float]
    final: Dict[str, float]
    final_fp64: Dict[str, float]

    @property
    def check(self):
        """CHECK DELTA_E: recomputed final minus initial."""
        return _sub(self.final, self.initial)

    @property
    def drift(self):
        """ENERGY DRIFT: initial + running minus the recomputed final."""
        return _sub(_add(self.initial, self.running), self.final)

    @property
    def gpu_drift(self):
        """GPU DRIFT: run-form recompute minus the FP64 recompute."""
        return _sub(self.final, self.final_fp64)

    def max_abs_drift(self):
        return max(abs(v) for v in self.drift.values())


def drift_check(box, system, initial, running, n_mol, capacity=None) -> DriftReport:
    """Drift of a finished run: `initial` = recompute_totals at the start,
    `running` = the running energy change (graspa_energy or dict)."""
    run = running.as_dict() if hasattr(running, "as_dict") else dict(running)
    final = recompute_totals(box, system, n_mol)
    if box.precision == engine.GRASPA_FP64:
        f64 = dict(final)
    else:
        f64 = recompute_totals_fp64(box, system, n_mol, capacity=capacity)
    return DriftReport(dict(initial), run, final, f64)


# ---------------------------------------------------------------- writer --

@dataclass
class RunRecord:
    """What the writer needs of one box's run."""
    system: object
    names: List[str]                        # species names (adsorbates)
    framework_name: str
    precision: int
    counters: np.ndarray                    # (n_species, GRASPA_N_COUNTERS): graspa_mc_stats layout
    n_mol: List[int]
    drift: DriftReport
    blocks: Optional[dict] = None           # graspa_batch_block_stats arrays
    seconds: float = 0.0
    cycles: int = 0
    moves: int = 0
    widom_w: Optional[np.ndarray] = None
    widom_n: Optional[np.ndarray] = None
    notes: List[str] = field(default_factory=list)


def _energy_block(out, title, e):
    out.append(" *** %s *** " % title)
    out.append("=" * 72)
    rows = [("VDW [Host-Host]:", e["hh_vdw"]), ("VDW [Host-Guest]:", e["hg_vdw"]), ("VDW [Guest-Guest]:", e["gg_vdw"]),
            ("Real Coulomb [Host-Host]:", e["hh_real"]), ("Real Coulomb [Host-Guest]:", e["hg_real"]),
            ("Real Coulomb [Guest-Guest]:", e["gg_real"]), ("Ewald [Host-Host]:", e["hh_ewald"]),
            ("Ewald [Host-Guest]:", e["hg_ewald"]), ("Ewald [Guest-Guest]:", e["gg_ewald"]), ("DNN Energy:", 0.0),
            ("Tail Correction Energy:", e["tail"]), ("Total Energy:", _total(e))]
    for label, v in rows:
        out.append("%-28s%.5f (%.5f [K], %.5f [kJ/mol])" % (label, v, v * ENERGY_TO_KELVIN, v * ENERGY_TO_KJ))
    out.append("=" * 72)


def _block_lengths(blocks):
    return np.maximum(np.asarray(blocks["cycles"], dtype=np.float64), 1.0)


def _print_values(out, sums, sq, nb):
    """Print_Values (print_statistics.cuh:456-476): per-block average and
    2-sigma error bar within the block, then the mean and 2-sigma spread of
    the block averages."""
    avgs = []
    for b in range(len(nb)):
        a = sums[b] / nb[b]
        s2 = sq[b] / nb[b]
        err = 2.0 * math.sqrt(s2 - a * a) if s2 - a * a > 0 else 0.0
        out.append("BLOCK [%d], Blocksize: %d, Average: %.5f, ErrorBar: %.5f" % (b, int(nb[b]), a, err))
        avgs.append(a)
    avgs = np.array(avgs)
    v = float(np.mean(avgs * avgs) - np.mean(avgs) ** 2)
    out.append("Overall: Average: %.5f, ErrorBar: %.5f" % (float(np.mean(avgs)), 2.0 * math.sqrt(v) if v > 0 else 0.0))


def heat_of_adsorption(rec: RunRecord):
    """Print_HeatOfAdsorption (print_statistics.cuh:483-560): per block and
    adsorbate i, Q_i = sum_j kB-scaled (<E N_j> - <E><N_j>) [cov(N)^-1]_ji - T,
    in kJ/mol; E = running energy change without host-host terms, E N_j =
    createmol + E N_j per cycle as the reference gathers it (axpy.cu:433-437)."""
    s = rec.system
    blk = rec.blocks
    ns = len(rec.names)
    nb = _block_lengths(blk)
    createmol = _total(rec.drift.initial) - rec.drift.initial["hh_vdw"] - rec.drift.initial["hh_real"] - \
        rec.drift.initial["hh_ewald"]
    nbk = len(nb)
    q = np.zeros((ns, nbk))
    for b in range(nbk):
        e = np.asarray(blk["e"][b])
        avg_e = (e.sum() - e[0] - e[3] - e[6]) / nb[b]
        n = np.array([blk["n"][b][c + 1] for c in range(ns)]) / nb[b]
        nn = np.array([[blk["nn"][b][c + 1][d + 1] for d in range(ns)] for c in range(ns)]) / nb[b]
        en = np.array([createmol * nb[b] + blk["en"][b][c + 1] for c in range(ns)]) / nb[b]
        cov = nn - np.outer(n, n)
        if not np.all(np.isfinite(cov)) or abs(np.linalg.det(cov)) <= 1e-300:
            q[:, b] = np.nan             # loadings that do not fluctuate: no fluctuation estimate
            continue
        inv = np.linalg.inv(cov)
        for i in range(ns):
            v = 0.0
            for j in range(ns):
                v += ENERGY_TO_KELVIN * (en[j] - avg_e * n[j]) * inv[j][i]
            q[i, b] = (v - s.temperature) * (0.01 / ENERGY_TO_KELVIN)
    return q


def widom_block(rec, w, n, rho):
    """ComputeFugacity (print_statistics.cuh:100-110): mean Rosenbluth
    weight, excess chemical potential [K], Henry coefficient [mol/kg/Pa] and
    the fugacity it implies, from the sum of W and the count of a block."""
    s = rec.system
    wr = w / n
    mu = ENERGY_TO_KELVIN * -(1.0 / s.beta) * math.log(wr) if wr > 0 else float("inf")
    henry = wr / (GAS_CONSTANT * s.temperature * rho)
    n_ads = rec.n_mol[0] if rec.n_mol else 0        # the reference uses component 1's molecule count
    fug = BOLTZMANN * s.temperature * n_ads / (s.volume * 1.0e-30 * wr) if wr > 0 else float("inf")
    return wr, mu, henry, fug


def write_output(rec: RunRecord, path_or_file):
    """Writes the run's output.txt sections; returns the text."""
    s = rec.system
    out = []
    out.append("=" * 42)
    out.append("====== graspa_amd GCMC run, box 0 ======")
    out.append("=" * 42)
    out.append("Precision form: %s" % ("plain FP64" if rec.precision == engine.GRASPA_FP64 else "fast (mixed)"))
    out.append("Framework: %s, %d atoms, volume %.5f A^3" % (rec.framework_name, len(s.fw_type), s.volume))
    out.append("Temperature: %.5f [K], Pressure: %.5f [Pa]" % (s.temperature, s.pressure_pa))
    out.extend(rec.notes)
    # move statistics (print_statistics.cuh:39-58, 202-213)
    for ci, name in enumerate(rec.names):
        c = rec.counters[ci]
        out.append("======================== MOVE STATISTICS FOR COMPONENT [%d] (%s) ========================" %
                   (ci + 1, name))
        out.append("=====================TRANSLATION MOVES=====================")
        out.append("Translation Performed: %d" % c[0])
        out.append("Translation Accepted: %d" % c[1])
        out.append("===========================================================")
        out.append("=====================ROTATION MOVES========================")
        out.append("Rotation Performed: %d" % c[2])
        out.append("Rotation Accepted: %d" % c[3])
        out.append("===========================================================")
        out.append("=====================SWAP MOVES=====================")
        out.append("Insertion Performed:   %d" % c[4])
        out.append("Insertion Accepted:    %d" % c[5])
        out.append("Deletion Performed:    %d" % c[6])
        out.append("Deletion Accepted:     %d" % c[7])
        if len(c) > 10:
            out.append("Reinsertion Performed: %d" % c[9])
            out.append("Reinsertion Accepted:  %d" % c[10])
        out.append("====================================================")
        if len(c) > engine.CNT_IDSW_ACC and c[engine.CNT_IDSW_TOT:engine.CNT_IDSW_TOT + len(rec.names)].any():
            # Print_IdentitySwap_Statistics (print_statistics.cuh:214-224)
            out.append("=====================IDENTITY SWAP MOVES=====================")
            for cj, other in enumerate(rec.names):
                if cj != ci:
                    out.append("Identity Swap Performed, FROM [%s (%d)] TO [%s (%d)]: %d (%d Accepted)" %
                               (name, ci + 1, other, cj + 1, c[engine.CNT_IDSW_TOT + cj], c[engine.CNT_IDSW_ACC + cj]))
            out.append("=============================================================")
        if c[8]:
            out.append("Widom Performed: %d" % c[8])
    out.append("=" * 96)
    out.append("Work took %.6f seconds" % rec.seconds)
    out.append("Cycles: %d, Moves: %d" % (rec.cycles, rec.moves))
    # energy summary (fxn_main.h:454-542)
    d = rec.drift
    out.append("======================== ENERGY SUMMARY (Simulation 0) =========================")
    _energy_block(out, "INITIAL STAGE", d.initial)
    _energy_block(out, "FINAL STAGE", d.final)
    _energy_block(out, "RUNNING DELTA_E (FINAL - INITIAL)", d.running)
    _energy_block(out, "CHECK DELTA_E (RECOMPUTED FINAL - INITIAL)", d.check)
    _energy_block(out, "ENERGY DRIFT (RUNNING FINAL - RECOMPUTED FINAL)", d.drift)
    _energy_block(out, "GPU DRIFT (RUN FORM FINAL - FP64 FINAL)", d.gpu_drift)
    out.append("=" * 80)
    blk = rec.blocks
    if blk is not None and int(np.sum(blk["cycles"])) > 0:
        nb = _block_lengths(blk)
        nblk = len(nb)
        # production averages of the energy (print_statistics.cuh:439-454)
        avg = _zero()
        sq = _zero()
        for b in range(nblk):
            for k, key in enumerate(TERMS):
                a = blk["e"][b][k] / nb[b]
                avg[key] += a / nblk
                sq[key] += a * a / nblk
        err = {k: 2.0 * math.sqrt(max(sq[k] - avg[k] * avg[k], 0.0)) for k in TERMS}
        out.append("======================== PRODUCTION PHASE AVERAGE ENERGIES (Simulation 0) =========================")
        _energy_block(out, "PRODUCTION PHASE AVERAGE ENERGY", _add(avg, d.initial))
        _energy_block(out, "PRODUCTION PHASE AVERAGE ENERGY ERRORBAR", err)
        out.append("=" * 80)
        rho = s.framework_density()
        # Rosenbluth summary (print_statistics.cuh:112-200)
        for ci, name in enumerate(rec.names):
            c = ci + 1
            wn = np.array([blk["widom_n"][b][c] for b in range(nblk)])
            if wn.sum() <= 0:
                continue
            out.append("=====================Rosenbluth Summary For Component [%d] (%s)=====================" % (c, name))
            out.append("There are %d blocks" % nblk)
            tot = {k: [0.0, 0.0] for k in ("wr", "mu", "henry", "fug")}
            for b in range(nblk):
                out.append("=====BLOCK %d=====" % b)
                out.append("Widom Performed: %.1f" % wn[b])
                if wn[b] > 0:
                    out.append("Framework Density: %.5f [kg/m^3]" % rho)
                    wr, mu, henry, fug = widom_block(rec, blk["widom_w"][b][c], wn[b], rho)
                    out.append("(Total) Averaged Rosenbluth Weight: %.10f" % wr)
                    out.append("(Total) Averaged Excess Mu: %.10f" % mu)
                    out.append("(Total) Averaged Henry Coefficient: %.10f" % henry)
                    out.append("(Total) Converted to Fugacity: %.10f [Pascal], Temp: %.5f [K]" % (fug, s.temperature))
                    out.append("Framework Density: %.5f [kg/m^3]" % rho)
                    out.append("(Insertion) Averaged Rosenbluth Weight: %.10f" % wr)
                    out.append("(Insertion) Averaged Excess Mu: %.10f" % mu)
                    out.append("(Insertion) Converted to Fugacity: %.10f [Pascal], Temp: %.5f [K]" % (fug, s.temperature))
                    for k, v in (("wr", wr), ("mu", mu), ("henry", henry), ("fug", fug)):
                        tot[k][0] += v
                        tot[k][1] += v * v
            def avg_err(k):
                a = tot[k][0] / nblk
                v = tot[k][1] / nblk - a * a
                return a, 2.0 * math.sqrt(v) if v > 0 else 0.0
            out.append("=========================AVERAGE========================")
            out.append("Averaged Rosenbluth Weight: %.10f +/- %.10f" % avg_err("wr"))
            out.append("Averaged Excess Chemical Potential: %.10f +/- %.10f" % avg_err("mu"))
            out.append("Averaged Henry Coefficient [mol/kg/Pa]: %.10g +/- %.10g" % avg_err("henry"))
            out.append("Averaged Fugacity: %.10f +/- %.10f" % avg_err("fug"))
        # heat of adsorption (print_statistics.cuh:483-560)
        out.append("============= BLOCK AVERAGES (HEAT OF ADSORPTION: kJ/mol) =========")
        q = heat_of_adsorption(rec)
        for ci, name in enumerate(rec.names):
            out.append("COMPONENT [%d] (%s)" % (ci + 1, name))
            for b in range(nblk):
                out.append("BLOCK [%d], Blocksize: %d, Average: %.5f" % (b, int(nb[b]), q[ci, b]))
            with np.errstate(invalid="ignore"):
                m = float(np.mean(q[ci]))
                v = float(np.mean(q[ci] ** 2) - m * m)
            out.append("Overall: Average: %.5f, ErrorBar: %.5f" % (m, 2.0 * math.sqrt(v) if v > 0 else 0.0))
            out.append("-----------------------------")
        out.append("=" * 62)
        # loadings (print_statistics.cuh:562-645)
        ncell = int(np.prod(s.unit_cells))
        cell_mass = s.fw_mass
        mw = [float(np.sum(c.masses)) for c in s.components]

        def comp_sums(c):
            return ([blk["n"][b][c] for b in range(nblk)], [blk["n2"][b][c] for b in range(nblk)])

        one = ([nb[b] for b in range(nblk)], [nb[b] for b in range(nblk)])   # framework: N = 1
        out.append("=================== BLOCK AVERAGES (LOADING: # MOLECULES)=============")
        out.append("COMPONENT [0] (%s)" % rec.framework_name)
        _print_values(out, one[0], one[1], nb)
        out.append("----------------------------------------------------------")
        for ci, name in enumerate(rec.names):
            out.append("COMPONENT [%d] (%s)" % (ci + 1, name))
            sm, sq2 = comp_sums(ci + 1)
            _print_values(out, sm, sq2, nb)
            out.append("----------------------------------------------------------")
        out.append("=" * 70)
        for title, scale_fn in (("mg/g", lambda w: 1000.0 * w / cell_mass), ("mol/kg", lambda w: 1000.0 / cell_mass)):
            out.append("=====================BLOCK AVERAGES (LOADING: %s)=============" % title)
            for ci, name in enumerate(rec.names):
                out.append("COMPONENT [%d] (%s)" % (ci + 1, name))
                out.append("Framework component 0, molar mass: %.5f" % (cell_mass / max(ncell, 1)))
                out.append("Framework total mass: %.5f" % cell_mass)
                f = scale_fn(mw[ci]) if cell_mass > 0 else 0.0
                sm, sq2 = comp_sums(ci + 1)
                _print_values(out, [x * f for x in sm], [x * f * f for x in sq2], nb)
                out.append("----------------------------------------------------------")
            out.append("=" * 62)
        out.append("=====================BLOCK AVERAGES (LOADING: g/L)=============")
        for ci, name in enumerate(rec.names):
            out.append("COMPONENT [%d] (%s)" % (ci + 1, name))
            out.append("Total Unit Cells %d " % ncell)
            gpl = mw[ci] * 10000.0 / 6.0221408 / s.volume
            sm, sq2 = comp_sums(ci + 1)
            _print_values(out, [x * gpl for x in sm], [x * gpl * gpl for x in sq2], nb)
            out.append("----------------------------------------------------------")
        out.append("=" * 62)
        out.append("=====================BLOCK AVERAGES (VOLUME Å^3)================")
        _print_values(out, [s.volume * nb[b] for b in range(nblk)], [s.volume ** 2 * nb[b] for b in range(nblk)], nb)
        out.append("----------------------------------------------------------")
        out.append("=" * 62)
    text = "\n".join(out) + "\n"
    if hasattr(path_or_file, "write"):
        path_or_file.write(text)
    else:
        with open(path_or_file, "w") as f:
            f.write(text)
    return text


def record_from_batch(batch, slot, box, system, initial, seconds=0.0, framework_name="framework", capacity=None):
    """RunRecord of slot `slot` of a graspa_batch after its production run:
    synchronises the batch, recomputes the final totals and reads the block
    statistics.  `initial` = recompute_totals of the box before the run."""
    batch.sync()
    st = batch.stats(slot)
    blocks = batch.block_stats(slot)
    n_mol = [int(x) for x in st.n_mol]
    drift = drift_check(box, system, initial, st.running, n_mol, capacity=capacity)
    return RunRecord(system=system, names=[c.name for c in system.components], framework_name=framework_name,
                     precision=box.precision, counters=np.asarray(st.counters), n_mol=n_mol, drift=drift,
                     blocks=blocks, seconds=seconds, cycles=int(np.sum(blocks["cycles"])), moves=int(st.moves),
                     widom_w=st.widom_w, widom_n=st.widom_n)
