"""Writes the golden fixtures for the oracle from gRASPA's own example decks.

Run in the build container, where the reference tree is mounted read-only at
/root/reference (it is not present where the tests run).  Each deck is read with
graspa_amd.deck and stored as a compact .npz (inputs); the energies gRASPA
printed for the same deck are stored in golden_values.json (expected outputs).

    python tests/golden/make_fixtures.py                 # decks and values
    python tests/golden/make_fixtures.py --values-only   # golden_values.json only
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from graspa_amd import deck  # noqa: E402

REF = "/root/reference/Examples"
NUM = r"(-?[0-9]+\.[0-9]+)"


def grab(text, pattern, n=1):
    m = re.findall(pattern, text)
    if len(m) < n:
        raise RuntimeError("pattern not found: " + pattern)
    return m[n - 1]


def bae_mixture():
    """Widom / Henry statistics of the Bae-Mixture deck (CO2 / methane in
    MATVAF, 3x3x2 cells, 296 K): the printed Rosenbluth weights, excess
    chemical potentials, Henry coefficients and fugacities of each block, the
    printed framework density and box volume, and the framework mass per
    unit cell summed from the CIF labels and pseudo_atoms.def masses."""
    d = os.path.join(REF, "Bae-Mixture")
    out = open(os.path.join(d, "output.txt")).read()
    names, mass, charge = deck.read_pseudo_atoms(os.path.join(d, "pseudo_atoms.def"))
    cell, _, _, _, fw_mass = deck.read_cif(os.path.join(d, "MATVAF_clean-Bae-PartialCharge-fixed-Original-N2-Charge.cif"),
                                           names, charge, mass, True, (1, 1, 1))
    blocks = re.findall(r"\(Total\) Averaged Rosenbluth Weight: " + NUM + r"\n\(Total\) Averaged Excess Mu: " + NUM
                        + r"\n\(Total\) Averaged Henry Coefficient: " + NUM
                        + r"\n\(Total\) Converted to Fugacity: " + NUM, out)
    return {
        "source": "Examples/Bae-Mixture/output.txt (Rosenbluth Summary), simulation.input, MATVAF CIF, pseudo_atoms.def",
        "temperature": 296.0, "unit_cells": [3, 3, 2],
        "cell_parameters": [12.9938, 13.0908, 22.3180, 90.0, 90.0, 102.5820],
        "framework_mass_per_cell": fw_mass,
        "framework_density": float(grab(out, r"Framework Density: " + NUM)),
        "volume": float(grab(out, r"Box Volume: +" + NUM)),
        "blocks": [[float(v) for v in b] for b in blocks]}


def co2_mfi_moves():
    """Move statistics of the CO2-MFI deck over its 5,000 cycles from an empty
    box (298 K, 1e4 Pa, CBMC 10 x 10) and the molecules left at the end (the
    final CO2 self energy over one molecule's)."""
    d = os.path.join(REF, "CO2-MFI")
    out = open(os.path.join(d, "output.txt")).read()
    g = {}
    for key, pat in (("translation", "Translation"), ("rotation", "Rotation"), ("insertion", "Insertion"),
                     ("deletion", "Deletion"), ("reinsertion", "Reinsertion")):
        g[key] = [int(grab(out, pat + r" Performed: +([0-9]+)")), int(grab(out, pat + r" Accepted: +([0-9]+)"))]
    self_final = float(re.findall(r"Component: 1, SelfAtomE: " + NUM, out)[-1])
    self_one = float(grab(out, r"DEBUG: comp: 1, IntraE: " + NUM + ", SelfE: " + NUM)[1])
    g["final_molecules"] = int(round(self_final / self_one))
    g["cycles"] = 5000
    g["seconds"] = float(grab(out, r"Work took " + NUM + " seconds"))
    g["source"] = "Examples/CO2-MFI/output.txt: move statistics, final self energy, Work took"
    return g


def values_only():
    """Adds the statistics entries to golden_values.json without rewriting the
    deck fixtures."""
    path = os.path.join(HERE, "golden_values.json")
    golden = json.load(open(path))
    golden["bae_mixture"] = bae_mixture()
    golden["co2_mfi_moves"] = co2_mfi_moves()
    with open(path, "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)


def main():
    golden = {}
    golden["bae_mixture"] = bae_mixture()
    golden["co2_mfi_moves"] = co2_mfi_moves()
    for b in (1, 2, 3, 4):
        d = os.path.join(REF, "Reference_NIST_SPCE", "Box-%d" % b)
        s = deck.read_deck(d)
        deck.save_system(os.path.join(HERE, "nist_box%d.npz" % b), s)
        out = open(os.path.join(d, "output.txt")).read()
        vdw, real = grab(out, r"Guest-Guest VDW: " + NUM + "; Real: " + NUM)
        fourier = grab(out, r"CPU Guest-Guest Fourier: " + NUM)
        selfe = grab(out, r"Component: 1, SelfAtomE: " + NUM)
        intra = grab(out, r"Component: 1, Intra-Molecular ExclusionE: " + NUM)
        t_intra, t_self = grab(out, r"DEBUG: comp: 1, IntraE: " + NUM + ", SelfE: " + NUM)
        tail = grab(out, r"TailE: " + NUM, 2)
        ggew = grab(out, r"GGEwaldE: " + NUM, 2)
        tail_pair = grab(out, r"TypeI: 0, TypeJ: 0, FF.size: 2, Energy: " + NUM)
        golden["nist_box%d" % b] = {
            "source": "Examples/Reference_NIST_SPCE/Box-%d/output.txt" % b,
            "gg_vdw": float(vdw), "gg_real": float(real), "gg_fourier": float(fourier),
            "self_guest": float(selfe), "intra_exclusion_guest": float(intra),
            "template_intra": float(t_intra), "template_self": float(t_self),
            "tail_total": float(tail), "gg_ewald": float(ggew), "tail_pair_00": float(tail_pair)}
    # TIP4P restart deck (RASPA-2 restart, RASPA-mode Ewald)
    d = os.path.join(REF, "Restart-Examples")
    s = deck.read_deck(d)
    deck.save_system(os.path.join(HERE, "tip4p_restart.npz"), s)
    out = open(os.path.join(d, "raspa2-long-output.txt")).read()
    vdw, real = grab(out, r"Guest-Guest VDW: " + NUM + "; Real: " + NUM)
    golden["tip4p_restart"] = {
        "source": "Examples/Restart-Examples/raspa2-long-output.txt",
        "gg_vdw": float(vdw), "gg_real": float(real),
        "gg_fourier": float(grab(out, r"CPU Guest-Guest Fourier: " + NUM)),
        "self_guest": float(grab(out, r"Component: 1, SelfAtomE: " + NUM)),
        "intra_exclusion_guest": float(grab(out, r"Component: 1, Intra-Molecular ExclusionE: " + NUM)),
        "template_intra": float(grab(out, r"DEBUG: comp: 1, IntraE: " + NUM)),
        "template_self": float(grab(out, r"DEBUG: comp: 1, IntraE: " + NUM + ", SelfE: " + NUM)[1]),
        "gg_ewald": float(grab(out, r"GGEwaldE: " + NUM, 1))}
    # CO2 in MFI-2x2x2: charged framework only
    d = os.path.join(REF, "CO2-MFI")
    s = deck.read_deck(d)
    deck.save_system(os.path.join(HERE, "co2_mfi.npz"), s)
    out = open(os.path.join(d, "output.txt")).read()
    golden["co2_mfi"] = {
        "source": "Examples/CO2-MFI/output.txt",
        "hh_fourier": float(grab(out, r"Host-Host Fourier: " + NUM)),
        "self_framework": float(grab(out, r"Component: 0, SelfAtomE: " + NUM)),
        "intra_exclusion_framework": float(grab(out, r"Component: 0, Intra-Molecular ExclusionE: " + NUM)),
        "template_intra": float(grab(out, r"DEBUG: comp: 1, IntraE: " + NUM)),
        "template_self": float(grab(out, r"DEBUG: comp: 1, IntraE: " + NUM + ", SelfE: " + NUM)[1]),
        "n_structure_factors": int(grab(out, r"CPU SF: ([0-9]+)")),
        "volume": float(grab(out, r"Box Volume: +" + NUM)),
        "beta": float(grab(out, r"Box Beta: +" + NUM)),
        "n_framework_atoms": 2304}
    with open(os.path.join(HERE, "golden_values.json"), "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if "--values-only" in sys.argv:
        values_only()
    else:
        main()
