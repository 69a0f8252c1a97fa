"""Writes the golden fixtures for the oracle from gRASPA's own example decks.

Run in the build container, where the reference tree is mounted read-only at
/root/reference (it is not present where the tests run).  Each deck is read with
graspa_amd.deck and stored as a compact .npz (inputs); the energies gRASPA
printed for the same deck are stored in golden_values.json (expected outputs).

    python tests/golden/make_fixtures.py
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


def main():
    golden = {}
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
    main()
