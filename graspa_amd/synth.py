"""Seeded synthetic GCMC / Widom systems.

Frameworks: an n^3 simple-cubic lattice (n = ceil(N^(1/3)), spacing L/n), a
seeded random subset of N sites, Gaussian jitter (sigma 0.3 A) wrapped into the
cell, three pseudo-atom types with epsilon/k ~ U(20,80) K and sigma ~ U(2.5,3.5) A,
charges ~ N(0, 0.3) e shifted to a neutral framework.  Adsorbates: rigid linear
CO2 (O-C-O, C +0.70 e, O -0.35 e, C-O 1.16 A; O 79 K / 3.05 A, C 27 K / 2.80 A),
N2 (N +-0.55 A at -0.482 e, massless centre +0.964 e; N 36 K / 3.31 A) and
all-trans united-atom n-butane (CH3 98 K / 3.75 A, CH2 46 K / 3.95 A, 1.54 A
bonds, 114 degree bends).  Lorentz-Berthelot mixing, truncated, no tail.

Each builder returns a :class:`graspa_amd.deck.System`; :func:`write_deck`
writes it as a gRASPA deck (simulation.input, force field, pseudo atoms,
molecule .def, P1 CIF with charges) so a job is a normal reference-style input.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .deck import Component, System, lorentz_berthelot

FW_MASS = (12.0, 16.0, 28.0)


def framework(n_atoms, L, seed, jitter=0.3):
    """Positions (n,3), charges (n,), types (n,) and per-type (eps_K, sigma)."""
    rng = np.random.default_rng(seed)
    eps = rng.uniform(20.0, 80.0, 3)
    sig = rng.uniform(2.5, 3.5, 3)
    n = int(math.ceil(n_atoms ** (1.0 / 3.0) - 1e-12))
    a = L / n
    grid = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    pick = rng.choice(len(grid), size=n_atoms, replace=False)
    pos = (grid[pick] + 0.5) * a + rng.normal(0.0, jitter, (n_atoms, 3))
    pos = np.mod(pos, L)
    types = rng.integers(0, 3, n_atoms).astype(np.int32)
    q = rng.normal(0.0, 0.3, n_atoms)
    q -= q.mean()
    return pos, q, types, eps, sig


def _co2():
    return dict(name="CO2", names=["O_co2", "C_co2"], eps=[79.0, 27.0], sig=[3.05, 2.80], charge=[-0.35, 0.70],
                mass=[15.9994, 12.0], atoms=[(0, (0.0, 0.0, -1.16)), (1, (0.0, 0.0, 0.0)), (0, (0.0, 0.0, 1.16))])


def _n2():
    return dict(name="N2", names=["N_n2", "N_com"], eps=[36.0, 0.0], sig=[3.31, 0.0], charge=[-0.482, 0.964],
                mass=[14.00674, 0.0], atoms=[(0, (0.0, 0.0, 0.55)), (1, (0.0, 0.0, 0.0)), (0, (0.0, 0.0, -0.55))])


def _butane():
    b, th = 1.54, math.radians(114.0)
    half = (math.pi - th) / 2.0
    pts = [(0.0, 0.0, 0.0)]
    for i in range(1, 4):
        x = pts[-1][0] + b * math.cos(half)
        z = b * math.sin(half) * (1 if i % 2 else 0)
        pts.append((x, 0.0, z))
    return dict(name="butane", names=["CH3", "CH2"], eps=[98.0, 46.0], sig=[3.75, 3.95], charge=[0.0, 0.0],
                mass=[15.035, 14.027], atoms=[(0, pts[0]), (1, pts[1]), (1, pts[2]), (0, pts[3])])


def _methane():
    # united-atom methane (TraPPE: eps/k 148 K, sigma 3.73 A), uncharged, one site
    return dict(name="methane", names=["CH4"], eps=[148.0], sig=[3.73], charge=[0.0], mass=[16.04246],
                atoms=[(0, (0.0, 0.0, 0.0))])


def _argon():
    # argon (eps/k 119.8 K, sigma 3.40 A), uncharged, one site
    return dict(name="argon", names=["Ar"], eps=[119.8], sig=[3.40], charge=[0.0], mass=[39.948],
                atoms=[(0, (0.0, 0.0, 0.0))])


SPECIES = {"CO2": _co2, "N2": _n2, "butane": _butane, "methane": _methane, "argon": _argon}


def build(n_fw=1536, L=24.0, species=("CO2",), mol_fraction=None, seed=0, T=300.0, P=1e5, cutoff=12.0,
          charges=True, K=10, O=10, probs=None, cycles=2000, fw_seed=None, ucells=1):
    """A cubic box with a synthetic framework and the given adsorbate species.
    ucells > 1 replicates the framework cell ucells times along each axis (the
    reference's UnitCells keyword), so the box is ucells * L wide."""
    fw_pos, fw_q, fw_t, fw_eps, fw_sig = framework(n_fw, L, seed if fw_seed is None else fw_seed)
    if ucells > 1:
        shifts = np.stack(np.meshgrid(*([np.arange(ucells)] * 3), indexing="ij"), -1).reshape(-1, 3) * L
        fw_pos = (fw_pos[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
        fw_q = np.tile(fw_q, len(shifts))
        fw_t = np.tile(fw_t, len(shifts))
        n_fw = len(fw_t)
        L = L * ucells
    names = ["Fw_a", "Fw_b", "Fw_c"]   # CIF labels lose trailing digits (read_data.cpp:1652)
    eps_k = list(fw_eps)
    sigma = list(fw_sig)
    pch = [0.0, 0.0, 0.0]
    pmass = list(FW_MASS)
    comps = []
    for sp_name in species:
        sp = SPECIES[sp_name]()
        base = len(names)
        names += sp["names"]
        eps_k += sp["eps"]
        sigma += sp["sig"]
        pch += sp["charge"] if charges else [0.0] * len(sp["charge"])
        pmass += sp["mass"]
        tpos = np.array([a[1] for a in sp["atoms"]], dtype=np.float64)
        types = np.array([base + a[0] for a in sp["atoms"]], dtype=np.int32)
        q = np.array([pch[t] for t in types])
        masses = np.array([pmass[t] for t in types])
        comps.append(Component(name=sp["name"], molsize=len(types), template_pos=tpos, types=types, charges=q,
                               masses=masses))
    nt = len(names)
    eps, sig, sh, tail, use_tail = lorentz_berthelot(np.array(eps_k), np.array(sigma), [False] * nt, [False] * nt,
                                                     cutoff * cutoff)
    s = System(cell=np.array([L, 0.0, 0.0, 0.0, L, 0.0, 0.0, 0.0, L]), pseudo_names=names,
               pseudo_mass=np.array(pmass), pseudo_charge=np.array(pch), eps=eps, sig=sig, shift=sh, tail=tail,
               use_tail=use_tail, cut_vdw=cutoff, cut_coul=cutoff, no_charges=not charges)
    s.fw_pos, s.fw_charge, s.fw_type = fw_pos, (fw_q if charges else np.zeros(n_fw)), fw_t
    s.fw_mass = float(sum(FW_MASS[t] for t in fw_t))
    s.components = comps
    s.temperature, s.pressure_pa, s.seed = T, P, seed
    s.n_trial_pos, s.n_trial_orient = K, O
    s.n_prod = cycles
    s.ewald_precision = 1e-6
    if charges:
        s.setup_ewald()
    fr = mol_fraction or [1.0 / len(comps)] * len(comps)
    for c, f in zip(comps, fr):
        c.mol_fraction = f
        pr = probs or dict(translation=1.0, rotation=1.0, swap=2.0)
        c.p_translation = pr.get("translation", 0.0)
        c.p_rotation = pr.get("rotation", 0.0)
        c.p_widom = pr.get("widom", 0.0)
        c.p_reinsertion = pr.get("reinsertion", 0.0)
        c.p_identity = pr.get("identity", 0.0)
        c.p_swap = pr.get("swap", 0.0)
    return s


def config0(seed=0, n_fw=1536, L=24.0, cycles=2000):
    """Plain FP64 GCMC of CO2, one box: translate:rotate:insert:delete 1:1:1:1."""
    return build(n_fw=n_fw, L=L, species=("CO2",), seed=seed, T=300.0, P=1e5, cycles=cycles,
                 probs=dict(translation=1.0, rotation=1.0, swap=2.0))


def config1(seed=0, n_fw=1536, L=24.0):
    """Widom batch: CO2 test insertions into the config-0 framework, empty box."""
    s = build(n_fw=n_fw, L=L, species=("CO2",), seed=seed, T=300.0, P=1e5, probs=dict(widom=1.0))
    s.use_max_step, s.max_step_per_cycle = True, 1
    return s


# TraPPE-UA n-butane bonded terms: bond 1.54 A, bend 114 deg with k/kB 62,500
# K/rad^2, dihedral cosine series (0, 355.03, -68.19, 791.32) K
BUTANE_BEND_K = 62500.0
BUTANE_TORSION_K = (0.0, 355.03, -68.19, 791.32)


def butane_chain():
    """The flexible n-butane model (bond, bend, dihedral) in internal units."""
    from .deck import K_TO_INTERNAL, Chain
    M = 4
    bond = np.full(M, 1.54)
    bond[0] = 0.0
    theta0 = np.zeros(M)
    theta0[2:] = math.radians(114.0)
    kb = np.zeros(M)
    kb[2:] = BUTANE_BEND_K / K_TO_INTERNAL
    tors = np.zeros((M, 4))
    tors[3] = np.array(BUTANE_TORSION_K) / K_TO_INTERNAL
    return Chain(bond=bond, theta0=theta0, k_bend=kb, torsion=tors)


def config2(seed=2000, n_fw=3000, L=30.0, P=1e4, flexible=False):
    """United-atom n-butane CBMC GCMC, uncharged, 16 x 8 trials.  flexible=True
    grows the chain bead by bead with the TraPPE bend and dihedral (the
    workload as stated); flexible=False is the rigid all-trans molecule, the
    form the reference can run."""
    s = build(n_fw=n_fw, L=L, species=("butane",), seed=seed, T=300.0, P=P, charges=False, K=16, O=8,
              probs=dict(translation=1.0, rotation=1.0, swap=2.0))
    if flexible:
        s.components[0].chain = butane_chain()
        s.components[0].rigid = False
    return s


def config3_box(b, seed_base=3000, cutoff=12.0):
    """Box b of the 256-box config: N ~ U{500..6000}, L ~ U(18,40), CO2/N2 50:50,
    T ~ U(250,350), P log-U(1e3,1e7).  A cell narrower than twice the cutoff
    is replicated (UnitCells 2 2 2), as a reference deck must do for the
    minimum-image convention, so N and L are those of the unit cell."""
    rng = np.random.default_rng(seed_base + b)
    n = int(rng.integers(500, 6001))
    L = float(rng.uniform(18.0, 40.0))
    T = float(rng.uniform(250.0, 350.0))
    P = float(10 ** rng.uniform(3.0, 7.0))
    uc = int(math.ceil(2.0 * cutoff / L - 1e-12))
    return build(n_fw=n, L=L, species=("CO2", "N2"), mol_fraction=[0.5, 0.5], seed=seed_base + b, T=T, P=P,
                 cycles=1000, probs=dict(translation=1.0, rotation=1.0, swap=2.0), ucells=max(uc, 1),
                 cutoff=cutoff)


def write_deck(system: System, dirpath: str, framework_name="synth", cycles=None):
    """Writes the system as a gRASPA deck directory."""
    s = system
    os.makedirs(dirpath, exist_ok=True)
    T = s.ntypes
    with open(os.path.join(dirpath, "force_field_mixing_rules.def"), "w") as f:
        f.write("# general rule for shifted vs truncated\ntruncated\n# general rule tailcorrections\nno\n")
        f.write("# number of defined interactions\n%d\n# type interaction, parameters\n" % T)
        for i, n in enumerate(s.pseudo_names):
            # the diagonal of the mixed table carries the per-atom values
            f.write("%s lennard-jones %.10f %.10f\n" % (n, s.eps[i, i] * 1.20272430057, s.sig[i, i]))
        f.write("# general mixing rule for Lennard-Jones\nLorentz-Berthelot\n")
    with open(os.path.join(dirpath, "pseudo_atoms.def"), "w") as f:
        f.write("#number of pseudo atoms\n%d\n#type print as chem oxi mass charge pol\n" % T)
        for i, n in enumerate(s.pseudo_names):
            f.write("%s yes %s %s 0 %.6f %.6f 0.0\n" % (n, n[0], n[0], s.pseudo_mass[i], s.pseudo_charge[i]))
    with open(os.path.join(dirpath, "force_field.def"), "w") as f:
        f.write("# rules to overwrite\n0\n# number of defined interactions\n0\n# mixing rules to overwrite\n0\n")
    for c in s.components:
        with open(os.path.join(dirpath, c.name + ".def"), "w") as f:
            f.write("# critical constants: Temperature [T], Pressure [Pa], and Acentric factor [-]\n0.0\n0.0\n0.0\n")
            f.write("#Number Of Atoms\n%d\n# Number of groups\n1\n# group\nrigid\n# number of atoms\n%d\n"
                    "# atomic positions\n" % (c.molsize, c.molsize))
            for i in range(c.molsize):
                x, y, z = c.template_pos[i]
                f.write("%d %s %.10f %.10f %.10f\n" % (i, s.pseudo_names[c.types[i]], x, y, z))
    inv = s.inv_cell
    with open(os.path.join(dirpath, framework_name + ".cif"), "w") as f:
        L = s.cell[0], s.cell[4], s.cell[8]
        f.write("data_%s\n_cell_length_a %.10f\n_cell_length_b %.10f\n_cell_length_c %.10f\n" % (framework_name, *L))
        f.write("_cell_angle_alpha 90\n_cell_angle_beta 90\n_cell_angle_gamma 90\n")
        f.write("_symmetry_space_group_name_H-M 'P 1'\n\nloop_\n_atom_site_label\n_atom_site_type_symbol\n"
                "_atom_site_fract_x\n_atom_site_fract_y\n_atom_site_fract_z\n_atom_site_charge\n")
        for p, q, t in zip(s.fw_pos, s.fw_charge, s.fw_type):
            fr = [inv[0] * p[0], inv[4] * p[1], inv[8] * p[2]]
            f.write("%s %s %.12f %.12f %.12f %.12f\n" % (s.pseudo_names[t], s.pseudo_names[t], *fr, q))
    with open(os.path.join(dirpath, "simulation.input"), "w") as f:
        f.write("NumberOfInitializationCycles 0\nNumberOfEquilibrationCycles 0\n")
        f.write("NumberOfProductionCycles %d\n" % (cycles if cycles is not None else s.n_prod))
        f.write("UseMaxStep %s\nMaxStepPerCycle %d\n" % ("yes" if s.use_max_step else "no", s.max_step_per_cycle))
        f.write("UseChargesFromCIFFile yes\nRestartFile no\nRandomSeed %d\n" % s.seed)
        f.write("NumberOfTrialPositions %d\nNumberOfTrialOrientations %d\n" % (s.n_trial_pos, s.n_trial_orient))
        f.write("NumberOfBlocks 5\nAdsorbateAllocateSpace %d\nNumberOfSimulations 1\nSingleSimulation yes\n"
                % s.allocate_space)
        f.write("InputFileType cif\nFrameworkName %s\nUnitCells 0 1 1 1\n" % framework_name)
        f.write("ChargeMethod %s\nTemperature %.6f\nPressure %.6f\n" % ("None" if s.no_charges else "Ewald",
                                                                        s.temperature, s.pressure_pa))
        f.write("OverlapCriteria %g\nCutOffVDW %g\nCutOffCoulomb %g\nEwaldPrecision %g\n\n" %
                (s.overlap, s.cut_vdw, s.cut_coul, s.ewald_precision))
        for i, c in enumerate(s.components):
            f.write("Component %d MoleculeName %s\n" % (i, c.name))
            f.write("            IdealGasRosenbluthWeight %g\n            FugacityCoefficient %g\n" %
                    (c.ideal_rosen, c.fugacity_coeff))
            f.write("            MolFraction %g\n" % c.mol_fraction)
            for key, v in (("TranslationProbability", c.p_translation), ("RotationProbability", c.p_rotation),
                           ("WidomProbability", c.p_widom), ("ReinsertionProbability", c.p_reinsertion),
                           ("IdentityChangeProbability", c.p_identity), ("SwapProbability", c.p_swap)):
                if v:
                    f.write("            %s %g\n" % (key, v))
            f.write("            CreateNumberOfMolecules 0\n\n")
