"""Reader for gRASPA simulation decks (the input side of the energy engine).

Reads the subset of ``simulation.input`` keywords, ``force_field_mixing_rules.def``,
``force_field.def``, ``pseudo_atoms.def``, molecule ``.def`` files, P1 CIF files and
RASPA-2 restart files that the GCMC/Widom hot path needs, with the reference's own
parsing rules (line-oriented substring matching, ``.def`` files parsed by line
number).  Citations are to gRASPA's src_clean/read_data.cpp.

The result is a :class:`System`: plain numpy arrays in the reference's internal
units (energy 10 J/mol, length Angstrom, charge e), ready to hand to the engine.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# read_data.cpp:611, 1200; data_struct.h:54-64
COULOMB_PREFACTOR = 138935.483496
K_TO_INTERNAL = 1.20272430057
ENERGY_TO_KELVIN = 1.2027242847
MASS_UNIT = 1.6605402e-27
LENGTH_UNIT = 1e-10
TIME_UNIT = 1e-12
BOLTZMANN = 1.380649e-23
AVOGADRO = 6.02214076e23
GAS_CONSTANT = 8.314462618


def beta_of(temperature: float) -> float:
    """fxn_main.h:117: 1/(k_B T) in internal units."""
    return 1.0 / (BOLTZMANN / (MASS_UNIT * LENGTH_UNIT ** 2 / TIME_UNIT ** 2) * temperature)


def pressure_internal(p_pa: float) -> float:
    """fxn_main.h:120: pascal to internal pressure units."""
    return p_pa / (MASS_UNIT / (LENGTH_UNIT * TIME_UNIT ** 2))


def _split(line: str) -> List[str]:
    """Split_Tab_Space: whitespace split."""
    return line.split()


@dataclass
class Chain:
    """Bonded model of a flexible linear chain.  No counterpart in the
    reference, which grows rigid molecules only (read_data.cpp:2104-2116).
    Entry i belongs to bead i: bond[i] to bead i - 1 (i >= 1), the harmonic
    bend U = k_bend[i]/2 (theta - theta0[i])^2 of the angle (i-2, i-1, i)
    (i >= 2) and the dihedral U = t0 + t1 (1 + cos phi) + t2 (1 - cos 2 phi)
    + t3 (1 + cos 3 phi) of (i-3, i-2, i-1, i) (i >= 3, phi = pi trans).
    Lengths in A, angles in rad, energies in internal units (10 J/mol)."""
    bond: np.ndarray
    theta0: np.ndarray
    k_bend: np.ndarray
    torsion: np.ndarray               # (M, 4)


@dataclass
class Component:
    """One adsorbate species (read_data.cpp:2044-2195 and 2298-2696)."""
    name: str
    molsize: int
    template_pos: np.ndarray          # (M, 3) from the .def file
    types: np.ndarray                 # (M,) pseudo-atom indices
    charges: np.ndarray               # (M,)
    masses: np.ndarray                # (M,)
    rigid: bool = True
    ideal_rosen: float = 1.0
    fugacity_coeff: float = 1.0
    mol_fraction: float = 1.0
    p_translation: float = 0.0
    p_rotation: float = 0.0
    p_widom: float = 0.0
    p_reinsertion: float = 0.0
    p_identity: float = 0.0
    p_swap: float = 0.0
    create_molecules: int = 0
    # flexible linear chain (None: rigid, as every reference molecule)
    chain: Optional[Chain] = None
    # initial configuration (restart file); (n_mol*M, 3) and MolIDs
    init_pos: Optional[np.ndarray] = None
    init_molid: Optional[np.ndarray] = None
    init_charge: Optional[np.ndarray] = None

    @property
    def has_charge(self) -> bool:
        return float(np.abs(self.charges).sum()) > 1e-6

    def cumulative_probabilities(self):
        """Move_Statistics::NormalizeProbabilities (data_struct.h:569-608)."""
        t, r, sr, w = self.p_translation, self.p_rotation, 0.0, self.p_widom
        re, idn, cbcf, sw = self.p_reinsertion, self.p_identity, 0.0, self.p_swap
        total = t + r + sr + w + re + idn + sw + cbcf
        if total > 1e-10:
            t, r, sr, w, sw, re, idn, cbcf = (v / total for v in (t, r, sr, w, sw, re, idn, cbcf))
        r += t
        sr += r
        w += sr
        re += w
        idn += re
        cbcf += idn
        sw += cbcf
        return dict(translation=t, rotation=r, special_rotation=sr, widom=w,
                    reinsertion=re, identity=idn, cbcf=cbcf, swap=sw, total=total)


@dataclass
class System:
    """A simulation box in the reference's internal units."""
    cell: np.ndarray                      # (9,) column vectors a,b,c
    pseudo_names: List[str]
    pseudo_mass: np.ndarray
    pseudo_charge: np.ndarray
    eps: np.ndarray                       # (T,T) internal units
    sig: np.ndarray                       # (T,T)
    shift: np.ndarray                     # (T,T)
    tail: np.ndarray                      # (T,T)
    use_tail: np.ndarray                  # (T,T) int
    cut_vdw: float = 12.0
    cut_coul: float = 12.0
    overlap: float = 1e6
    no_charges: bool = True
    ewald_precision: float = 1e-6
    alpha: float = 0.0
    kmax: tuple = (0, 0, 0)
    recip_cut: float = 0.0
    lammps_ewald: bool = False
    prefactor: float = COULOMB_PREFACTOR
    fw_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    fw_charge: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fw_type: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    fw_mass: float = 0.0
    unit_cells: tuple = (1, 1, 1)
    components: List[Component] = field(default_factory=list)
    temperature: float = 300.0
    pressure_pa: float = 1e5
    seed: int = 0
    n_init: int = 0
    n_equil: int = 0
    n_prod: int = 0
    use_max_step: bool = False
    max_step_per_cycle: int = 1
    n_trial_pos: int = 10
    n_trial_orient: int = 10
    vdw_real_bias: bool = True
    allocate_space: int = 10240
    n_blocks: int = 5                     # NumberOfBlocks (block averages)
    print_every: int = 0                  # PrintEvery (cycles between progress lines)
    movies_every: int = 0                 # MoviesEvery (cycles between LAMMPS snapshots, 0 = none)
    save_output: bool = False             # SaveOutputToFile

    # -- derived geometry (maths.cuh:28-55) --
    @property
    def inv_cell(self) -> np.ndarray:
        x = self.cell
        m11, m21, m31 = x[0], x[3], x[6]
        m12, m22, m32 = x[1], x[4], x[7]
        m13, m23, m33 = x[2], x[5], x[8]
        det = self.volume
        r = np.zeros(9)
        r[0] = +(m22 * m33 - m32 * m23) / det
        r[3] = -(m21 * m33 - m31 * m23) / det
        r[6] = +(m21 * m32 - m31 * m22) / det
        r[1] = -(m12 * m33 - m32 * m13) / det
        r[4] = +(m11 * m33 - m31 * m13) / det
        r[7] = -(m11 * m32 - m31 * m12) / det
        r[2] = +(m12 * m23 - m22 * m13) / det
        r[5] = -(m11 * m23 - m21 * m13) / det
        r[8] = +(m11 * m22 - m21 * m12) / det
        return r

    @property
    def volume(self) -> float:
        x = self.cell
        m11, m21, m31 = x[0], x[3], x[6]
        m12, m22, m32 = x[1], x[4], x[7]
        m13, m23, m33 = x[2], x[5], x[8]
        return float(+m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31))

    @property
    def cubic(self) -> bool:
        return (abs(self.cell[3]) + abs(self.cell[6]) + abs(self.cell[7])) <= 1e-10

    @property
    def ntypes(self) -> int:
        return len(self.pseudo_names)

    @property
    def beta(self) -> float:
        return beta_of(self.temperature)

    @property
    def pressure(self) -> float:
        return pressure_internal(self.pressure_pa)

    @property
    def nvec(self) -> int:
        kx, ky, kz = self.kmax
        return (kx + 1) * (2 * ky + 1) * (2 * kz + 1)

    def setup_ewald(self):
        """read_Ewald_Parameters_from_input (read_data.cpp:609-709), RASPA-2 branch:
        alpha and kmax from the precision and the Coulomb cutoff."""
        if self.lammps_ewald:
            lx, ly, lz = self.cell[0], self.cell[4], self.cell[8]
            xy = xz = yz = 0.0
            ux = 2 * math.pi / lx
            uy = 2 * math.pi * (-xy) / lx / ly
            uz = 2 * math.pi * (xy * yz - ly * xz) / lx / ly / lz
            vy = 2 * math.pi / ly
            vz = 2 * math.pi * (-yz) / ly / lz
            wz = 2 * math.pi / lz
            kx, ky, kz = self.kmax
            kvx = kx * ux
            kvy = kx * uy + ky * vy
            kvz = kx * uz + ky * vz + kz * wz
            self.recip_cut = max(kvx * kvx, kvy * kvy, kvz * kvz) * 1.00001
            return
        p, rc = self.ewald_precision, self.cut_coul
        tol = math.sqrt(abs(math.log(p * rc)))
        alpha = math.sqrt(abs(math.log(p * rc * tol))) / rc
        tol1 = math.sqrt(-math.log(p * rc * (2.0 * tol * alpha) ** 2))
        self.alpha = alpha
        kmax = [int(_cpp_round(0.25 + self.cell[i] * alpha * tol1 / math.pi)) for i in (0, 4, 8)]
        self.kmax = tuple(kmax)
        self.recip_cut = (1.05 * float(max(kmax))) ** 2

    def pseudo_counts(self, n_mol_per_comp=None) -> np.ndarray:
        """NumberOfPseudoAtoms over framework + adsorbates (tail corrections)."""
        n = np.zeros(self.ntypes, dtype=np.int64)
        for t in self.fw_type:
            n[t] += 1
        for ci, c in enumerate(self.components):
            nm = 0 if n_mol_per_comp is None else n_mol_per_comp[ci]
            for t in c.types:
                n[t] += nm
        return n

    @property
    def has_tail(self) -> bool:
        return bool(np.any(self.use_tail))

    def framework_density(self) -> float:
        """FrameworkDensity (print_statistics.cuh:70-98), kg/m^3."""
        return self.fw_mass * 1.0e-3 / (AVOGADRO * self.volume * 1.0e-30)


def _cpp_round(x: float) -> float:
    """std::round: half away from zero."""
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def lorentz_berthelot(eps_k, sig, shifted, tails, cut_vdw2):
    """ForceField_Processing (read_data.cpp:1179-1247): mixed epsilon (internal
    units), sigma, shift and tail tables."""
    n = len(eps_k)
    eps = np.zeros((n, n))
    sg = np.zeros((n, n))
    sh = np.zeros((n, n))
    tail = np.zeros((n, n))
    use_tail = np.zeros((n, n), dtype=np.int32)
    for i in range(n):
        for j in range(n):
            e = math.sqrt(eps_k[i] * eps_k[j]) / K_TO_INTERNAL
            s = 0.5 * (sig[i] + sig[j])
            eps[i, j] = e
            sg[i, j] = s
            if shifted[i] and shifted[j]:
                sh[i, j] = _shift_value(e, s, cut_vdw2)
            if tails[i] and tails[j]:
                tail[i, j] = tail_value(e, s, cut_vdw2)
                use_tail[i, j] = 1
    return eps, sg, sh, tail, use_tail


def _shift_value(eps, sig, rc2):
    """Get_Shifted_Value (read_data.cpp:737-748)."""
    arg2 = sig * sig
    temp = rc2 / arg2
    rri3 = 1.0 / (temp * temp * temp + 0.0)
    return 4.0 * eps * (rri3 * (rri3 - 1.0))


def tail_value(eps, sig, rc2):
    """GetTailCorrectionValue (read_data.cpp:837-846)."""
    arg2 = sig * sig * sig
    rr = math.sqrt(rc2)
    term1 = arg2 ** 4 / (9.0 * rr ** 9)
    term2 = arg2 ** 2 / (3.0 * rr ** 3)
    return 16.0 * 3.14159265358979323846 / 2.0 * eps * (term1 - term2)


def cell_from_parameters(a, b, c, alpha_deg, beta_deg, gamma_deg, ucells=(1, 1, 1)) -> np.ndarray:
    """CheckFrameworkCIF (read_data.cpp:1507-1547): cell matrix from lengths
    and angles, times the number of unit cells."""
    d2r = 180.0 / 3.14159265358979323846
    al, be, ga = alpha_deg / d2r, beta_deg / d2r, gamma_deg / d2r
    ax = a
    by = b * math.sin(ga)
    tempd = (math.cos(al) - math.cos(ga) * math.cos(be)) / math.sin(ga)
    cz = c * math.sqrt(1 - math.cos(be) ** 2 - tempd ** 2)
    bx = b * math.cos(ga)
    cx = c * math.cos(be)
    cy = c * tempd
    nx, ny, nz = ucells
    return np.array([nx * ax, 0.0, 0.0, ny * bx, ny * by, 0.0, nz * cx, nz * cy, nz * cz])


# --------------------------------------------------------------------------
# file readers
# --------------------------------------------------------------------------

def read_pseudo_atoms(path):
    """PseudoAtomParser (read_data.cpp:1324-1369)."""
    names, mass, charge = [], [], []
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(_split(lines[1])[0])
    for i in range(n):
        t = _split(lines[3 + i])
        names.append(t[0])
        mass.append(float(t[5]))
        charge.append(float(t[6]))
    return names, np.array(mass), np.array(charge)


def read_mixing_rules(path):
    """ForceFieldParser (read_data.cpp:772-834): shifted/truncated, tail,
    per-atom epsilon (K) and sigma."""
    with open(path) as f:
        lines = f.read().splitlines()
    shifted = _split(lines[1])[0] == "shifted"
    tail = _split(lines[3])[0] == "yes"
    n = int(_split(lines[5])[0])
    names, eps, sig = [], [], []
    for i in range(n):
        t = _split(lines[7 + i])
        if len(t) > 1 and t[1].lower() != "lennard-jones":
            # lennard-jones-1264 rows carry {C12, C6, C4, C10} (read_data.cpp:820-834),
            # not epsilon / sigma: reading them as such would give wrong energies
            raise NotImplementedError("%s interaction %s in %s: this engine evaluates 12-6 Lennard-Jones only"
                                      % (t[0], t[1], os.path.basename(path)))
        names.append(t[0])
        eps.append(float(t[2]))
        sig.append(float(t[3]))
    return names, np.array(eps), np.array(sig), shifted, tail


def apply_force_field_overrides(path, names, use_tail, tail, eps, sig, cut_vdw2):
    """OverWriteTailCorrection (read_data.cpp:1132-1172): per-pair tail
    corrections switched on in force_field.def."""
    if not os.path.exists(path):
        return
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(_split(lines[1])[0])
    for k in range(n):
        t = _split(lines[3 + k])
        if len(t) != 4:
            raise ValueError("Tail Correction Overwrite need 4 terms for each entry")
        if t[3] == "yes":
            i, j = names.index(t[0]), names.index(t[1])
            v = tail_value(eps[i, j], sig[i, j], cut_vdw2)
            tail[i, j] = tail[j, i] = v
            use_tail[i, j] = use_tail[j, i] = 1


def read_molecule_def(path, pseudo_names, pseudo_charge, pseudo_mass):
    """MoleculeDefinitionParser (read_data.cpp:2044-2195): rigid molecules only."""
    with open(path) as f:
        lines = f.read().splitlines()
    M = int(_split(lines[5])[0])
    if _split(lines[9])[0].lower() != "rigid":
        raise ValueError("Currently Not allowing flexible molecule")
    pos, types = [], []
    for i in range(M):
        t = _split(lines[13 + i])
        if len(t) == 5:
            pos.append([float(t[2]), float(t[3]), float(t[4])])
        elif len(t) == 2 and M == 1:
            pos.append([0.0, 0.0, 0.0])
        else:
            raise ValueError("Flexible molecules not implemented")
        types.append(pseudo_names.index(t[1]))
    types = np.array(types, dtype=np.int32)
    charges = np.array([pseudo_charge[t] for t in types])
    if charges.sum() > 1e-10:
        raise ValueError("Molecule not neutral, bad")
    masses = np.array([pseudo_mass[t] for t in types])
    return np.array(pos, dtype=np.float64), types, charges, masses


def read_cif(path, pseudo_names, pseudo_charge, pseudo_mass, use_cif_charges, ucells):
    """CheckFrameworkCIF (read_data.cpp:1474-1849): P1 CIF, supercell by
    replication, labels with trailing digits stripped."""
    with open(path) as f:
        lines = f.read().splitlines()
    vals = {}
    for ln in lines:
        for key in ("_cell_length_a", "_cell_length_b", "_cell_length_c",
                    "_cell_angle_alpha", "_cell_angle_beta", "_cell_angle_gamma"):
            if key in ln:
                vals[key] = float(_split(ln)[1])
    cell = cell_from_parameters(vals["_cell_length_a"], vals["_cell_length_b"], vals["_cell_length_c"],
                                vals["_cell_angle_alpha"], vals["_cell_angle_beta"], vals["_cell_angle_gamma"],
                                ucells)
    cols = {}
    start = end = None
    count = 0
    found = False
    for i, ln in enumerate(lines):
        if "_atom_site" in ln:
            if start is None:
                start = i
            found = True
            for key, idx in (("_atom_site_label", 0), ("_atom_site_fract_x", 1), ("_atom_site_fract_y", 2),
                             ("_atom_site_fract_z", 3), ("_atom_site_charge", 4)):
                if key in ln:
                    cols[idx] = count
            count += 1
        elif found:
            end = i - 1
            break
    fpos, charge, types = [], [], []
    mass = 0.0
    if end is not None:
        for ln in lines[end + 1:]:
            t = _split(ln)
            if len(t) < 4:
                break
            label = t[cols[0]].rstrip("0123456789")
            ti = pseudo_names.index(label)
            q = float(t[cols[4]]) if (use_cif_charges and 4 in cols) else float(pseudo_charge[ti])
            fpos.append([float(t[cols[1]]), float(t[cols[2]]), float(t[cols[3]])])
            charge.append(q)
            types.append(ti)
            mass += pseudo_mass[ti]
    nx, ny, nz = ucells
    out_pos, out_q, out_t = [], [], []
    for ix in range(int(nx)):
        for jy in range(int(ny)):
            for kz in range(int(nz)):
                for a in range(len(fpos)):
                    fx = (fpos[a][0] + ix) * (1.0 / nx)
                    fy = (fpos[a][1] + jy) * (1.0 / ny)
                    fz = (fpos[a][2] + kz) * (1.0 / nz)
                    out_pos.append([fx * cell[0] + fy * cell[3] + fz * cell[6],
                                    fx * cell[1] + fy * cell[4] + fz * cell[7],
                                    fx * cell[2] + fy * cell[5] + fz * cell[8]])
                    out_q.append(charge[a])
                    out_t.append(types[a])
    return (cell, np.array(out_pos, dtype=np.float64).reshape(-1, 3), np.array(out_q, dtype=np.float64),
            np.array(out_t, dtype=np.int32), mass * nx * ny * nz)


def read_restart(path, system: System):
    """RestartFileParser (read_data.cpp:3000-3221): molecule counts, atom
    positions (molecules unwrapped about their first atom), MolIDs, charges."""
    with open(path) as f:
        lines = f.read().splitlines()
    prev = 0
    for ci, comp in enumerate(system.components):
        start = None
        nmol = 0
        for i, ln in enumerate(lines):
            if ("Component: %d" % ci) in ln:
                nmol = int(_split(ln)[3])
                start = i + 2
                break
        if nmol == 0:
            comp.init_pos = np.zeros((0, 3))
            comp.init_molid = np.zeros(0, dtype=np.int64)
            comp.init_charge = np.zeros(0)
            continue
        M = comp.molsize
        interval = nmol * M
        pos = np.zeros((interval, 3))
        molid = np.zeros(interval, dtype=np.int64)
        charge = np.zeros(interval)
        inv = system.inv_cell
        first = None
        for a in range(interval):
            t = _split(lines[start + a])
            p = np.array([float(t[3]), float(t[4]), float(t[5])])
            atomid = int(t[2])
            molid[a] = int(t[1]) - prev
            if atomid == 0:
                first = p.copy()
            else:
                d = p - first
                d = _pbc(d, system.cell, inv, system.cubic)
                p = first + d
            pos[a] = p
        for a in range(interval):
            t = _split(lines[start + 3 * interval + a])
            charge[a] = float(t[3])
        comp.init_pos = pos
        comp.init_molid = molid
        comp.init_charge = charge
        prev += nmol


def _pbc(d, cell, inv, cubic):
    """maths.cuh:427-450 (for the restart unwrapping only)."""
    d = d.copy()
    if cubic:
        for k, idx in ((0, 0), (1, 4), (2, 8)):
            v = d[k] * inv[idx] + (0.5 if d[k] >= 0.0 else -0.5)
            d[k] -= float(int(v)) * cell[idx]
        return d
    s = [inv[0] * d[0] + inv[3] * d[1] + inv[6] * d[2],
         inv[1] * d[0] + inv[4] * d[1] + inv[7] * d[2],
         inv[2] * d[0] + inv[5] * d[1] + inv[8] * d[2]]
    s = [x - float(int(x + (0.5 if x >= 0.0 else -0.5))) for x in s]
    return np.array([cell[0] * s[0] + cell[3] * s[1] + cell[6] * s[2],
                     cell[1] * s[0] + cell[4] * s[1] + cell[7] * s[2],
                     cell[2] * s[0] + cell[5] * s[1] + cell[8] * s[2]])


# Keywords of simulation.input for methods this engine does not run
# (read_data.cpp:233-709): a deck that switches one of them on is refused
# rather than run without it.
_UNSUPPORTED_NUMERIC = ("GibbsParticleXferProbability", "GibbsVolumeChangeProbability",
                        "NPTVolumeChangeProbability")
_UNSUPPORTED_YES = ("RunTMMC", "UseDNNforHostGuest", "BlockPockets", "TURN_OFF_CBMC_SWAP", "SeparateFrameworkComponents",
                    "DifferentFrameworks")
_UNSUPPORTED_COMPONENT = ("CBCFProbability", "RotationSpecialProbability", "GibbsSwapProbability")


def _reject_unsupported(get, comp_blocks):
    # UseLJ1264 yes|true switches every pair to the 12-6-4 polynomial with the
    # GENERIC2_HC overrides of force_field.def (maths.cuh:454-476,
    # read_data.cpp:506-514): refused, since the engine's pair terms are 12-6
    t = get("UseLJ1264")
    if t and len(t) > 1 and t[1].lower() in ("yes", "true"):
        raise NotImplementedError("UseLJ1264 %s: the 12-6-4 polynomial pair potential is not supported by this "
                                  "engine (12-6 Lennard-Jones only)" % t[1])
    for key in _UNSUPPORTED_NUMERIC:
        t = get(key)
        if t and len(t) > 1 and float(t[1]) > 0.0:
            raise NotImplementedError("%s %s: this engine runs GCMC / Widom in one box only" % (key, t[1]))
    for key in _UNSUPPORTED_YES:
        t = get(key)
        if t and len(t) > 1 and t[1].lower() == "yes":
            raise NotImplementedError("%s yes is not supported by this engine" % key)
    t = get("NumberOfSimulations")
    if t and len(t) > 1 and int(t[1]) > 1:
        raise NotImplementedError("NumberOfSimulations %s: one box per deck (run several decks as a batch)" % t[1])
    t = get("InputFileType")
    if t and len(t) > 1 and t[1].lower() != "cif":
        raise NotImplementedError("InputFileType %s: frameworks are read from P1 CIF files" % t[1])
    t = get("NumberofFrameworkComponents")
    if t and len(t) > 1 and int(t[1]) > 1:
        raise NotImplementedError("NumberofFrameworkComponents %s: one framework component" % t[1])
    for blk in comp_blocks:
        for key in _UNSUPPORTED_COMPONENT:
            if key in blk and float(blk[key]) > 0.0:
                raise NotImplementedError("%s %s (component %s) is not supported by this engine"
                                          % (key, blk[key], blk["MoleculeName"]))


def read_deck(dirpath: str) -> System:
    """Read a gRASPA deck directory into a :class:`System`."""
    sim = os.path.join(dirpath, "simulation.input")
    with open(sim) as f:
        lines = f.read().splitlines()
    comp_blocks = []
    cur = None
    global_lines = []
    for ln in lines:
        t = _split(ln)
        if not t:
            continue
        if ln.lstrip().startswith("Component ") and "MoleculeName" in ln:
            cur = {"MoleculeName": t[3]}
            comp_blocks.append(cur)
            continue
        if cur is not None and len(t) >= 2:
            cur.setdefault(t[0], t[1])
        elif cur is None:
            global_lines.append(ln)

    def get(key, default=None):
        # read_simulation_input (read_data.cpp:233-381): every line containing the
        # key assigns, so the last matching line before the components wins
        found = default
        for ln in global_lines:
            if key in ln:
                found = _split(ln)
        return found

    names_ff, eps_k, sig_a, shifted, tails = read_mixing_rules(os.path.join(dirpath, "force_field_mixing_rules.def"))
    pnames, pmass, pcharge = read_pseudo_atoms(os.path.join(dirpath, "pseudo_atoms.def"))
    if pnames != names_ff:
        raise ValueError("Order of pseudo-atom and force field definition don't match!")
    cut_vdw = float(get("CutOffVDW", ["", "12.0"])[1])
    cut_coul = float(get("CutOffCoulomb", ["", "12.0"])[1])
    eps, sig, sh, tail, use_tail = lorentz_berthelot(eps_k, sig_a, [shifted] * len(eps_k), [tails] * len(eps_k),
                                                     cut_vdw * cut_vdw)
    apply_force_field_overrides(os.path.join(dirpath, "force_field.def"), names_ff, use_tail, tail, eps, sig,
                                cut_vdw * cut_vdw)
    cm = get("ChargeMethod")
    no_charges = not (cm is not None and cm[1].lower() == "ewald")
    fwname = get("FrameworkName")[1]
    uc = get("UnitCells 0")
    ucells = (float(uc[2]), float(uc[3]), float(uc[4])) if uc else (1.0, 1.0, 1.0)
    use_cif_q = False
    t = get("UseChargesFromCIFFile")
    if t is not None and t[1].lower() == "yes":
        use_cif_q = True
    cell, fpos, fq, ft, fmass = read_cif(os.path.join(dirpath, fwname + ".cif"), pnames, pcharge, pmass,
                                         use_cif_q, ucells)
    s = System(cell=cell, pseudo_names=pnames, pseudo_mass=pmass, pseudo_charge=pcharge,
               eps=eps, sig=sig, shift=sh, tail=tail, use_tail=use_tail,
               cut_vdw=cut_vdw, cut_coul=cut_coul, no_charges=no_charges)
    s.fw_pos, s.fw_charge, s.fw_type, s.fw_mass = fpos, fq, ft, fmass
    s.unit_cells = tuple(int(v) for v in ucells)
    t = get("OverlapCriteria")
    if t:
        s.overlap = float(t[1])
    t = get("EwaldPrecision")
    if t:
        s.ewald_precision = float(t[1])
    t = get("Ewald_UseLAMMPS_Setup")
    if t and t[1].lower() == "yes":
        s.lammps_ewald = True
        s.alpha = float(get("Ewald_Alpha")[1])
        kv3 = get("Ewald_kvectors")
        s.kmax = (int(kv3[1]), int(kv3[2]), int(kv3[3]))
    s.setup_ewald()
    t = get("Temperature")
    if t:
        s.temperature = float(t[1])
    t = get("Pressure")
    if t:
        s.pressure_pa = float(t[1])
    for key, attr, conv in (("RandomSeed", "seed", int), ("NumberOfInitializationCycles", "n_init", int),
                            ("NumberOfEquilibrationCycles", "n_equil", int),
                            ("NumberOfProductionCycles", "n_prod", int),
                            ("MaxStepPerCycle", "max_step_per_cycle", int),
                            ("NumberOfTrialPositions", "n_trial_pos", int),
                            ("NumberOfTrialOrientations", "n_trial_orient", int),
                            ("AdsorbateAllocateSpace", "allocate_space", int)):
        t = get(key)
        if t:
            setattr(s, attr, conv(t[1]))
    for key, attr in (("NumberOfBlocks", "n_blocks"), ("PrintEvery", "print_every"),
                      ("MoviesEvery", "movies_every")):
        t = get(key)
        if t:
            setattr(s, attr, int(t[1]))
    t = get("SaveOutputToFile")
    s.save_output = bool(t and t[1].lower() == "yes")
    _reject_unsupported(get, comp_blocks)
    t = get("UseMaxStep")
    s.use_max_step = bool(t and t[1].lower() == "yes")
    t = get("CBMCBiasingMethod")
    if t and t[1] == "LJ_Biasing":
        s.vdw_real_bias = False
    for blk in comp_blocks:
        pos, types, charges, masses = read_molecule_def(os.path.join(dirpath, blk["MoleculeName"] + ".def"),
                                                        pnames, pcharge, pmass)
        c = Component(name=blk["MoleculeName"], molsize=len(types), template_pos=pos, types=types,
                      charges=charges, masses=masses)
        for key, attr in (("IdealGasRosenbluthWeight", "ideal_rosen"), ("FugacityCoefficient", "fugacity_coeff"),
                          ("MolFraction", "mol_fraction"), ("TranslationProbability", "p_translation"),
                          ("RotationProbability", "p_rotation"), ("WidomProbability", "p_widom"),
                          ("ReinsertionProbability", "p_reinsertion"), ("IdentityChangeProbability", "p_identity"),
                          ("SwapProbability", "p_swap")):
            if key in blk:
                setattr(c, attr, float(blk[key]))
        if "CreateNumberOfMolecules" in blk:
            c.create_molecules = int(blk["CreateNumberOfMolecules"])
        s.components.append(c)
    t = get("RestartFile")
    if t and t[1].lower() == "yes":
        read_restart(os.path.join(dirpath, "RestartInitial", "System_0", "restartfile"), s)
    return s


# --------------------------------------------------------------------------
# compact binary form of a System (numpy .npz, no pickles)
# --------------------------------------------------------------------------

_SCALARS = ("cut_vdw", "cut_coul", "overlap", "no_charges", "ewald_precision", "alpha", "recip_cut", "lammps_ewald",
            "prefactor", "fw_mass", "temperature", "pressure_pa", "seed", "n_init", "n_equil", "n_prod",
            "use_max_step", "max_step_per_cycle", "n_trial_pos", "n_trial_orient", "vdw_real_bias",
            "allocate_space", "n_blocks", "print_every", "movies_every", "save_output")
_COMP_SCALARS = ("molsize", "rigid", "ideal_rosen", "fugacity_coeff", "mol_fraction", "p_translation", "p_rotation",
                 "p_widom", "p_reinsertion", "p_identity", "p_swap", "create_molecules")


def save_system(path: str, s: System):
    d = {"cell": s.cell, "pseudo_names": np.array(s.pseudo_names), "pseudo_mass": s.pseudo_mass,
         "pseudo_charge": s.pseudo_charge, "eps": s.eps, "sig": s.sig, "shift": s.shift, "tail": s.tail,
         "use_tail": s.use_tail, "kmax": np.array(s.kmax), "unit_cells": np.array(s.unit_cells),
         "fw_pos": s.fw_pos, "fw_charge": s.fw_charge, "fw_type": s.fw_type, "n_comp": np.array(len(s.components))}
    for k in _SCALARS:
        d[k] = np.array(getattr(s, k))
    for i, c in enumerate(s.components):
        p = "c%d_" % i
        d[p + "name"] = np.array(c.name)
        for k in ("template_pos", "types", "charges", "masses"):
            d[p + k] = getattr(c, k)
        for k in _COMP_SCALARS:
            d[p + k] = np.array(getattr(c, k))
        for k in ("init_pos", "init_molid", "init_charge"):
            v = getattr(c, k)
            if v is not None:
                d[p + k] = v
    np.savez_compressed(path, **d)


def load_system(path: str) -> System:
    z = np.load(path, allow_pickle=False)
    s = System(cell=z["cell"].astype(np.float64), pseudo_names=[str(x) for x in z["pseudo_names"]],
               pseudo_mass=z["pseudo_mass"], pseudo_charge=z["pseudo_charge"], eps=z["eps"], sig=z["sig"],
               shift=z["shift"], tail=z["tail"], use_tail=z["use_tail"])
    for k in _SCALARS:
        if k in z.files:        # files written before a field existed keep its default
            setattr(s, k, type(getattr(s, k))(z[k].item()))
    s.kmax = tuple(int(v) for v in z["kmax"])
    s.unit_cells = tuple(int(v) for v in z["unit_cells"])
    s.fw_pos, s.fw_charge, s.fw_type = z["fw_pos"].reshape(-1, 3), z["fw_charge"], z["fw_type"].astype(np.int32)
    for i in range(int(z["n_comp"])):
        p = "c%d_" % i
        c = Component(name=str(z[p + "name"]), molsize=int(z[p + "molsize"]), template_pos=z[p + "template_pos"],
                      types=z[p + "types"].astype(np.int32), charges=z[p + "charges"], masses=z[p + "masses"])
        for k in _COMP_SCALARS:
            setattr(c, k, type(getattr(c, k))(z[p + k].item()))
        for k in ("init_pos", "init_molid", "init_charge"):
            if p + k in z.files:
                setattr(c, k, z[p + k])
        s.components.append(c)
    return s
