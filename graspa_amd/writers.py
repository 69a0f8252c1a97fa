"""Restart, movie and per-component data files of a box (the output side of
the deck reader).

Formats follow the reference's writers (src_clean/write_data.h) so that its
reader and other tools read them: the RASPA-2 style restart file
(create_Restart_file, write_data.h:245-263: cell info, maximum MC-move
changes, per-component maximum translation/rotation, then per-adsorbate atom
positions, velocities, forces, charges, scaling and fixed flags), the LAMMPS
data file of the movie (create_movie_file, write_data.h:210-227: box, masses,
pair coefficients, atoms with each molecule wrapped into the box about its
first atom) and the AllData component dump (write_data.h:265-290).
Positions are written with 15 significant digits as the reference does, so a
restart round trip keeps them to ~1e-15 relative.

`molecules` arguments are lists over adsorbate species of (pos (n*M, 3),
charge (n*M,), type (n*M,), molid (n*M,)) as Box.get_molecules returns them;
MolIDs equal the molecule slot (the engine keeps a deleted slot's MolIDs for
the molecule moved into it, as the reference does).
"""
from __future__ import annotations

import os

import numpy as np


def _g(x, digits=6):
    """C++ ostream default formatting of a double (precision `digits`, %g)."""
    return "%.*g" % (digits, float(x))


def write_restart(path, system, molecules, max_translation=None, max_rotation=None):
    """RASPA-2 restart file of the box (create_Restart_file,
    write_data.h:155-263); directories are created as needed."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    c = system.cell
    out = []
    out.append("Cell info:")
    out.append("=" * 72)
    out.append("number-of-unit-cells: 1 1 1")
    for tag in ("unit-cell-vector", "cell-vector"):
        out.append("%s-a: %s %s %s" % (tag, _g(c[0]), _g(c[1]), _g(c[2])))
        out.append("%s-b: %s %s %s" % (tag, _g(c[3]), _g(c[4]), _g(c[5])))
        out.append("%s-c: %s %s %s" % (tag, _g(c[6]), _g(c[7]), _g(c[8])))
        out.append("")
    out.append("cell-lengths: %s %s %s " % (_g(c[0]), _g(c[4]), _g(c[8])))
    out.append("cell-angles: 90 90 90 ")
    out.append("")
    out.append("")
    out.append("Maximum changes for MC-moves:")
    out.append("=" * 72)
    out.append("Maximum-volume-change: 0.006250")
    out.append("Maximum-Gibbs-volume-change: 0.025000")
    out.append("Maximum-box-shape-change: 0.100000 0.100000 0.100000, 0.100000 0.100000 0.100000, "
               "0.100000 0.100000 0.100000")
    out.append("")
    out.append("")
    out.append("Acceptance targets for MC-moves:")
    out.append("=" * 72)
    out.append("Target-volume-change: 0.500000")
    out.append("Target-box-shape-change: 0.500000")
    out.append("Target-Gibbs-volume-change: 0.500000")
    out.append("")
    out.append("")
    n_tot = sum(len(m[3]) // max(system.components[i].molsize, 1) for i, m in enumerate(molecules))
    out.append("Components: %d (Adsorbates %d, Cations 0)" % (len(system.components), n_tot))
    out.append("=" * 72)
    for i, comp in enumerate(system.components):
        out.append("Components 0 (%s) " % comp.name)
        out.append("")
        mt = max_translation[i] if max_translation is not None else (0.0, 0.0, 0.0)
        mr = max_rotation[i] if max_rotation is not None else (0.0, 0.0, 0.0)
        out.append("Maximum-translation-change component %d: %s %s %s" % (i, _g(mt[0]), _g(mt[1]), _g(mt[2])))
        out.append("Maximum-translation-in-plane-change component %d: 0.000000,0.000000,0.000000" % i)
        out.append("Maximum-rotation-change component %d: %s %s %s" % (i, _g(mr[0]), _g(mr[1]), _g(mr[2])))
        out.append("")
    out.append("Reactions: 0")
    prev = 0
    for i, comp in enumerate(system.components):
        pos, q, _t, molid = molecules[i]
        M = comp.molsize
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        molid = np.asarray(molid, dtype=np.int64)
        n = len(molid) // M
        out.append("")
        out.append("Component: %d   Adsorbate %d molecules of %s" % (i, n, comp.name))
        out.append("-" * 72)
        idx = [(prev + int(molid[j]), j - int(molid[j]) * M) for j in range(len(molid))]
        for j, (m, a) in enumerate(idx):
            out.append("Adsorbate-atom-position: %d %d %s  %s  %s" % (m, a, _g(pos[j, 0], 15), _g(pos[j, 1], 15),
                                                                    _g(pos[j, 2], 15)))
        for m, a in idx:
            out.append("Adsorbate-atom-velocity: %d %d 0  0  0" % (m, a))
        for m, a in idx:
            out.append("Adsorbate-atom-force: %d %d 0  0  0" % (m, a))
        for j, (m, a) in enumerate(idx):
            out.append("Adsorbate-atom-charge: %d %d %s" % (m, a, _g(q[j])))
        for m, a in idx:
            out.append("Adsorbate-atom-scaling: %d %d 1" % (m, a))
        for m, a in idx:
            out.append("Adsorbate-atom-fixed: %d %d 0 0 0" % (m, a))
        prev += n
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def _wrap(p, cell, inv):
    """WrapInBox: position into [0, 1) cell fractions."""
    s = inv.reshape(3, 3).T @ p            # inv stored column-wise like the cell
    s -= np.floor(s)
    return cell.reshape(3, 3).T @ s


def write_lammps_movie(path, system, molecules):
    """LAMMPS data file of the box (WriteBox_LAMMPS / WriteAtoms_LAMMPS,
    write_data.h:9-117): framework atoms as molecule 1, each adsorbate
    molecule wrapped into the box about its first atom."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    c = np.asarray(system.cell, dtype=np.float64)
    inv = np.asarray(system.inv_cell, dtype=np.float64)
    T = system.ntypes
    n_atoms = len(system.fw_type) + sum(len(m[3]) for m in molecules)
    out = ["# LAMMPS data file of a graspa_amd box", "# Total atoms: %d" % n_atoms, "%d atoms" % n_atoms,
           "0 bonds", "0 angles", "0 dihedrals", "0 impropers",
           "0 %s xlo xhi" % _g(c[0]), "0 %s ylo yhi" % _g(c[4]), "0 %s zlo zhi" % _g(c[8]),
           "%s %s %s xy xz yz" % (_g(c[3]), _g(c[6]), _g(c[7])), "",
           "%d atom types" % T, "0 bond types", "0 angle types", "0 dihedral types", "0 improper types", "",
           "Masses", ""]
    for t in range(T):
        out.append("%d %s # %s" % (t + 1, _g(system.pseudo_mass[t]), system.pseudo_names[t]))
    out += ["", "Pair Coeffs", ""]
    eps, sig = np.asarray(system.eps), np.asarray(system.sig)
    for t in range(T):
        out.append("%d %s %s # %s" % (t + 1, _g(eps[t, t] * 0.00239006), _g(sig[t, t]), system.pseudo_names[t]))
    out += ["", "Atoms", ""]
    k = 0
    fw = np.asarray(system.fw_pos, dtype=np.float64).reshape(-1, 3)
    for j in range(len(system.fw_type)):
        p = _wrap(fw[j], c, inv)
        t = int(system.fw_type[j])
        k += 1
        out.append("%d 1 %d %s %s %s %s # framework %s" % (k, t + 1, _g(system.fw_charge[j], 10), _g(p[0], 10),
                                                          _g(p[1], 10), _g(p[2], 10), system.pseudo_names[t]))
    mol_base = 1
    for i, comp in enumerate(system.components):
        pos, q, ty, molid = molecules[i]
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        M = comp.molsize
        shift = np.zeros(3)
        for j in range(len(molid)):
            if j % M == 0:
                w = _wrap(pos[j], c, inv)
                shift = w - pos[j]
            p = pos[j] + shift
            t = int(ty[j])
            k += 1
            out.append("%d %d %d %s %s %s %s # %s %s" % (k, mol_base + int(molid[j]) + 1, t + 1, _g(q[j], 10),
                                                        _g(p[0], 10), _g(p[1], 10), _g(p[2], 10), comp.name,
                                                        system.pseudo_names[t]))
        mol_base += len(molid) // max(M, 1)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_all_data(dirpath, system, molecules):
    """AllData/Component_<i>.data dumps (WriteAllData, write_data.h:265-290):
    MolID, atom index, position, charge, scaling and type per atom."""
    os.makedirs(dirpath, exist_ok=True)
    for i, comp in enumerate(system.components):
        pos, q, ty, molid = molecules[i]
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        M = comp.molsize
        lines = ["x y z charge scale scaleCoul Type"]
        for j in range(len(molid)):
            m = int(molid[j])
            lines.append(" %d %d %s  %s  %s %s 1 1 %d" % (m, j - m * M, _g(pos[j, 0]), _g(pos[j, 1]), _g(pos[j, 2]),
                                                        _g(q[j]), int(ty[j])))
        with open(os.path.join(dirpath, "Component_%d.data" % (i + 1)), "w") as f:
            f.write("\n".join(lines) + "\n")


def box_molecules(box):
    """The adsorbate molecules of an engine Box, in the writers' layout."""
    return [box.get_molecules(ci) for ci in range(1, len(box.system.components) + 1)]
