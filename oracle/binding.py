"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (graspa_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class OrBox(C.Structure):
    _fields_ = [("cell", C.c_double * 9), ("inv_cell", C.c_double * 9), ("cubic", C.c_int),
                ("volume", C.c_double), ("ntypes", C.c_int), ("eps", _f64p), ("sig", _f64p),
                ("shift", _f64p), ("tail", _f64p), ("use_tail", _i32p), ("cut_vdw2", C.c_double),
                ("cut_coul2", C.c_double), ("overlap", C.c_double), ("no_charges", C.c_int),
                ("alpha", C.c_double), ("prefactor", C.c_double), ("kmax", C.c_int * 3),
                ("recip_cut", C.c_double), ("lammps_ewald", C.c_int)]


class OrAtoms(C.Structure):
    _fields_ = [("molsize", C.c_int), ("n_mol", C.c_int), ("capacity", C.c_int), ("pos", _f64p),
                ("charge", _f64p), ("type", _i32p), ("molid", _i64p)]


class OrEnergy(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("hh_vdw", "hg_vdw", "gg_vdw", "hh_real", "hg_real", "gg_real",
                                          "hh_ewald", "hg_ewald", "gg_ewald", "tail")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}

    def total(self):
        return sum(getattr(self, n) for n, _ in self._fields_)


class OrTrace(C.Structure):
    _fields_ = [("cycle", C.c_int), ("step", C.c_int), ("move", C.c_int), ("comp", C.c_int), ("mol", C.c_int),
                ("accepted", C.c_int), ("overlap", C.c_int), ("_pad", C.c_int), ("pacc", C.c_double),
                ("rosen", C.c_double), ("de", OrEnergy)]


class OrMC(C.Structure):
    _fields_ = [("p_translation", C.c_double * 8), ("p_rotation", C.c_double * 8), ("p_widom", C.c_double * 8),
                ("p_swap", C.c_double * 8), ("mol_fraction", C.c_double * 8), ("fugacity_coeff", C.c_double * 8),
                ("ideal_rosen", C.c_double * 8), ("excl_intra", C.c_double * 8), ("excl_self", C.c_double * 8),
                ("max_translation", (C.c_double * 3) * 8), ("max_rotation", (C.c_double * 3) * 8),
                ("pseudo_per_mol", (C.c_int64 * 32) * 8), ("has_charge", C.c_int * 8),
                ("n_trial_pos", C.c_int), ("n_trial_orient", C.c_int), ("vdw_real_bias", C.c_int),
                ("use_max_step", C.c_int), ("max_step_per_cycle", C.c_int), ("has_tail", C.c_int),
                ("temperature", C.c_double), ("pressure_pa", C.c_double), ("beta", C.c_double),
                ("pressure", C.c_double), ("seed", C.c_uint), ("pool_size", C.c_int), ("production", C.c_int),
                ("n_pseudo", C.c_int64 * 32), ("total_molecules", C.c_int64),
                ("trans_total", C.c_int64 * 8), ("trans_acc", C.c_int64 * 8), ("rot_total", C.c_int64 * 8),
                ("rot_acc", C.c_int64 * 8), ("ins_total", C.c_int64 * 8), ("ins_acc", C.c_int64 * 8),
                ("del_total", C.c_int64 * 8), ("del_acc", C.c_int64 * 8), ("widom_total", C.c_int64 * 8),
                ("widom_rosen_sum", C.c_double * 8), ("widom_rosen_n", C.c_double * 8), ("running", OrEnergy),
                ("p_reinsertion", C.c_double * 8), ("reins_total", C.c_int64 * 8), ("reins_acc", C.c_int64 * 8),
                ("p_identity", C.c_double * 8), ("idsw_total", (C.c_int64 * 8) * 8),
                ("idsw_acc", (C.c_int64 * 8) * 8),
                ("flexible", C.c_int * 8), ("bond", (C.c_double * 8) * 8), ("bend_theta0", (C.c_double * 8) * 8),
                ("bend_k", (C.c_double * 8) * 8), ("torsion", ((C.c_double * 4) * 8) * 8)]


class OrWidomOut(C.Structure):
    _fields_ = [("rosen", C.c_double), ("energy_total", C.c_double), ("vdw", C.c_double), ("real", C.c_double),
                ("ewald", C.c_double), ("first_bead", C.c_int), ("orientation", C.c_int), ("ok", C.c_int),
                ("_pad", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_box_finalize.argtypes = [C.POINTER(OrBox)]
        L.or_pbc.argtypes = [_f64p, _f64p, _f64p, C.c_int]
        L.or_vdw.argtypes = [C.POINTER(OrBox), C.c_int, C.c_int, C.c_double, C.c_double, _f64p]
        L.or_coulomb_real.argtypes = [C.POINTER(OrBox), C.c_double, C.c_double, C.c_double, C.c_double]
        L.or_coulomb_real.restype = C.c_double
        L.or_nvec.argtypes = [C.POINTER(OrBox)]
        L.or_total_vdw_real.argtypes = [C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, C.POINTER(OrEnergy)]
        L.or_ewald_total.argtypes = [C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, C.POINTER(OrEnergy), _f64p, _f64p,
                                     _f64p, _f64p, _i32p]
        L.or_exclusion_rigid.argtypes = [C.POINTER(OrBox), C.POINTER(OrAtoms), _f64p, _f64p]
        L.or_tail_total.argtypes = [C.POINTER(OrBox), _i64p]
        L.or_tail_total.restype = C.c_double
        L.or_tail_delta.argtypes = [C.POINTER(OrBox), _i64p, _i64p, C.c_int]
        L.or_tail_delta.restype = C.c_double
        L.or_site_energies.argtypes = [C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, C.c_int, _f64p, _i32p, _f64p,
                                       C.c_int, C.c_int64, C.c_int, C.c_int64, _f64p, _u8p]
        L.or_ewald_delta.argtypes = [C.POINTER(OrBox), _f64p, _f64p, C.c_int, _f64p, _f64p, C.c_int, _f64p, _f64p,
                                     _f64p, _f64p, _f64p]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = C.c_int
        L.or_rotate_quaternion.argtypes = [_f64p, C.c_double, C.c_double, C.c_double]
        L.or_flex_trial.argtypes = [C.POINTER(OrMC), C.c_int, C.c_int, _f64p, C.c_uint64, C.c_uint64, C.c_uint64,
                                    C.c_int, _f64p]
        L.or_rand_uniform.restype = C.c_double
        L.or_srand.argtypes = [C.c_uint]
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_philox_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32]
        L.or_philox_uniform.restype = C.c_double
        L.or_mc_setup.argtypes = [C.POINTER(OrMC)]
        L.or_mc_cycles.argtypes = [C.POINTER(OrMC), C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, _f64p, _f64p,
                                   C.c_int, C.c_int, C.POINTER(OrTrace), C.c_int]
        L.or_batch_cycles.argtypes = [C.POINTER(OrMC), C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, _f64p, _f64p,
                                      C.c_uint64, C.c_uint64, _i64p, C.c_int, C.c_int, C.POINTER(OrTrace), C.c_int]
        L.or_mc_create_molecules.argtypes = [C.POINTER(OrMC), C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, _f64p,
                                             _f64p, C.c_int, C.c_int, C.c_int64, C.POINTER(OrEnergy)]
        L.or_widom_batch.argtypes = [C.POINTER(OrBox), C.POINTER(OrAtoms), C.c_int, C.c_int, _f64p, _f64p, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                     C.c_uint64, C.c_int64, C.c_int, C.POINTER(OrWidomOut)]
        _lib = L
    return _lib


def _p(a, t=_f64p):
    return a.ctypes.data_as(t)


class OracleSystem:
    """Oracle view of a :class:`graspa_amd.deck.System` (holds the arrays)."""

    def __init__(self, system, capacity=None):
        self.system = system
        s = system
        T = s.ntypes
        self._eps = np.ascontiguousarray(s.eps, dtype=np.float64).reshape(T * T)
        self._sig = np.ascontiguousarray(s.sig, dtype=np.float64).reshape(T * T)
        self._shift = np.ascontiguousarray(s.shift, dtype=np.float64).reshape(T * T)
        self._tail = np.ascontiguousarray(s.tail, dtype=np.float64).reshape(T * T)
        self._use_tail = np.ascontiguousarray(s.use_tail, dtype=np.int32).reshape(T * T)
        b = OrBox()
        for i in range(9):
            b.cell[i] = float(s.cell[i])
        b.ntypes = T
        b.eps, b.sig, b.shift, b.tail = _p(self._eps), _p(self._sig), _p(self._shift), _p(self._tail)
        b.use_tail = _p(self._use_tail, _i32p)
        b.cut_vdw2 = s.cut_vdw * s.cut_vdw
        b.cut_coul2 = s.cut_coul * s.cut_coul
        b.overlap = s.overlap
        b.no_charges = int(s.no_charges)
        b.alpha = s.alpha
        b.prefactor = s.prefactor
        for i in range(3):
            b.kmax[i] = int(s.kmax[i])
        b.recip_cut = s.recip_cut
        b.lammps_ewald = int(s.lammps_ewald)
        lib().or_box_finalize(C.byref(b))
        self.box = b
        # component 0: framework
        self._arrays = []
        comps = (OrAtoms * (1 + len(s.components)))()
        nfw = len(s.fw_type)
        self._set_atoms(comps[0], nfw if nfw else 1, 1 if nfw else 0, 1, s.fw_pos, s.fw_charge, s.fw_type,
                        np.zeros(nfw, dtype=np.int64))
        for ci, c in enumerate(s.components):
            cap = capacity if capacity is not None else max(s.allocate_space // max(c.molsize, 1), 1)
            if c.init_pos is not None and len(c.init_pos):
                nmol = len(c.init_pos) // c.molsize
                pos, q = c.init_pos, (c.init_charge if c.init_charge is not None else np.tile(c.charges, nmol))
                types, molid = np.tile(c.types, nmol), c.init_molid
            else:
                nmol = 0
                pos, q, types, molid = c.template_pos, c.charges, c.types, np.zeros(c.molsize, dtype=np.int64)
            self._set_atoms(comps[1 + ci], c.molsize, nmol, max(cap, nmol, 1), pos, q, types, molid)
        self.comps = comps
        self.ncomp = 1 + len(s.components)

    def _set_atoms(self, a, molsize, nmol, cap, pos, q, types, molid):
        n_cap = cap * molsize
        P = np.zeros((n_cap, 3))
        Q = np.zeros(n_cap)
        Ty = np.zeros(n_cap, dtype=np.int32)
        Id = np.zeros(n_cap, dtype=np.int64)
        n = len(pos)
        P[:n] = pos
        Q[:n] = q
        Ty[:n] = types
        Id[:n] = molid
        self._arrays += [P, Q, Ty, Id]
        a.molsize, a.n_mol, a.capacity = molsize, nmol, cap
        a.pos, a.charge, a.type, a.molid = _p(P), _p(Q), _p(Ty, _i32p), _p(Id, _i64p)

    def comp_arrays(self, ci):
        """numpy views (pos, charge, type, molid) of component ci (0 = framework)."""
        a = self.comps[ci]
        n = a.capacity * a.molsize
        pos = np.ctypeslib.as_array(a.pos, shape=(n * 3,)).reshape(n, 3)
        q = np.ctypeslib.as_array(a.charge, shape=(n,))
        t = np.ctypeslib.as_array(a.type, shape=(n,))
        m = np.ctypeslib.as_array(a.molid, shape=(n,))
        return pos, q, t, m

    # ---- totals ----
    def total_vdw_real(self):
        e = OrEnergy()
        lib().or_total_vdw_real(C.byref(self.box), self.comps, self.ncomp, C.byref(e))
        return e

    def ewald_total(self):
        nv = lib().or_nvec(C.byref(self.box))
        sk_fw = np.zeros(2 * nv)
        sk_ads = np.zeros(2 * nv)
        self_e = np.zeros(self.ncomp)
        intra = np.zeros(self.ncomp)
        nact = C.c_int(0)
        e = OrEnergy()
        lib().or_ewald_total(C.byref(self.box), self.comps, self.ncomp, C.byref(e), _p(sk_fw), _p(sk_ads),
                             _p(self_e), _p(intra), C.byref(nact))
        return e, sk_fw, sk_ads, self_e, intra, nact.value

    def exclusion(self, ci):
        i, s = C.c_double(0), C.c_double(0)
        lib().or_exclusion_rigid(C.byref(self.box), C.byref(self.comps[ci]), C.byref(i), C.byref(s))
        return i.value, s.value

    def site_energies(self, pos, types, charges, self_comp=-1, self_molid=-1, excl_comp=-1, excl_molid=-1):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        t = np.ascontiguousarray(types, dtype=np.int32)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        out = np.zeros((n, 4))
        ov = np.zeros(n, dtype=np.uint8)
        lib().or_site_energies(C.byref(self.box), self.comps, self.ncomp, n, _p(pos), _p(t, _i32p), _p(q), self_comp,
                               self_molid, excl_comp, excl_molid, _p(out), _p(ov, _u8p))
        return out, ov

    def ewald_delta(self, sk_same, sk_cross, old_pos, old_q, new_pos, new_q):
        op = np.ascontiguousarray(old_pos, dtype=np.float64).reshape(-1, 3)
        npos = np.ascontiguousarray(new_pos, dtype=np.float64).reshape(-1, 3)
        oq = np.ascontiguousarray(old_q, dtype=np.float64)
        nq = np.ascontiguousarray(new_q, dtype=np.float64)
        same, cross = C.c_double(0), C.c_double(0)
        sk_new = np.zeros_like(sk_same)
        lib().or_ewald_delta(C.byref(self.box), _p(sk_same), _p(sk_cross), len(op), _p(op), _p(oq), len(npos), _p(npos),
                             _p(nq), C.byref(same), C.byref(cross), _p(sk_new))
        return same.value, cross.value, sk_new

    def widom_batch(self, comp, sk_ads, sk_fw, seed, n_insert, first_index=0, stream=0, tail_de=0.0):
        s = self.system
        c = s.components[comp - 1]
        intra, selfe = self.exclusion(comp) if (c.has_charge and not s.no_charges) else (0.0, 0.0)
        out = (OrWidomOut * n_insert)()
        rc = lib().or_widom_batch(C.byref(self.box), self.comps, self.ncomp, comp, _p(sk_ads), _p(sk_fw), s.beta,
                                  s.n_trial_pos, s.n_trial_orient, int(s.vdw_real_bias), intra, selfe, tail_de,
                                  seed, stream, first_index, n_insert, out)
        if rc != 0:
            raise RuntimeError("or_widom_batch failed")
        return out


def widom_records_to_numpy(out):
    n = len(out)
    arr = np.zeros(n, dtype=[("rosen", "f8"), ("energy", "f8"), ("vdw", "f8"), ("real", "f8"), ("ewald", "f8"),
                             ("first_bead", "i4"), ("orientation", "i4"), ("ok", "i4")])
    for i, o in enumerate(out):
        arr[i] = (o.rosen, o.energy_total, o.vdw, o.real, o.ewald, o.first_bead, o.orientation, o.ok)
    return arr


def make_mc(system, oracle_sys, production=1, pool_size=333334):
    """An or_mc set up from a System, as Initialize() would (main.cpp:58-343)."""
    s = system
    mc = OrMC()
    for ci, c in enumerate(s.components):
        k = ci + 1
        cp = c.cumulative_probabilities()
        mc.p_translation[k] = cp["translation"]
        mc.p_rotation[k] = cp["rotation"]
        mc.p_widom[k] = cp["widom"]
        mc.p_reinsertion[k] = cp["reinsertion"]
        mc.p_identity[k] = cp["identity"]
        mc.p_swap[k] = cp["swap"]
        mc.mol_fraction[k] = c.mol_fraction
        mc.fugacity_coeff[k] = c.fugacity_coeff
        mc.ideal_rosen[k] = c.ideal_rosen
        if c.has_charge and not s.no_charges:
            mc.excl_intra[k], mc.excl_self[k] = oracle_sys.exclusion(k)
        for d in range(3):
            mc.max_translation[k][d] = s.cell[4 * d] * 0.1
            mc.max_rotation[k][d] = 30.0 / (180 / 3.1415)
        for t in c.types:
            mc.pseudo_per_mol[k][int(t)] += 1
        mc.has_charge[k] = int(c.has_charge)
        ch = getattr(c, "chain", None)
        if ch is not None:
            mc.flexible[k] = 1
            for i in range(c.molsize):
                mc.bond[k][i] = float(ch.bond[i])
                mc.bend_theta0[k][i] = float(ch.theta0[i])
                mc.bend_k[k][i] = float(ch.k_bend[i])
                for t in range(4):
                    mc.torsion[k][i][t] = float(ch.torsion[i][t])
    mc.n_trial_pos = s.n_trial_pos
    mc.n_trial_orient = s.n_trial_orient
    mc.vdw_real_bias = int(s.vdw_real_bias)
    mc.use_max_step = int(s.use_max_step)
    mc.max_step_per_cycle = s.max_step_per_cycle
    mc.has_tail = int(s.has_tail)
    mc.temperature = s.temperature
    mc.pressure_pa = s.pressure_pa
    mc.beta = s.beta
    mc.pressure = s.pressure
    mc.seed = s.seed
    mc.pool_size = pool_size
    mc.production = production
    nm = [oracle_sys.comps[1 + i].n_mol for i in range(len(s.components))]
    n = s.pseudo_counts(nm)
    for t in range(min(32, len(n))):
        mc.n_pseudo[t] = int(n[t])
    mc.total_molecules = 1 + sum(nm)
    return mc
