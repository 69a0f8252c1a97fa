"""Python view of the engine's C-ABI (include/graspa_engine.h) over ctypes.

:class:`Box` mirrors the per-move energy calls of gRASPA's Monte Carlo driver
(the boundary the engine replaces); every method is a thin call into
``graspa_amd/libgraspa_engine.so``.  There is no CPU fallback: if the library
is missing or no HIP device is present, the calls fail with an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import deck as _deck

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgraspa_engine.so")

GRASPA_FP64 = 0
GRASPA_MIXED = 1
MOVE_NAMES = {0: "translation", 1: "rotation", 2: "insertion", 3: "deletion", 4: "widom", 5: "none"}

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


class EngineError(RuntimeError):
    pass


class BoxParams(C.Structure):
    _fields_ = [("cell", C.c_double * 9), ("ntypes", C.c_int), ("epsilon", _f64p), ("sigma", _f64p), ("shift", _f64p),
                ("cutoff_vdw", C.c_double), ("cutoff_coulomb", C.c_double), ("overlap_criteria", C.c_double),
                ("use_charges", C.c_int), ("alpha", C.c_double), ("kmax", C.c_int * 3),
                ("reciprocal_cutoff", C.c_double), ("lammps_ewald", C.c_int), ("prefactor", C.c_double),
                ("precision", C.c_int), ("device", C.c_int)]


class Energy(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("hh_vdw", "hg_vdw", "gg_vdw", "hh_real", "hg_real", "gg_real",
                                          "hh_ewald", "hg_ewald", "gg_ewald", "tail")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}

    def total(self):
        return sum(getattr(self, n) for n, _ in self._fields_)


class WidomResult(C.Structure):
    _fields_ = [("rosen", C.c_double), ("energy", C.c_double), ("vdw", C.c_double), ("real", C.c_double),
                ("ewald", C.c_double), ("first_bead", C.c_int), ("orientation", C.c_int), ("ok", C.c_int),
                ("_pad", C.c_int)]


class WidomSummary(C.Structure):
    _fields_ = [("sum_w", C.c_double), ("sum_w2", C.c_double), ("sum_we", C.c_double), ("n", C.c_double),
                ("n_first_bead", C.c_double), ("n_ok", C.c_double), ("screened_pairs", C.c_double),
                ("evaluated_pairs", C.c_double)]


class EvalShape(C.Structure):
    """Launch shape of the Widom pair-energy kernel (graspa_eval_shape)."""
    _fields_ = [("waves", C.c_int), ("per_cu", C.c_int), ("framework_lds", C.c_int)]

    def as_dict(self):
        return {"waves": self.waves, "per_cu": self.per_cu, "framework_lds": self.framework_lds}


class ChainModel(C.Structure):
    """graspa_chain: bonded model of a flexible linear chain."""
    _fields_ = [("bond", C.c_double * 4), ("theta0", C.c_double * 4), ("k_bend", C.c_double * 4),
                ("torsion", (C.c_double * 4) * 4)]


class Species(C.Structure):
    _fields_ = [("p_translation", C.c_double), ("p_rotation", C.c_double), ("p_widom", C.c_double),
                ("p_swap", C.c_double), ("mol_fraction", C.c_double), ("fugacity_coeff", C.c_double),
                ("ideal_rosen", C.c_double), ("excl_intra", C.c_double), ("excl_self", C.c_double),
                ("max_translation", C.c_double * 3), ("max_rotation", C.c_double * 3),
                ("pseudo_per_mol", C.c_int64 * 32), ("p_reinsertion", C.c_double), ("p_identity", C.c_double)]


class MCParams(C.Structure):
    _fields_ = [("n_species", C.c_int), ("species", Species * 8), ("n_trial_pos", C.c_int),
                ("n_trial_orient", C.c_int), ("vdw_real_bias", C.c_int), ("use_max_step", C.c_int),
                ("max_step_per_cycle", C.c_int), ("has_tail", C.c_int), ("tail", _f64p), ("use_tail", _i32p),
                ("beta", C.c_double), ("pressure", C.c_double), ("seed", C.c_uint), ("pool_size", C.c_int),
                ("n_pseudo", C.c_int64 * 32)]


class BatchBox(C.Structure):
    _fields_ = [("box", C.c_void_p), ("params", C.POINTER(MCParams)), ("seed", C.c_uint64), ("stream", C.c_uint32),
                ("n_block", C.c_int)]


class Trace(C.Structure):
    _fields_ = [("cycle", C.c_int), ("step", C.c_int), ("move", C.c_int), ("comp", C.c_int), ("mol", C.c_int),
                ("accepted", C.c_int), ("overlap", C.c_int), ("_pad", C.c_int), ("pacc", C.c_double),
                ("rosen", C.c_double), ("de", Energy)]


NBLOCK, MAX_COMPS = 16, 8
N_COUNTERS = 25          # move counters per species (GRASPA_N_COUNTERS)
CNT_IDSW_TOT = 11        # identity swaps to species s attempted at [CNT_IDSW_TOT + s], accepted at
CNT_IDSW_ACC = 18        # [CNT_IDSW_ACC + s] (GRASPA_CNT_IDSW_TOT / _ACC)
MOVE_REINSERTION = 6
MOVE_IDENTITY_SWAP = 7


class BlockStats(C.Structure):
    """graspa_block_stats: per-block sums of the current production run."""
    _fields_ = [("n_block", C.c_int64), ("block_size", C.c_int64), ("cycles", C.c_int64 * NBLOCK),
                ("n", (C.c_double * MAX_COMPS) * NBLOCK), ("n2", (C.c_double * MAX_COMPS) * NBLOCK),
                ("nn", ((C.c_double * MAX_COMPS) * MAX_COMPS) * NBLOCK), ("en", (C.c_double * MAX_COMPS) * NBLOCK),
                ("e", (C.c_double * 10) * NBLOCK), ("widom_w", (C.c_double * MAX_COMPS) * NBLOCK),
                ("widom_n", (C.c_double * MAX_COMPS) * NBLOCK)]

    def arrays(self):
        """numpy views: dict of arrays indexed [block][component ...]."""
        nb = int(self.n_block)
        out = {"n_block": nb, "block_size": int(self.block_size),
               "cycles": np.ctypeslib.as_array(self.cycles).copy()[:nb]}
        for k in ("n", "n2", "nn", "en", "e", "widom_w", "widom_n"):
            out[k] = np.ctypeslib.as_array(getattr(self, k)).copy()[:nb]
        return out


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "graspa_last_error": (C.c_char_p, []),
    "graspa_version": (C.c_char_p, []),
    "graspa_device_count": (C.c_int, [_i32p]),
    "graspa_box_create": (C.c_int, [C.POINTER(BoxParams), C.POINTER(C.c_void_p)]),
    "graspa_box_destroy": (None, [C.c_void_p]),
    "graspa_box_set_framework": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, _i32p]),
    "graspa_box_add_component": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, _i32p, _f64p, _i32p]),
    "graspa_box_set_molecules": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, _f64p, _i32p, _i64p]),
    "graspa_box_get_molecules": (C.c_int, [C.c_void_p, C.c_int, _i32p, _f64p, _f64p, _i32p, _i64p]),
    "graspa_box_init_ewald": (C.c_int, [C.c_void_p, C.POINTER(Energy)]),
    "graspa_box_get_structure_factors": (C.c_int, [C.c_void_p, _f64p, _f64p]),
    "graspa_box_set_structure_factors": (C.c_int, [C.c_void_p, _f64p, _f64p]),
    "graspa_box_nvec": (C.c_int, [C.c_void_p, _i32p, _i32p]),
    "graspa_site_energies": (C.c_int, [C.c_void_p, C.c_int, _f64p, _i32p, _f64p, C.c_int, C.c_int64, C.c_int,
                                       C.c_int64, _f64p, _u8p]),
    "graspa_trial_energies": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, _i32p, _f64p, C.c_int, C.c_int64,
                                        C.c_int, C.c_int64, _f64p, _u8p]),
    "graspa_single_body_delta": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int, _f64p, _f64p, _i32p, _f64p,
                                           C.POINTER(Energy), _i32p]),
    "graspa_ewald_delta": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, C.c_int, _f64p, _f64p, _f64p, _f64p]),
    "graspa_ewald_accept": (C.c_int, [C.c_void_p]),
    "graspa_commit_move": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p]),
    "graspa_commit_insert": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p, _i32p, C.c_int64]),
    "graspa_commit_delete": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "graspa_total_energy": (C.c_int, [C.c_void_p, C.POINTER(Energy)]),
    "graspa_widom_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                     C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_int,
                                     C.POINTER(WidomResult), C.POINTER(WidomSummary)]),
    "graspa_widom_enqueue": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_int]),
    "graspa_widom_enqueue_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_int]),
    "graspa_widom_totals": (C.c_int, [C.c_void_p, C.POINTER(WidomSummary), C.c_int]),
    "graspa_widom_graph_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                            C.c_double, C.c_uint64, C.c_uint32, C.c_int64, C.c_int]),
    "graspa_widom_graph_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "graspa_widom_set_shape": (C.c_int, [C.c_void_p, C.POINTER(EvalShape)]),
    "graspa_widom_get_shape": (C.c_int, [C.c_void_p, C.POINTER(EvalShape)]),
    "graspa_widom_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                        C.c_double, C.c_uint64, C.c_int, C.c_int, C.POINTER(EvalShape), _f64p,
                                        C.c_int, _i32p]),
    "graspa_mc_create": (C.c_int, [C.c_void_p, C.POINTER(MCParams), C.POINTER(C.c_void_p)]),
    "graspa_mc_destroy": (None, [C.c_void_p]),
    "graspa_mc_cycles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(Trace), C.c_int, _i64p]),
    "graspa_mc_create_molecules": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int64, _i32p, C.POINTER(Energy)]),
    "graspa_mc_stats": (C.c_int, [C.c_void_p, C.POINTER(Energy), _i32p, _i64p, _f64p, _f64p]),
    "graspa_mc_cycle_times": (C.c_int, [C.c_void_p, _f64p, C.c_int]),
    "graspa_timing_enable": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "graspa_timing_read": (C.c_int, [C.c_void_p, _f64p, _i64p]),
    "graspa_box_sync": (C.c_int, [C.c_void_p]),
    "graspa_glibc_rand_sequence": (C.c_int, [C.c_uint, C.c_int, _i32p]),
    "graspa_coulomb_table": (C.c_int, [C.c_double, C.c_double, _f64p, C.c_int]),
    "graspa_free_map": (C.c_int, [C.POINTER(BoxParams), C.c_int, _f64p, _i32p, C.c_int, C.POINTER(C.c_uint32),
                        C.c_int, _i32p]),
    "graspa_dead_map": (C.c_int, [C.POINTER(BoxParams), C.c_int, _f64p, _i32p, C.c_int, C.POINTER(C.c_uint32),
                                  C.c_int, _i32p]),
    "graspa_batch_create": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "graspa_batch_destroy": (None, [C.c_void_p]),
    "graspa_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, _i64p]),
    "graspa_batch_box_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Energy), _i32p, _i64p, _f64p, _f64p, _f64p,
                                         _f64p, _i64p, _i64p, _i64p]),
    "graspa_batch_cycle_times": (C.c_int, [C.c_void_p, C.c_int, _f64p, C.c_int]),
    "graspa_batch_trace": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "graspa_batch_kernel_time": (C.c_int, [C.c_void_p, _f64p, _i64p]),
    "graspa_batch_sync": (C.c_int, [C.c_void_p]),
    "graspa_batch_max_steps": (C.c_int, [C.c_void_p, C.c_int, _f64p, _f64p]),
    "graspa_box_set_chain": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "graspa_batch_block_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BlockStats)]),
    "graspa_batch_set_box": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "graspa_batch_set_cycles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "graspa_batch_launch": (C.c_int, [C.c_void_p, C.c_int, _i32p]),
    "graspa_batch_launch_async": (C.c_int, [C.c_void_p, C.c_int]),
    "graspa_batch_wait": (C.c_int, [C.c_void_p, _i32p]),
    "graspa_batch_set_time_slice": (C.c_int, [C.c_void_p, C.c_double]),
    "graspa_batch_busy": (C.c_int, [C.c_void_p, _f64p]),
    "graspa_box_footprint": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "graspa_batch_progress": (C.c_int, [C.c_void_p, _i64p, _i64p, _i64p]),
}


def lib():
    """Loads the engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError("libgraspa_engine.so not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().graspa_last_error().decode()
        raise EngineError("%s failed (%d): %s" % (what, rc, msg))


def _p(a, t=_f64p):
    return a.ctypes.data_as(t)


def box_params(system, precision=GRASPA_FP64, device=0, keep=None):
    """graspa_box_params of a System; `keep` (a list) holds the arrays the
    struct points into."""
    s = system
    T = s.ntypes
    eps = np.ascontiguousarray(s.eps, dtype=np.float64).reshape(T * T)
    sig = np.ascontiguousarray(s.sig, dtype=np.float64).reshape(T * T)
    sh = np.ascontiguousarray(s.shift, dtype=np.float64).reshape(T * T)
    if keep is not None:
        keep += [eps, sig, sh]
    p = BoxParams()
    for i in range(9):
        p.cell[i] = float(s.cell[i])
    p.ntypes = T
    p.epsilon, p.sigma, p.shift = _p(eps), _p(sig), _p(sh)
    p.cutoff_vdw, p.cutoff_coulomb, p.overlap_criteria = s.cut_vdw, s.cut_coul, s.overlap
    p.use_charges = 0 if s.no_charges else 1
    p.alpha = s.alpha
    for i in range(3):
        p.kmax[i] = int(s.kmax[i])
    p.reciprocal_cutoff = s.recip_cut
    p.lammps_ewald = int(s.lammps_ewald)
    p.prefactor = s.prefactor
    p.precision = precision
    p.device = device
    return p


def _voxel_map(fn, system, site_type):
    keep = []
    p = box_params(system, keep=keep)
    xyz = np.ascontiguousarray(system.fw_pos, dtype=np.float64).reshape(-1)
    ty = np.ascontiguousarray(system.fw_type, dtype=np.int32)
    nv = np.zeros(3, dtype=np.int32)
    f = getattr(lib(), fn)
    n = f(C.byref(p), len(ty), _p(xyz), _p(ty, _i32p), site_type, None, 0, _p(nv, _i32p))
    if n < 0:
        raise EngineError("%s failed" % fn)
    w = np.zeros(n, dtype=np.uint32)
    f(C.byref(p), len(ty), _p(xyz), _p(ty, _i32p), site_type, w.ctypes.data_as(C.POINTER(C.c_uint32)), n,
      _p(nv, _i32p))
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")[: int(nv[0]) * int(nv[1]) * int(nv[2])]
    return bits.reshape(int(nv[2]), int(nv[1]), int(nv[0])).astype(bool), tuple(int(x) for x in nv)


def dead_map(system, site_type):
    """The certain-overlap voxel map of the system's framework for probe sites
    of site_type (host computation): (bits as a bool array [nz, ny, nx], nv)."""
    return _voxel_map("graspa_dead_map", system, site_type)


def free_map(system, site_type):
    """The certainly-free voxel map (same voxels as dead_map): voxels whose
    every point is beyond the overlap distance of every framework atom."""
    return _voxel_map("graspa_free_map", system, site_type)


class Box:
    """One simulation box resident on a GPU."""

    def __init__(self, system, precision=GRASPA_FP64, device=0, capacity=None):
        s = system
        self.system = s
        T = s.ntypes
        self._eps = np.ascontiguousarray(s.eps, dtype=np.float64).reshape(T * T)
        self._sig = np.ascontiguousarray(s.sig, dtype=np.float64).reshape(T * T)
        self._shift = np.ascontiguousarray(s.shift, dtype=np.float64).reshape(T * T)
        p = BoxParams()
        for i in range(9):
            p.cell[i] = float(s.cell[i])
        p.ntypes = T
        p.epsilon, p.sigma, p.shift = _p(self._eps), _p(self._sig), _p(self._shift)
        p.cutoff_vdw, p.cutoff_coulomb, p.overlap_criteria = s.cut_vdw, s.cut_coul, s.overlap
        p.use_charges = 0 if s.no_charges else 1
        p.alpha = s.alpha
        for i in range(3):
            p.kmax[i] = int(s.kmax[i])
        p.reciprocal_cutoff = s.recip_cut
        p.lammps_ewald = int(s.lammps_ewald)
        p.prefactor = s.prefactor
        p.precision = precision
        p.device = device
        h = C.c_void_p()
        _check(lib().graspa_box_create(C.byref(p), C.byref(h)), "graspa_box_create")
        self.h = h
        self.precision = precision
        n = len(s.fw_type)
        fx = np.ascontiguousarray(s.fw_pos, dtype=np.float64).reshape(-1)
        fq = np.ascontiguousarray(s.fw_charge, dtype=np.float64)
        ft = np.ascontiguousarray(s.fw_type, dtype=np.int32)
        _check(lib().graspa_box_set_framework(h, n, _p(fx), _p(fq), _p(ft, _i32p)), "set_framework")
        self.comp_ids = []
        for c in s.components:
            cap = capacity if capacity is not None else max(s.allocate_space // max(c.molsize, 1), 1)
            tx = np.ascontiguousarray(c.template_pos, dtype=np.float64).reshape(-1)
            tt = np.ascontiguousarray(c.types, dtype=np.int32)
            tq = np.ascontiguousarray(c.charges, dtype=np.float64)
            cid = C.c_int()
            _check(lib().graspa_box_add_component(h, c.molsize, cap, _p(tx), _p(tt, _i32p), _p(tq), C.byref(cid)),
                   "add_component")
            self.comp_ids.append(cid.value)
            ch = getattr(c, "chain", None)
            if ch is not None:
                m = ChainModel()
                for i in range(min(4, c.molsize)):
                    m.bond[i], m.theta0[i], m.k_bend[i] = float(ch.bond[i]), float(ch.theta0[i]), float(ch.k_bend[i])
                    for t in range(4):
                        m.torsion[i][t] = float(ch.torsion[i][t])
                _check(lib().graspa_box_set_chain(h, cid.value, C.byref(m)), "set_chain")
            if c.init_pos is not None and len(c.init_pos):
                nmol = len(c.init_pos) // c.molsize
                q = c.init_charge if c.init_charge is not None else np.tile(c.charges, nmol)
                self.set_molecules(cid.value, c.init_pos, q, np.tile(c.types, nmol), c.init_molid)

    def close(self):
        if self.h:
            lib().graspa_box_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state ----
    def set_molecules(self, comp, pos, charge, types, molid):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        q = np.ascontiguousarray(charge, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.int32)
        m = np.ascontiguousarray(molid, dtype=np.int64)
        M = self.system.components[comp - 1].molsize
        _check(lib().graspa_box_set_molecules(self.h, comp, len(t) // M, _p(pos), _p(q), _p(t, _i32p), _p(m, _i64p)),
               "set_molecules")

    def get_molecules(self, comp):
        n = C.c_int()
        _check(lib().graspa_box_get_molecules(self.h, comp, C.byref(n), None, None, None, None), "get_molecules")
        M = self.system.components[comp - 1].molsize
        na = n.value * M
        pos = np.zeros((na, 3))
        q = np.zeros(na)
        t = np.zeros(na, dtype=np.int32)
        m = np.zeros(na, dtype=np.int64)
        _check(lib().graspa_box_get_molecules(self.h, comp, C.byref(n), _p(pos), _p(q), _p(t, _i32p), _p(m, _i64p)),
               "get_molecules")
        return pos, q, t, m

    def nvec(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().graspa_box_nvec(self.h, C.byref(a), C.byref(b)), "nvec")
        return a.value, b.value

    def init_ewald(self):
        e = Energy()
        _check(lib().graspa_box_init_ewald(self.h, C.byref(e)), "init_ewald")
        return e

    def structure_factors(self):
        nv, _ = self.nvec()
        f, a = np.zeros(2 * nv), np.zeros(2 * nv)
        _check(lib().graspa_box_get_structure_factors(self.h, _p(f), _p(a)), "get_structure_factors")
        return f, a

    def set_structure_factors(self, sk_fw, sk_ads):
        f = np.ascontiguousarray(sk_fw, dtype=np.float64)
        a = np.ascontiguousarray(sk_ads, dtype=np.float64)
        _check(lib().graspa_box_set_structure_factors(self.h, _p(f), _p(a)), "set_structure_factors")

    # ---- per-move energies ----
    def site_energies(self, pos, types, charges, self_comp=-1, self_molid=-1, excl_comp=-1, excl_molid=-1):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(types, dtype=np.int32)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        n = len(t)
        out = np.zeros((n, 4))
        ov = np.zeros(n, dtype=np.uint8)
        _check(lib().graspa_site_energies(self.h, n, _p(pos), _p(t, _i32p), _p(q), self_comp, self_molid, excl_comp,
                                          excl_molid, _p(out), _p(ov, _u8p)), "site_energies")
        return out, ov

    def trial_energies(self, pos, chain_types, chain_charges, n_trials, self_comp=-1, self_molid=-1):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(chain_types, dtype=np.int32)
        q = np.ascontiguousarray(chain_charges, dtype=np.float64)
        out = np.zeros((n_trials, 4))
        ov = np.zeros(n_trials, dtype=np.uint8)
        _check(lib().graspa_trial_energies(self.h, n_trials, len(t), _p(pos), _p(t, _i32p), _p(q), self_comp,
                                           self_molid, -1, -1, _p(out), _p(ov, _u8p)), "trial_energies")
        return out, ov

    def single_body_delta(self, comp, molid, old_pos, new_pos, types, charges):
        op = np.ascontiguousarray(old_pos, dtype=np.float64).reshape(-1)
        np_ = np.ascontiguousarray(new_pos, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(types, dtype=np.int32)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        e = Energy()
        ov = C.c_int()
        _check(lib().graspa_single_body_delta(self.h, comp, molid, len(t), _p(op), _p(np_), _p(t, _i32p), _p(q),
                                              C.byref(e), C.byref(ov)), "single_body_delta")
        return e, ov.value

    def ewald_delta(self, old_pos, old_q, new_pos, new_q):
        op = np.ascontiguousarray(old_pos, dtype=np.float64).reshape(-1)
        oq = np.ascontiguousarray(old_q, dtype=np.float64)
        npos = np.ascontiguousarray(new_pos, dtype=np.float64).reshape(-1)
        nq = np.ascontiguousarray(new_q, dtype=np.float64)
        same, cross = C.c_double(), C.c_double()
        _check(lib().graspa_ewald_delta(self.h, len(oq), _p(op), _p(oq), len(nq), _p(npos), _p(nq), C.byref(same),
                                        C.byref(cross)), "ewald_delta")
        return same.value, cross.value

    def ewald_accept(self):
        _check(lib().graspa_ewald_accept(self.h), "ewald_accept")

    def commit_move(self, comp, mol, pos):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        _check(lib().graspa_commit_move(self.h, comp, mol, _p(p)), "commit_move")

    def commit_insert(self, comp, pos, charges, types, molid):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        q = np.ascontiguousarray(charges, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.int32)
        _check(lib().graspa_commit_insert(self.h, comp, _p(p), _p(q), _p(t, _i32p), molid), "commit_insert")

    def commit_delete(self, comp, mol):
        _check(lib().graspa_commit_delete(self.h, comp, mol), "commit_delete")

    def total_energy(self):
        e = Energy()
        _check(lib().graspa_total_energy(self.h, C.byref(e)), "total_energy")
        return e

    def widom_batch(self, comp, n_insert, seed, stream=0, first_index=0, excl=0.0, tail_de=0.0, results=True):
        s = self.system
        res = (WidomResult * n_insert)() if results else None
        summ = WidomSummary()
        _check(lib().graspa_widom_batch(self.h, comp, s.n_trial_pos, s.n_trial_orient, s.beta, int(s.vdw_real_bias),
                                        excl, tail_de, seed, stream, first_index, n_insert, res, C.byref(summ)),
               "widom_batch")
        return res, summ

    def widom_enqueue(self, comp, n_insert, seed, stream=0, first_index=0, excl=0.0, tail_de=0.0):
        """Queues one Widom batch on the box's stream and returns at once; its
        summary goes into the device-side running totals (widom_totals)."""
        s = self.system
        _check(lib().graspa_widom_enqueue(self.h, comp, s.n_trial_pos, s.n_trial_orient, s.beta,
                                          int(s.vdw_real_bias), excl, tail_de, seed, stream, first_index, n_insert),
               "widom_enqueue")

    def widom_enqueue_noalloc(self, comp, n_insert, seed, stream=0, first_index=0, excl=0.0, tail_de=0.0):
        """widom_enqueue with the pipeline's scratch in the library's fixed
        per-device block: allocates nothing (graspa_widom_enqueue_noalloc)."""
        s = self.system
        _check(lib().graspa_widom_enqueue_noalloc(self.h, comp, s.n_trial_pos, s.n_trial_orient, s.beta,
                                                  int(s.vdw_real_bias), excl, tail_de, seed, stream, first_index,
                                                  n_insert),
               "widom_enqueue_noalloc")

    def footprint(self):
        """Device bytes the box holds now (graspa_box_footprint)."""
        v = C.c_uint64()
        _check(lib().graspa_box_footprint(self.h, C.byref(v)), "box_footprint")
        return int(v.value)

    def widom_graph(self, comp, n_insert, seed, stream=0, first_index=0, excl=0.0, tail_de=0.0):
        """Captures one queued Widom batch as a HIP graph kept in the box
        (graspa_widom_graph_create); replay r runs first_index + r n_insert."""
        s = self.system
        _check(lib().graspa_widom_graph_create(self.h, comp, s.n_trial_pos, s.n_trial_orient, s.beta,
                                               int(s.vdw_real_bias), excl, tail_de, seed, stream, first_index,
                                               n_insert), "widom_graph_create")

    def widom_graph_launch(self, n_replays=1):
        """Queues replays of the captured batch (graspa_widom_graph_launch)."""
        _check(lib().graspa_widom_graph_launch(self.h, int(n_replays)), "widom_graph_launch")

    def widom_totals(self, reset=False):
        """Waits for the queued batches; summary fields summed over every batch
        since the last reset."""
        summ = WidomSummary()
        _check(lib().graspa_widom_totals(self.h, C.byref(summ), int(reset)), "widom_totals")
        return summ

    def set_eval_shape(self, waves=12, per_cu=0, framework_lds=1, default=False):
        """Launch shape of the pair-energy kernel; default=True restores the built-in one."""
        shp = None if default else C.byref(EvalShape(waves, per_cu, framework_lds))
        _check(lib().graspa_widom_set_shape(self.h, shp), "widom_set_shape")

    def eval_shape(self):
        shp = EvalShape()
        _check(lib().graspa_widom_get_shape(self.h, C.byref(shp)), "widom_get_shape")
        return shp.as_dict()

    def autotune_widom(self, comp, n_insert, seed, reps=3, excl=0.0, tail_de=0.0, cap=64):
        """Times the Widom batch in every launch shape that fits and keeps the
        fastest in the box.  Returns (chosen shape, [(shape, ms per batch)])."""
        s = self.system
        tried = (EvalShape * cap)()
        ms = np.zeros(cap)
        n = C.c_int(0)
        _check(lib().graspa_widom_autotune(self.h, comp, s.n_trial_pos, s.n_trial_orient, s.beta,
                                           int(s.vdw_real_bias), excl, tail_de, seed, n_insert, reps, tried,
                                           ms.ctypes.data_as(_f64p), cap, C.byref(n)), "widom_autotune")
        k = min(cap, n.value)
        return self.eval_shape(), [(tried[i].as_dict(), float(ms[i])) for i in range(k)]

    # ---- measurement ----
    def timing_enable(self, family, on=True):
        _check(lib().graspa_timing_enable(self.h, family.encode(), int(on)), "timing_enable")

    def timing_read(self):
        t, n = C.c_double(), C.c_int64()
        _check(lib().graspa_timing_read(self.h, C.byref(t), C.byref(n)), "timing_read")
        return t.value, n.value

    def sync(self):
        _check(lib().graspa_box_sync(self.h), "sync")


def widom_results_to_numpy(res):
    n = len(res)
    arr = np.zeros(n, dtype=[("rosen", "f8"), ("energy", "f8"), ("vdw", "f8"), ("real", "f8"), ("ewald", "f8"),
                             ("first_bead", "i4"), ("orientation", "i4"), ("ok", "i4")])
    raw = np.ctypeslib.as_array(res)
    for name in arr.dtype.names:
        arr[name] = raw[name if name != "energy" else "energy"]
    return arr


def exclusion_energy(system, comp_index):
    """Rigid intra + self exclusion of one molecule of a species
    (ewald_preparation.h:351-366), host FP64 from the species template."""
    import math
    c = system.components[comp_index - 1]
    if system.no_charges or not c.has_charge:
        return 0.0, 0.0
    cell, inv, cubic = system.cell, system.inv_cell, system.cubic
    pos = c.template_pos if (c.init_pos is None or not len(c.init_pos)) else c.init_pos[:c.molsize]
    q = c.charges
    intra = 0.0
    for i in range(c.molsize - 1):
        for j in range(i + 1, c.molsize):
            d = _deck._pbc(np.asarray(pos[i]) - np.asarray(pos[j]), cell, inv, cubic)
            r = math.sqrt(float(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
            intra += system.prefactor * q[i] * q[j] * math.erf(system.alpha * r) / r
    pself = system.prefactor * system.alpha / math.sqrt(math.pi)
    selfe = sum(pself * qi * qi for qi in q)
    return intra, selfe


def make_mc_params(system, box, pool_size=333334):
    """graspa_mc_params from a System, as Initialize() sets them up (main.cpp:58-343)."""
    s = system
    p = MCParams()
    p.n_species = len(s.components)
    for i, c in enumerate(s.components):
        cp = c.cumulative_probabilities()
        sp = p.species[i]
        sp.p_translation, sp.p_rotation, sp.p_widom, sp.p_swap = (cp["translation"], cp["rotation"], cp["widom"],
                                                                  cp["swap"])
        sp.p_reinsertion = cp["reinsertion"]
        sp.p_identity = cp["identity"]
        sp.mol_fraction, sp.fugacity_coeff, sp.ideal_rosen = c.mol_fraction, c.fugacity_coeff, c.ideal_rosen
        sp.excl_intra, sp.excl_self = exclusion_energy(s, i + 1)
        for d in range(3):
            sp.max_translation[d] = s.cell[4 * d] * 0.1
            sp.max_rotation[d] = 30.0 / (180 / 3.1415)
        for t in c.types:
            sp.pseudo_per_mol[int(t)] += 1
    p.n_trial_pos, p.n_trial_orient = s.n_trial_pos, s.n_trial_orient
    p.vdw_real_bias = int(s.vdw_real_bias)
    p.use_max_step, p.max_step_per_cycle = int(s.use_max_step), s.max_step_per_cycle
    p.has_tail = int(s.has_tail)
    box._tail = np.ascontiguousarray(s.tail, dtype=np.float64).reshape(-1)
    box._use_tail = np.ascontiguousarray(s.use_tail, dtype=np.int32).reshape(-1)
    p.tail = _p(box._tail)
    p.use_tail = _p(box._use_tail, _i32p)
    p.beta, p.pressure = s.beta, s.pressure
    p.seed = s.seed
    p.pool_size = pool_size
    nm = [len(c.init_pos) // c.molsize if (c.init_pos is not None) else 0 for c in s.components]
    n = s.pseudo_counts(nm)
    for t in range(min(32, len(n))):
        p.n_pseudo[t] = int(n[t])
    return p


class MonteCarlo:
    """GCMC driver in the reference's random-number order (parity mode)."""

    def __init__(self, box, pool_size=333334):
        self.box = box
        self.params = make_mc_params(box.system, box, pool_size)
        h = C.c_void_p()
        _check(lib().graspa_mc_create(box.h, C.byref(self.params), C.byref(h)), "graspa_mc_create")
        self.h = h

    def cycles(self, first_cycle, n_cycles, production=True, max_trace=0):
        tr = (Trace * max_trace)() if max_trace else None
        n = C.c_int64()
        _check(lib().graspa_mc_cycles(self.h, first_cycle, n_cycles, int(production), tr, max_trace, C.byref(n)),
               "graspa_mc_cycles")
        return n.value, tr

    def create_molecules(self, comp, n, max_attempts=10 ** 7):
        """CreateMolecule_InOneBox (axpy.cu:300-377) for component `comp`:
        returns the energy change of creating n molecules."""
        got = C.c_int()
        e = Energy()
        _check(lib().graspa_mc_create_molecules(self.h, comp, n, max_attempts, C.byref(got), C.byref(e)),
               "graspa_mc_create_molecules")
        return e

    def cycle_times(self, n):
        out = np.zeros(n)
        k = lib().graspa_mc_cycle_times(self.h, _p(out), n)
        return out[:max(k, 0)]

    def stats(self):
        ns = len(self.box.system.components)
        e = Energy()
        nm = np.zeros(ns, dtype=np.int32)
        cnt = np.zeros(N_COUNTERS * ns, dtype=np.int64)
        ww = np.zeros(ns)
        wn = np.zeros(ns)
        _check(lib().graspa_mc_stats(self.h, C.byref(e), _p(nm, _i32p), _p(cnt, _i64p), _p(ww), _p(wn)), "mc_stats")
        return e, nm, cnt.reshape(ns, N_COUNTERS), ww, wn

    def close(self):
        if self.h:
            lib().graspa_mc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchStats:
    """State of one box of a batch (graspa_batch_box_stats)."""

    def __init__(self, running, n_mol, counters, widom_w, widom_n, load_sum, load_sum2, n_load, moves, cycles):
        self.running, self.n_mol, self.counters = running, n_mol, counters
        self.widom_w, self.widom_n = widom_w, widom_n
        self.load_sum, self.load_sum2, self.n_load = load_sum, load_sum2, n_load
        self.moves, self.cycles = moves, cycles

    def loading(self):
        """Mean molecules per box of each species over the production cycles."""
        return self.load_sum / max(self.n_load, 1)


class Batch:
    """Device-resident GCMC of many boxes (graspa_batch_*): box i draws its
    moves from the Philox stream (seed_i, stream_i, move)."""

    def __init__(self, boxes, seeds, streams=None, trace_cap=0, cycle_cap=0, params=None, n_slots=None):
        boxes = list(boxes)
        n = max(len(boxes), n_slots or 0)
        streams = list(range(len(boxes))) if streams is None else list(streams)
        self.boxes = boxes + [None] * (n - len(boxes))
        self.params = (list(params) if params is not None else [make_mc_params(b.system, b) for b in boxes])
        self.params += [None] * (n - len(self.params))
        arr = (BatchBox * n)()
        for i, b in enumerate(boxes):
            arr[i].box = b.h
            arr[i].params = C.pointer(self.params[i])
            arr[i].seed = int(seeds[i])
            arr[i].stream = int(streams[i])
            arr[i].n_block = int(getattr(boxes[i].system, "n_blocks", 5))
        self._arr = arr
        self.trace_cap = trace_cap
        self.n = n
        h = C.c_void_p()
        _check(lib().graspa_batch_create(C.cast(arr, C.c_void_p), n, trace_cap, cycle_cap, C.byref(h)),
               "graspa_batch_create")
        self.h = h

    def set_box(self, slot, box, params=None, seed=0, stream=0):
        """Puts `box` into `slot` (None empties it)."""
        if box is None:
            _check(lib().graspa_batch_set_box(self.h, slot, None), "batch_set_box")
            self.boxes[slot], self.params[slot] = None, None
            return
        params = params if params is not None else make_mc_params(box.system, box)
        bb = BatchBox()
        bb.box, bb.params, bb.seed, bb.stream = box.h, C.pointer(params), int(seed), int(stream)
        bb.n_block = int(getattr(box.system, "n_blocks", 5))
        _check(lib().graspa_batch_set_box(self.h, slot, C.byref(bb)), "batch_set_box")
        self.boxes[slot], self.params[slot] = box, params

    def set_cycles(self, slot, n_cycles, production=True):
        _check(lib().graspa_batch_set_cycles(self.h, slot, n_cycles, int(production)), "batch_set_cycles")

    def launch(self, moves_per_launch=0):
        """One launch; returns the number of boxes still short of their cycles."""
        r = C.c_int()
        _check(lib().graspa_batch_launch(self.h, moves_per_launch, C.byref(r)), "batch_launch")
        return r.value

    def launch_async(self, moves_per_launch=0):
        """Queues one launch and returns at once (graspa_batch_launch_async)."""
        _check(lib().graspa_batch_launch_async(self.h, moves_per_launch), "batch_launch_async")

    def set_time_slice(self, us):
        """Per-box move budgets of about `us` microseconds per launch (0: off)."""
        _check(lib().graspa_batch_set_time_slice(self.h, float(us)), "batch_set_time_slice")

    def wait(self):
        """Finishes the launch in flight; returns the boxes still short of their cycles."""
        r = C.c_int()
        _check(lib().graspa_batch_wait(self.h, C.byref(r)), "batch_wait")
        return r.value

    def progress(self):
        m, c, t = (np.zeros(self.n, dtype=np.int64) for _ in range(3))
        _check(lib().graspa_batch_progress(self.h, _p(m, _i64p), _p(c, _i64p), _p(t, _i64p)), "batch_progress")
        self._last_progress = (m, c, t)
        return m, c, t

    def total_moves(self):
        return int(self.progress()[0].sum())

    def slot_done(self, slot):
        m, c, t = getattr(self, "_last_progress", None) or self.progress()
        return self.boxes[slot] is not None and c[slot] >= t[slot]

    def run(self, n_cycles, production=True, moves_per_launch=0):
        m = C.c_int64()
        _check(lib().graspa_batch_run(self.h, n_cycles, int(production), moves_per_launch, C.byref(m)),
               "graspa_batch_run")
        return m.value

    def stats(self, i):
        ns = len(self.boxes[i].system.components)
        e = Energy()
        nm = np.zeros(ns, dtype=np.int32)
        cnt = np.zeros(N_COUNTERS * ns, dtype=np.int64)
        ww, wn, ls, ls2 = np.zeros(ns), np.zeros(ns), np.zeros(ns), np.zeros(ns)
        nl, mv, cy = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().graspa_batch_box_stats(self.h, i, C.byref(e), _p(nm, _i32p), _p(cnt, _i64p), _p(ww), _p(wn),
                                            _p(ls), _p(ls2), C.byref(nl), C.byref(mv), C.byref(cy)), "batch_stats")
        return BatchStats(e, nm, cnt.reshape(ns, N_COUNTERS), ww, wn, ls, ls2, nl.value, mv.value, cy.value)

    def block_stats(self, i):
        """Per-block sums of slot i's current production run (graspa_batch_block_stats)."""
        b = BlockStats()
        _check(lib().graspa_batch_block_stats(self.h, i, C.byref(b)), "batch_block_stats")
        return b.arrays()

    def cycle_times(self, i, n):
        out = np.zeros(n)
        k = lib().graspa_batch_cycle_times(self.h, i, _p(out), n)
        if k < 0:
            _check(k, "batch_cycle_times")
        return out[:k]

    def trace(self, i):
        buf = (Trace * max(self.trace_cap, 1))()
        k = lib().graspa_batch_trace(self.h, i, C.cast(buf, C.c_void_p), self.trace_cap)
        if k < 0:
            _check(k, "batch_trace")
        return buf[:k]

    def busy_seconds(self):
        """Device-clock seconds each slot's box has spent in launches."""
        out = np.zeros(self.n)
        _check(lib().graspa_batch_busy(self.h, _p(out)), "batch_busy")
        return out

    def kernel_time(self):
        t, n = C.c_double(), C.c_int64()
        _check(lib().graspa_batch_kernel_time(self.h, C.byref(t), C.byref(n)), "batch_kernel_time")
        return t.value, n.value

    def sync(self):
        _check(lib().graspa_batch_sync(self.h), "batch_sync")

    def max_steps(self, i):
        """(max translation, max rotation) of box i, [n_species, 3] each."""
        ns = len(self.boxes[i].system.components)
        mt, mr = np.zeros(3 * ns), np.zeros(3 * ns)
        _check(lib().graspa_batch_max_steps(self.h, i, _p(mt), _p(mr)), "batch_max_steps")
        return mt.reshape(ns, 3), mr.reshape(ns, 3)

    def close(self):
        if getattr(self, "h", None):
            lib().graspa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
