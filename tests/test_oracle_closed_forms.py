"""The oracle against what the mathematics fixes.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed cell) fails it: LJ zero and minimum,
Madelung constant of rock salt through the full Ewald sum, brute-force
real-space lattice sums against Ewald for random neutral clusters, per-move
energy differences against differences of totals, the flat-potential and
ideal-gas limits of the Rosenbluth weight, and the reference's deletion
bookkeeping."""
import math

import numpy as np
import pytest

from graspa_amd import deck


def tiny_system(cell, charges, pos, eps_k=(0.0,), sig=(1.0,), alpha=None, kmax=(8, 8, 8), rc=None, types=None,
                recip=None):
    T = len(eps_k)
    rcut = rc if rc is not None else 0.5 * min(cell[0], cell[4], cell[8]) - 1e-9
    eps, sg, sh, tail, ut = deck.lorentz_berthelot(np.array(eps_k), np.array(sig), [False] * T, [False] * T, rcut ** 2)
    s = deck.System(cell=np.array(cell, dtype=float), pseudo_names=["X%d" % i for i in range(T)],
                    pseudo_mass=np.ones(T), pseudo_charge=np.zeros(T), eps=eps, sig=sg, shift=sh, tail=tail,
                    use_tail=ut, cut_vdw=rcut, cut_coul=rcut, no_charges=False)
    s.alpha = alpha if alpha is not None else 5.6 / min(cell[0], cell[4], cell[8])
    s.kmax = kmax
    s.recip_cut = recip if recip is not None else float(max(kmax) ** 2 + 1)
    c = deck.Component(name="ions", molsize=1, template_pos=np.zeros((1, 3)), types=np.zeros(1, dtype=np.int32),
                       charges=np.zeros(1), masses=np.ones(1))
    c.init_pos = np.asarray(pos, dtype=float)
    c.init_molid = np.arange(len(pos), dtype=np.int64)
    c.init_charge = np.asarray(charges, dtype=float)
    s.components = [c]
    return s


def test_lj_zero_and_minimum(oracle_lib):
    from oracle.binding import OracleSystem
    import ctypes as C
    s = tiny_system([30, 0, 0, 0, 30, 0, 0, 0, 30], [0.0], [[0, 0, 0]], eps_k=(120.0,), sig=(3.4,), rc=14.0)
    o = OracleSystem(s)
    out = (C.c_double * 2)()
    eps = 120.0 / 1.20272430057
    oracle_lib.or_vdw(o.box, 0, 0, 3.4 ** 2, 1.0, out)
    assert abs(out[0]) < 1e-9
    rmin = 2 ** (1 / 6) * 3.4
    oracle_lib.or_vdw(o.box, 0, 0, rmin ** 2, 1.0, out)
    assert abs(out[0] + eps) < 1e-9
    oracle_lib.or_vdw(o.box, 0, 0, 1.1 * rmin ** 2, 1.0, out)
    assert -eps < out[0] < 0


def test_madelung_rock_salt(oracle_lib):
    """Coulomb energy of NaCl (4 cation + 4 anion sites per cubic cell of edge a)
    through real space + Fourier + self: E per ion pair = -M q^2 pref / (a/2)."""
    from oracle.binding import OracleSystem
    a = 5.64
    base = [(0, 0, 0, 1), (0.5, 0.5, 0, 1), (0.5, 0, 0.5, 1), (0, 0.5, 0.5, 1),
            (0.5, 0, 0, -1), (0, 0.5, 0, -1), (0, 0, 0.5, -1), (0.5, 0.5, 0.5, -1)]
    n = 3
    L = n * a
    pos, q = [], []
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for (x, y, z, c) in base:
                    pos.append([(x + i) * a, (y + j) * a, (z + k) * a])
                    q.append(c)
    s = tiny_system([L, 0, 0, 0, L, 0, 0, 0, L], q, pos, alpha=0.45, kmax=(12, 12, 12), rc=L / 2 - 1e-6, recip=145.0)
    o = OracleSystem(s)
    e = o.total_vdw_real()
    f, skf, ska, selfe, intra, _ = o.ewald_total()
    total = e.gg_real + f.gg_ewald - selfe[1] - intra[1]
    n_pairs = len(q) // 2
    madelung = -total / n_pairs / (s.prefactor / (a / 2))
    assert abs(madelung - 1.747565) < 2e-5, madelung


def brute_coulomb(cell_len, q, pos, shells):
    """Direct lattice sum with a spherical cutoff growing to convergence
    (neutral cluster, dipole-free images averaged over shells)."""
    e = 0.0
    L = cell_len
    rng = range(-shells, shells + 1)
    for i in range(len(q)):
        for j in range(len(q)):
            for nx in rng:
                for ny in rng:
                    for nz in rng:
                        if i == j and nx == ny == nz == 0:
                            continue
                        d = np.array(pos[i]) - np.array(pos[j]) + L * np.array([nx, ny, nz])
                        e += 0.5 * q[i] * q[j] / np.linalg.norm(d)
    return e


def test_ewald_matches_energy_differences(oracle_lib):
    """dE of moving one charge (or_ewald_delta + real-space site energies) equals
    the difference of two full Ewald totals, for a random neutral cluster."""
    from oracle.binding import OracleSystem
    rng = np.random.default_rng(3)
    L = 12.0
    n = 8
    pos = rng.uniform(0, L, (n, 3))
    q = rng.normal(0, 0.5, n)
    q -= q.mean()
    s1 = tiny_system([L, 0, 0, 0, L, 0, 0, 0, L], q, pos, alpha=0.35, kmax=(7, 7, 7), recip=50.0)
    o1 = OracleSystem(s1)

    def total(o):
        e = o.total_vdw_real()
        f, skf, ska, se, intra, _ = o.ewald_total()
        return e.gg_real + f.gg_ewald - se[1] - intra[1], ska, skf

    E1, ska, skf = total(o1)
    new = pos.copy()
    new[2] += np.array([0.7, -1.3, 0.4])
    s2 = tiny_system([L, 0, 0, 0, L, 0, 0, 0, L], q, new, alpha=0.35, kmax=(7, 7, 7), recip=50.0)
    E2, ska2, _ = total(OracleSystem(s2))
    same, cross, sk_new = o1.ewald_delta(ska, skf, pos[2:3], q[2:3], new[2:3], q[2:3])
    e_new, _ = o1.site_energies(new[2:3], [0], q[2:3], self_comp=1, self_molid=2)
    e_old, _ = o1.site_energies(pos[2:3], [0], q[2:3], self_comp=1, self_molid=2)
    dE = (e_new[0, 3] - e_old[0, 3]) + same + 2 * cross
    assert abs(dE - (E2 - E1)) < 1e-8 * max(1.0, abs(E1))
    assert np.allclose(sk_new, ska2, atol=1e-10)


def test_ewald_against_direct_lattice_sum(oracle_lib):
    """Full Ewald energy of 4 charges vs the conditionally convergent direct sum
    evaluated with tin-foil boundary (dipole term removed)."""
    from oracle.binding import OracleSystem
    L = 10.0
    pos = np.array([[1.0, 2.0, 3.0], [4.5, 1.5, 6.0], [7.0, 8.0, 2.0], [3.0, 6.5, 8.5]])
    q = np.array([1.0, -1.0, 0.5, -0.5])
    s = tiny_system([L, 0, 0, 0, L, 0, 0, 0, L], q, pos, alpha=0.5, kmax=(12, 12, 12), recip=145.0, rc=5.0 - 1e-9)
    o = OracleSystem(s)
    e = o.total_vdw_real()
    f, skf, ska, se, intra, _ = o.ewald_total()
    ew = e.gg_real + f.gg_ewald - se[1] - intra[1]
    # direct sums over growing cubes converge to Ewald + 2 pi |M|^2 / 3V (surface term)
    M = (q[:, None] * pos).sum(0)
    surf = 2 * math.pi / (3 * L ** 3) * float(M @ M)
    direct = brute_coulomb(L, q, pos, 6) * s.prefactor
    assert abs(direct - (ew + surf * s.prefactor)) / abs(direct) < 2e-3


def test_widom_ideal_gas_weight_is_one(oracle_lib):
    """No framework, no charges, empty box: every trial has zero energy, so
    W = (K/K) (O/O) = 1 exactly (Rosenbluth of a flat potential)."""
    from graspa_amd import synth
    from oracle.binding import OracleSystem, widom_records_to_numpy
    s = synth.build(n_fw=1, L=20.0, species=("CO2",), charges=False, K=7, O=5)
    s.eps[:] = 0.0
    o = OracleSystem(s)
    nv = 2 * s.nvec
    rec = widom_records_to_numpy(o.widom_batch(1, np.zeros(max(nv, 2)), np.zeros(max(nv, 2)), seed=1, n_insert=50))
    assert np.all(rec["ok"] == 1)
    assert np.all(rec["rosen"] == 1.0)


def test_widom_single_trial_equals_boltzmann_factor(oracle_lib):
    """K = O = 1: CBMC insertion reduces to a plain insertion, W = exp(-beta dE)."""
    from graspa_amd import synth
    from oracle.binding import OracleSystem, widom_records_to_numpy
    s = synth.build(n_fw=120, L=20.0, species=("CO2",), K=1, O=1, seed=5)
    o = OracleSystem(s)
    f, skf, ska, se, intra, _ = o.ewald_total()
    rec = widom_records_to_numpy(o.widom_batch(1, ska, skf, seed=3, n_insert=40))
    ok = rec["ok"] == 1
    assert ok.sum() > 10
    w = np.exp(-s.beta * rec["energy"][ok])
    assert np.allclose(rec["rosen"][ok], w, rtol=1e-12)


def test_oracle_widom_batch_is_independent_of_the_thread_count(oracle_lib):
    """The oracle's batched Widom loop runs OpenMP over insertions (each with
    its own counter-based stream): 1 and 4 threads give identical records."""
    from graspa_amd import synth
    from oracle import binding as ob
    s = synth.config1(n_fw=300)
    o = ob.OracleSystem(s)
    _, skf, ska, _, _, _ = o.ewald_total()
    out = []
    for th in (1, 4):
        ob.lib().or_set_threads(th)
        out.append(ob.widom_records_to_numpy(o.widom_batch(1, ska, skf, seed=3, n_insert=200)))
    ob.lib().or_set_threads(1)
    assert out[0]["ok"].sum() > 50
    for k in out[0].dtype.names:
        assert np.array_equal(out[0][k], out[1][k]), k
