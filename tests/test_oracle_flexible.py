"""Flexible n-butane growth in the oracle, pinned by what the bonded model
fixes.  The reference has no flexible molecules (read_data.cpp:2104-2116), so
nothing of its own can pin this; each check below is one a plausible mistake
fails:

- the trial generator of a bead must sample the bonded Boltzmann densities:
  bend angle theta from sin(theta) exp(-beta k/2 (theta - theta0)^2) (a dropped
  sin(theta) Jacobian moves <theta> by 9 standard errors here, k for k/2
  changes <(theta - theta0)^2> by half), dihedral phi from exp(-beta U(phi))
  with phi = 180 degrees trans (a cis/trans convention error flips <cos phi>),
  bead 2's azimuth and bead 1's direction uniform;
- a Widom insertion of the flexible chain must give <W> = <exp(-beta U_ext)>
  over ideal chains placed uniformly (U_ext the framework energy), the ratio
  of the configurational integrals with and without the framework; the
  right side is estimated from chains built here in numpy by inverse-CDF
  sampling of the same densities and energies from a numpy LJ sum (a dropped
  1/O per growth step is a factor 8^3 off);
- GCMC of chains that interact with the framework but not with each other
  is an ideal adsorbed gas: the loading is Poisson with mean
  beta y P V <exp(-beta U_ext)> -- this pins the insertion and deletion
  (retrace) weights and their detailed balance.
"""
import ctypes as C
import math

import numpy as np
import pytest

from graspa_amd import synth

BETA_K = 1.0 / 300.0   # beta in 1/K at 300 K


def _grid_moments(x, logp, funcs):
    p = np.exp(logp - logp.max())
    w = p / np.trapezoid(p, x)
    return [np.trapezoid(f(x) * w, x) for f in funcs]


def _angle(a, b, c):
    u, v = a - b, c - b
    cs = (u * v).sum(-1) / np.linalg.norm(u, axis=-1) / np.linalg.norm(v, axis=-1)
    return np.arccos(np.clip(cs, -1.0, 1.0))


def _dihedral(a, b, c, d):
    b1, b2, b3 = b - a, c - b, d - c
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    x = (n1 * n2).sum(-1)
    y = (np.cross(n1, n2) * (b2 / np.linalg.norm(b2, axis=-1)[..., None])).sum(-1)
    return np.arctan2(y, x) % (2 * math.pi)


def _trials(mc, bead, prev, n, seed=7):
    from oracle import binding as ob
    out = np.zeros((n, 3))
    p = np.ascontiguousarray(prev, dtype=np.float64).ravel()
    o = np.zeros(3)
    for t in range(n):
        ob.lib().or_flex_trial(C.byref(mc), 1, bead, ob._p(p), seed, 3, t // 8, t % 8, ob._p(o))
        out[t] = o
    return out


def _setup(n_fw=60, L=24.0, P=1e4):
    from oracle import binding as ob
    s = synth.config2(n_fw=n_fw, L=L, P=P, flexible=True)
    o = ob.OracleSystem(s, capacity=256)
    mc = ob.make_mc(s, o, production=1)
    return s, o, mc


def test_bonded_distributions_of_the_trial_generator(oracle_lib):
    s, o, mc = _setup()
    ch = s.components[0].chain
    beta = s.beta
    prev = np.asarray(s.components[0].template_pos)
    n = 60000
    # bead 3: bend angle at bead 2 and dihedral 0-1-2-3
    p3 = _trials(mc, 3, prev[:3], n)
    assert np.allclose(np.linalg.norm(p3 - prev[2], axis=1), ch.bond[3], atol=1e-12)
    th = _angle(prev[1][None], prev[2][None], p3)
    ph = _dihedral(prev[0][None], prev[1][None], prev[2][None], p3)
    k, t0 = ch.k_bend[3], ch.theta0[3]
    xs = np.linspace(1e-6, math.pi - 1e-6, 200001)
    lp = np.log(np.sin(xs)) - beta * 0.5 * k * (xs - t0) ** 2
    m_th, m_th2 = _grid_moments(xs, lp, [lambda x: x, lambda x: (x - t0) ** 2])
    se = th.std() / math.sqrt(n)
    assert abs(th.mean() - m_th) < 4 * se, (th.mean(), m_th, se)
    assert abs(((th - t0) ** 2).mean() / m_th2 - 1.0) < 0.03
    c = ch.torsion[3]
    ps = np.linspace(0.0, 2 * math.pi, 200001)
    up = c[0] + c[1] * (1 + np.cos(ps)) + c[2] * (1 - np.cos(2 * ps)) + c[3] * (1 + np.cos(3 * ps))
    funcs = [np.cos, lambda x: np.cos(2 * x), lambda x: np.cos(3 * x), np.sin]
    want = _grid_moments(ps, -beta * up, funcs)
    for f, w in zip(funcs, want):
        v = f(ph)
        assert abs(v.mean() - w) < 4 * v.std() / math.sqrt(n) + 1e-12, (f, v.mean(), w)
    # trans dominates at 300 K
    assert want[0] < -0.5
    # bead 2: bend at bead 1 and a uniform azimuth about the 0-1 axis
    p2 = _trials(mc, 2, prev[:2], n // 2)
    th2 = _angle(prev[0][None], prev[1][None], p2)
    assert abs(th2.mean() - m_th) < 4 * th2.std() / math.sqrt(n // 2)
    ax = (prev[1] - prev[0]) / np.linalg.norm(prev[1] - prev[0])
    perp = (p2 - prev[1]) - np.outer((p2 - prev[1]) @ ax, ax)
    r = np.linalg.norm(perp, axis=1).mean()
    assert np.all(np.abs(perp.mean(0)) < 4 * r / math.sqrt(n // 2))
    # bead 1: uniform on the bond sphere
    p1 = _trials(mc, 1, prev[:1], n // 2) - prev[0]
    u = p1 / ch.bond[1]
    assert np.allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-12)
    m = n // 2
    assert np.all(np.abs(u.mean(0)) < 4 / math.sqrt(3 * m))
    assert np.all(np.abs((u * u).mean(0) - 1 / 3) < 4 * math.sqrt(4 / 45 / m))


def _ideal_chains(s, n, rng):
    """n ideal butane chains, built independently of the oracle: bend angles
    and dihedrals by inverse-CDF sampling of their densities on a grid, a
    uniform random rotation and a uniform first bead."""
    ch = s.components[0].chain
    beta = s.beta

    def sample(x, logp, m):
        p = np.exp(logp - logp.max())
        cdf = np.cumsum(p)
        cdf /= cdf[-1]
        return np.interp(rng.random(m), cdf, x)

    xs = np.linspace(1e-6, math.pi - 1e-6, 100001)
    th = sample(xs, np.log(np.sin(xs)) - beta * 0.5 * ch.k_bend[2] * (xs - ch.theta0[2]) ** 2, 2 * n).reshape(n, 2)
    ps = np.linspace(0.0, 2 * math.pi, 100001)
    c = ch.torsion[3]
    up = c[0] + c[1] * (1 + np.cos(ps)) + c[2] * (1 - np.cos(2 * ps)) + c[3] * (1 + np.cos(3 * ps))
    phi = sample(ps, -beta * up, n)
    b = ch.bond[1]
    pos = np.zeros((n, 4, 3))
    pos[:, 1, 2] = b
    # bead 2 in the xz plane at angle th[:,0] from the 0-1 axis at bead 1
    pos[:, 2, 0] = b * np.sin(th[:, 0])
    pos[:, 2, 2] = b - b * np.cos(th[:, 0])
    A, B, Cc = pos[:, 0], pos[:, 1], pos[:, 2]
    bc = (Cc - B) / np.linalg.norm(Cc - B, axis=1)[:, None]
    nn = np.cross(B - A, bc)
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    mm = np.cross(nn, bc)
    t3 = th[:, 1]
    d = (-b * np.cos(t3))[:, None] * bc + (b * np.sin(t3) * np.cos(phi))[:, None] * mm + \
        (b * np.sin(t3) * np.sin(phi))[:, None] * nn
    pos[:, 3] = Cc + d
    # uniform rotation (random unit quaternion) about bead 0, then translation
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    w, x, y, z = q.T
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
                  np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
                  np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    pos = np.einsum("nij,naj->nai", R, pos)
    pos += (rng.random((n, 3)) * s.cell[0])[:, None, :]
    return pos


def _external_energy(s, pos):
    """Framework LJ energy (truncated 12-6, minimum image) of chains (n, 4, 3);
    inf when a pair exceeds the overlap criterion or comes within 0.1 A."""
    fw = np.asarray(s.fw_pos)
    t_fw = np.asarray(s.fw_type)
    L = s.cell[0]
    types = np.asarray(s.components[0].types)
    U = np.zeros(len(pos))
    bad = np.zeros(len(pos), dtype=bool)
    for a in range(4):
        t = int(types[a])
        eps = np.asarray(s.eps)[t_fw, t]
        sig2 = np.asarray(s.sig)[t_fw, t] ** 2
        d = fw[None, :, :] - pos[:, a, None, :]
        d -= L * np.round(d / L)
        rr = (d * d).sum(-1)
        x = (sig2[None, :] / rr) ** 3
        e = 4.0 * eps[None, :] * (x * x - x)
        inside = rr < s.cut_vdw ** 2
        U += np.where(inside, e, 0.0).sum(1)
        bad |= (inside & ((e > s.overlap) | (rr < 0.01))).any(1)
    return np.where(bad, np.inf, U)


def _boltzmann_average(s, n=120000, seed=3):
    rng = np.random.default_rng(seed)
    vals = []
    for _ in range(n // 20000):
        pos = _ideal_chains(s, 20000, rng)
        vals.append(np.exp(-s.beta * _external_energy(s, pos)))
    v = np.concatenate(vals)
    return v.mean(), v.std() / math.sqrt(len(v))


def _no_guest_guest(s):
    """Switches off the butane-butane interactions (framework-butane stay)."""
    eps = np.array(s.eps, dtype=np.float64)
    sh = np.array(s.shift, dtype=np.float64)
    g = [int(t) for t in set(s.components[0].types)]
    for a in g:
        for b in g:
            eps[a, b] = 0.0
            sh[a, b] = 0.0
    s.eps, s.shift = eps, sh


def test_widom_weight_equals_ratio_of_configurational_integrals(oracle_lib):
    from oracle import binding as ob
    s, _, _ = _setup(n_fw=60)
    s.components[0].p_translation = s.components[0].p_rotation = s.components[0].p_swap = 0.0
    s.components[0].p_widom = 1.0
    o = ob.OracleSystem(s, capacity=16)
    mc = ob.make_mc(s, o, production=1)
    _, skf, ska, _, _, _ = o.ewald_total()
    n_cyc = 1500                     # 20 Widom insertions per cycle
    tr = (ob.OrTrace * (20 * n_cyc))()
    mi = C.c_int64(0)
    n = ob.lib().or_batch_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 21, 0, C.byref(mi), 0,
                                 n_cyc, tr, 20 * n_cyc)
    w = np.array([t.rosen for t in tr[:n]])
    assert all(t.move == 4 for t in tr[:n])
    want, want_se = _boltzmann_average(s)
    se = w.std() / math.sqrt(len(w))
    assert 0.2 < want < 50.0     # the framework matters, and the integral is well sampled
    assert abs(w.mean() - want) < 4.0 * math.hypot(se, want_se), (w.mean(), se, want, want_se)


def test_gcmc_of_non_interacting_chains_is_poisson_with_the_widom_mean(oracle_lib):
    from oracle import binding as ob
    s, _, _ = _setup(n_fw=60, P=2e6)
    _no_guest_guest(s)
    o = ob.OracleSystem(s, capacity=256)
    mc = ob.make_mc(s, o, production=1)
    _, skf, ska, _, _, _ = o.ewald_total()
    want_w, want_se = _boltzmann_average(s)
    z = s.beta * s.pressure * s.volume * want_w
    assert 2.0 < z < 20.0
    n_cyc = 5000
    tr = (ob.OrTrace * 256)()
    mi = C.c_int64(0)
    nm = np.zeros(n_cyc)
    for cyc in range(n_cyc):
        ob.lib().or_batch_cycles(C.byref(mc), o.box, o.comps, o.ncomp, ob._p(ska), ob._p(skf), 33, 0, C.byref(mi),
                                 cyc, 1, tr, 256)
        nm[cyc] = o.comps[1].n_mol
    x = nm[300:]
    blocks = x[: len(x) // 20 * 20].reshape(20, -1).mean(1)
    err = blocks.std(ddof=1) / math.sqrt(20)
    assert mc.ins_acc[1] > 2000 and mc.del_acc[1] > 2000 and mc.trans_acc[1] > 0
    tol = 4.0 * math.hypot(err, z * want_se / want_w)
    assert abs(x.mean() - z) < tol, (x.mean(), z, err)
    assert abs(x.var() / x.mean() - 1.0) < 0.15, (x.var(), x.mean())
