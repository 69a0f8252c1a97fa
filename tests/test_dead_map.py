"""The certain-overlap voxel maps of the overlap screens (host computation).

A marked voxel must be one where every point has a framework pair that the
reference flags as an overlap: inside the VDW cutoff with an LJ energy above
OverlapCriteria, or closer than 0.1 A (VDW_Coulomb.cu:1299-1300).  Checked
with an independent numpy evaluation at random points and at the corners of
marked voxels, in a cubic and a triclinic cell; a map that tested the voxel
centre instead of its farthest corner, or used the wrong image or cell
transform, fails these checks."""
import numpy as np
import pytest


def _lj(system, ta, ts, rr):
    T = system.ntypes
    eps = np.asarray(system.eps).reshape(T, T)[ta, ts]
    sig = np.asarray(system.sig).reshape(T, T)[ta, ts]
    sh = np.asarray(system.shift).reshape(T, T)[ta, ts]
    x = (sig * sig / rr) ** 3
    return 4.0 * eps * (x * x - x) - sh


def _overlapping(system, pts, site_type):
    """For each point: does some framework pair overlap (reference rule)?"""
    cell = np.asarray(system.cell, dtype=float).reshape(3, 3).T     # columns a, b, c
    inv = np.linalg.inv(cell)
    fw = np.asarray(system.fw_pos, dtype=float)
    ty = np.asarray(system.fw_type)
    out = np.zeros(len(pts), dtype=bool)
    for k, p in enumerate(pts):
        d = fw - p
        s = d @ inv.T
        s -= np.round(s)
        d = s @ cell.T
        rr = (d * d).sum(1)
        inside = rr < system.cut_vdw ** 2
        e = _lj(system, ty, site_type, np.maximum(rr, 1e-12))
        out[k] = bool(np.any(inside & ((e > system.overlap) | (rr < 0.01))))
    return out


def _systems():
    from graspa_amd import synth
    cub = synth.config1()
    tri = synth.config1(n_fw=400)
    tri.cell = tri.cell.copy()
    tri.cell[3], tri.cell[6], tri.cell[7] = 3.5, -2.0, 1.5
    tri.setup_ewald()
    return [("cubic", cub), ("triclinic", tri)]


@pytest.mark.parametrize("name", ["cubic", "triclinic"])
def test_marked_voxels_overlap_everywhere(name):
    from graspa_amd import engine
    s = dict(_systems())[name]
    cell = np.asarray(s.cell, dtype=float).reshape(3, 3).T
    rng = np.random.default_rng(3)
    for site_type in (3, 4):
        m, nv = engine.dead_map(s, site_type)
        marked = np.argwhere(m)                 # (z, y, x)
        assert len(marked) > 0.1 * m.size       # dense frameworks: a large part of the volume
        pick = marked[rng.choice(len(marked), 300, replace=False)]
        idx = pick[:, ::-1].astype(float)       # (x, y, z)
        nva = np.array(nv, dtype=float)
        pts = []
        for corner in range(8):                 # every corner of each picked voxel
            off = np.array([(corner >> a) & 1 for a in range(3)], dtype=float)
            pts.append((idx + off) / nva)
        pts.append((idx + rng.uniform(0, 1, idx.shape)) / nva)
        frac = np.concatenate(pts)
        cart = frac @ cell.T
        assert _overlapping(s, cart, site_type).all()


def test_unmarked_space_exists_and_map_depends_on_type():
    from graspa_amd import engine, synth
    s = synth.config1(n_fw=300)
    m3, _ = engine.dead_map(s, 3)
    m4, _ = engine.dead_map(s, 4)
    assert 0.05 < m3.mean() < 0.95
    assert not np.array_equal(m3, m4)


@pytest.mark.parametrize("name", ["cubic", "triclinic"])
def test_free_voxels_hold_no_close_pair(name):
    """A certainly-free voxel (the first-bead screen skips trials there) has no
    point within the overlap distance of any framework atom: checked at the
    corners and random points of free voxels with an independent numpy
    minimum-image distance against the closed-form overlap radius; the free
    and certain-overlap maps never share a voxel.  A map that tested the voxel
    centre (or dropped the margin) would put points near a voxel face inside
    the overlap radius."""
    from graspa_amd import engine
    s = dict(_systems())[name]
    cell = np.asarray(s.cell, dtype=float).reshape(3, 3).T
    inv = np.linalg.inv(cell)
    fw = np.asarray(s.fw_pos, dtype=float)
    ty = np.asarray(s.fw_type)
    T = s.ntypes
    rng = np.random.default_rng(5)
    for site_type in (3, 4):
        f, nv = engine.free_map(s, site_type)
        d, nv2 = engine.dead_map(s, site_type)
        assert nv == nv2 and not np.any(f & d)
        free = np.argwhere(f)
        assert len(free) > 0.01 * f.size
        # overlap radius per framework type: E_LJ(r) = OverlapCriteria or r = 0.1 A
        eps = np.asarray(s.eps).reshape(T, T)[:, site_type]
        sig = np.asarray(s.sig).reshape(T, T)[:, site_type]
        sh = np.asarray(s.shift).reshape(T, T)[:, site_type]
        c = (s.overlap + sh) / (4.0 * eps)
        r2_ov = np.maximum(0.01, sig * sig / np.cbrt(0.5 * (1.0 + np.sqrt(1.0 + 4.0 * c))))
        pick = free[rng.choice(len(free), 300, replace=False)][:, ::-1].astype(float)
        nva = np.array(nv, dtype=float)
        pts = [(pick + np.array([(k >> a) & 1 for a in range(3)], dtype=float)) / nva for k in range(8)]
        pts.append((pick + rng.uniform(0, 1, pick.shape)) / nva)
        cart = np.concatenate(pts) @ cell.T
        for p in cart:
            dd = fw - p
            sf = dd @ inv.T
            sf -= np.round(sf)
            dd = sf @ cell.T
            rr = (dd * dd).sum(1)
            assert np.all(rr >= r2_ov[ty]), (p, rr.min())
        assert not _overlapping(s, cart[:200], site_type).any()
