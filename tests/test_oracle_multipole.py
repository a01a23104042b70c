"""The gravity walk's cell-term options (the reference's "multipole" parameter, Parameters.cpp:309)
in the oracle, pinned by what the mathematics fixes: the monopole term is the point-mass force of
the cell's centre of mass and errs as (size/d)^2 against its members' exact sum; the fast
expansion (FastMultipoleForces, NeighbourSearch.h:478-760, first order) reproduces the per-target
cell terms at its expansion point, its tensor is the gradient of their field (symmetric, traceless
outside the sources) and its potential slope the field itself, and its error at a distance delta
from the expansion point falls as delta^2.  The tree twin's walk with each option stays within the
tree's error of the exact sum, and with every cell opened all options give the same forces.
CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden, vec


def _cells(seed=4, n=7):
    """Cells at distances 4-9 from the origin (the expansion point) with generic moments."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    com = u * rng.uniform(4.0, 9.0, size=(n, 1))
    q = rng.normal(scale=0.05, size=(n, 5))
    m = rng.uniform(0.5, 2.0, size=n)
    return com, q, m


def _per_target(com, q, m, rp, quad):
    a, p = np.zeros(3), 0.0
    for c in range(len(m)):
        if quad:
            ac, pc = O.cell_quad_force(com[c], q[c], m[c], rp)
        else:
            ac, pc = O.cell_mono_force(com[c], m[c], rp)
        a += ac
        p += pc
    return a, p


def _tensor(coef):
    xx, xy, yy, xz, yz, zz = coef[4:]
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def test_monopole_term_is_point_mass_force():
    """ComputeCellMonopoleForces: the pull of the cell's mass at its centre of mass; the same as
    the quadrupole term with zero moments; against a cluster's exact sum it errs as (size/d)^2."""
    com = np.array([0.3, -0.2, 0.5])
    rp = np.array([-1.0, 0.4, 2.0])
    a, p = O.cell_mono_force(com, 2.5, rp)
    dr = com - rp
    d = np.linalg.norm(dr)
    assert a == pytest.approx(2.5 * dr / d ** 3, rel=1e-15)
    assert p == pytest.approx(2.5 / d, rel=1e-15)
    aq, pq = O.cell_quad_force(com, np.zeros(5), 2.5, rp)
    assert aq == pytest.approx(a, rel=1e-15) and pq == pytest.approx(p, rel=1e-15)
    rng = np.random.default_rng(8)
    rs = rng.normal(scale=0.1, size=(10, 3))
    ms = 0.5 + rng.random(10)
    c = (ms[:, None] * rs).sum(0) / ms.sum()
    errs = []
    Rs = np.array([2.0, 4.0, 8.0, 16.0])
    for R in Rs:
        rp = c + R * np.array([0.6, -0.48, 0.64])
        dr = rs - rp
        ex = (ms[:, None] * dr / np.linalg.norm(dr, axis=1)[:, None] ** 3).sum(0)
        am, _ = O.cell_mono_force(c, ms.sum(), rp)
        errs.append(np.linalg.norm(am - ex) / np.linalg.norm(ex))
    assert np.polyfit(np.log(Rs), np.log(errs), 1)[0] == pytest.approx(-2.0, abs=0.2)


@pytest.mark.parametrize("quad", [0, 1])
def test_fast_expansion_at_its_point_is_the_per_target_sum(quad):
    com, q, m = _cells()
    coef, a, p = O.fast_multipole(com, q, m, quad, np.zeros((1, 3)))
    ae, pe = _per_target(com, q, m, np.zeros(3), quad)
    assert a[0] == pytest.approx(ae, rel=1e-13)
    assert p[0] == pytest.approx(pe, rel=1e-13)
    assert coef[1:4] == pytest.approx(ae, rel=1e-13)


@pytest.mark.parametrize("quad", [0, 1])
def test_fast_tensor_is_the_field_gradient(quad):
    """Central differences of the per-target field about the expansion point give the tensor
    column by column; the potential's slope is the acceleration (a = grad gpot); the tensor is
    traceless (the field is harmonic outside the cells).  A wrong sign, a swapped component or a
    dropped quadrupole update fails one of these."""
    com, q, m = _cells(seed=11)
    coef, _, _ = O.fast_multipole(com, q, m, quad, np.zeros((1, 3)))
    T = _tensor(coef)
    eps = 1e-4
    for k in range(3):
        e = np.zeros(3)
        e[k] = eps
        ap, pp = _per_target(com, q, m, e, quad)
        am, pm = _per_target(com, q, m, -e, quad)
        assert (ap - am) / (2 * eps) == pytest.approx(T[:, k], rel=1e-6, abs=1e-8 * np.abs(T).max())
        assert (pp - pm) / (2 * eps) == pytest.approx(coef[1 + k], rel=1e-7)
    assert abs(np.trace(T)) < 1e-12 * np.abs(T).max()


@pytest.mark.parametrize("quad", [0, 1])
def test_fast_series_error_is_second_order(quad):
    com, q, m = _cells(seed=5)
    rng = np.random.default_rng(2)
    dirs = rng.normal(size=(8, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    deltas = np.array([0.4, 0.2, 0.1, 0.05])
    errs = []
    for dl in deltas:
        pts = dl * dirs
        _, a, _ = O.fast_multipole(com, q, m, quad, pts)
        w = 0.0
        for i, x in enumerate(pts):
            ae, _ = _per_target(com, q, m, x, quad)
            w = max(w, np.linalg.norm(a[i] - ae) / np.linalg.norm(ae))
        errs.append(w)
    assert np.polyfit(np.log(deltas), np.log(errs), 1)[0] == pytest.approx(2.0, abs=0.15)


MODES = ["quadrupole", "monopole", "fast_monopole", "fast_quadrupole"]


def _twin_gravity(d, mode, **over):
    p = O.params_from_dict(dict(d["meta"]["params"], multipole=mode, **over))
    st = {"h": d["s0_h"], "zeta": d["s0_zeta"], "hfactor": d["s0_hfactor"]}
    tw = O.TreeTwin(p, d["in_r"], d["in_m"], d["s0_h"])
    return tw.gravity(d["in_r"], d["in_m"], st)


def test_walk_options_within_tree_error():
    """The reference's exact kernel-softened sum (its bruteforce run) against the tree twin with
    each cell-term option: quadrupole below monopole, the fast (first-order Taylor) options at
    the monopole's level."""
    d = load_golden("sphere4k_m4_bf")
    exact = vec(d, "s0", ("atx", "aty", "atz"))
    rms = {}
    for mode in MODES:
        at, _, nc, _ = _twin_gravity(d, mode)
        e = np.linalg.norm(at - exact, axis=1) / np.linalg.norm(exact, axis=1)
        rms[mode] = np.sqrt(np.mean(e ** 2))
        assert nc.sum() > 0
    assert rms["quadrupole"] < 5e-3          # opening angle thetamaxsqd = 0.25 here
    assert rms["quadrupole"] < 0.2 * rms["monopole"]
    assert rms["fast_quadrupole"] < rms["fast_monopole"]
    for mode in MODES[1:]:
        assert rms[mode] < 2.5e-2, (mode, rms[mode])


def test_walk_options_agree_with_every_cell_opened():
    d = load_golden("sphere4k_m4_bf")
    ref = _twin_gravity(d, "quadrupole", thetamaxsqd=1e-30)
    assert ref[2].sum() == 0
    for mode in MODES[1:]:
        at, gp, nc, nn = _twin_gravity(d, mode, thetamaxsqd=1e-30)
        assert np.array_equal(at, ref[0]) and np.array_equal(gp, ref[1]), mode
        assert np.array_equal(nn, ref[3])
