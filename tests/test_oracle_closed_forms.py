"""Pins of the oracle's functions that the golden fixtures leave open, by what the mathematics
fixes: kernel normalisation, derivative and potential relations, continuity at the knots and the
Newtonian limits (SmoothingKernel.h:132-408); brute-force neighbour sets and tree moments; the
N-body direct sums' two-body closed form, Newton's third law and their Newtonian limit."""
import numpy as np
import pytest
from scipy import integrate

import oracle as O

W0, W1, WOMEGA, WZETA, WGRAV, WPOT, W0S2, WOMEGAS2, WZETAS2 = range(9)
KERNELS = [("m4", 2.0), ("quintic", 3.0)]


def params(kernel, tab=0):
    return O.params_from_dict({"kernel": kernel, "tabulated_kernel": tab})


def K(p, w, s):
    return O.kernel_eval(p, w, float(s))


@pytest.mark.parametrize("kernel,rng", KERNELS)
def test_kernel_normalised(kernel, rng):
    p = params(kernel)
    val, _ = integrate.quad(lambda s: 4.0 * np.pi * s * s * K(p, W0, s), 0.0, rng,
                            points=[1.0, 2.0], epsabs=1e-13, epsrel=1e-13, limit=200)
    # the reference's 3D norms use its truncated 1/pi (Constants.h), so 1 to ~1e-14
    assert val == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("kernel,rng", KERNELS)
def test_kernel_derivative_and_potential_relations(kernel, rng):
    """w1 = dw0/ds, wgrav = -dwpot/ds, womega = -3 w0 - s w1, wzeta = wpot - s wgrav."""
    p = params(kernel)
    eps = 1e-6
    for s in np.linspace(0.05, rng - 0.05, 37):
        d0 = (K(p, W0, s + eps) - K(p, W0, s - eps)) / (2 * eps)
        assert K(p, W1, s) == pytest.approx(d0, rel=1e-6, abs=1e-9)
        dp = (K(p, WPOT, s + eps) - K(p, WPOT, s - eps)) / (2 * eps)
        assert K(p, WGRAV, s) == pytest.approx(-dp, rel=1e-6, abs=1e-9)
        assert K(p, WOMEGA, s) == pytest.approx(-3 * K(p, W0, s) - s * K(p, W1, s),
                                                rel=1e-12, abs=1e-12)
        # wzeta = wpot - s wgrav for M4; the reference's quintic wzeta (SmoothingKernel.h
        # quintic section) lacks the 12/359 factor of its wgrav/wpot and has the opposite
        # sign, i.e. carries -359/12 relative to that convention; kept for parity.  Its
        # constants are given to 8-11 digits (about 1e-7 of its scale of 34).
        zrel = K(p, WPOT, s) - s * K(p, WGRAV, s)
        tol = 1e-12
        if kernel == "quintic":
            zrel *= -359.0 / 12.0
            tol = 3e-6
        assert K(p, WZETA, s) == pytest.approx(zrel, rel=0, abs=tol)


@pytest.mark.parametrize("kernel,rng", KERNELS)
def test_kernel_knots_and_newtonian_limit(kernel, rng):
    p = params(kernel)
    knots = [1.0, 2.0] + ([3.0] if rng == 3.0 else [])
    for s in knots:
        for w in (W0, W1, WGRAV, WPOT):
            lo, hi = K(p, w, s - 1e-12), K(p, w, s + 1e-12)
            if kernel == "quintic" and s == 3.0 and w in (WGRAV, WPOT):
                # the reference's quintic gravity kernels carry the factor 12/359 where 1/30
                # would make them continuous: inside the kernel edge they are 360/359 of the
                # Newtonian value (kept for parity, pinned to 1e-12 by the bruteforce fixture)
                assert lo == pytest.approx(hi * 360.0 / 359.0, rel=1e-10), (w, s)
                continue
            assert lo == pytest.approx(hi, rel=1e-8, abs=1e-8), (w, s)
    for s in (rng, rng + 0.5, 7.0):
        assert K(p, WGRAV, s) == pytest.approx(1.0 / s ** 2, rel=1e-12)
        assert K(p, WPOT, s) == pytest.approx(1.0 / s, rel=1e-12)
        assert K(p, W0, s) == 0.0 and K(p, W1, s) == 0.0


@pytest.mark.parametrize("kernel,rng", KERNELS)
def test_tabulated_kernel_is_nearest_lower_entry(kernel, rng):
    """SmoothingKernel.h:611-657: entry k holds the analytic value at s = k*range/1000 and a
    lookup returns the entry below s (no interpolation)."""
    pa, pt = params(kernel), params(kernel, tab=1)
    for k in (0, 1, 137, 500, 999):
        s0 = k * rng / 1000.0
        for frac in (0.1, 0.5, 0.9):
            s = (k + frac) * rng / 1000.0
            for w in (W0, W1, WGRAV, WPOT):
                if k == 0 and w == WGRAV:
                    continue
                assert K(pt, w, s) == pytest.approx(K(pa, w, s0), rel=1e-12, abs=1e-14)
    assert K(pt, W0, rng) == 0.0


def test_neighbour_sets_brute_force():
    rng = np.random.default_rng(1)
    N = 400
    r = rng.random((N, 3))
    h = 0.03 + 0.05 * rng.random(N)
    p = params("m4")
    d2 = ((r[:, None, :] - r[None, :, :]) ** 2).sum(-1)
    kr2 = (2.0 * h) ** 2
    gather = d2 < kr2[:, None]
    sg = gather | (d2 < kr2[None, :])
    for mode, ref in ((0, gather), (1, sg)):
        off, ids = O.neighbour_sets(p, r, h, mode)
        for i in range(N):
            got = np.sort(ids[off[i]:off[i + 1]])
            assert np.array_equal(got, np.flatnonzero(ref[i])), (mode, i)


def test_tree_twin_moments_brute_force():
    rng = np.random.default_rng(2)
    N = 300
    r = rng.random((N, 3))
    m = 0.5 + rng.random(N)
    h = 0.01 + 0.02 * rng.random(N)
    t = O.TreeTwin(params("m4"), r, m, h).arrays()
    perm = t["perm"]
    rs, ms, hs = r[perm], m[perm], h[perm]
    for n in range(N - 1):
        f, l = t["first"][n], t["last"][n]
        mm = ms[f:l + 1]
        com = (mm[:, None] * rs[f:l + 1]).sum(0) / mm.sum()
        d = rs[f:l + 1] - com
        q = np.array([(mm * (3 * d[:, 0] ** 2 - (d ** 2).sum(1))).sum(),
                      (mm * 3 * d[:, 0] * d[:, 1]).sum(),
                      (mm * (3 * d[:, 1] ** 2 - (d ** 2).sum(1))).sum(),
                      (mm * 3 * d[:, 2] * d[:, 0]).sum(),
                      (mm * 3 * d[:, 2] * d[:, 1]).sum()])
        assert t["mass"][n] == pytest.approx(mm.sum(), rel=1e-13)
        assert np.allclose(t["com"][n], com, rtol=0, atol=1e-13)
        assert np.allclose(t["quad"][n], q, rtol=1e-9, atol=1e-12)
        assert np.array_equal(t["bbmin"][n], rs[f:l + 1].min(0))
        assert np.array_equal(t["bbmax"][n], rs[f:l + 1].max(0))
        assert t["hmax"][n] == hs[f:l + 1].max()


def test_nbody_direct_two_body_closed_form():
    r = np.array([[0.1, -0.2, 0.3], [0.4, 0.2, 0.3]])
    v = np.array([[0.0, 0.1, 0.0], [0.3, -0.1, 0.2]])
    m = np.array([2.0, 3.0])
    a, ad, gp = O.nbody_direct(r, v, m)
    dr = r[1] - r[0]
    d = np.linalg.norm(dr)
    rh = dr / d
    dv = v[1] - v[0]
    assert np.allclose(a[0], m[1] * rh / d ** 2, rtol=1e-14, atol=0)
    assert np.allclose(a[1], -m[0] * rh / d ** 2, rtol=1e-14, atol=0)
    assert gp[0] == pytest.approx(m[1] / d, rel=1e-14)
    assert np.allclose(ad[0], m[1] * (dv - 3 * np.dot(dv, rh) * rh) / d ** 3, rtol=1e-13,
                       atol=1e-15)


def test_nbody_direct_third_law_and_newtonian_limit():
    rng = np.random.default_rng(4)
    N = 64
    r = rng.normal(size=(N, 3))
    v = rng.normal(size=(N, 3))
    m = 0.1 + rng.random(N)
    a, ad, gp = O.nbody_direct(r, v, m)
    scale = np.abs(m[:, None] * a).max()
    assert np.abs((m[:, None] * a).sum(0)).max() < 1e-13 * N * scale
    assert np.abs((m[:, None] * ad).sum(0)).max() < 1e-12 * N * np.abs(m[:, None] * ad).max()
    # kernel-smoothed sums equal the point-mass sums once every pair is beyond the kernel
    p = params("m4")
    h = np.full(N, 1e-3)
    a2, ad2, gp2 = O.nbody_direct_smoothed(p, r, v, m, h)
    assert np.allclose(a2, a, rtol=1e-12, atol=0)
    assert np.allclose(gp2, gp, rtol=1e-12, atol=0)
    assert np.allclose(ad2, ad, rtol=1e-9, atol=1e-12)


def test_direct_plummer_two_body_closed_form():
    """orc_direct_plummer on two and three bodies: the Plummer-softened pair acceleration
    m dr / (r^2 + eps^2)^(3/2) and potential -m / sqrt(r^2 + eps^2), superposed over the
    sources, with the target itself excluded."""
    r = np.array([[0.1, -0.2, 0.3], [0.4, 0.2, 0.35], [-0.5, 0.05, 0.0]])
    m = np.array([2.0, 3.0, 0.7])
    for eps in (0.0, 0.01, 0.3):
        a, pot = O.direct_plummer(r[:2], m[:2], eps, [0, 1])
        dr = r[1] - r[0]
        s2 = dr @ dr + eps * eps
        assert np.allclose(a[0], m[1] * dr / s2 ** 1.5, rtol=1e-14, atol=0)
        assert np.allclose(a[1], -m[0] * dr / s2 ** 1.5, rtol=1e-14, atol=0)
        assert pot[0] == pytest.approx(-m[1] / np.sqrt(s2), rel=1e-14)
        assert pot[1] == pytest.approx(-m[0] / np.sqrt(s2), rel=1e-14)
        a3, pot3 = O.direct_plummer(r, m, eps, [2])
        want_a = sum(m[j] * (r[j] - r[2]) / ((r[j] - r[2]) @ (r[j] - r[2]) + eps * eps) ** 1.5
                     for j in (0, 1))
        want_p = sum(-m[j] / np.sqrt((r[j] - r[2]) @ (r[j] - r[2]) + eps * eps) for j in (0, 1))
        assert np.allclose(a3[0], want_a, rtol=1e-14, atol=0)
        assert pot3[0] == pytest.approx(want_p, rel=1e-14)


def test_direct_plummer_unsoftened_limit_is_nbody_direct():
    """eps -> 0: the Plummer sum becomes the unsoftened star-star sum, whose oracle is pinned
    against the reference (Nbody::CalculateDirectGravForces fixture); the softened sum departs
    from it as O(eps^2 / r^2)."""
    rng = np.random.default_rng(5)
    N = 200
    r = rng.normal(size=(N, 3))
    m = 0.1 + rng.random(N)
    a0, _, gp0 = O.nbody_direct(r, np.zeros((N, 3)), m)
    t = np.arange(N)
    a, pot = O.direct_plummer(r, m, 0.0, t)
    assert np.allclose(a, a0, rtol=1e-12, atol=1e-14 * np.abs(a0).max())
    assert np.allclose(-pot, gp0, rtol=1e-12, atol=0)
    d2min = min(((r[i] - r[j]) ** 2).sum() for i in range(N) for j in range(i))
    prev = None
    for eps in (1e-2, 1e-3, 1e-4):
        a, pot = O.direct_plummer(r, m, eps, t)
        err = np.abs(-pot - gp0).max() / np.abs(gp0).max()
        assert err < 2.0 * eps * eps / d2min
        if prev is not None:
            assert err < prev * 0.02       # quadratic in eps
        prev = err


def _cluster_exact(rs, ms, rp):
    dr = rs - rp
    d = np.sqrt((dr ** 2).sum(1))
    return (ms[:, None] * dr / d[:, None] ** 3).sum(0), (ms / d).sum()


@pytest.mark.parametrize("kind,order", [("generic", 3), ("symmetric_pair", 4)])
def test_cell_quadrupole_far_field_order(kind, order):
    """The quadrupole cell term (NeighbourSearch.h:423-475) against the exact sum over the
    cell's members: the monopole alone errs as (d/R)^2; adding the quadrupole leaves the
    octupole, (d/R)^3, or (d/R)^4 for a symmetric pair whose octupole vanishes.  A dropped
    term, a wrong sign or a transposed moment leaves a (d/R)^2 error and fails the slope."""
    rng = np.random.default_rng(6)
    if kind == "generic":
        rs = rng.normal(scale=0.1, size=(12, 3)) + np.array([0.3, -0.2, 0.1])
        ms = 0.5 + rng.random(12)
    else:
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        rs = np.array([0.3, -0.2, 0.1]) + np.outer([0.1, -0.1], axis)
        ms = np.array([1.5, 1.5])
    # the moments as the tree twin builds them for the node holding every member (the root),
    # pinned against brute force in test_tree_twin_moments_brute_force
    tw = O.TreeTwin(params("m4"), rs, ms, np.full(len(ms), 1e-3)).arrays()
    com, q, mc = tw["com"][0], tw["quad"][0], tw["mass"][0]
    assert mc == pytest.approx(ms.sum(), rel=1e-14)
    size = np.sqrt((((rs - com) ** 2).sum(1)).max())
    dirs = rng.normal(size=(16, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    Rs = size * np.array([8.0, 16.0, 32.0, 64.0])
    eq, ep, em = [], [], []
    for R in Rs:
        wq, wp, wm = 0.0, 0.0, 0.0
        for u in dirs:
            rp = com + R * u
            ax, px = _cluster_exact(rs, ms, rp)
            aq, pq = O.cell_quad_force(com, q, mc, rp)
            am, _ = O.cell_quad_force(com, np.zeros(5), mc, rp)
            # the cell term accumulates the acceleration the cell exerts on the point
            wq = max(wq, np.linalg.norm(aq - ax) / np.linalg.norm(ax))
            wp = max(wp, abs(pq - px) / px)
            wm = max(wm, np.linalg.norm(am - ax) / np.linalg.norm(ax))
        eq.append(wq); ep.append(wp); em.append(wm)
    slope = lambda e: np.polyfit(np.log(Rs), np.log(e), 1)[0]
    assert slope(em) == pytest.approx(-2.0, abs=0.25)
    assert slope(eq) < -(order - 0.2)
    assert slope(ep) < -(order - 0.2)
    assert eq[0] < 0.2 * em[0]
