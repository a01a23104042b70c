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
