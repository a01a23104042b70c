"""Ewald periodic gravity in the oracle, pinned by what the mathematics fixes (the reference
needs the GSL library for periodic gravity, Ewald.cpp:103-108, and this image has none, so no
reference run can be made): the simple-cubic lattice constant and the uniform background's
harmonic term near the origin, the half-box symmetry of the full periodic field, independence of
the forces from the real/Fourier splitting parameter, the interpolation and wrap rules of the
lookup, and a perfect lattice's vanishing periodic force.  CPU only."""
import numpy as np
import pytest

import oracle as O

# potential of a unit mass in a simple-cubic periodic lattice with a neutralising background,
# minus 1/r, at r -> 0 (the Ewald/Madelung constant of the simple-cubic lattice)
XI_SC = -2.837297479480620


def _periodic_params(L=1.0):
    prm = {"kernel": "m4", "tabulated_kernel": 0, "self_gravity": 1}
    for k in range(3):
        prm[f"boundary_lhs[{k}]"] = prm[f"boundary_rhs[{k}]"] = "periodic"
        prm[f"boxmin[{k}]"] = 0.0
        prm[f"boxmax[{k}]"] = L
    return O.params_from_dict(prm)


@pytest.mark.parametrize("L", [1.0, 2.5])
def test_potential_near_origin_is_lattice_constant_plus_background_term(L):
    """table_pot(x) = (XI + pi/(4 mult^2))/L + (2 pi/3) x^2/L^3 + O(x^4): the lattice constant,
    the k = 0 constant the reference leaves out (pi/(alpha^2 L^3), alpha = 2 mult/L) and the
    uniform background's harmonic potential."""
    e = O.ewald_table([L, L, L])
    T = e.table
    dx = e.spacing[0]
    const = (XI_SC + np.pi / 4.0) / L
    for i in (1, 2):
        x = i * dx
        want = const + (2.0 * np.pi / 3.0) * x * x / L ** 3
        assert T[0, 0, 0, i] == pytest.approx(want, abs=1e-4 * i ** 4 / L)
    # the same along the other axes and the diagonal direction's harmonic term
    assert T[0, 0, 1, 0] == pytest.approx(T[0, 0, 0, 1], rel=1e-12)
    assert T[0, 1, 0, 0] == pytest.approx(T[0, 0, 0, 1], rel=1e-12)


def test_acceleration_near_origin_is_background_pull():
    """The images plus background pull a unit mass at x from its source with
    a = (4 pi/3) x/L^3 away from it, i.e. table_acc_x(|dr|) = -(4 pi/3)|dr|/L^3 (the caller
    multiplies by sgn(dr)); Richardson extrapolation removes the cubic term."""
    L = 1.0
    e = O.ewald_table([L, L, L])
    T = e.table
    x = e.spacing[0]
    lin = (8.0 * T[1, 0, 0, 1] - T[1, 0, 0, 2]) / 6.0
    assert lin == pytest.approx(-(4.0 * np.pi / 3.0) * x / L ** 3, rel=2e-4)
    # transverse components vanish on the axis
    assert abs(T[2, 0, 0, 3]) < 1e-12 and abs(T[3, 0, 0, 3]) < 1e-12


def test_full_periodic_field_vanishes_on_the_half_box_faces():
    """The full field of a unit mass and its images (table + the nearest term x/r^3) is odd and
    periodic, so its normal component is zero on every half-box face."""
    e = O.ewald_table([1.0, 1.0, 1.0])
    T = e.table
    h = e.spacing
    ih = int(round(0.5 / h[0]))
    for j in range(0, ih + 1, 3):
        for k in range(0, ih + 1, 4):
            x, y, z = ih * h[0], j * h[1], k * h[2]
            r3 = (x * x + y * y + z * z) ** 1.5
            assert T[1, k, j, ih] + x / r3 == pytest.approx(0.0, abs=1e-9)


def test_forces_do_not_depend_on_the_splitting_parameter():
    """Real-space and Fourier parts split at alpha = 2 mult/L: the forces must not depend on
    mult, the potential only through the left-out constant pi/(4 mult^2 L).  A wrong sign or
    factor in either series breaks this."""
    a = O.ewald_table([1.0, 1.0, 1.0], ewald_mult=1.0).table
    b = O.ewald_table([1.0, 1.0, 1.0], ewald_mult=1.6).table
    for c in (1, 2, 3):
        assert np.max(np.abs(a[c] - b[c])) < 1e-10 * np.max(np.abs(a[c]))
    d = (a[0] - b[0]).ravel()[1:]          # the origin node is zero in both
    assert np.allclose(d, np.pi / 4.0 * (1.0 - 1.0 / 1.6 ** 2), rtol=0, atol=1e-10)


def test_non_cubic_box_keeps_half_box_symmetry_and_splitting_independence():
    L = [1.0, 1.5, 0.8]
    a = O.ewald_table(L, ewald_mult=1.0)
    b = O.ewald_table(L, ewald_mult=1.4)
    for c in (1, 2, 3):
        assert np.max(np.abs(a.table[c] - b.table[c])) < 1e-8 * np.max(np.abs(a.table[c]))
    T, h = a.table, a.spacing
    iy = int(round(0.75 / h[1]))
    x, y, z = 2 * h[0], iy * h[1], 3 * h[2]
    r3 = (x * x + y * y + z * z) ** 1.5
    assert T[2, 3, iy, 2] + y / r3 == pytest.approx(0.0, abs=1e-8)


def test_lookup_interpolates_and_wraps():
    """CalculatePeriodicCorrection: at a node the correction is m times the node's values with
    the signs of dr; between nodes the acceleration is trilinear, the potential the lower
    corner's; a separation beyond half a box swaps its Newtonian term for the nearest image's."""
    e = O.ewald_table([1.0, 1.0, 1.0])
    T, h = e.table, e.spacing
    m = 0.7
    eps = 1.0 + 1e-9          # just above the node, so the lower corner is the node itself
    a, p = O.ewald_corr(e, m, [-3 * eps * h[0], 2 * eps * h[1], -5 * eps * h[2]])
    assert a == pytest.approx([-m * T[1, 5, 2, 3], m * T[2, 5, 2, 3], -m * T[3, 5, 2, 3]],
                              rel=1e-7)
    assert p == pytest.approx(m * T[0, 5, 2, 3], rel=1e-14)
    dr = np.array([3.25 * h[0], 2.5 * h[1], 5.75 * h[2]])
    a, p = O.ewald_corr(e, m, dr)
    w = np.array([0.25, 0.5, 0.75])
    acc = np.zeros(3)
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                wt = (w[0] if cx else 1 - w[0]) * (w[1] if cy else 1 - w[1]) * \
                     (w[2] if cz else 1 - w[2])
                acc += wt * T[1:, 5 + cz, 2 + cy, 3 + cx]
    assert a == pytest.approx(m * acc, rel=1e-12)
    assert p == pytest.approx(m * T[0, 5, 2, 3], rel=1e-12)
    # wrap: dr + L e_x gives the wrapped correction plus the Newtonian swap
    d1 = np.array([0.3, -0.1, 0.2])
    d2 = d1 - np.array([1.0, 0.0, 0.0])
    a1, p1 = O.ewald_corr(e, m, d1)
    a2, p2 = O.ewald_corr(e, m, d2)
    n1 = m * d1 / np.linalg.norm(d1) ** 3
    n2 = m * d2 / np.linalg.norm(d2) ** 3
    assert a2 + n2 == pytest.approx(a1 + n1, rel=1e-12, abs=1e-14)
    assert p2 + m / np.linalg.norm(d2) == pytest.approx(p1 + m / np.linalg.norm(d1), rel=1e-12)


def test_perfect_lattice_feels_no_periodic_force():
    """Every particle of a perfect cubic lattice in a periodic box sits at a symmetry point:
    the exact periodic force is zero.  The nearest-image sum alone leaves O(0.1) forces from
    the half-box cut; with the Ewald corrections only the table's interpolation error remains.
    Momentum is conserved pair by pair (the correction is odd)."""
    n = 8
    g = (np.arange(n) + 0.5) / n
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    N = len(r)
    m = np.full(N, 1.0 / N)
    p = _periodic_params()
    e = O.ewald_table([1.0, 1.0, 1.0])
    st = {"h": np.full(N, 0.1), "zeta": np.zeros(N), "hfactor": np.full(N, 1e4)}
    a, gp = O.direct_gravity_periodic(p, e, r, m, st)
    a0, _ = O.direct_gravity(p, r, m, st)
    assert np.abs(a).max() < 1e-3 * np.abs(a0).max()
    assert np.ptp(gp) < 1e-12 * np.abs(gp).max()
    rng = np.random.default_rng(3)
    rj = (r + 0.02 * rng.normal(size=r.shape)) % 1.0
    a, _ = O.direct_gravity_periodic(p, e, rj, m, st)
    assert np.abs((m[:, None] * a).sum(0)).max() < 1e-14
