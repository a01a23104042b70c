"""The gravity walk's relative opening criteria (the reference's "gravity_mac" = gadget2 or eigenmac,
Tree.h:413-432, KDTree.cpp:900-903, 1058-1079) in the oracle, pinned by what the mathematics
fixes: the gadget2 test's threshold distance d^4 = m rmax^2 / (macerror |a|); the eigenmac test's
lambda = 2 sqrt(-p/3) bounds the quadrupole's largest eigenvalue magnitude and equals it for an
axisymmetric quadrupole, and its threshold distance is (0.5 lambda / macerror)^(2/3) times the
group factor; in the twin's walk a zero acceleration opens every cell (the exact sum), a huge
macerror gives the geometric walk back, and the error falls as macerror does.  CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden, vec


def _p(mac, macerror=1e-4):
    return O.params_from_dict({"gravity_mac": mac, "macerror": macerror})


def test_gadget2_threshold():
    p = _p("gadget2", 1e-3)
    half = np.array([0.1, 0.2, 0.05])
    m, amin = 2.0, 0.7
    rmax2 = (half ** 2).sum()
    d4 = m * rmax2 / (p.macerror * amin)
    d2 = np.sqrt(d4)
    q = np.zeros(5)
    assert O.cell_mac_opens(p, half, m, q, d2 * (1 - 1e-9), amin)
    assert not O.cell_mac_opens(p, half, m, q, d2 * (1 + 1e-9), amin)
    # twice the mass or half the acceleration moves the threshold by 2^(1/4) in distance
    assert O.cell_mac_opens(p, half, 2 * m, q, d2 * 1.4, amin)
    assert not O.cell_mac_opens(p, half, 2 * m, q, d2 * 1.42, amin)
    assert O.cell_mac_opens(p, half, m, q, d2 * 1.4, 0.5 * amin)
    # a zero acceleration (the state before any force) opens every massive cell
    assert O.cell_mac_opens(p, half, m, q, 1e30, 0.0)


def _lambda_from_threshold(p, q, gfac=1.0):
    """The eigenmac threshold distance^2 found by bisection, turned back into lambda."""
    lo, hi = 0.0, 1e12
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if O.cell_mac_opens(p, np.ones(3), 1.0, q, mid, gfac):
            lo = mid
        else:
            hi = mid
    return 2.0 * p.macerror * (lo / gfac) ** 1.5


def _qmatrix(q):
    return np.array([[q[0], q[1], q[3]], [q[1], q[2], q[4]], [q[3], q[4], -(q[0] + q[2])]])


def test_eigenmac_lambda_bounds_the_largest_eigenvalue():
    p = _p("eigenmac", 1e-3)
    # axisymmetric: eigenvalues (a, a, -2a), lambda = 2|a| exactly
    for a in (0.3, -0.8):
        q = np.array([a, 0.0, a, 0.0, 0.0])
        assert _lambda_from_threshold(p, q) == pytest.approx(2 * abs(a), rel=1e-9)
    rng = np.random.default_rng(9)
    for _ in range(20):
        q = rng.normal(size=5)
        lam = _lambda_from_threshold(p, q)
        ev = np.linalg.eigvalsh(_qmatrix(q))
        assert lam >= np.abs(ev).max() * (1 - 1e-9)
        # -p = tr(Q^2)/2, so lambda = 2 sqrt(tr(Q^2)/6)
        assert lam == pytest.approx(2 * np.sqrt((ev ** 2).sum() / 6), rel=1e-9)
    # the group factor scales the threshold distance^2 linearly
    q = np.array([0.5, 0.1, -0.2, 0.3, 0.05])
    assert _lambda_from_threshold(p, q, 3.0) == pytest.approx(_lambda_from_threshold(p, q),
                                                              rel=1e-9)
    # a cell without a quadrupole (lambda = 0) is never opened by it
    assert not O.cell_mac_opens(p, np.ones(3), 1.0, np.zeros(5), 1e-30, 5.0)


def _walk(d, mac, macerror, prev=True, **over):
    p = O.params_from_dict(dict(d["meta"]["params"], gravity_mac=mac, macerror=macerror, **over))
    st = {"h": d["s0_h"], "zeta": d["s0_zeta"], "hfactor": d["s0_hfactor"]}
    tw = O.TreeTwin(p, d["in_r"], d["in_m"], d["s0_h"])
    ap = vec(d, "s0", ("atx", "aty", "atz")) if prev else None
    gp = d["s0_gpot"] if prev else None
    return tw.gravity(d["in_r"], d["in_m"], st, ap, gp)


@pytest.mark.parametrize("mac", ["gadget2", "eigenmac"])
def test_relative_criteria_in_the_walk(mac):
    """With the reference's exact accelerations and potentials as the stocked state: a huge
    macerror reproduces the geometric walk bit for bit, smaller ones open more and the error
    against the exact sum falls; gadget2 keeps the RMS error below macerror."""
    d = load_golden("sphere4k_m4_bf")
    exact = vec(d, "s0", ("atx", "aty", "atz"))
    geo = _walk(d, "geometric", 1e-4)
    big = _walk(d, mac, 1e30)
    for k in range(4):
        assert np.array_equal(big[k], geo[k])
    prev = None
    for me in (1e-2, 1e-3, 1e-4, 1e-5):
        at, _, nc, nn = _walk(d, mac, me)
        e = np.linalg.norm(at - exact, axis=1) / np.linalg.norm(exact, axis=1)
        rms = np.sqrt(np.mean(e ** 2))
        if prev is not None:
            assert rms <= prev * 1.0001
            assert nn.sum() >= prev_nn
        if mac == "gadget2":
            assert rms < me
        prev, prev_nn = rms, nn.sum()
    assert prev < 0.05 * np.sqrt(np.mean((np.linalg.norm(geo[0] - exact, axis=1)
                                          / np.linalg.norm(exact, axis=1)) ** 2))


def test_gadget2_from_a_zero_state_is_the_exact_sum():
    """Before any force (atree = 0) the gadget2 walk opens every cell: the same forces as the
    walk with every cell opened geometrically."""
    d = load_golden("sphere4k_m4_bf")
    at, gp, nc, nn = _walk(d, "gadget2", 1e-4, prev=False)
    assert nc.sum() == 0
    at0, gp0, _, nn0 = _walk(d, "geometric", 1e-4, thetamaxsqd=1e-30)
    assert np.array_equal(nn, nn0)
    assert np.array_equal(at, at0) and np.array_equal(gp, gp0)


@pytest.mark.parametrize("mac", ["gadget2", "eigenmac"])
def test_sim_setup_takes_two_force_passes(mac):
    """SphSimulation::PostInitialConditionsSetup with a relative criterion: forces from the
    geometric walk first, then the tree stocked with them and the forces recomputed.  The
    second pass's forces are closer to the exact sum than the geometric walk's."""
    d = load_golden("sphere4k_m4_bf")
    prm = dict(d["meta"]["params"], neib_search="kdtree")
    out = {}
    for m in ("geometric", mac):
        s = O.Sim(O.params_from_dict(dict(prm, gravity_mac=m)), d["in_r"], d["in_v"],
                  d["in_m"], d["in_u"])
        s.setup()
        out[m] = (s.get("atree"), s.get("h"))
    exact = vec(d, "s0", ("atx", "aty", "atz"))
    assert np.array_equal(out[mac][1], out["geometric"][1])      # no second h-rho pass
    err = {m: np.sqrt(np.mean((np.linalg.norm(out[m][0] - exact, axis=1)
                               / np.linalg.norm(exact, axis=1)) ** 2)) for m in out}
    assert err[mac] < 0.5 * err["geometric"]
