"""Stars beside the gas: the oracle's star-gas restatements pinned against the compiled
reference's own gas + star runs (hybrid1k_soft: star sums kernel-smoothed, nbody_softening = 1;
hybrid1k_hard: point-mass star-star sums, nbody_softening = 0; both with the brute-force search,
whose star-gas lists are exact, and the conservative SPH-star zeta terms on), and against the
invariants the star-gas pair terms obey.  CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden, vec

CASES = [("hybrid1k_soft", 1), ("hybrid1k_hard", 0)]


def relmax(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _stars(d, soft, prefix="star_s0"):
    return O.stars(vec(d, prefix), d["star_in_m"], d["star_in_h"], softening=soft)


@pytest.mark.parametrize("case,soft", CASES)
def test_star_sums_after_setup_match_reference(case, soft):
    """HydroTree::UpdateAllStarGasForces (brute-force lists: softened inside k (h_s + h_j)/2,
    point masses outside) plus the star-star sums of the reference's setup: every star's
    acceleration, jerk and potential to 1e-12 of their maxima."""
    d = load_golden(case)
    p = O.params_from_dict(d["meta"]["params"])
    rs, vs = vec(d, "star_s0"), vec(d, "star_s0", ("vx", "vy", "vz"))
    ms, hs = d["star_in_m"], d["star_in_h"]
    assert np.array_equal(rs, d["star_in_r"])
    st = _stars(d, soft)
    r, h = vec(d, "s0"), d["s0_h"]
    ag, pg = O.stars_from_gas(p, st, r, d["in_m"], h)
    # both branches of the star-gas sums occur in this run
    dist = np.linalg.norm(r[None, :, :] - rs[:, None, :], axis=2)
    near = dist < 0.5 * 2.0 * hs[:, None] + 0.5 * 2.0 * h[None, :]
    assert 0 < near.sum() < near.size
    if soft:
        a2, ad2, p2 = O.nbody_direct_smoothed(p, rs, vs, ms, hs)
    else:
        a2, ad2, p2 = O.nbody_direct(rs, vs, ms)
    assert relmax(ag + a2, vec(d, "star_s0", ("ax", "ay", "az"))) < 1e-12
    assert relmax(pg + p2, d["star_s0_gpot"]) < 1e-12
    assert relmax(ad2, vec(d, "star_s0", ("adotx", "adoty", "adotz"))) < 1e-12


@pytest.mark.parametrize("case,soft", CASES)
def test_gas_star_pull_after_setup_matches_reference(case, soft):
    """GradhSph::ComputeStarGravForces: the gas's tree acceleration and potential after the
    setup are the exact gas sums (brute-force search) plus the stars' pull, to 1e-12."""
    d = load_golden(case)
    p = O.params_from_dict(d["meta"]["params"])
    r = vec(d, "s0")
    st = {k: d[f"s0_{k}"] for k in ("h", "zeta", "hfactor")}
    at, gp = O.direct_gravity(p, r, d["in_m"], st)
    a_s, p_s = O.gas_from_stars(p, r, d["s0_h"], _stars(d, soft))
    assert relmax(at + a_s, vec(d, "s0", ("atx", "aty", "atz"))) < 1e-12
    assert relmax(gp + p_s, d["s0_gpot"]) < 1e-12
    # the stars' pull is a sizeable part of the gas's gravity in these runs
    assert np.max(np.abs(a_s)) > 0.1 * np.max(np.abs(at))


@pytest.mark.parametrize("case,soft", CASES)
def test_star_gas_pair_terms_obey_newtons_third_law(case, soft):
    """For the M4 kernel the two directions agree on which pairs are softened
    (|dr| < h_s + h_j, i.e. s = 2|dr|/(h_s + h_j) < 2, beyond which wgrav = 1/s^2 and
    wpot = 1/s), so the star-gas forces cancel pair by pair: m_s a_s(gas) = -sum_j m_j a_j(s),
    and the potential energies agree.  A dropped mass factor, sign or softening length fails."""
    d = load_golden(case)
    p = O.params_from_dict(d["meta"]["params"])
    assert p.kernel == 0
    r, m, h = vec(d, "s0"), d["in_m"], d["s0_h"]
    ms = d["star_in_m"]
    st = _stars(d, soft)
    a_star, p_star = O.stars_from_gas(p, st, r, m, h)
    a_gas, p_gas = O.gas_from_stars(p, r, h, st)
    f_star = (ms[:, None] * a_star).sum(0)
    f_gas = (m[:, None] * a_gas).sum(0)
    scale = (ms[:, None] * np.abs(a_star)).sum()
    assert np.max(np.abs(f_star + f_gas)) < 1e-12 * scale
    assert (ms * p_star).sum() == pytest.approx((m * p_gas).sum(), rel=1e-12)


def test_star_zeta_terms_vanish_beyond_the_kernel_and_match_closed_form():
    """The stars' zeta terms: a star beyond the kernel's reach leaves zeta as without stars
    (wzeta = 0 for s >= 2); one at a gas particle's centre adds m_s (2/(h+h_s))^2 wzeta(0)
    times dh/drho / omega, wzeta(0) = 7/5 for M4, with nbody_softening = 1, or m_s (4/h^2)
    wzeta(0) times the same with nbody_softening = 0."""
    prm = {"kernel": "m4", "tabulated_kernel": 0, "gas_eos": "isothermal"}
    p = O.params_from_dict(prm)
    rng = np.random.default_rng(7)
    N = 400
    r = rng.uniform(-1, 1, size=(N, 3))
    m = np.full(N, 1.0 / N)
    h0 = np.full(N, 0.2)
    u = np.ones(N)
    ref = O.compute_h(p, r, m, h0, u)
    far = O.stars(np.array([[50.0, 0.0, 0.0]]), np.array([0.3]), np.array([0.1]))
    got = O.compute_h_stars(p, r, m, h0, u, far)
    assert np.array_equal(got["zeta"], ref["zeta"])
    assert np.array_equal(got["h"], ref["h"])
    i = 17
    for soft in (1, 0):
        hs = 0.05
        st = O.stars(r[i:i + 1].copy(), np.array([0.3]), np.array([hs]), softening=soft)
        got = O.compute_h_stars(p, r, m, h0, u, st)
        h, rho, iom = ref["h"][i], ref["rho"][i], ref["invomega"][i]
        assert got["h"][i] == h and got["invomega"][i] == iom
        deriv = -h / (3.0 * rho)
        isq = (2.0 / (h + hs)) ** 2 if soft else 4.0 / h ** 2
        want = ref["zeta"][i] + deriv * 0.3 * isq * 1.4 * iom
        assert got["zeta"][i] == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("case,soft", CASES)
def test_hybrid_run_matches_reference(case, soft):
    """The oracle's whole gas + star run (SphSimulation setup and MainLoop with the N-body
    half: star KDK, star-gas sums both ways, zeta terms, combined timestep) against the
    reference: h-rho state after setup to 1e-12, timestep and time to 1e-13, and the gas and
    star trajectories after three steps."""
    d = load_golden(case)
    p = O.params_from_dict(d["meta"]["params"])
    s = O.Sim(p, d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    s.set_stars(d["star_in_r"], d["star_in_v"], d["star_in_m"], d["star_in_h"], softening=soft,
                nbody_mult=d["meta"]["params"]["nbody_mult"])
    s.setup()
    for f in ("h", "rho", "invomega", "zeta", "sound", "pressure"):
        assert relmax(s.get(f), d[f"s0_{f}"]) < 1e-12, f
    assert relmax(s.get("a"), vec(d, "s0", ("ax", "ay", "az"))) < 1e-12
    assert relmax(s.get("star_a"), vec(d, "star_s0", ("ax", "ay", "az"))) < 1e-12
    assert s.timestep == pytest.approx(d["diag"][0, 2], rel=1e-13)
    assert s.timestep == pytest.approx(d["star_s0_dt"][0], rel=1e-13)
    n = d["meta"]["nsteps"]
    for _ in range(n):
        s.step()
    assert s.t == pytest.approx(d["diag"][n, 1], rel=1e-13)
    assert np.max(np.abs(s.get("r") - vec(d, f"s{n}"))) < 1e-11
    assert relmax(s.get("v"), vec(d, f"s{n}", ("vx", "vy", "vz"))) < 1e-10
    assert relmax(s.get("u"), d[f"s{n}_u"]) < 1e-11
    assert np.max(np.abs(s.get("star_r") - vec(d, f"star_s{n}"))) < 1e-11
    assert relmax(s.get("star_v"), vec(d, f"star_s{n}", ("vx", "vy", "vz"))) < 1e-10
    assert relmax(s.get("star_a"), vec(d, f"star_s{n}", ("ax", "ay", "az"))) < 1e-10
    assert relmax(s.get("star_gpot"), d[f"star_s{n}_gpot"]) < 1e-10
