"""The CPU oracle pinned against the compiled reference's own outputs (golden fixtures),
closed forms of the kernels and brute-force neighbour sets.  CPU only."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden, vec

CASES_HYDRO = ["box16_m4", "box16_m4_tab"]


def _setup_and_run(d):
    p = O.params_from_dict(d["meta"]["params"])
    s = O.Sim(p, d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    s.setup()
    return p, s


def relmax(a, b):
    scale = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / scale


@pytest.mark.parametrize("case", CASES_HYDRO + ["sphere4k_m4_bf", "cloud4k_quintic_bf"])
def test_setup_h_rho_matches_reference(case):
    """grad-h h-rho after the reference's three setup passes (SphSimulation.cpp:204-565)."""
    d = load_golden(case)
    p, s = _setup_and_run(d)
    for f in ("h", "rho", "invomega", "zeta", "sound", "pressure"):
        assert relmax(s.get(f), d[f"s0_{f}"]) < 1e-12, f


@pytest.mark.parametrize("case", CASES_HYDRO)
def test_hydro_forces_and_trajectory_match_reference(case):
    d = load_golden(case)
    p, s = _setup_and_run(d)
    assert relmax(s.get("a"), vec(d, "s0", ("ax", "ay", "az"))) < 1e-12
    assert relmax(s.get("dudt"), d["s0_dudt"]) < 1e-10
    assert s.timestep == pytest.approx(d["diag"][0, 2], rel=1e-13)
    n = d["meta"]["nsteps"]
    for _ in range(n):
        s.step()
    assert np.max(np.abs(s.get("r") - vec(d, f"s{n}"))) < 1e-12
    assert relmax(s.get("v"), vec(d, f"s{n}", ("vx", "vy", "vz"))) < 1e-10
    assert relmax(s.get("u"), d[f"s{n}_u"]) < 1e-12
    assert s.t == pytest.approx(d["diag"][n, 1], rel=1e-13)


@pytest.mark.parametrize("case", ["sphere4k_m4_bf", "cloud4k_quintic_bf"])
def test_direct_gravity_matches_reference_bruteforce(case):
    """neib_search=bruteforce makes the reference's walk an exact direct sum
    (BruteForceTree.cpp:214), pinning the kernel-softened pair terms and self potential."""
    d = load_golden(case)
    p = O.params_from_dict(d["meta"]["params"])
    st = {k: d[f"s0_{k}"] for k in ("h", "zeta", "hfactor")}
    at, gp = O.direct_gravity(p, vec(d, "s0"), d["in_m"], st)
    assert relmax(at, vec(d, "s0", ("atx", "aty", "atz"))) < 1e-12
    assert relmax(gp, d["s0_gpot"]) < 1e-12
    a, dudt, div_v = O.hydro_forces(p, 1, vec(d, "s0"), vec(d, "s0", ("vx", "vy", "vz")),
                                    d["in_m"], {k: d[f"s0_{k}"] for k in
                                                ("h", "rho", "invomega", "pressure", "sound",
                                                 "hfactor")})
    assert relmax(a + at, vec(d, "s0", ("ax", "ay", "az"))) < 1e-12


def test_tree_twin_error_below_reference_tree_error():
    """Tree gravity of the device-walk twin vs the exact sum, against the reference
    KD-tree's own error at the same opening angle (treeerror.py:21-34 metric)."""
    bf = load_golden("sphere4k_m4_bf")
    kd = load_golden("sphere4k_m4_kd")
    p = O.params_from_dict(bf["meta"]["params"])
    r = vec(bf, "s0")
    st = {k: bf[f"s0_{k}"] for k in ("h", "zeta", "hfactor")}
    exact = vec(bf, "s0", ("atx", "aty", "atz"))
    tw = O.TreeTwin(p, r, bf["in_m"], st["h"])
    at, gp, nc, nn = tw.gravity(r, bf["in_m"], st)
    e_twin = np.linalg.norm(at - exact, axis=1) / np.linalg.norm(exact, axis=1)
    e_ref = np.linalg.norm(vec(kd, "s0", ("atx", "aty", "atz")) - exact, axis=1) / \
        np.linalg.norm(exact, axis=1)
    assert np.sqrt(np.mean(e_twin ** 2)) < np.sqrt(np.mean(e_ref ** 2))
    assert np.max(e_twin) < np.max(e_ref)


def test_reference_energy_bookkeeping():
    """Diagnostics restated from SimAnalysis.hpp:52-185 reproduce the reference's Etot."""
    d = load_golden("sphere4k_m4_bf")
    p = O.params_from_dict(d["meta"]["params"])
    out = O.diagnostics(p, vec(d, "s0"), vec(d, "s0", ("vx", "vy", "vz")), d["in_m"],
                        d["s0_u"], d["s0_gpot"])
    assert out[0] == pytest.approx(d["diag"][0, 3], rel=1e-12)
    assert out[3] == pytest.approx(d["diag"][0, 6], rel=1e-12)


@pytest.mark.parametrize("case", ["blob16_m4_blk", "cloud4k_quintic_blk"])
def test_block_timesteps_match_reference(case):
    """Block timesteps (Simulation.cpp:1764-2197, SphLeapfrogKDK.cpp:76-329) against the
    reference run step by step: the integer time, level_max, level_step, nresync, dt_max and
    the number of particles ending a step agree exactly at every step (including the steps
    where the reference adds or removes a level and rescales the integer time), and after the
    last step every particle's level, nstep, nlast and levelneib agree exactly and r, v, u,
    h to rounding."""
    d = load_golden(case)
    p, s = _setup_and_run(d)
    diag = d["diag"]
    n0 = s.block_state()
    assert n0[:4] == tuple(int(x) for x in diag[0, 14:18])
    assert n0[4] == pytest.approx(diag[0, 18], rel=1e-14)
    assert s.timestep == pytest.approx(diag[0, 2], rel=1e-14)
    for f in ("level", "levelneib", "nstep", "nlast"):
        assert np.array_equal(s.get(f), d[f"s0_{f}"]), f
    nsteps = d["meta"]["nsteps"]
    for k in range(1, nsteps + 1):
        s.step()
        n, lmax, lstep, nres, dtmax, nfin = s.block_state()
        assert (n, lmax, lstep, nres) == tuple(int(x) for x in diag[k, 14:18]), k
        assert dtmax == pytest.approx(diag[k, 18], rel=1e-13), k
        assert nfin == int(diag[k, 19]), k
        assert s.t == pytest.approx(diag[k, 1], rel=1e-13), k
        assert s.timestep == pytest.approx(diag[k, 2], rel=1e-13), k
    for f in ("level", "levelneib", "nstep", "nlast"):
        assert np.array_equal(s.get(f), d[f"s{nsteps}_{f}"]), f
    assert np.allclose(s.get("tlast"), d[f"s{nsteps}_tlast"], rtol=1e-13, atol=0)
    assert np.max(np.abs(s.get("r") - vec(d, f"s{nsteps}"))) < 1e-11
    assert relmax(s.get("v"), vec(d, f"s{nsteps}", ("vx", "vy", "vz"))) < 1e-9
    assert relmax(s.get("u"), d[f"s{nsteps}_u"]) < 1e-11
    assert relmax(s.get("h"), d[f"s{nsteps}_h"]) < 1e-11


@pytest.mark.parametrize("case,smoothed", [("plummer512_direct", 0),
                                           ("plummer512_smoothed", 1)])
def test_nbody_direct_sums_match_reference(case, smoothed):
    """Star-star direct sums of the reference's pure N-body run after setup: unsoftened
    (Nbody::CalculateDirectGravForces, Nbody.cpp:233-293) and kernel-smoothed over
    2/(h_i + h_j) (NbodyLeapfrogKDK::CalculateDirectSmoothedGravForces, :78-150), a, jerk and
    potential to 1e-12 of their maxima."""
    d = load_golden(case)
    r, v, m, h = d["in_r"], d["in_v"], d["in_m"], d["in_h"]
    assert np.array_equal(vec(d, "s0"), r)
    if smoothed:
        p = O.params_from_dict(d["meta"]["params"])
        a, ad, gp = O.nbody_direct_smoothed(p, r, v, m, h)
        # every star has several partners inside the kernel, so the smoothed branch is exercised
        dr = r[:, None, :] - r[None, :, :]
        s = np.sqrt((dr ** 2).sum(-1)) * 2.0 / (h[:, None] + h[None, :])
        assert (s < 2.0).sum() - len(m) > 5 * len(m)
    else:
        a, ad, gp = O.nbody_direct(r, v, m)
    assert relmax(a, vec(d, "s0", ("ax", "ay", "az"))) < 1e-12
    assert relmax(ad, vec(d, "s0", ("adotx", "adoty", "adotz"))) < 1e-12
    assert relmax(gp, d["s0_gpot"]) < 1e-12


DIAG_CASES = [("box16_m4", 0), ("box16_m4", 10), ("box16_m4_tab", 10), ("sphere4k_m4_bf", 0),
              ("sphere4k_m4_bf", 3), ("cloud4k_quintic_bf", 3), ("blob16_m4_blk", 24),
              ("cloud4k_quintic_blk", 12)]


def diag_scales(r, v, m):
    """Magnitudes against which the momentum and angular momentum (sums that cancel to near
    zero) are compared."""
    pm = (m[:, None] * np.abs(v)).sum()
    lm = (m * np.linalg.norm(r, axis=1) * np.linalg.norm(v, axis=1)).sum()
    return pm, lm


def check_diag_row(out, row, r, v, m, tol=1e-12):
    """out = {Etot, ketot, utot, gpetot, mom[3], angmom[3], mtot} against a reference .diag row
    (columns Nsteps t timestep Etot ketot utot gpetot momx momy momz angx angy angz mtot)."""
    pm, lm = diag_scales(r, v, m)
    for k, name in enumerate(("Etot", "ketot", "utot", "gpetot")):
        assert out[k] == pytest.approx(row[3 + k], rel=tol, abs=tol * abs(row[3])), name
    assert np.allclose(out[4:7], row[7:10], rtol=0, atol=tol * pm), "mom"
    assert np.allclose(out[7:10], row[10:13], rtol=0, atol=tol * lm), "angmom"
    assert out[10] == pytest.approx(row[13], rel=1e-14), "mtot"


@pytest.mark.parametrize("case,step", DIAG_CASES)
def test_diagnostics_match_reference_rows(case, step):
    """Simulation::CalculateDiagnostics (SimAnalysis.hpp:52-185) restated: every column of
    the reference's diagnostics (energy terms, momentum, angular momentum, mass) from the
    reference's own particle state at that step."""
    d = load_golden(case)
    p = O.params_from_dict(d["meta"]["params"])
    r, v = vec(d, f"s{step}"), vec(d, f"s{step}", ("vx", "vy", "vz"))
    out = O.diagnostics(p, r, v, d["in_m"], d[f"s{step}_u"], d[f"s{step}_gpot"])
    check_diag_row(out, d["diag"][step], r, v, d["in_m"])
