"""The per-method entry points a reference-side MainLoop drives the step with
(TimeIntegration<3>, Simulation::ComputeBlockTimesteps, NeighbourSearch::GetGatherNeighbourList):
- the point gather list against the brute-force set (bit-exact, periodic images included) and
  its overflow code;
- SphSimulation::MainLoop with block timesteps composed from them (AdvanceParticles,
  CheckBoundaries, BuildTree, UpdateAllSphProperties, the force pass, the CheckTimesteps loop,
  ComputeBlockTimesteps, EndTimestep) against the reference's block state at every step of the
  hot-blob fixture, and bit for bit against the whole-loop entry point gdf_main_loop;
- SetActiveParticles and CorrectionTerms against their definitions on the device's own
  block-timestep state."""
import numpy as np
import pytest

from conftest import load_golden, vec

pytestmark = pytest.mark.gpu


def gpu_sim(d, precision="fp64"):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(d["meta"]["params"], precision=precision)
    s.set_particles(d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    return s


def brute_gather(r, rp, rs, box=None):
    dr = r - rp
    if box is not None:
        lo, hi = box
        size, half = hi - lo, 0.5 * (hi - lo)
        dr = np.where(dr > half, dr - size, np.where(dr < -half, dr + size, dr))
    d2 = (dr[:, 0] * dr[:, 0] + dr[:, 1] * dr[:, 1]) + dr[:, 2] * dr[:, 2]
    return np.nonzero(d2 < rs * rs)[0]


@pytest.mark.parametrize("case", ["box16_m4", "sphere4k_m4_bf"])
def test_gather_neighbour_list_matches_brute_force(case):
    d = load_golden(case)
    s = gpu_sim(d)
    s.setup()
    r = s.get("r")
    prm = d["meta"]["params"]
    box = None
    if prm.get("boundary_lhs[0]") == "periodic":
        box = (np.array([prm[f"boxmin[{k}]"] for k in range(3)]),
               np.array([prm[f"boxmax[{k}]"] for k in range(3)]))
    rng = np.random.default_rng(17)
    h = s.get("h")
    for q in range(24):
        # at particles (self included), between them, and near the box faces
        rp = r[rng.integers(len(r))] + (0.0 if q % 3 == 0 else rng.normal(scale=0.02, size=3))
        if box is not None and q % 4 == 3:
            rp = box[0] + 0.01 * rng.random(3)
        rs = (1.0 + 2.0 * rng.random()) * float(np.median(h))
        want = brute_gather(r, rp, rs, box)
        got = s.gather_neighbour_list(rp, rs)
        assert np.array_equal(got, want), q
        if len(want) > 1:
            assert s.gather_neighbour_list(rp, rs, nmax=len(want) - 1) is None


def test_block_mainloop_composed_from_entry_points():
    """SphSimulation::MainLoop's sequence (SphSimulation.cpp:574-880) driven through the
    per-method entry points over the 24 steps of the hot-blob fixture (4 levels, three block
    cycles): the integer time, level_max, level_step, nresync and dt_max follow the reference
    at every step, and the final state equals gdf_main_loop's bit for bit."""
    d = load_golden("blob16_m4_blk")
    prm = d["meta"]["params"]
    ldm = int(prm.get("level_diff_max", 1))
    whole = gpu_sim(d)
    whole.setup()
    s = gpu_sim(d)
    s.setup()
    diag = d["diag"]
    nsteps = d["meta"]["nsteps"]
    n, _, lstep, _, _, _ = s.block_state()
    t, timestep, _ = s.time()
    rounds = 0
    for k in range(1, nsteps + 1):
        n += 1
        t += timestep
        s.advance_particles_block(n, t, timestep)
        s.check_boundaries()
        s.build_tree()
        while True:
            s.density()
            if prm.get("self_gravity", 0):
                s.all_forces()
            else:
                s.hydro_forces()
            rounds += 1
            if s.check_timesteps(ldm, lstep, n, timestep) == 0:
                break
        s.compute_block_timesteps()
        n, lmax, lstep, nres, dtmax, _ = s.block_state()
        _, timestep, _ = s.time()
        s.end_timestep_block(n, t, timestep)
        assert (n, lmax, lstep, nres) == tuple(int(x) for x in diag[k, 14:18]), k
        assert dtmax == pytest.approx(diag[k, 18], rel=1e-13), k
        assert t == pytest.approx(diag[k, 1], rel=1e-13), k
        assert timestep == pytest.approx(diag[k, 2], rel=1e-13), k
    assert rounds > nsteps          # the CheckTimesteps loop repeated some force passes
    whole.step(nsteps)
    for f in ("r", "v", "u", "h", "rho", "a", "level", "levelneib", "nstep", "nlast",
              "tlast", "dtstep"):
        assert np.array_equal(s.get(f), whole.get(f)), f
    for f in ("level", "levelneib", "nstep", "nlast"):
        assert np.array_equal(s.get(f), d[f"s{nsteps}_{f}"]), f


def test_set_active_and_correction_terms_definitions():
    """On the device's own block state after 5 steps of the hot blob: SetActiveParticles(n)
    marks exactly the particles with n - nlast == nstep; CorrectionTerms(n) changes exactly
    their velocities, by dt (a - a0) / 2 with dt their own step length, and their internal
    energy by dt (dudt - dudt0) / 2."""
    d = load_golden("blob16_m4_blk")
    s = gpu_sim(d)
    s.setup()
    s.step(5)
    n = s.block_state()[0] + 1
    nlast, nstep = s.get("nlast"), s.get("nstep")
    s.set_active_particles(n)
    ends = (n - nlast) == nstep
    assert np.array_equal(s.get("active") == 1, ends)
    assert 0 < ends.sum() < len(ends)
    v0, a, a0, dt = s.get("v"), s.get("a"), s.get("a0"), s.get("dtstep")
    u0, dudt, dudt0 = s.get("u"), s.get("dudt"), s.get("dudt0")
    t, timestep, _ = s.time()
    s.correction_terms(n, t, timestep)
    v1, u1 = s.get("v"), s.get("u")
    assert np.array_equal(v1[~ends], v0[~ends])
    assert np.array_equal(u1[~ends], u0[~ends])
    e = ends
    assert np.allclose(v1[e], v0[e] + 0.5 * dt[e, None] * (a[e] - a0[e]), rtol=1e-14, atol=1e-14)
    assert np.allclose(u1[e], u0[e] + 0.5 * (dudt[e] - dudt0[e]) * dt[e], rtol=1e-14, atol=0)
