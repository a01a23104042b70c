"""The hydro-only box and the cold sphere at their full benchmark sizes.

Periodic jittered-lattice box, 32^3 = 32,768 particles (M4, adiabatic, global timestep, 10
steps): the FP64 device run against the oracle's run of the same steps (the oracle is pinned
step by step by the reference's 16^3 fixtures), and the mixed path against the FP64 path.

Self-gravitating cold sphere, 262,144 particles (isothermal, M4, quadrupole tree gravity at
theta = 0.5 with the reference's kernel softening by h, GradhSph.cpp:555-565): the FP32 mixed
path against the FP64 device path after the setup (accelerations within 1e-4 of the largest)
and over 50 steps (positions, and energies against the FP64 run's own drift)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def gpu(g, precision):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(g["params"], precision=precision)
    s.set_particles(g["r"], g["v"], g["m"], g["u"])
    return s


def test_box_32cubed_ten_steps_match_oracle():
    from gandalf_amd import ic as ICS
    g = ICS.jittered_lattice_box(n_side=32, seed=1)
    s = gpu(g, "fp64")
    s.setup()
    o = O.Sim(O.params_from_dict(g["params"]), g["r"], g["v"], g["m"], g["u"])
    o.setup()
    assert relmax(s.get("h"), o.get("h")) < 1e-12
    a64 = s.get("a")
    assert relmax(a64, o.get("a")) < 1e-12
    m = gpu(g, "mixed")
    m.setup()
    e = np.linalg.norm(m.get("a") - a64, axis=1)
    assert e.max() < 1e-4 * np.max(np.linalg.norm(a64, axis=1))
    s.step(10)
    for _ in range(10):
        o.step()
    assert np.max(np.abs(s.get("r") - o.get("r"))) < 1e-12
    assert relmax(s.get("v"), o.get("v")) < 1e-9
    assert relmax(s.get("u"), o.get("u")) < 1e-12
    assert s.time()[0] == pytest.approx(o.t, rel=1e-12)


def test_cold_sphere_262k_mixed_against_fp64():
    from gandalf_amd import ic as ICS
    g = ICS.cold_sphere(N=262144, seed=2)
    s64, s32 = gpu(g, "fp64"), gpu(g, "mixed")
    s64.setup()
    s32.setup()
    a64, a32 = s64.get("a"), s32.get("a")
    amax = np.max(np.linalg.norm(a64, axis=1))
    assert np.max(np.linalg.norm(a32 - a64, axis=1)) < 1e-4 * amax
    E0 = {k: s.diagnostics()[0] for k, s in (("64", s64), ("32", s32))}
    s64.step(50)
    s32.step(50)
    assert s64.time()[2] == s32.time()[2] == 50
    r64, r32 = s64.get("r"), s32.get("r")
    assert np.max(np.abs(r32 - r64)) < 1e-4
    drift = {k: abs(s.diagnostics()[0] - E0[k]) / abs(E0[k]) for k, s in (("64", s64), ("32", s32))}
    assert drift["32"] < max(2.0 * drift["64"], 1e-5), drift
