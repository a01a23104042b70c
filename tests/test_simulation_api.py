"""The reference's Python-facing surface (SimulationBase, Parameters, SphSnapshotBase;
src/Simulation.i:239-248) over the device path.  Parameter semantics run on the CPU; the
simulation runs (python and file initial conditions, snapshots, restarts) on the GPU, against
the reference's golden run of the periodic box."""
import os

import numpy as np
import pytest

from conftest import load_golden, vec


def test_parameters_defaults_and_types():
    from gandalf_amd.simulation import GandalfError, Parameters
    p = Parameters()
    assert p.GetParameter("h_fac") == "1.2" and p.GetParameter("Nleafmax") == "6"
    assert p.GetParameter("out_file_form") == "su" and p.GetParameter("tabulated_kernel") == "1"
    p.SetParameter("Nlevels", "8")
    p.SetParameter("thetamaxsqd", 0.25)
    p.SetParameter("kernel", "quintic")
    assert p.intparams["Nlevels"] == 8 and p.floatparams["thetamaxsqd"] == 0.25
    assert p.stringparams["kernel"] == "quintic"
    p.SetParameter("no_such_key", 3)                  # reported and ignored
    with pytest.raises(GandalfError, match="the parameter no_such_key is not recognized"):
        p.GetParameter("no_such_key")


def test_parameters_file(tmp_path):
    from gandalf_amd import ic
    from gandalf_amd.simulation import Parameters
    g = ic.jittered_lattice_box(n_side=4, seed=1)
    fn = tmp_path / "run.dat"
    ic.write_param_file(fn, g["params"], "in.col", "RUN")
    p = Parameters()
    p.ReadParamsFile(fn)
    assert p.stringparams["run_id"] == "RUN" and p.stringparams["in_file"] == "in.col"
    assert p.stringparams["boundary_lhs[0]"] == "periodic"
    assert p.floatparams["boxmax[0]"] == 1.0 and p.intparams["tabulated_kernel"] == 0


def test_factory_and_setparam_guards():
    from gandalf_amd.simulation import GandalfError, SimulationBase
    with pytest.raises(GandalfError):
        SimulationBase.SimulationFactory(2, "gradhsph")
    with pytest.raises(GandalfError):
        SimulationBase.SimulationFactory(3, "meshlessfv")
    s = SimulationBase.SimulationFactory(3, "gradhsph")
    with pytest.raises(GandalfError, match="Not possible to change the number of dimensions"):
        s.SetParam("ndim", 2)
    with pytest.raises(GandalfError, match="Cannot change the type of simulation"):
        s.SetParam("sim", "nbody")
    with pytest.raises(GandalfError, match='"python" as "ic"'):
        s.PreSetupForPython()


def _box_sim(d, **over):
    from gandalf_amd.simulation import SimulationBase
    s = SimulationBase.SimulationFactory(3, "gradhsph")
    for k, v in d["meta"]["params"].items():
        if k not in ("ndim", "sim"):
            s.SetParam(k, v)
    for k, v in over.items():
        s.SetParam(k, v)
    return s


@pytest.mark.gpu
def test_python_ic_run_reproduces_reference(tmp_path, monkeypatch):
    """ic = python: PreSetupForPython, ImportArray per quantity, SetupSimulation, then
    InteractiveRun: h and rho after the setup and the state after 10 steps are the reference's
    (box16_m4 golden run); the snapshots written meanwhile come back as snapshot objects whose
    arrays equal the files they name."""
    from gandalf_amd.simulation import SphSnapshotBase
    monkeypatch.chdir(tmp_path)
    d = load_golden("box16_m4")
    N = len(d["in_m"])
    s = _box_sim(d, ic="python", Nhydro=N, run_id="PY", tsnapfirst=0.05, dt_snap=0.05,
                 out_file_form="su", noutputstep=1000, nrestartstep=1000, litesnap=1,
                 dt_litesnap=0.02, tlitesnapfirst=0.0)
    s.PreSetupForPython()
    for k, q in enumerate("xyz"):
        s.ImportArray(d["in_r"][:, k], q)
        s.ImportArray(d["in_v"][:, k], "v" + q)
    s.ImportArray(d["in_m"], "m")
    s.ImportArray(d["in_u"], "u")
    s.SetupSimulation()
    assert np.max(np.abs(s.sim.get("h") - d["s0_h"])) / np.max(d["s0_h"]) < 1e-12
    assert s.diag["Etot"] == pytest.approx(d["diag"][0, 3], rel=1e-12)
    snaps = s.InteractiveRun(10)
    assert s.Nsteps == 10 and s.t == pytest.approx(d["diag"][10, 1], rel=1e-12)
    assert np.max(np.abs(s.sim.get("r") - vec(d, "s10"))) < 1e-12
    assert len(snaps) >= 2
    for snap in snaps:
        assert os.path.exists(snap.filename) and snap.filename.startswith("PY.su.")
        back = SphSnapshotBase(snap.filename)
        back.ReadSnapshot("su", s)
        for name in ("x", "vy", "h", "rho", "u"):
            a = snap.ExtractArray(name)[1]
            b = back.ExtractArray(name, "sph")[1]
            assert np.array_equal(a, b), (snap.filename, name)
    with open("PY.restart") as f:
        assert f.read().split() == ["su", snaps[-1].filename]
    # lite snapshots every dt_litesnap (Simulation.cpp:527-540): single precision, N particles,
    # the step count of the state each was written in, one per lite time passed
    lites = sorted(f for f in os.listdir(".") if f.startswith("PY.slite."))
    assert len(lites) == s.Noutlitesnap and len(lites) >= 3
    steps = []
    for name in lites:
        b = open(name, "rb").read()
        assert np.frombuffer(b[20:24], np.int32)[0] == 4
        assert np.frombuffer(b[36:40], np.int32)[0] == N
        assert len(b) == 1236 + 5 * 20 + 5 * 20 + 4 * 7 * N
        steps.append(int(np.frombuffer(b[244:252], np.int64)[0]))
    assert steps == sorted(steps) and steps[-1] <= 10


@pytest.mark.gpu
def test_file_ic_and_restart(tmp_path, monkeypatch):
    """ic = file from the reference's own su snapshot after 10 steps: the run continues from
    its time and step count.  A run that wrote a restart snapshot is resumed by a new
    simulation with restart set: same time, step count and snapshot counters, and the next
    steps follow the first run's to the smoothing-length tolerance (the restart recomputes h
    from the file's state, as the reference's does)."""
    monkeypatch.chdir(tmp_path)
    d = load_golden("box16_m4")
    with open("ref.su", "wb") as f:
        f.write(bytes(d["s10_su"]))
    a = _box_sim(d, ic="file", in_file="ref.su", in_file_form="su", run_id="A",
                 out_file_form="su", tsnapfirst=1e30, nrestartstep=3, noutputstep=1000)
    a.SetupSimulation()
    assert a.Nsteps == 10 and a.t == float(d["diag"][10, 1])
    a.Run(3)
    assert os.path.exists("A.restart") and os.path.exists("A.su.tmp")
    t_r, n_r = a.t, a.Nsteps
    b = _box_sim(d, ic="file", in_file="ref.su", in_file_form="su", run_id="A", restart=1,
                 out_file_form="su", tsnapfirst=1e30, nrestartstep=1000, noutputstep=1000)
    b.SetupSimulation()
    assert b.t == t_r and b.Nsteps == n_r == 13
    a.Run(4)
    b.Run(4)
    assert a.Nsteps == b.Nsteps == 17
    assert np.max(np.abs(a.sim.get("r") - b.sim.get("r"))) < 1e-5
    assert b.diag["Etot"] == pytest.approx(a.diag["Etot"], rel=1e-5)
