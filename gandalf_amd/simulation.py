"""The reference's Python-facing simulation surface over the device path.

Mirrors what GANDALF's SWIG layer exposes (src/Simulation.i:239-248): ``SimulationBase``
(src/Headers/Simulation.h:113-140, src/Common/Simulation.cpp) with ``SimulationFactory``,
``SetParam`` / ``GetParam``, ``PreSetupForPython``, ``ImportArray``, ``SetupSimulation``,
``MainLoop``, ``Run``, ``InteractiveRun``, ``Output``, ``RestartSnapshot``, snapshot reads and
writes, and ``SphSnapshotBase`` (src/Headers/SphSnapshot.h:125-128) with
``CopyDataFromSimulation`` and ``ExtractArray``.  Parameter names and defaults are the
reference's (src/Common/Parameters.cpp:160-575) for the keys this path reads.  Errors raise
``GandalfError`` where the reference raises ``ExceptionHandler`` errors.

The particle state lives on the GPU (one device context, gandalf_amd.gpu.GpuSim); this layer
keeps only the simulation clock and the output bookkeeping on the host, as the reference's
Simulation object does.  Supported: ndim = 3, sim = "gradhsph" (or "sph") and "nbody";
ic = "file" (column, su, sf), "python" and restarts from ``<run_id>.restart``; lite snapshots
(slite) are written when litesnap is set.
"""
from __future__ import annotations

import os
import time

import numpy as np

from .gpu import FIELDS3, GdfError, GpuSim


class GandalfError(GdfError):
    """An error the reference raises through ExceptionHandler::raise (Exception.cpp:49-68)."""


# (name, default) per type, the reference's defaults (Parameters.cpp:160-575), for the keys the
# device path and the run control read
_INT = {
    "ndim": 3, "Nstepsmax": 99999999, "noutputstep": 128, "ndiagstep": 1024,
    "nrestartstep": 512, "litesnap": 0, "dimensionless": 0, "Nlevels": 1,
    "level_diff_max": 1, "tabulated_kernel": 1, "hydro_forces": 1, "self_gravity": 0,
    "Nleafmax": 6, "ntreebuildstep": 1, "ntreestockstep": 1, "Nhydro": 0, "Nstar": 0,
    "nbody_softening": 1, "sub_systems": 0, "sink_particles": 0, "com_frame": 0,
    # this build's own keys: restart (the reference's "gandalf -r" switch), the h-rho
    # iteration's fixed-point iterations and the initial candidate batches (0 = reference)
    "restart": 0, "hrho_iteration_max": 0, "cand_batches_init": 0,
}
_FLOAT = {
    "tend": 1.0, "tmax_wallclock": 9.99e20, "dt_snap": 0.2, "tsnapfirst": 0.2,
    "dt_litesnap": 0.2, "tlitesnapfirst": 0.0, "dt_python": 8.0, "accel_mult": 0.3,
    "courant_mult": 0.15, "nbody_mult": 0.1, "energy_mult": 0.4, "h_fac": 1.2,
    "h_converge": 0.01, "gamma_eos": 1.66666666666666, "temp0": 1.0, "mu_bar": 1.0,
    "rho_bary": 1.0e-14, "alpha_visc": 1.0, "beta_visc": 2.0, "thetamaxsqd": 0.1,
    "boxmin[0]": -9.9e30, "boxmin[1]": -9.9e30, "boxmin[2]": -9.9e30,
    "boxmax[0]": 9.9e30, "boxmax[1]": 9.9e30, "boxmax[2]": 9.9e30,
    # this build's own keys: Plummer softening of the N-body tree walk, a fixed timestep
    "nbody_eps": 0.0, "fixed_dt": 0.0,
}
_STRING = {
    "sim": "sph", "nbody": "hermite4", "ic": "box", "run_id": "", "in_file": "",
    "in_file_form": "su", "out_file_form": "su", "sph_integration": "lfkdk", "kernel": "m4",
    "gas_eos": "energy_eqn", "avisc": "mon97", "acond": "none",
    "time_dependent_avisc": "none", "neib_search": "kdtree", "gravity_mac": "geometric",
    "multipole": "quadrupole",
    "boundary_lhs[0]": "open", "boundary_rhs[0]": "open", "boundary_lhs[1]": "open",
    "boundary_rhs[1]": "open", "boundary_lhs[2]": "open", "boundary_rhs[2]": "open",
    # this build's own key: the device path's pair arithmetic (fp64 or mixed)
    "gpu_precision": "fp64",
}
_FORMATS = ("column", "su", "sf")
# formats written only (the reference reads no lite snapshot)
_WRITE_FORMATS = _FORMATS + ("slite", "seren_lite")


class Parameters:
    """Parameters (src/Headers/Parameters.h): typed maps with the reference's defaults."""

    def __init__(self):
        self.intparams = dict(_INT)
        self.floatparams = dict(_FLOAT)
        self.stringparams = dict(_STRING)

    def SetParameter(self, key, value):
        """Parameters::SetParameter (Parameters.cpp:582-600): the value is parsed for the
        key's type; an unknown key is reported and ignored."""
        value = str(value)
        if key in self.intparams:
            self.intparams[key] = int(float(value))
        elif key in self.floatparams:
            self.floatparams[key] = float(value)
        elif key in self.stringparams:
            self.stringparams[key] = value
        else:
            print(f"Warning: parameter {key} was not recognized")

    def GetParameter(self, key):
        """Parameters::GetParameter (Parameters.cpp:551-573)."""
        for d in (self.intparams, self.floatparams, self.stringparams):
            if key in d:
                return str(d[key]) if d is not self.floatparams else f"{d[key]:g}"
        raise GandalfError(f"Error: the parameter {key} is not recognized ")

    def ReadParamsFile(self, filename):
        """Parameters::ReadParamsFile (Parameters.cpp:75-152): ``comment : key = value`` per
        line; lines without both separators are skipped."""
        with open(filename) as f:
            for line in f:
                if ":" not in line or "=" not in line:
                    continue
                rest = line.split(":", 1)[1]
                key, value = rest.split("=", 1)
                self.SetParameter(key.strip(), value.strip())

    def as_dict(self):
        out = {}
        for d in (self.intparams, self.floatparams, self.stringparams):
            out.update(d)
        return out


class SimulationBase:
    """SimulationBase / Simulation<3> (src/Common/Simulation.cpp) driving the device path."""

    def __init__(self, params: Parameters | None = None):
        self.simparams = params if params is not None else Parameters()
        self.ParametersProcessed = False
        self.setup = False
        self.sim: GpuSim | None = None
        self.t = 0.0
        self.timestep = 0.0
        self.Nsteps = 0
        self.n = 0
        self.Nblocksteps = 0
        self.nresync = 1
        self.Noutsnap = 0
        self.Noutlitesnap = 0
        self.tsnaplast = 0.0
        self.tsnapnext = 0.0
        self.tlitesnaplast = 0.0
        self.tlitesnapnext = 0.0
        self.nlastrestart = 0
        self.restart = False
        self.diag = None
        self.diag0 = None
        self._imported = {}

    # ---- construction and parameters (Simulation.cpp:58-160, 287-370) ----
    @staticmethod
    def SimulationFactory(ndim, simtype, params: Parameters | None = None):
        if int(ndim) != 3:
            raise GandalfError("Error: the device path runs ndim = 3 only")
        if simtype not in ("sph", "gradhsph", "nbody"):
            raise GandalfError(f"Error: simulation type {simtype} is not supported by the "
                               "device path")
        sim = SimulationBase(params)
        sim.simparams.stringparams["sim"] = simtype
        sim.simparams.intparams["ndim"] = 3
        return sim

    def SetParam(self, key, value):
        if self.ParametersProcessed:
            raise GandalfError("Error: the non-return point for setting parameters has been "
                               "reached!")
        if key == "ndim":
            raise GandalfError("Error: Not possible to change the number of dimensions!")
        if key == "sim":
            raise GandalfError("Error: Cannot change the type of simulation afterwards!")
        self.simparams.SetParameter(key, value)

    def GetParam(self, key):
        return self.simparams.GetParameter(key)

    def GetIntAndFloatParameterKeys(self):
        return list(self.simparams.intparams) + list(self.simparams.floatparams)

    def ProcessParameters(self):
        """Simulation::ProcessParameters: the device context for these parameters."""
        if self.ParametersProcessed:
            raise GandalfError("Error: ProcessParameters has been already called!")
        p = self.simparams
        prm = p.as_dict()
        if prm["sim"] == "sph":
            prm["sim"] = "gradhsph"
        prec = p.stringparams["gpu_precision"]
        if prec not in ("fp64", "mixed"):
            raise GandalfError(f"Error: gpu_precision {prec} is not fp64 or mixed")
        self.sim = GpuSim(prm, precision=prec)
        self.run_id = p.stringparams["run_id"]
        self.tend = p.floatparams["tend"]
        self.dt_snap = p.floatparams["dt_snap"]
        self.tsnapnext = p.floatparams["tsnapfirst"]
        self.litesnap = p.intparams["litesnap"]
        self.dt_litesnap = p.floatparams["dt_litesnap"]
        self.tlitesnapnext = p.floatparams["tlitesnapfirst"]
        self.dt_python = p.floatparams["dt_python"]
        self.Nstepsmax = p.intparams["Nstepsmax"]
        self.noutputstep = p.intparams["noutputstep"]
        self.ndiagstep = p.intparams["ndiagstep"]
        self.nrestartstep = p.intparams["nrestartstep"]
        self.restart = bool(p.intparams["restart"]) or self.restart
        self.ParametersProcessed = True

    # ---- initial conditions (SimulationIC.hpp:50-110, Simulation.cpp:1290-1620) ----
    def PreSetupForPython(self):
        """Simulation::PreSetupForPython (Simulation.cpp:1290-1316)."""
        if self.simparams.stringparams["ic"] != "python":
            raise GandalfError('Error: you should call this function only if you are using '
                               '"python" as "ic" parameter')
        self.ProcessParameters()
        nbody = self.simparams.stringparams["sim"] == "nbody"
        self.Nimport = self.simparams.intparams["Nstar" if nbody else "Nhydro"]
        self._imported = {}

    def ImportArray(self, values, quantity, type="sph"):
        """Simulation::ImportArray (Simulation.cpp:1442-1620): one named per-particle quantity
        (x y z vx vy vz m h u rho) of the particles to be set up."""
        if not self.ParametersProcessed or self.setup:
            raise GandalfError("Error: ImportArray needs PreSetupForPython before the setup")
        nbody = self.simparams.stringparams["sim"] == "nbody"
        want = "star" if nbody else "sph"
        if type not in (want, "default" if not nbody else want):
            raise GandalfError(f"Error: particle type {type} cannot be imported here")
        values = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if values.size != self.Nimport:
            raise GandalfError(f"Error: the array you are passing has a size of {values.size}, "
                               f"but memory has been allocated for {self.Nimport} particles")
        if quantity not in ("x", "y", "z", "vx", "vy", "vz", "m", "h", "u", "rho"):
            raise GandalfError(f"Error: the quantity {quantity} is not recognised")
        self._imported[quantity] = values.copy()

    def _generate_ic(self):
        """Simulation::GenerateIC (SimulationIC.hpp:50-110): a restart reads the snapshot named
        in <run_id>.restart; else ic = file reads in_file; ic = python takes the imports."""
        p = self.simparams
        if self.restart:
            fn = f"{self.run_id}.restart"
            if os.path.exists(fn):
                with open(fn) as f:
                    form, name = f.read().split()[:2]
                self.ReadSnapshotFile(name, form)
                return
            self.restart = False
        ic = p.stringparams["ic"]
        if ic == "file":
            self.ReadSnapshotFile(p.stringparams["in_file"], p.stringparams["in_file_form"])
        elif ic == "python":
            imp = self._imported
            miss = [q for q in ("x", "y", "z", "m") if q not in imp]
            if miss:
                raise GandalfError(f"Error: arrays {miss} were not imported")
            n = self.Nimport
            zero = np.zeros(n)
            r = np.stack([imp["x"], imp["y"], imp["z"]], axis=1)
            v = np.stack([imp.get("vx", zero), imp.get("vy", zero), imp.get("vz", zero)], axis=1)
            self.sim.set_particles(r, v, imp["m"], imp.get("u", zero), h=imp.get("h"))
            if "rho" in imp and self.simparams.stringparams["sim"] != "nbody":
                self.sim.set("rho", imp["rho"])
        else:
            raise GandalfError(f"Error: ic = {ic} is not available on the device path "
                               "(use file or python)")

    def SetupSimulation(self):
        """SimulationBase::SetupSimulation (Simulation.cpp:638-695)."""
        if self.setup:
            raise GandalfError("This simulation has been already set up")
        if self.simparams.stringparams["ic"] == "python":
            if not self.ParametersProcessed:
                raise GandalfError("Error: you are attempting to setup a simulation with "
                                   "initial conditions generated from Python. Before setting "
                                   "up the simulation, you need to import the initial "
                                   "conditions")
        else:
            if self.ParametersProcessed:
                raise GandalfError("The parameters of the simulation have been already "
                                   "processed. It means that you shouldn't be calling this "
                                   "function, please consult the documentation.")
            self.ProcessParameters()
        self._generate_ic()
        if self.restart:
            # SimulationIO.hpp:2553-2555: the next snapshot follows the restored one
            self.tsnapnext = self.tsnaplast + self.dt_snap
            self.tlitesnapnext = self.tlitesnaplast + self.dt_litesnap
        self.sim.setup()
        self.t, self.timestep, self.Nsteps = self.sim.time()
        self._sync_block_clock()
        self.CalculateDiagnostics()
        self.diag0 = self.diag
        self.setup = True
        self.Output()

    # ---- the step and the run loops (Simulation.cpp:382-494) ----
    def _sync_block_clock(self):
        n, _, _, nres, _, _ = self.sim.block_state()
        if int(self.simparams.intparams["Nlevels"]) > 1:
            self.n, self.nresync = n, max(nres, 1)
        else:
            self.n, self.nresync = self.Nsteps, 1

    def MainLoop(self):
        """One SphSimulation::MainLoop / NbodySimulation::MainLoop on the device."""
        if not self.setup:
            raise GandalfError("Error: MainLoop before SetupSimulation")
        self.sim.step(1)
        self.t, self.timestep, self.Nsteps = self.sim.time()
        self._sync_block_clock()
        if self.n % self.nresync == 0:
            self.Nblocksteps += 1

    def Run(self, Nadvance=-1):
        """SimulationBase::Run: steps until tend or the target step count."""
        target = self.Nstepsmax if Nadvance < 0 else self.Nsteps + Nadvance
        while self.t < self.tend and self.Nsteps < target:
            self.MainLoop()
            self.Output()
        self.CalculateDiagnostics()

    def InteractiveRun(self, Nadvance=-1):
        """SimulationBase::InteractiveRun: as Run, for at most dt_python seconds of CPU time;
        returns a snapshot object for every snapshot written meanwhile."""
        target = self.Nstepsmax if Nadvance < 0 else self.Nsteps + Nadvance
        snaps = []
        t0 = time.process_time()
        while self.t < self.tend and self.Nsteps < target and \
                time.process_time() - t0 < self.dt_python:
            self.MainLoop()
            if self.t >= self.tsnapnext:
                self.CalculateDiagnostics()
            fn = self.Output()
            if fn:
                s = SphSnapshotBase.SphSnapshotFactory(fn, self, 3)
                s.CopyDataFromSimulation()
                snaps.append(s)
        if self.t >= self.tend or self.Nsteps >= target:
            self.CalculateDiagnostics()
        return snaps

    # ---- output (Simulation.cpp:500-632) ----
    def Output(self):
        """SimulationBase::Output: a snapshot <run_id>.<form>.<NNNNN> and the restart log when
        t passes tsnapnext, diagnostics every ndiagstep block steps and a restart snapshot
        every nrestartstep steps.  Returns the snapshot's name, or "" if none was written."""
        if self.Nsteps % self.noutputstep == 0:
            print(f"t : {self.t}    dt : {self.timestep}    Nsteps : {self.Nsteps}")
        filename = ""
        form = self.simparams.stringparams["out_file_form"]
        # a lite snapshot for movies (Simulation.cpp:527-540)
        if self.litesnap == 1 and self.t >= self.tlitesnapnext:
            self.Noutlitesnap += 1
            self.tlitesnaplast = self.tlitesnapnext
            self.tlitesnapnext += self.dt_litesnap
            self.WriteSnapshotFile(f"{self.run_id}.slite.{self.Noutlitesnap:05d}", "slite")
        if self.t >= self.tsnapnext:
            self.Noutsnap += 1
            self.tsnaplast = self.tsnapnext
            self.tsnapnext += self.dt_snap
            filename = f"{self.run_id}.{form}.{self.Noutsnap:05d}"
            self.WriteSnapshotFile(filename, form)
            with open(f"{self.run_id}.restart", "w") as f:
                f.write(f"{form}\n{filename}\n")
        if self.Nblocksteps % self.ndiagstep == 0 and self.n % self.nresync == 0:
            self.CalculateDiagnostics()
        if self.n % self.nresync == 0 and self.Nsteps - self.nlastrestart >= self.nrestartstep:
            self.RestartSnapshot()
            self.nlastrestart = self.Nsteps
        return filename

    def RestartSnapshot(self):
        """SimulationBase::RestartSnapshot (Simulation.cpp:604-631)."""
        form = self.simparams.stringparams["out_file_form"]
        filename = f"{self.run_id}.{form}.tmp"
        self.WriteSnapshotFile(filename, form)
        with open(f"{self.run_id}.restart", "w") as f:
            f.write(f"{form}\n{filename}\n")

    def WriteSnapshotFile(self, filename, form):
        """Simulation::WriteSnapshotFile (SimulationIO.hpp:60-150)."""
        if form not in _WRITE_FORMATS:
            raise GandalfError(f"Error: snapshot format {form} is not available")
        self.sim.snapshot_counters(self.Noutsnap, self.Noutlitesnap, self.tsnaplast,
                                   self.tlitesnaplast)
        self.sim.write_snapshot(filename, form)
        return True

    def ReadSnapshotFile(self, filename, form):
        """Simulation::ReadSnapshotFile: the state, the time and step count, and on a restart
        the snapshot counters."""
        if form not in _FORMATS:
            raise GandalfError(f"Error: snapshot format {form} is not available")
        if not os.path.exists(filename):
            raise GandalfError(f"Error: snapshot file {filename} not found")
        self.sim.read_snapshot(filename, form)
        self.t, self.timestep, self.Nsteps = self.sim.time()
        self.Noutsnap, self.Noutlitesnap, self.tsnaplast, self.tlitesnaplast = \
            self.sim.snapshot_counters()
        return True

    def CalculateDiagnostics(self):
        """Simulation::CalculateDiagnostics (SimAnalysis.hpp:52-185), FP64 on the device."""
        out = self.sim.diagnostics()
        self.diag = {"Etot": out[0], "ketot": out[1], "utot": out[2], "gpetot": out[3],
                     "mom": out[4:7].copy(), "angmom": out[7:10].copy(), "mtot": out[10]}
        return self.diag

    def GetInitialEnergy(self):
        return self.diag0["Etot"] if self.diag0 else 0.0


class SphSnapshotBase:
    """SphSnapshotBase (src/Headers/SphSnapshot.h, src/Common/SphSnapshot.cpp): the particle
    arrays of one moment, copied from a running simulation or read from a snapshot file."""

    _SPH = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "h", "rho", "u", "dudt",
            "gpot", "iorig")
    _STAR = ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "m", "h", "gpot")

    def __init__(self, filename="", sim: SimulationBase | None = None, ndim=3):
        self.filename = filename
        self.simulation = sim
        self.ndim = ndim
        self.t = 0.0
        self.data = {}
        self.allocated = False

    @staticmethod
    def SphSnapshotFactory(filename, sim, ndim):
        return SphSnapshotBase(filename, sim, ndim)

    def CopyDataFromSimulation(self):
        """SphSnapshotBase::CopyDataFromSimulation (SphSnapshot.cpp:230-372): the gas (or the
        stars of an N-body run) in the reference's particle order."""
        g = self.simulation.sim
        nbody = self.simulation.simparams.stringparams["sim"] == "nbody"
        names = self._STAR if nbody else self._SPH
        arr = {}
        for f in ("r", "v", "a"):
            x = g.get(f)
            pre = {"r": "", "v": "v", "a": "a"}[f]
            for k, ax in enumerate("xyz"):
                arr[f"{pre}{ax}"] = x[:, k].copy()
        for f in ("m", "h", "gpot") + (() if nbody else ("rho", "u", "dudt")):
            arr[f] = g.get(f)
        if not nbody:
            arr["iorig"] = np.arange(g.N, dtype=np.float64)
        self.data = {("star" if nbody else "sph"): {k: arr[k] for k in names}}
        self.t = self.simulation.t
        self.allocated = True

    def ReadSnapshot(self, form, sim: SimulationBase):
        """Read the snapshot file into a device context of the simulation's parameters and
        copy its arrays (the reference reads into its own buffers)."""
        tmp = SimulationBase.SimulationFactory(3, sim.simparams.stringparams["sim"],
                                               sim.simparams)
        tmp.ParametersProcessed = False
        tmp.ProcessParameters()
        tmp.ReadSnapshotFile(self.filename, form)
        self.simulation = tmp
        self.CopyDataFromSimulation()

    def GetRealType(self, type):
        if type == "default":
            return next(iter(self.data)) if self.data else "sph"
        return type

    def ExtractArray(self, name, type="default", RequestedUnit="default"):
        """SphSnapshotBase::ExtractArray (SphSnapshot.cpp:480-560): (unit info, array, scaling
        factor).  The device path runs in dimensionless units: the scaling factor is 1."""
        if not self.allocated:
            raise GandalfError("Error: requested a snapshot that is not allocated")
        type = self.GetRealType(type)
        if type not in ("sph", "star", "binary", "dust"):
            raise GandalfError(f"Error: the type {type} was not recognized!")
        if type not in self.data or name not in self.data[type]:
            raise GandalfError(f"Warning: the selected array: {name} has not been recognized")
        unit = {"name": name, "label": name, "unit": "", "scaling_factor": 1.0}
        return unit, self.data[type][name], 1.0
