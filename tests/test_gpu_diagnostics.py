"""Device diagnostics (gdf_calculate_diagnostics, Simulation::CalculateDiagnostics,
SimAnalysis.hpp:52-185) against every column of the reference's own diagnostics rows: from the
reference's particle state loaded onto the device, and from the device's own trajectory where
that trajectory reproduces the reference's (the hydro box, FP64)."""
import numpy as np
import pytest

from conftest import load_golden, vec
from test_oracle_golden import DIAG_CASES, check_diag_row

pytestmark = pytest.mark.gpu


def gpu_sim(d, precision="fp64"):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(d["meta"]["params"], precision=precision)
    s.set_particles(d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    return s


@pytest.mark.parametrize("case,step", DIAG_CASES)
def test_device_diagnostics_of_reference_state(case, step):
    """The reference's state at a step, loaded through the C ABI: the device reduction gives
    the reference's Etot, ketot, utot, gpetot, momentum, angular momentum and mass to 1e-12."""
    from gandalf_amd.gpu import GpuSim
    d = load_golden(case)
    r, v = vec(d, f"s{step}"), vec(d, f"s{step}", ("vx", "vy", "vz"))
    s = GpuSim(d["meta"]["params"])
    s.set_particles(r, v, d["in_m"], d[f"s{step}_u"])
    s.set("gpot", d[f"s{step}_gpot"])
    check_diag_row(s.diagnostics(), d["diag"][step], r, v, d["in_m"])


@pytest.mark.parametrize("case,smoothed", [("plummer512_direct", 0),
                                           ("plummer512_smoothed", 1)])
def test_device_diagnostics_of_reference_star_state(case, smoothed):
    """Pure N-body run: Etot = ketot + gpetot of the stars (no internal energy)."""
    from gandalf_amd.gpu import GpuSim
    d = load_golden(case)
    n = d["meta"]["nsteps"]
    for step in (0, n):
        r, v = vec(d, f"s{step}"), vec(d, f"s{step}", ("vx", "vy", "vz"))
        s = GpuSim(d["meta"]["params"])
        s.set_particles(r, v, d["in_m"], np.zeros(len(d["in_m"])), h=d["in_h"])
        s.set("gpot", d[f"s{step}_gpot"])
        check_diag_row(s.diagnostics(), d["diag"][step], r, v, d["in_m"])


@pytest.mark.parametrize("case", ["box16_m4", "box16_m4_tab"])
def test_device_diagnostics_along_trajectory(case):
    """Setup and every step of the device's own run: its diagnostics follow the reference's
    rows (the trajectory itself agrees to ~1e-12)."""
    d = load_golden(case)
    s = gpu_sim(d)
    s.setup()
    m = d["in_m"]
    out = s.diagnostics()
    check_diag_row(out, d["diag"][0], s.get("r"), s.get("v"), m, tol=1e-11)
    for k in range(1, d["meta"]["nsteps"] + 1):
        s.step()
        check_diag_row(s.diagnostics(), d["diag"][k], s.get("r"), s.get("v"), m, tol=1e-10)
