"""The gravity walk's cell-term options ("multipole": quadrupole, monopole, fast_monopole,
fast_quadrupole) on the HIP path against the oracle's tree twin: the same interaction sets (per
group counts equal, the fast options' extra acceptance test included) and the forces to rounding
in FP64; the mixed path within the FP32 tolerance of the FP64 path; periodic boxes keep their
Ewald corrections with the fast options."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden, vec

pytestmark = pytest.mark.gpu

MODES = ["monopole", "fast_monopole", "fast_quadrupole"]
TOL = 1e-12


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def gpu_sim(d, precision="fp64", **over):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(dict(d["meta"]["params"], **over), precision=precision)
    s.set_particles(d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    return s


@pytest.mark.parametrize("case", ["sphere4k_m4_bf", "cloud4k_quintic_bf"])
@pytest.mark.parametrize("mode", MODES)
def test_multipole_matches_tree_twin(case, mode):
    d = load_golden(case)
    print(f"case {case} multipole {mode}", flush=True)
    s = gpu_sim(d, multipole=mode)
    s.setup()
    p = O.params_from_dict(dict(d["meta"]["params"], multipole=mode))
    st = {k: s.get(k) for k in ("h", "zeta", "hfactor")}
    r, m = s.get("r"), s.get("m")
    at, gp, nc, nn = O.TreeTwin(p, r, m, st["h"]).gravity(r, m, st)
    gnc, gnn = s.gravity_group_counts()
    assert np.array_equal(gnc, nc)
    assert np.array_equal(gnn, nn)
    assert relmax(s.get("atree"), at) < TOL
    assert relmax(s.get("gpot"), gp) < TOL
    exact = vec(d, "s0", ("atx", "aty", "atz"))
    e = np.linalg.norm(s.get("atree") - exact, axis=1) / np.linalg.norm(exact, axis=1)
    assert np.sqrt(np.mean(e ** 2)) < 2.5e-2


@pytest.mark.parametrize("mode", ["fast_monopole", "fast_quadrupole"])
def test_fast_multipole_mixed_precision(mode):
    d = load_golden("sphere4k_m4_bf")
    s64 = gpu_sim(d, multipole=mode)
    s64.setup()
    s32 = gpu_sim(d, "mixed", multipole=mode)
    s32.setup()
    a64, a32 = s64.get("a"), s32.get("a")
    e = np.linalg.norm(a32 - a64, axis=1) / np.max(np.linalg.norm(a64, axis=1))
    assert np.max(e) < 1e-4


def test_fast_multipole_in_a_periodic_box():
    """A clump in a periodic box: the fast expansion of the cells plus every cell's and pair's
    Ewald correction against the oracle's exact periodic sum, to the tree's error."""
    from test_gpu_ewald import gpu_box, oracle_periodic, periodic_gravity_blob
    g = periodic_gravity_blob()
    for prec in ("fp64", "mixed"):
        s = gpu_box(g, prec, multipole="fast_quadrupole")
        s.setup()
        at, _ = oracle_periodic(s, g)
        e = np.linalg.norm(s.get("atree") - at, axis=1) / np.max(np.linalg.norm(at, axis=1))
        assert np.sqrt(np.mean(e ** 2)) < 1e-2, prec
