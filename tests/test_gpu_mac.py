"""The gravity walk's relative opening criteria (gravity_mac = gadget2, eigenmac) on the HIP path
against the oracle's restatement of the reference's setup and main loop: the setup's two force
passes (geometric first, then the chosen criterion on the tree stocked with those forces) and
three global steps, each criterion reading the walk group's atree / gpot stocked at the step's
tree build; FP64 to rounding, the mixed path within 1e-4 of FP64."""
import numpy as np
import pytest

import oracle as O
from conftest import load_golden

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def params(d, mac, **over):
    # the gravity walk through the oracle's tree twin (same cells and walk groups)
    return dict(d["meta"]["params"], neib_search="kdtree", gravity_mac=mac, **over)


def gpu_sim(prm, d, precision="fp64"):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(prm, precision=precision)
    s.set_particles(d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    return s


@pytest.mark.parametrize("mac,macerror", [("gadget2", 1e-4), ("eigenmac", 1e-4),
                                          ("gadget2", 1e-3)])
def test_relative_mac_setup_and_steps_match_oracle(mac, macerror):
    d = load_golden("sphere4k_m4_bf")
    print(f"criterion {mac} macerror {macerror}", flush=True)
    prm = params(d, mac, macerror=macerror)
    s = gpu_sim(prm, d)
    s.setup()
    assert s.opening_criterion() == mac
    o = O.Sim(O.params_from_dict(prm), d["in_r"], d["in_v"], d["in_m"], d["in_u"])
    o.setup()
    for f in ("h", "atree", "gpot", "a"):
        assert relmax(s.get(f), o.get(f)) < TOL, f
    for _ in range(3):
        s.step(1)
        o.step()
    assert np.max(np.abs(s.get("r") - o.get("r"))) < 1e-12
    assert relmax(s.get("v"), o.get("v")) < 1e-10
    assert relmax(s.get("atree"), o.get("atree")) < 1e-10


def test_relative_mac_mixed_precision():
    d = load_golden("sphere4k_m4_bf")
    prm = params(d, "gadget2")
    s64 = gpu_sim(prm, d)
    s64.setup()
    s32 = gpu_sim(prm, d, "mixed")
    s32.setup()
    a64, a32 = s64.get("a"), s32.get("a")
    e = np.linalg.norm(a32 - a64, axis=1) / np.max(np.linalg.norm(a64, axis=1))
    assert np.max(e) < 1e-4


def test_opening_criterion_switch():
    """NeighbourSearch::SetOpeningCriterion: switching a geometric context to gadget2 and
    building the tree again stocks the current forces; an unknown criterion is refused."""
    d = load_golden("sphere4k_m4_bf")
    prm = params(d, "geometric")
    s = gpu_sim(prm, d)
    s.setup()
    at_geo = s.get("atree")
    s.set_opening_criterion("gadget2")
    assert s.opening_criterion() == "gadget2"
    s.build_tree()
    s.density()
    s.all_forces()
    p = O.params_from_dict(dict(prm, gravity_mac="gadget2"))
    st = {k: s.get(k) for k in ("h", "zeta", "hfactor")}
    r, m = s.get("r"), s.get("m")
    at, gp, nc, nn = O.TreeTwin(p, r, m, st["h"]).gravity(r, m, st, at_geo, None)
    gnc, gnn = s.gravity_group_counts()
    assert np.array_equal(gnc, nc) and np.array_equal(gnn, nn)
    assert relmax(s.get("atree"), at) < TOL
    from gandalf_amd.gpu import lib
    assert lib().gdf_set_opening_criterion(s._h, 3) < 0
    assert s.opening_criterion() == "gadget2"
