"""The tree build's launches as a captured HIP graph (cfg.graph_capture): every replay, with the
step's own eager launches and host reads between replays, leaves the particle state bit for bit
as the eager build does, for global and block timesteps in FP64 and mixed precision."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "a", "h", "rho", "u", "gpot")


def pair(case, precision="fp64"):
    from gandalf_amd.gpu import GpuSim
    d = load_golden(case)
    out = []
    for cap in (1, 0):
        s = GpuSim(dict(d["meta"]["params"], graph_capture=cap), precision=precision)
        s.set_particles(d["in_r"], d["in_v"], d["in_m"], d["in_u"])
        out.append(s)
    return d, out


@pytest.mark.parametrize("case,precision", [("sphere4k_m4_bf", "fp64"), ("box16_m4", "mixed"),
                                            ("blob16_m4_blk", "fp64")])
def test_graph_replays_match_eager_builds(case, precision):
    d, (g, e) = pair(case, precision)
    g.setup()
    e.setup()
    replays0, state = g.graph_state()
    assert state == 1, state
    assert e.graph_state() == (0, 0)
    for step in range(6):
        g.step(1)
        e.step(1)
        for f in FIELDS:
            assert np.array_equal(g.get(f), e.get(f)), (step, f)
        assert g.time() == e.time()
    replays, state = g.graph_state()
    assert state == 1 and replays >= replays0 + 4     # the first build of each buffer arrangement
                                                   # is captured, the others are replays


def test_pools_refuse_graph_capture():
    """Serving drives several batches from several host threads, while the captured graphs are
    for a context driven by one thread alone: a pool asked for them is refused."""
    from gandalf_amd import serving as S
    from gandalf_amd.gpu import GdfError, GpuPool
    with pytest.raises(GdfError):
        GpuPool(dict(S.class_params("sphere"), graph_capture=1), "mixed", max_particles=1 << 14)
