"""Serving pools (gdf_pool): independent requests batched in one device context.

Each request's results are the same alone and in the mix: the batch sorts every request into
its own subtree and walks from it, so a request's arithmetic does not depend on the others and
its final state is bit-identical whether it ran alone or beside requests admitted before and
after it.  Admission is refused when a request does not fit the particle budget and succeeds
once resident requests have finished; a pool run agrees with the standalone context of the
same request (same physics, different tree resolution) to the mixed-precision tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("r", "v", "a", "h", "u", "rho", "gpot")


def ics(kind, sizes, seed0):
    from gandalf_amd import serving as S
    return [S.make_request_ic(kind, n, seed0 + k) for k, n in enumerate(sizes)]


SIZES = {"box": (1000, 512, 1728), "sphere": (1500, 1024, 2048),
         "plummer": (2048, 1024, 3000)}


@pytest.mark.parametrize("kind", ["box", "sphere", "plummer"])
def test_results_same_alone_and_in_mix(kind):
    """Request A run alone, and A with B admitted after A's first step and C after its second
    (both larger or smaller than A): A's final fields agree bit for bit."""
    from gandalf_amd import serving as S
    from gandalf_amd.gpu import GpuPool
    a, b, c = ics(kind, SIZES[kind], 40)
    params = S.class_params(kind)
    nsteps = 4
    alone = GpuPool(params, "mixed", max_particles=1 << 16)
    ia = alone.admit(a["r"], a["v"], a["m"], a["u"], nsteps)
    for _ in range(nsteps):
        alone.step()
    assert alone.status(ia)[:2] == (2, nsteps)
    mix = GpuPool(params, "mixed", max_particles=1 << 16)
    ja = mix.admit(a["r"], a["v"], a["m"], a["u"], nsteps)
    mix.step()
    jb = mix.admit(b["r"], b["v"], b["m"], b["u"], nsteps)
    mix.step()
    jc = mix.admit(c["r"], c["v"], c["m"], c["u"], nsteps)
    assert mix.resident_particles == sum(x["r"].shape[0] for x in (a, b, c))
    for _ in range(nsteps):
        mix.step()
    # A leaves after two of these steps, B after three, C after all four
    assert [mix.status(j)[:2] for j in (ja, jb, jc)] == [(2, nsteps)]*3
    assert mix.resident_particles == 0
    # the N-body class has no gas: no density, smoothing length or internal energy
    for f in (("r", "v", "a", "gpot") if kind == "plummer" else FIELDS):
        x, y = alone.get(ia, f), mix.get(ja, f)
        assert np.array_equal(x, y), (kind, f, float(np.max(np.abs(x - y))))
    assert alone.status(ia)[3] == mix.status(ja)[3]


def test_admission_by_capacity_and_retirement():
    from gandalf_amd import serving as S
    from gandalf_amd.gpu import GpuPool
    a, b = ics("sphere", (1500, 1200), 60)
    pool = GpuPool(S.class_params("sphere"), "mixed", max_particles=2000)
    ia = pool.admit(a["r"], a["v"], a["m"], a["u"], 2)
    assert ia is not None and pool.resident_particles == 1500
    assert pool.admit(b["r"], b["v"], b["m"], b["u"], 2) is None     # does not fit: queued
    assert pool.resident_particles == 1500
    assert pool.step() == (1, 1500)
    assert pool.step() == (1, 1500)
    assert pool.status(ia)[0] == 2 and pool.resident_particles == 0
    ib = pool.admit(b["r"], b["v"], b["m"], b["u"], 2)
    assert ib is not None and pool.resident_particles == 1200
    assert pool.step() == (1, 1200)


@pytest.mark.parametrize("kind", ["box", "sphere"])
def test_pool_agrees_with_standalone_context(kind):
    """The batch's tree uses 19 bits per axis under the request prefix (the standalone context
    21), so neighbour order and multipole cells may differ: after 3 steps positions agree to
    1e-6 of the system size and accelerations to the mixed-precision / tree-error level."""
    from gandalf_amd import serving as S
    from gandalf_amd.gpu import GpuPool, GpuSim
    (a,) = ics(kind, SIZES[kind][:1], 70)
    pool = GpuPool(S.class_params(kind), "mixed", max_particles=1 << 16)
    ia = pool.admit(a["r"], a["v"], a["m"], a["u"], 3)
    s = GpuSim(a["params"], precision="mixed")
    s.set_particles(a["r"], a["v"], a["m"], a["u"])
    s.setup()
    for _ in range(3):
        pool.step()
    s.step(3)
    assert np.max(np.abs(pool.get(ia, "r") - s.get("r"))) < 1e-6
    ap, asd = pool.get(ia, "a"), s.get("a")
    e = np.linalg.norm(ap - asd, axis=1) / np.max(np.linalg.norm(asd, axis=1))
    assert np.max(e) < (1e-4 if kind == "box" else 5e-3)


def test_serving_loop_small_stream():
    """Six requests of the three classes arriving at 50/s, 3 steps each: every request is
    served, each of its steps timed, and the particle-steps add up."""
    from gandalf_amd import serving as S
    reqs = S.request_stream(nreq=6, rate=50.0, nmin=512, nmax=2048, nsteps=3, seed=9)
    srv = S.Server(budget_bytes=1 << 30)
    out = srv.run(reqs)
    srv.close()
    assert all(q.finished > 0 and len(q.step_latency) == 3 for q in reqs)
    assert out["particle_steps"] == sum(3 * q.n for q in reqs)
    assert out["step_latency_ms"]["p99"] >= out["step_latency_ms"]["p50"] > 0


def test_failed_growth_leaves_resident_requests_untouched(monkeypatch):
    """Admitting a request grows the batch's arrays; if that allocation fails part way
    (forced from the host: GDF_FAULT_POOL_GROW makes the grow report an allocation failure
    after allocating), the admission fails with GDF_ERR_NOMEM, the resident request keeps its
    state, and the pool steps on to the same result as a pool that never saw the failure."""
    from gandalf_amd import serving as S
    from gandalf_amd.gpu import GdfError, GpuPool
    a, b = ics("sphere", (1500, 2000), 80)
    params = S.class_params("sphere")

    def run(fail_admission):
        pool = GpuPool(params, "mixed", max_particles=1 << 16)
        ia = pool.admit(a["r"], a["v"], a["m"], a["u"], 3)
        pool.step()
        if fail_admission:
            monkeypatch.setenv("GDF_FAULT_POOL_GROW", "1")
            with pytest.raises(GdfError, match="status -2: .*injected allocation failure.*"
                                               "resident requests are unchanged"):
                pool.admit(b["r"], b["v"], b["m"], b["u"], 3)
            monkeypatch.delenv("GDF_FAULT_POOL_GROW")
            assert pool.resident_particles == 1500
        pool.step()
        pool.step()
        assert pool.status(ia)[:2] == (2, 3)
        return {f: pool.get(ia, f) for f in FIELDS}

    clean, failed = run(False), run(True)
    for f in FIELDS:
        assert np.array_equal(clean[f], failed[f]), f
