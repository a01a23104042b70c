"""gdf_calculate_diagnostics_noalloc (device arrays the caller owns, no device allocation)
against gdf_calculate_diagnostics on the same particles: two calls in a row, the second on a
smaller input, so the shared scratch array holds the first call's partials when the second
runs.  Tolerance: the two sum the same terms in different particle orders (the context holds
its particles in Morton order), so they agree to FP64 rounding, 1e-13 of each term's scale."""
import ctypes

import numpy as np
import pytest

from conftest import load_golden, vec
from test_oracle_golden import diag_scales

pytestmark = pytest.mark.gpu

GDF_ERR_INVALID = -1


def context_diagnostics(prm, r, v, m, u, gpot):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(prm)
    s.set_particles(r, v, m, u)
    s.set("gpot", gpot)
    return s.diagnostics()


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime the library itself links (the
    caller's own allocations: the entry point under test allocates nothing)."""

    def __init__(self):
        from gandalf_amd.gpu import lib
        lib()                                   # loads libamdhip64 with the library
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_size_t(a.nbytes), 1) == 0      # host to device
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p,
                                  ctypes.c_size_t(out.nbytes), 2) == 0    # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc_diagnostics(r, v, m, u, gpot, hydro, grav):
    from gandalf_amd.gpu import lib
    D = DeviceArrays()
    try:
        ins = [D.put(a) for a in (r, v, m, u, gpot)]
        out = D.put(np.full(11, np.nan))
        rc = lib().gdf_calculate_diagnostics_noalloc(len(m), *ins, int(hydro), int(grav), out,
                                                     None)
        assert D.hip.hipDeviceSynchronize() == 0
        return rc, D.get(out, 11)
    finally:
        D.free()


def check_same(a, b, r, v, m):
    pm, lm = diag_scales(r, v, m)
    for k in range(4):
        assert a[k] == pytest.approx(b[k], rel=1e-13, abs=1e-13 * abs(b[0])), k
    assert np.allclose(a[4:7], b[4:7], rtol=0, atol=1e-13 * pm)
    assert np.allclose(a[7:10], b[7:10], rtol=0, atol=1e-13 * lm)
    assert a[10] == pytest.approx(b[10], rel=1e-14)


def test_noalloc_matches_context_entry_point_twice():
    d = load_golden("sphere4k_m4_bf")
    prm = d["meta"]["params"]
    r, v = vec(d, "s3"), vec(d, "s3", ("vx", "vy", "vz"))
    m, u, gpot = d["in_m"], d["s3_u"], d["s3_gpot"]
    for n in (len(m), 1000):         # the second call is smaller: fewer partial blocks
        print(f"noalloc diagnostics: sphere4k_m4_bf state at step 3, first {n} particles",
              flush=True)
        sl = slice(0, n)
        want = context_diagnostics(prm, r[sl], v[sl], m[sl], u[sl], gpot[sl])
        rc, got = noalloc_diagnostics(r[sl], v[sl], m[sl], u[sl], gpot[sl], 1, 1)
        assert rc == 0
        check_same(got, want, r[sl], v[sl], m[sl])


def test_noalloc_rejects_sizes_beyond_maximum():
    from gandalf_amd.gpu import lib
    L = lib()
    for n in (0, (1 << 30) + 1):
        rc = L.gdf_calculate_diagnostics_noalloc(n, None, None, None, None, None, 1, 1, None,
                                                 None)
        assert rc == GDF_ERR_INVALID
        assert b"N must be in [1, 2^30]" in L.gdf_last_error(None)
    rc = L.gdf_calculate_diagnostics_noalloc(8, None, None, None, None, None, 1, 1, None, None)
    assert rc == GDF_ERR_INVALID
    assert b"device-accessible" in L.gdf_last_error(None)
