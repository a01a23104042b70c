"""Column snapshots (Simulation::WriteColumnSnapshotFile / ReadColumnSnapshotFile,
SimulationIO.hpp:444-548, :193-272).  The writer, given the state the reference held after 10
steps of the periodic box, reproduces the reference's own snapshot file byte for byte (CPU);
a snapshot read into a device context and written back is the same file (GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden, vec

LIB = os.path.join(ROOT, "gandalf_amd", "libgandalf_gpu.so")


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.c_void_p)


@pytest.mark.skipif(not os.path.exists(LIB), reason="libgandalf_gpu.so not built")
def test_column_writer_matches_reference_file(tmp_path):
    d = load_golden("box16_m4")
    n = d["meta"]["nsteps"]
    ref = bytes(d[f"s{n}_column"])
    lib = C.CDLL(LIB)
    f = lib.gdf_format_column
    f.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double] + [C.c_void_p] * 6
    f.restype = C.c_int
    r = vec(d, f"s{n}")
    v = vec(d, f"s{n}", ("vx", "vy", "vz"))
    out = tmp_path / "snap.col"
    rc = f(str(out).encode(), r.shape[0], 0, float(d["diag"][n, 1]), _p(r), _p(v),
           _p(d["in_m"]), _p(d[f"s{n}_h"]), _p(d[f"s{n}_rho"]), _p(d[f"s{n}_u"]))
    assert rc == 0
    assert out.read_bytes() == ref


@pytest.mark.gpu
def test_column_snapshot_roundtrip_through_device(tmp_path):
    from gandalf_amd.gpu import GpuSim
    d = load_golden("box16_m4")
    n = d["meta"]["nsteps"]
    src = tmp_path / "ref.col"
    src.write_bytes(bytes(d[f"s{n}_column"]))
    tab = np.loadtxt(src, skiprows=4)
    s = GpuSim(d["meta"]["params"])
    s.read_snapshot(src)
    assert s.N == tab.shape[0]
    assert np.array_equal(s.get("r"), tab[:, 0:3])
    assert np.array_equal(s.get("v"), tab[:, 3:6])
    assert np.array_equal(s.get("h"), tab[:, 7]) and np.array_equal(s.get("rho"), tab[:, 8])
    assert s.time()[0] == float(open(src).read().split()[3])
    out = tmp_path / "out.col"
    s.write_snapshot(out)
    assert out.read_bytes() == src.read_bytes()
