"""ctypes binding of libgandalf_gpu (include/gandalf_gpu.h).

The only compute path is the HIP library: if libgandalf_gpu.so is missing or no GPU is
visible, construction raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgandalf_gpu.so")

GDF_KERNEL = {"m4": 0, "quintic": 1}
GDF_EOS = {"energy_eqn": 0, "isothermal": 1, "barotropic": 2}
PREC = {"fp64": 0, "mixed": 1}
# the reference's "multipole" parameter values (Parameters.cpp:309), as GDF_MULTIPOLE_*
MULTIPOLE = {"quadrupole": 0, "monopole": 1, "fast_monopole": 2, "fast_quadrupole": 3}
# and of its "gravity_mac" parameter (Parameters.cpp:308), as GDF_MAC_*
GRAVITY_MAC = {"geometric": 0, "gadget2": 1, "eigenmac": 2}
FIELDS3 = ("r", "v", "a", "atree", "r0", "v0", "a0", "adot")


GDF_ERR_CAPACITY = -5       # status of the C ABI: capacity exhausted / request does not fit


class GdfError(RuntimeError):
    """Non-zero status from the C ABI (the reference raises ExceptionHandler errors)."""


class Config(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("tabulated_kernel", C.c_int),
        ("h_fac", C.c_double), ("h_converge", C.c_double),
        ("gas_eos", C.c_int), ("gamma_eos", C.c_double), ("temp0", C.c_double),
        ("mu_bar", C.c_double), ("rho_bary", C.c_double),
        ("alpha_visc", C.c_double), ("beta_visc", C.c_double),
        ("hydro_forces", C.c_int), ("self_gravity", C.c_int),
        ("thetamaxsqd", C.c_double), ("Nleafmax", C.c_int), ("periodic", C.c_int),
        ("boxmin", C.c_double * 3), ("boxmax", C.c_double * 3),
        ("courant_mult", C.c_double), ("accel_mult", C.c_double), ("energy_mult", C.c_double),
        ("precision", C.c_int), ("group_max", C.c_int),
        ("nbody", C.c_int), ("nbody_direct", C.c_int), ("nbody_eps", C.c_double),
        ("nbody_mult", C.c_double), ("fixed_dt", C.c_double),
        ("Nlevels", C.c_int), ("level_diff_max", C.c_int),
        ("hrho_iteration_max", C.c_int), ("cand_batches_init", C.c_int),
        ("restart", C.c_int),
        ("nbody_softening", C.c_int), ("conservative_sph_star_gravity", C.c_int),
        ("n_ewald_grid", C.c_int), ("gr_bhewaldseriesn", C.c_int),
        ("ewald_mult", C.c_double), ("efratio", C.c_double),
        ("ntreebuildstep", C.c_int), ("ntreestockstep", C.c_int),
        ("graph_capture", C.c_int), ("multipole", C.c_int),
        ("gravity_mac", C.c_int), ("macerror", C.c_double),
    ]


_lib = None

# C prototypes of every entry point in include/gandalf_gpu.h
_vp, _i, _d, _i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
PROTOTYPES = {
    "gdf_config_default": (_i, [C.POINTER(Config)]),
    "gdf_create": (_i, [C.POINTER(Config), C.POINTER(_vp)]),
    "gdf_destroy": (None, [_vp]),
    "gdf_last_error": (C.c_char_p, [_vp]),
    "gdf_set_particles": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "gdf_get_field": (_i, [_vp, C.c_char_p, _vp]),
    "gdf_set_field": (_i, [_vp, C.c_char_p, _vp]),
    "gdf_num_particles": (_i, [_vp]),
    "gdf_build_tree": (_i, [_vp]),
    "gdf_update_all_sph_properties": (_i, [_vp, _i]),
    "gdf_update_all_sph_hydro_forces": (_i, [_vp]),
    "gdf_update_all_sph_forces": (_i, [_vp]),
    "gdf_maximum_smoothing_length": (_i, [_vp, C.POINTER(_d)]),
    "gdf_advance_particles": (_i, [_vp, _d]),
    "gdf_check_boundaries": (_i, [_vp]),
    "gdf_end_timestep": (_i, [_vp, _d]),
    "gdf_compute_global_timestep": (_i, [_vp, C.POINTER(_d)]),
    "gdf_calculate_diagnostics": (_i, [_vp, _vp]),
    "gdf_setup_simulation": (_i, [_vp]),
    "gdf_main_loop": (_i, [_vp, _i]),
    "gdf_time": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i64)]),
    "gdf_block_state": (_i, [_vp, _vp, _vp]),
    "gdf_write_snapshot": (_i, [_vp, C.c_char_p, C.c_char_p]),
    "gdf_read_snapshot": (_i, [_vp, C.c_char_p, C.c_char_p]),
    "gdf_format_column": (_i, [C.c_char_p, _i, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gdf_format_su": (_i, [C.c_char_p, _i, _i, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gdf_format_sf": (_i, [C.c_char_p, _i, _i, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gdf_format_slite": (_i, [C.c_char_p, _i, _i, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gdf_pool_create": (_i, [C.POINTER(Config), _i64, C.POINTER(_vp)]),
    "gdf_pool_destroy": (None, [_vp]),
    "gdf_pool_last_error": (C.c_char_p, [_vp]),
    "gdf_pool_admit": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i, C.POINTER(_i)]),
    "gdf_pool_step": (_i, [_vp, _vp]),
    "gdf_pool_status": (_i, [_vp, _i, _vp, C.POINTER(_d)]),
    "gdf_pool_get_field": (_i, [_vp, _i, C.c_char_p, _vp]),
    "gdf_pool_release": (_i, [_vp, _i]),
    "gdf_pool_resident_particles": (_i64, [_vp]),
    "gdf_neighbour_sets": (_i64, [_vp, _i, _vp, _vp, _i64]),
    "gdf_tree_arrays": (_i, [_vp] + [_vp] * 15 + [C.POINTER(_i)]),
    "gdf_interaction_counts": (_i, [_vp, _vp]),
    "gdf_gravity_group_counts": (_i, [_vp, _vp, _vp]),
    "gdf_pass_times": (_i, [_vp, _vp]),
    "gdf_set_timing": (_i, [_vp, _i]),
    "gdf_nbody_direct_forces": (_i, [_vp, _i]),
    "gdf_calculate_diagnostics_noalloc": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp]),
    "gdf_advance_particles_block": (_i, [_vp, _i, _d, _d]),
    "gdf_set_active_particles": (_i, [_vp, _i]),
    "gdf_correction_terms": (_i, [_vp, _i, _d, _d]),
    "gdf_check_timesteps": (_i, [_vp, _i, _i, _i, _d, C.POINTER(_i)]),
    "gdf_compute_block_timesteps": (_i, [_vp]),
    "gdf_end_timestep_block": (_i, [_vp, _i, _d, _d]),
    "gdf_get_gather_neighbour_list": (_i, [_vp, _vp, _d, _i, _vp]),
    "gdf_snapshot_counters": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "gdf_build_tree_step": (_i, [_vp, _i, _i, _i, _i]),
    "gdf_graph_state": (_i, [_vp, C.POINTER(_i64), C.POINTER(_i)]),
    "gdf_set_opening_criterion": (_i, [_vp, _i]),
    "gdf_opening_criterion": (_i, [_vp]),
    "gdf_set_stars": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "gdf_num_stars": (_i, [_vp]),
    "gdf_get_star_field": (_i, [_vp, C.c_char_p, _vp]),
    "gdf_update_all_star_gas_forces": (_i, [_vp]),
    "gdf_nbody_star_forces": (_i, [_vp, _i]),
    "gdf_nbody_advance_particles": (_i, [_vp, _d]),
    "gdf_nbody_correction_terms": (_i, [_vp, _d]),
    "gdf_nbody_end_timestep": (_i, [_vp, _d]),
}


def lib():
    """Load libgandalf_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GdfError(f"{LIB_PATH} is missing: build it with python -m gandalf_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def config_from_params(prm: dict, precision: str = "fp64", group_max: int = 64) -> Config:
    """gdf_config from a reference-style parameter dict (defaults: Parameters.cpp:160-575)."""
    c = Config()
    lib().gdf_config_default(C.byref(c))
    c.kernel = GDF_KERNEL[prm.get("kernel", "m4")]
    c.tabulated_kernel = int(prm.get("tabulated_kernel", 1))
    c.h_fac = float(prm.get("h_fac", 1.2))
    c.h_converge = float(prm.get("h_converge", 0.01))
    c.gas_eos = GDF_EOS[prm.get("gas_eos", "energy_eqn")]
    c.gamma_eos = float(prm.get("gamma_eos", 1.66666666666666))
    c.temp0 = float(prm.get("temp0", 1.0))
    c.mu_bar = float(prm.get("mu_bar", 1.0))
    c.rho_bary = float(prm.get("rho_bary", 1.0e-14))
    c.alpha_visc = float(prm.get("alpha_visc", 1.0))
    c.beta_visc = float(prm.get("beta_visc", 2.0))
    c.hydro_forces = int(prm.get("hydro_forces", 1))
    c.self_gravity = int(prm.get("self_gravity", 0))
    c.thetamaxsqd = float(prm.get("thetamaxsqd", 0.1))
    c.Nleafmax = int(prm.get("Nleafmax", 6))
    c.periodic = 1 if prm.get("boundary_lhs[0]", "open") == "periodic" else 0
    for k in range(3):
        c.boxmin[k] = float(prm.get(f"boxmin[{k}]", -9.9e30))
        c.boxmax[k] = float(prm.get(f"boxmax[{k}]", 9.9e30))
    c.courant_mult = float(prm.get("courant_mult", 0.15))
    c.accel_mult = float(prm.get("accel_mult", 0.3))
    c.energy_mult = float(prm.get("energy_mult", 0.4))
    c.precision = PREC[precision]
    c.group_max = int(group_max)
    # pure N-body (sim = nbody): tree walk with Plummer softening nbody_eps (an extension), or
    # the reference's direct sums selected by nbody_softening
    c.nbody = 1 if prm.get("sim", "gradhsph") == "nbody" else 0
    c.nbody_eps = float(prm.get("nbody_eps", 0.0))
    c.nbody_softening = int(prm.get("nbody_softening", 1))
    c.conservative_sph_star_gravity = int(prm.get("conservative_sph_star_gravity", 1))
    c.n_ewald_grid = int(prm.get("nEwaldGrid", 16))
    c.gr_bhewaldseriesn = int(prm.get("gr_bhewaldseriesn", 10))
    c.ewald_mult = float(prm.get("ewald_mult", 1.0))
    c.efratio = float(prm.get("EFratio", 1.0))
    c.ntreebuildstep = int(prm.get("ntreebuildstep", 1))
    c.ntreestockstep = int(prm.get("ntreestockstep", 1))
    c.graph_capture = int(prm.get("graph_capture", 0))
    c.multipole = MULTIPOLE[prm.get("multipole", "quadrupole")]
    c.gravity_mac = GRAVITY_MAC[prm.get("gravity_mac", "geometric")]
    c.macerror = float(prm.get("macerror", 1.0e-4))
    if c.nbody and c.nbody_eps <= 0.0:
        c.nbody_direct = 2 if c.nbody_softening else 1
    c.nbody_mult = float(prm.get("nbody_mult", 0.1))
    c.fixed_dt = float(prm.get("fixed_dt", 0.0))
    c.Nlevels = int(prm.get("Nlevels", 1))
    c.level_diff_max = int(prm.get("level_diff_max", 1))
    c.hrho_iteration_max = int(prm.get("hrho_iteration_max", 0))
    c.cand_batches_init = int(prm.get("cand_batches_init", 0))
    c.restart = int(prm.get("restart", 0))
    return c


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class GpuSim:
    """One device context (one HIP stream, device-resident particle state)."""

    def __init__(self, params: dict, precision: str = "fp64", group_max: int = 64):
        self.params = dict(params)
        self.cfg = config_from_params(params, precision, group_max)
        h = C.c_void_p()
        rc = lib().gdf_create(C.byref(self.cfg), C.byref(h))
        if rc != 0:
            raise GdfError(f"gdf_create failed ({rc})")
        self._h = h
        self.N = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().gdf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            msg = lib().gdf_last_error(self._h).decode()
            raise GdfError(f"status {rc}: {msg}")
        return rc

    # ---- particle data ----
    def set_particles(self, r, v, m, u, h=None):
        r, v, m, u = _d(r), _d(v), _d(m), _d(u)
        hh = None if h is None else _d(h)
        self.N = r.shape[0]
        self._chk(lib().gdf_set_particles(self._h, self.N, _p(r), _p(v), _p(m), _p(hh), _p(u)))

    def get(self, name):
        w = 3 if name in FIELDS3 else 1
        out = np.empty(self.N * w)
        self._chk(lib().gdf_get_field(self._h, name.encode(), _p(out)))
        return out.reshape(self.N, 3) if w == 3 else out

    def set(self, name, arr):
        self._chk(lib().gdf_set_field(self._h, name.encode(), _p(_d(arr))))

    # ---- hot path ----
    def build_tree(self):
        self._chk(lib().gdf_build_tree(self._h))

    def build_tree_step(self, rebuild_tree, n, ntreebuildstep, ntreestockstep):
        """HydroTree::BuildTree's schedule: True after a rebuild, False after a refit."""
        return bool(self._chk(lib().gdf_build_tree_step(self._h, int(bool(rebuild_tree)), int(n),
                                                        int(ntreebuildstep),
                                                        int(ntreestockstep))))

    def set_opening_criterion(self, mac):
        """NeighbourSearch::SetOpeningCriterion: "geometric", "gadget2" or "eigenmac"."""
        self._chk(lib().gdf_set_opening_criterion(self._h, GRAVITY_MAC[mac]))

    def opening_criterion(self):
        names = {v: k for k, v in GRAVITY_MAC.items()}
        return names[self._chk(lib().gdf_opening_criterion(self._h))]

    def graph_state(self):
        """(replays of the captured tree-build graphs, state: 1 in use, 0 off, negative: why a
        capture was given up)"""
        r, u = C.c_int64(), C.c_int()
        self._chk(lib().gdf_graph_state(self._h, C.byref(r), C.byref(u)))
        return r.value, u.value

    def density(self, nopbc=False):
        self._chk(lib().gdf_update_all_sph_properties(self._h, int(nopbc)))

    def hydro_forces(self):
        self._chk(lib().gdf_update_all_sph_hydro_forces(self._h))

    def all_forces(self):
        self._chk(lib().gdf_update_all_sph_forces(self._h))

    def advance(self, dt):
        self._chk(lib().gdf_advance_particles(self._h, dt))

    def end_timestep(self, dt):
        self._chk(lib().gdf_end_timestep(self._h, dt))

    def global_timestep(self):
        d = C.c_double()
        self._chk(lib().gdf_compute_global_timestep(self._h, C.byref(d)))
        return d.value

    def diagnostics(self):
        out = np.zeros(11)
        self._chk(lib().gdf_calculate_diagnostics(self._h, _p(out)))
        return out

    def setup(self):
        self._chk(lib().gdf_setup_simulation(self._h))

    def step(self, n=1):
        self._chk(lib().gdf_main_loop(self._h, int(n)))

    def time(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_int64()
        self._chk(lib().gdf_time(self._h, C.byref(t), C.byref(dt), C.byref(n)))
        return t.value, dt.value, n.value

    def write_snapshot(self, path, fmt="column"):
        self._chk(lib().gdf_write_snapshot(self._h, str(path).encode(), fmt.encode()))

    def read_snapshot(self, path, fmt="column"):
        self._chk(lib().gdf_read_snapshot(self._h, str(path).encode(), fmt.encode()))
        self.N = lib().gdf_num_particles(self._h)

    def snapshot_counters(self, nout=None, nlite=None, tlast=None, tlitelast=None):
        """(Noutsnap, Noutlitesnap, tsnaplast, tlitesnaplast) the snapshot writers use; values
        given replace them (all four together, or none)."""
        g = np.zeros(2, dtype=np.int64)
        gd = np.zeros(2)
        st = sd = None
        if nout is not None:
            st = np.array([nout, nlite], dtype=np.int64)
            sd = np.array([tlast, tlitelast], dtype=np.float64)
        self._chk(lib().gdf_snapshot_counters(self._h, _p(st), _p(sd), _p(g), _p(gd)))
        return int(g[0]), int(g[1]), float(gd[0]), float(gd[1])

    def block_state(self):
        """(n, level_max, level_step, nresync, dt_max, particles that ended a step last)"""
        iv = np.zeros(5, dtype=np.int64)
        dv = np.zeros(1)
        self._chk(lib().gdf_block_state(self._h, _p(iv), _p(dv)))
        return int(iv[0]), int(iv[1]), int(iv[2]), int(iv[3]), float(dv[0]), int(iv[4])

    def set_timing(self, on=True):
        self._chk(lib().gdf_set_timing(self._h, int(on)))

    def pass_times(self):
        out = np.zeros(9)
        self._chk(lib().gdf_pass_times(self._h, _p(out)))
        return dict(zip(("tree", "density", "hydro", "gravity", "integrate", "k_gravity",
                         "k_gravity_launches", "k_hydro", "k_hydro_launches"), out))

    def interaction_counts(self):
        out = np.zeros(6, dtype=np.int64)
        self._chk(lib().gdf_interaction_counts(self._h, _p(out)))
        return dict(zip(("gather_pairs", "sg_pairs", "grav_pp", "grav_pc", "density_evals",
                         "sg_tests"), (int(x) for x in out)))

    # ---- parity taps ----
    def neighbour_sets(self, mode):
        off = np.zeros(self.N + 1, dtype=np.int64)
        tot = self._chk(lib().gdf_neighbour_sets(self._h, mode, _p(off), None, 0))
        ids = np.zeros(max(tot, 1), dtype=np.int32)
        got = self._chk(lib().gdf_neighbour_sets(self._h, mode, _p(off), _p(ids), tot))
        return off, ids[:got]

    def tree_arrays(self):
        N = self.N
        nn, ni = 2 * N - 1, max(N - 1, 1)
        a = dict(perm=np.zeros(N, np.int32), keys=np.zeros(N, np.uint64),
                 left=np.zeros(ni, np.int32), right=np.zeros(ni, np.int32),
                 first=np.zeros(ni, np.int32), last=np.zeros(ni, np.int32),
                 mass=np.zeros(nn), com=np.zeros((nn, 3)), quad=np.zeros((nn, 5)),
                 bbmin=np.zeros((nn, 3)), bbmax=np.zeros((nn, 3)), hmax=np.zeros(nn),
                 group_first=np.zeros(N, np.int32), group_count=np.zeros(N, np.int32),
                 group_node=np.zeros(N, np.int32))
        ng = C.c_int()
        order = ("perm", "keys", "left", "right", "first", "last", "mass", "com", "quad",
                 "bbmin", "bbmax", "hmax", "group_first", "group_count", "group_node")
        self._chk(lib().gdf_tree_arrays(self._h, *[_p(a[k]) for k in order], C.byref(ng)))
        g = ng.value
        for k in ("group_first", "group_count", "group_node"):
            a[k] = a[k][:g]
        return a

    def nbody_direct(self, smoothed=False):
        """Star-star direct sums (Nbody::CalculateDirectGravForces, or the kernel-smoothed
        NbodyLeapfrogKDK variant): fills a, adot and gpot."""
        self._chk(lib().gdf_nbody_direct_forces(self._h, int(bool(smoothed))))

    # ---- stars beside the gas (Nbody<3>::stardata of an SphSimulation) ----
    STAR_FIELDS3 = ("r", "v", "a", "adot", "r0", "v0", "a0")

    def set_stars(self, r, v, m, h):
        r, v, m, h = _d(r).reshape(-1, 3), _d(v).reshape(-1, 3), _d(m), _d(h)
        self._chk(lib().gdf_set_stars(self._h, r.shape[0], _p(r), _p(v), _p(m), _p(h)))

    @property
    def Nstar(self):
        return lib().gdf_num_stars(self._h)

    def star(self, name):
        n = self.Nstar
        w = 3 if name in self.STAR_FIELDS3 else 1
        out = np.empty(n * w)
        self._chk(lib().gdf_get_star_field(self._h, name.encode(), _p(out)))
        return out.reshape(n, 3) if w == 3 else out

    def update_all_star_gas_forces(self):
        self._chk(lib().gdf_update_all_star_gas_forces(self._h))

    def nbody_star_forces(self, zero_first=False):
        self._chk(lib().gdf_nbody_star_forces(self._h, int(bool(zero_first))))

    def nbody_advance_particles(self, t):
        self._chk(lib().gdf_nbody_advance_particles(self._h, float(t)))

    def nbody_correction_terms(self, t):
        self._chk(lib().gdf_nbody_correction_terms(self._h, float(t)))

    def nbody_end_timestep(self, t):
        self._chk(lib().gdf_nbody_end_timestep(self._h, float(t)))

    # ---- per-method entry points (TimeIntegration / NeighbourSearch of the reference) ----
    def advance_particles_block(self, n, t, timestep):
        self._chk(lib().gdf_advance_particles_block(self._h, int(n), float(t), float(timestep)))

    def set_active_particles(self, n):
        self._chk(lib().gdf_set_active_particles(self._h, int(n)))

    def correction_terms(self, n, t, timestep):
        self._chk(lib().gdf_correction_terms(self._h, int(n), float(t), float(timestep)))

    def check_timesteps(self, level_diff_max, level_step, n, timestep):
        out = C.c_int(0)
        self._chk(lib().gdf_check_timesteps(self._h, int(level_diff_max), int(level_step),
                                            int(n), float(timestep), C.byref(out)))
        return out.value

    def compute_block_timesteps(self):
        self._chk(lib().gdf_compute_block_timesteps(self._h))

    def end_timestep_block(self, n, t, timestep):
        self._chk(lib().gdf_end_timestep_block(self._h, int(n), float(t), float(timestep)))

    def check_boundaries(self):
        self._chk(lib().gdf_check_boundaries(self._h))

    def gather_neighbour_list(self, rp, rsearch, nmax=None):
        """Import-order ids within rsearch of rp, or None when more than nmax (the reference's
        -1 overflow code)."""
        nmax = self.N if nmax is None else int(nmax)
        out = np.zeros(max(nmax, 1), dtype=np.int32)
        n = lib().gdf_get_gather_neighbour_list(self._h, _p(_d(rp)), float(rsearch), nmax,
                                                 _p(out))
        if n == -1:
            return None
        self._chk(n)
        return out[:n].copy()

    def gravity_group_counts(self):
        a = self.tree_arrays()
        g = a["group_first"].size
        nc = np.zeros(g, np.int32)
        nn = np.zeros(g, np.int32)
        self._chk(lib().gdf_gravity_group_counts(self._h, _p(nc), _p(nn)))
        return nc, nn


class GpuPool:
    """A serving batch (gdf_pool): independent requests of one configuration class resident in
    one device context, advanced together one step at a time."""

    def __init__(self, params: dict, precision: str = "mixed", max_particles: int = 1 << 22,
                 group_max: int = 64):
        self.params = dict(params)
        cfg = config_from_params(self.params, precision, group_max)
        h = C.c_void_p()
        rc = lib().gdf_pool_create(C.byref(cfg), int(max_particles), C.byref(h))
        if rc != 0:
            raise GdfError(f"gdf_pool_create failed ({rc})")
        self._h = h
        self.max_particles = int(max_particles)

    def close(self):
        if getattr(self, "_h", None):
            lib().gdf_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise GdfError(f"status {rc}: {lib().gdf_pool_last_error(self._h).decode()}")
        return rc

    def admit(self, r, v, m, u, nsteps):
        """Request id, or None when it does not fit beside the resident requests."""
        r, v, m, u = _d(r), _d(v), _d(m), _d(u)
        rid = C.c_int(-1)
        rc = lib().gdf_pool_admit(self._h, r.shape[0], _p(r), _p(v), _p(m), _p(u), int(nsteps),
                                  C.byref(rid))
        if rc == GDF_ERR_CAPACITY:
            return None
        self._chk(rc)
        return rid.value

    def step(self):
        """One step of every resident request: (requests stepped, particle-steps)."""
        out = np.zeros(2, dtype=np.int64)
        self._chk(lib().gdf_pool_step(self._h, _p(out)))
        return int(out[0]), int(out[1])

    def status(self, rid):
        """(state 0 unknown / 1 resident / 2 finished, steps done, particles, time)"""
        out = np.zeros(3, dtype=np.int64)
        t = C.c_double()
        self._chk(lib().gdf_pool_status(self._h, int(rid), _p(out), C.byref(t)))
        return int(out[0]), int(out[1]), int(out[2]), t.value

    def get(self, rid, name):
        n = self.status(rid)[2]
        w = 3 if name in FIELDS3 else 1
        out = np.empty(n * w)
        self._chk(lib().gdf_pool_get_field(self._h, int(rid), name.encode(), _p(out)))
        return out.reshape(n, 3) if w == 3 else out

    def release(self, rid):
        self._chk(lib().gdf_pool_release(self._h, int(rid)))

    @property
    def resident_particles(self):
        return int(lib().gdf_pool_resident_particles(self._h))
