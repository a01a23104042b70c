"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, the smoke check and the benchmark's cpu_baseline leg, never by the
product package.  Builds liboracle.so from gdf_oracle.c on first use if it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

W0, W1, WOMEGA, WZETA, WGRAV, WPOT, W0_S2, WOMEGA_S2, WZETA_S2 = range(9)


class Params(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("tabulated", C.c_int),
        ("h_fac", C.c_double), ("h_converge", C.c_double),
        ("eos", C.c_int), ("gamma", C.c_double), ("temp0", C.c_double),
        ("mu_bar", C.c_double), ("rho_bary", C.c_double),
        ("energy_integration", C.c_int),
        ("alpha_visc", C.c_double), ("beta_visc", C.c_double),
        ("hydro_forces", C.c_int), ("self_gravity", C.c_int),
        ("thetamaxsqd", C.c_double), ("nleafmax", C.c_int), ("periodic", C.c_int),
        ("boxmin", C.c_double * 3), ("boxmax", C.c_double * 3),
        ("courant_mult", C.c_double), ("accel_mult", C.c_double), ("energy_mult", C.c_double),
        ("nlevels", C.c_int), ("level_diff_max", C.c_int), ("grav_direct", C.c_int),
        ("multipole", C.c_int), ("gravity_mac", C.c_int), ("macerror", C.c_double),
    ]


class Tree(C.Structure):
    _fields_ = [
        ("N", C.c_int), ("perm", C.POINTER(C.c_int32)), ("keys", C.POINTER(C.c_uint64)),
        ("left", C.POINTER(C.c_int32)), ("right", C.POINTER(C.c_int32)),
        ("first", C.POINTER(C.c_int32)), ("last", C.POINTER(C.c_int32)),
        ("parent", C.POINTER(C.c_int32)),
        ("mass", C.POINTER(C.c_double)), ("com", C.POINTER(C.c_double)),
        ("quad", C.POINTER(C.c_double)), ("bbmin", C.POINTER(C.c_double)),
        ("bbmax", C.POINTER(C.c_double)), ("hmax", C.POINTER(C.c_double)),
        ("cdistsqd", C.POINTER(C.c_double)),
        ("group_first", C.POINTER(C.c_int32)), ("group_count", C.POINTER(C.c_int32)),
        ("group_node", C.POINTER(C.c_int32)), ("ngroups", C.c_int),
    ]


class Ewald(C.Structure):
    """orc_ewald: the correction table of a fully periodic box."""
    _fields_ = [("Ng", C.c_int * 3), ("dI", C.c_double * 3), ("L", C.c_double * 3),
                ("field", C.c_void_p)]


def ewald_table(L, nseries=10, ngrid=16, ewald_mult=1.0, efratio=1.0):
    """Ewald::Ewald's correction field: an Ewald handle whose .table is (4, Ng_z, Ng_y, Ng_x)
    (potential, a_x, a_y, a_z) and .spacing the node spacing per axis."""
    Lv = _d(np.asarray(L, dtype=np.float64).reshape(3))
    Ng = np.zeros(3, dtype=np.int32)
    dI = np.zeros(3)
    lib().orc_ewald_table(_p(Lv), int(nseries), int(ngrid), float(ewald_mult), float(efratio),
                          None, _p(Ng), _p(dI))
    field = np.zeros(4 * int(np.prod(Ng)))
    lib().orc_ewald_table(_p(Lv), int(nseries), int(ngrid), float(ewald_mult), float(efratio),
                          _p(field), _p(Ng), _p(dI))
    e = Ewald()
    for k in range(3):
        e.Ng[k] = int(Ng[k]); e.dI[k] = float(dI[k]); e.L[k] = float(Lv[k])
    e.field = _p(field)
    e.table = field.reshape(4, int(Ng[2]), int(Ng[1]), int(Ng[0]))
    e.spacing = 1.0 / dI
    e._keep = field
    return e


def ewald_corr(e, m, dr):
    """Ewald::CalculatePeriodicCorrection: (acceleration, potential) correction of mass m at
    dr = r_source - r_target."""
    a = np.zeros(3); pot = np.zeros(1)
    lib().orc_ewald_corr(C.byref(e), float(m), _p(_d(np.asarray(dr, dtype=np.float64))), _p(a),
                         _p(pot))
    return a, float(pot[0])


def direct_gravity_periodic(p: Params, e, r, m, st):
    """Exact periodic gravity: nearest-image pair terms plus every pair's Ewald correction."""
    r, m = _d(r), _d(m)
    N = r.shape[0]
    at = np.zeros((N, 3)); gp = np.zeros(N)
    lib().orc_direct_gravity_periodic(C.byref(p), C.byref(e), N, _p(r), _p(m),
                                      _p(_d(st["h"])), _p(_d(st["zeta"])),
                                      _p(_d(st["hfactor"])), _p(at), _p(gp))
    return at, gp


class Stars(C.Structure):
    """orc_stars: stars beside the gas (positions (n,3), masses, smoothing lengths)."""
    _fields_ = [("n", C.c_int), ("r", C.c_void_p), ("m", C.c_void_p), ("h", C.c_void_p),
                ("softening", C.c_int), ("conservative", C.c_int)]


def stars(r, m, h, softening=1, conservative=1):
    """An orc_stars view of the arrays; the arrays are kept on the returned object."""
    r, m, h = _d(r).reshape(-1, 3), _d(m), _d(h)
    st = Stars(r.shape[0], _p(r), _p(m), _p(h), int(softening), int(conservative))
    st._keep = (r, m, h)
    return st


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        src = os.path.join(_HERE, "gdf_oracle.c")
        if (not os.path.exists(_LIB)) or os.path.getmtime(_LIB) < os.path.getmtime(src):
            subprocess.check_call(["make", "-s", "-f", os.path.join(_HERE, "Makefile")])
        L = C.CDLL(_LIB)
        vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        pp = C.POINTER(Params)
        L.orc_kernel_eval.argtypes = [pp, i, d]; L.orc_kernel_eval.restype = d
        L.orc_h_guess.argtypes = [pp, i, vp]; L.orc_h_guess.restype = d
        L.orc_neighbour_sets.argtypes = [pp, i, vp, vp, i, vp, vp, i64]
        L.orc_neighbour_sets.restype = i64
        L.orc_compute_h.argtypes = [pp, i, vp, vp, i] + [vp] * 8
        L.orc_compute_h.restype = i
        L.orc_hydro_forces.argtypes = [pp, i, i] + [vp] * 12
        L.orc_direct_gravity.argtypes = [pp, i] + [vp] * 7
        L.orc_direct_plummer.argtypes = [i, vp, vp, d, vp, i, vp, vp]
        L.orc_nbody_direct.argtypes = [i] + [vp] * 6
        L.orc_nbody_direct_smoothed.argtypes = [pp, i] + [vp] * 7
        L.orc_tree_alloc.argtypes = [C.POINTER(Tree), i]
        L.orc_tree_free.argtypes = [C.POINTER(Tree)]
        L.orc_tree_build.argtypes = [pp, C.POINTER(Tree), vp, vp, vp, i]
        L.orc_tree_build.restype = i
        L.orc_tree_refit_h.argtypes = [pp, C.POINTER(Tree), vp]
        L.orc_tree_gravity.argtypes = [pp, C.POINTER(Tree)] + [vp] * 11
        L.orc_cell_quad_force.argtypes = [vp, vp, d, vp, vp, vp]
        L.orc_cell_mono_force.argtypes = [vp, d, vp, vp, vp]
        L.orc_cell_mac_opens.argtypes = [pp, vp, d, vp, d, d]
        L.orc_fast_multipole.argtypes = [i, vp, vp, vp, i, i, vp, vp, vp, vp]
        L.orc_timesteps.argtypes = [pp, i] + [vp] * 7
        L.orc_timesteps.restype = d
        L.orc_diagnostics.argtypes = [pp, i] + [vp] * 6
        L.orc_sim_create.argtypes = [pp, i, vp, vp, vp, vp]
        L.orc_sim_create.restype = vp
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_setup.argtypes = [vp]; L.orc_sim_setup.restype = i
        L.orc_sim_step.argtypes = [vp]; L.orc_sim_step.restype = i
        L.orc_sim_get.argtypes = [vp, C.c_char_p, vp]; L.orc_sim_get.restype = i
        L.orc_sim_time.argtypes = [vp]; L.orc_sim_time.restype = d
        L.orc_sim_timestep.argtypes = [vp]; L.orc_sim_timestep.restype = d
        L.orc_sim_block_state.argtypes = [vp, vp, vp]
        L.orc_ewald_table.argtypes = [vp, i, i, d, d, vp, vp, vp]
        L.orc_ewald_table.restype = i
        L.orc_ewald_corr.argtypes = [C.POINTER(Ewald), d, vp, vp, vp]
        L.orc_direct_gravity_periodic.argtypes = [pp, C.POINTER(Ewald), i] + [vp] * 7
        L.orc_compute_h_stars.argtypes = [pp, i, vp, vp, i, C.POINTER(Stars)] + [vp] * 8
        L.orc_compute_h_stars.restype = i
        L.orc_gas_from_stars.argtypes = [pp, i, vp, vp, C.POINTER(Stars), vp, vp]
        L.orc_stars_from_gas.argtypes = [pp, C.POINTER(Stars), i, vp, vp, vp, vp, vp]
        L.orc_sim_set_stars.argtypes = [vp, i, vp, vp, vp, vp, i, i, d]
        L.orc_sim_set_stars.restype = i
        _lib = L
    return _lib


# the reference's "multipole" parameter values (Parameters.cpp:309)
MULTIPOLE = {"quadrupole": 0, "monopole": 1, "fast_monopole": 2, "fast_quadrupole": 3}
# and of its "gravity_mac" parameter (Parameters.cpp:308)
GRAVITY_MAC = {"geometric": 0, "gadget2": 1, "eigenmac": 2}


def params_from_dict(prm: dict) -> Params:
    """Oracle parameters from a reference-style parameter dict (gandalf_amd.ic)."""
    p = Params()
    p.kernel = 0 if prm.get("kernel", "m4") == "m4" else 1
    p.tabulated = int(prm.get("tabulated_kernel", 1))
    p.h_fac = float(prm.get("h_fac", 1.2))
    p.h_converge = float(prm.get("h_converge", 0.01))
    eos = prm.get("gas_eos", "energy_eqn")
    p.eos = {"energy_eqn": 0, "isothermal": 1, "barotropic": 2}[eos]
    p.gamma = float(prm.get("gamma_eos", 1.66666666666666))
    p.temp0 = float(prm.get("temp0", 1.0))
    p.mu_bar = float(prm.get("mu_bar", 1.0))
    p.rho_bary = float(prm.get("rho_bary", 1.0e-14))
    p.energy_integration = 1 if eos == "energy_eqn" else 0
    p.alpha_visc = float(prm.get("alpha_visc", 1.0))
    p.beta_visc = float(prm.get("beta_visc", 2.0))
    p.hydro_forces = int(prm.get("hydro_forces", 1))
    p.self_gravity = int(prm.get("self_gravity", 0))
    p.thetamaxsqd = float(prm.get("thetamaxsqd", 0.1))
    p.nleafmax = int(prm.get("Nleafmax", 6))
    p.periodic = 1 if prm.get("boundary_lhs[0]", "open") == "periodic" else 0
    for k in range(3):
        p.boxmin[k] = float(prm.get(f"boxmin[{k}]", -9.9e30))
        p.boxmax[k] = float(prm.get(f"boxmax[{k}]", 9.9e30))
    p.courant_mult = float(prm.get("courant_mult", 0.15))
    p.accel_mult = float(prm.get("accel_mult", 0.3))
    p.energy_mult = float(prm.get("energy_mult", 0.4))
    p.nlevels = int(prm.get("Nlevels", 1))
    p.level_diff_max = int(prm.get("level_diff_max", 1))
    p.grav_direct = 1 if prm.get("neib_search", "kdtree") == "bruteforce" else 0
    p.multipole = MULTIPOLE[prm.get("multipole", "quadrupole")]
    p.gravity_mac = GRAVITY_MAC[prm.get("gravity_mac", "geometric")]
    p.macerror = float(prm.get("macerror", 1.0e-4))
    return p


def kernel_eval(p: Params, which: int, s: float) -> float:
    return lib().orc_kernel_eval(C.byref(p), which, s)


def h_guess(p: Params, r) -> float:
    r = _d(r)
    return lib().orc_h_guess(C.byref(p), r.shape[0], _p(r))


def neighbour_sets(p: Params, r, h, mode: int):
    """CSR (offsets, ids) of the exact neighbour sets; mode 0 gather, 1 scatter-gather,
    2 inclusive search radius k*h."""
    r, h = _d(r), _d(h)
    N = r.shape[0]
    off = np.zeros(N + 1, dtype=np.int64)
    tot = lib().orc_neighbour_sets(C.byref(p), N, _p(r), _p(h), mode, _p(off), None, 0)
    ids = np.zeros(max(abs(tot), 1), dtype=np.int32)
    lib().orc_neighbour_sets(C.byref(p), N, _p(r), _p(h), mode, _p(off), _p(ids), abs(tot))
    return off, ids[:abs(tot)]


def compute_h(p: Params, r, m, h, u, nopbc=False):
    r, m = _d(r), _d(m)
    N = r.shape[0]
    h = _d(h).copy(); u = _d(u).copy()
    out = {k: np.zeros(N) for k in ("rho", "invomega", "zeta", "hfactor", "sound", "pressure")}
    rc = lib().orc_compute_h(C.byref(p), N, _p(r), _p(m), int(nopbc), _p(h), _p(out["rho"]),
                             _p(out["invomega"]), _p(out["zeta"]), _p(out["hfactor"]), _p(u),
                             _p(out["sound"]), _p(out["pressure"]))
    if rc != 0:
        raise RuntimeError("oracle h-rho iteration failed")
    out["h"] = h
    out["u"] = u
    return out


def compute_h_stars(p: Params, r, m, h, u, st, nopbc=False):
    """compute_h with the stars' zeta terms (GradhSph.cpp:282-307)."""
    r, m = _d(r), _d(m)
    N = r.shape[0]
    h = _d(h).copy(); u = _d(u).copy()
    out = {k: np.zeros(N) for k in ("rho", "invomega", "zeta", "hfactor", "sound", "pressure")}
    rc = lib().orc_compute_h_stars(C.byref(p), N, _p(r), _p(m), int(nopbc), C.byref(st), _p(h),
                                   _p(out["rho"]), _p(out["invomega"]), _p(out["zeta"]),
                                   _p(out["hfactor"]), _p(u), _p(out["sound"]),
                                   _p(out["pressure"]))
    if rc != 0:
        raise RuntimeError("oracle h-rho iteration failed")
    out["h"] = h
    out["u"] = u
    return out


def gas_from_stars(p: Params, r, h, st):
    """GradhSph::ComputeStarGravForces for every gas particle: (a, gpot) of the stars' pull."""
    r, h = _d(r), _d(h)
    N = r.shape[0]
    a = np.zeros((N, 3)); gp = np.zeros(N)
    lib().orc_gas_from_stars(C.byref(p), N, _p(r), _p(h), C.byref(st), _p(a), _p(gp))
    return a, gp


def stars_from_gas(p: Params, st, r, m, h):
    """HydroTree::UpdateAllStarGasForces (brute-force lists): (a, gpot) of every star."""
    r, m, h = _d(r), _d(m), _d(h)
    a = np.zeros((st.n, 3)); gp = np.zeros(st.n)
    lib().orc_stars_from_gas(C.byref(p), C.byref(st), r.shape[0], _p(r), _p(m), _p(h), _p(a),
                             _p(gp))
    return a, gp


def hydro_forces(p: Params, variant, r, v, m, st):
    r, v, m = _d(r), _d(v), _d(m)
    N = r.shape[0]
    a = np.zeros((N, 3)); dudt = np.zeros(N); div_v = np.zeros(N)
    lib().orc_hydro_forces(C.byref(p), variant, N, _p(r), _p(v), _p(m), _p(_d(st["h"])),
                           _p(_d(st["rho"])), _p(_d(st["invomega"])), _p(_d(st["pressure"])),
                           _p(_d(st["sound"])), _p(_d(st["hfactor"])), _p(a), _p(dudt),
                           _p(div_v))
    return a, dudt, div_v


def direct_gravity(p: Params, r, m, st):
    r, m = _d(r), _d(m)
    N = r.shape[0]
    at = np.zeros((N, 3)); gp = np.zeros(N)
    lib().orc_direct_gravity(C.byref(p), N, _p(r), _p(m), _p(_d(st["h"])), _p(_d(st["zeta"])),
                             _p(_d(st["hfactor"])), _p(at), _p(gp))
    return at, gp


def direct_plummer(r, m, eps, targets):
    r, m = _d(r), _d(m)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    a = np.zeros((t.size, 3)); pot = np.zeros(t.size)
    lib().orc_direct_plummer(r.shape[0], _p(r), _p(m), eps, _p(t), t.size, _p(a), _p(pot))
    return a, pot


def nbody_direct(r, v, m):
    """Nbody::CalculateDirectGravForces: unsoftened star-star sums (a, adot, gpot)."""
    r, v, m = _d(r), _d(v), _d(m)
    N = r.shape[0]
    a = np.zeros((N, 3)); ad = np.zeros((N, 3)); gp = np.zeros(N)
    lib().orc_nbody_direct(N, _p(r), _p(v), _p(m), _p(a), _p(ad), _p(gp))
    return a, ad, gp


def nbody_direct_smoothed(p: Params, r, v, m, h):
    """NbodyLeapfrogKDK::CalculateDirectSmoothedGravForces (a, adot, gpot)."""
    r, v, m, h = _d(r), _d(v), _d(m), _d(h)
    N = r.shape[0]
    a = np.zeros((N, 3)); ad = np.zeros((N, 3)); gp = np.zeros(N)
    lib().orc_nbody_direct_smoothed(C.byref(p), N, _p(r), _p(v), _p(m), _p(h), _p(a), _p(ad),
                                    _p(gp))
    return a, ad, gp


def cell_mac_opens(p: Params, half, mc, q, dmac2, gfac) -> bool:
    """The gadget2 / eigenmac opening test of one cell (p.gravity_mac selects it)."""
    return bool(lib().orc_cell_mac_opens(C.byref(p), _p(_d(half)), float(mc), _p(_d(q)),
                                         float(dmac2), float(gfac)))


def cell_mono_force(com, mc, rp):
    """The oracle's monopole cell term at one point: (acceleration, potential)."""
    a = np.zeros(3)
    pot = np.zeros(1)
    lib().orc_cell_mono_force(_p(_d(com)), float(mc), _p(_d(rp)), _p(a), _p(pot))
    return a, float(pot[0])


def fast_multipole(dcom, q, m, quad, d):
    """The fast multipole expansion of cells (centres dcom relative to the expansion point,
    moments q, masses m) evaluated at points d relative to it: (coef[10], acc, pot)."""
    dcom = _d(np.atleast_2d(dcom)); q = _d(np.atleast_2d(q)); m = _d(np.atleast_1d(m))
    d = _d(np.atleast_2d(d))
    coef = np.zeros(10)
    acc = np.zeros((d.shape[0], 3)); pot = np.zeros(d.shape[0])
    lib().orc_fast_multipole(dcom.shape[0], _p(dcom), _p(q), _p(m), int(quad), d.shape[0],
                             _p(d), _p(coef), _p(acc), _p(pot))
    return coef, acc, pot


def cell_quad_force(com, q, mc, rp):
    """The oracle's quadrupole cell term at one point: (acceleration, potential)."""
    a = np.zeros(3)
    pot = np.zeros(1)
    lib().orc_cell_quad_force(_p(_d(com)), _p(_d(q)), float(mc), _p(_d(rp)), _p(a), _p(pot))
    return a, float(pot[0])


class TreeTwin:
    """CPU twin of the device tree; arrays exposed as numpy copies."""

    def __init__(self, p: Params, r, m, h, group_max=64):
        self.p = p
        r, m, h = _d(r), _d(m), _d(h)
        N = r.shape[0]
        self._t = Tree()
        lib().orc_tree_alloc(C.byref(self._t), N)
        lib().orc_tree_build(C.byref(p), C.byref(self._t), _p(r), _p(m), _p(h), group_max)
        self.N = N

    def __del__(self):
        try:
            lib().orc_tree_free(C.byref(self._t))
        except Exception:
            pass

    def arr(self, name, n, dtype):
        ptr = getattr(self._t, name)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype).copy()

    def arrays(self):
        N = self.N
        nn, ni = 2 * N - 1, max(N - 1, 1)
        g = self._t.ngroups
        return dict(
            perm=self.arr("perm", N, np.int32), keys=self.arr("keys", N, np.uint64),
            left=self.arr("left", ni, np.int32), right=self.arr("right", ni, np.int32),
            first=self.arr("first", ni, np.int32), last=self.arr("last", ni, np.int32),
            parent=self.arr("parent", nn, np.int32), mass=self.arr("mass", nn, np.float64),
            com=self.arr("com", 3 * nn, np.float64).reshape(nn, 3),
            quad=self.arr("quad", 5 * nn, np.float64).reshape(nn, 5),
            bbmin=self.arr("bbmin", 3 * nn, np.float64).reshape(nn, 3),
            bbmax=self.arr("bbmax", 3 * nn, np.float64).reshape(nn, 3),
            hmax=self.arr("hmax", nn, np.float64), cdistsqd=self.arr("cdistsqd", nn, np.float64),
            group_first=self.arr("group_first", g, np.int32),
            group_count=self.arr("group_count", g, np.int32),
            group_node=self.arr("group_node", g, np.int32))

    def refit_h(self, h):
        h = _d(h)
        lib().orc_tree_refit_h(C.byref(self.p), C.byref(self._t), _p(h))

    def gravity(self, r, m, st, atree_prev=None, gpot_prev=None):
        """Tree gravity; atree_prev / gpot_prev: the particles' tree acceleration and potential
        at the tree stock, which the gadget2 / eigenmac criteria read (zero when omitted)."""
        r, m = _d(r), _d(m)
        N = r.shape[0]
        at = np.zeros((N, 3)); gp = np.zeros(N)
        nc = np.zeros(self._t.ngroups, dtype=np.int32); nn = np.zeros_like(nc)
        ap = _d(np.zeros((N, 3)) if atree_prev is None else atree_prev)
        gq = _d(np.zeros(N) if gpot_prev is None else gpot_prev)
        lib().orc_tree_gravity(C.byref(self.p), C.byref(self._t), _p(r), _p(m),
                               _p(_d(st["h"])), _p(_d(st["zeta"])), _p(_d(st["hfactor"])),
                               _p(at), _p(gp), _p(nc), _p(nn), _p(ap), _p(gq))
        return at, gp, nc, nn


def timesteps(p: Params, st):
    N = st["h"].shape[0]
    dt = np.zeros(N)
    dmin = lib().orc_timesteps(C.byref(p), N, _p(_d(st["h"])), _p(_d(st["sound"])),
                               _p(_d(st["div_v"])), _p(_d(st["a"])), _p(_d(st["u"])),
                               _p(_d(st["dudt"])), _p(dt))
    return dmin, dt


def diagnostics(p: Params, r, v, m, u, gpot):
    out = np.zeros(11)
    lib().orc_diagnostics(C.byref(p), r.shape[0], _p(_d(r)), _p(_d(v)), _p(_d(m)), _p(_d(u)),
                          _p(_d(gpot)), _p(out))
    return out


class Sim:
    """Full-step oracle driver (reference MainLoop with global or block timesteps)."""

    FIELDS3 = ("r", "v", "a", "atree")

    def __init__(self, p: Params, r, v, m, u):
        self.p = p
        r, v, m, u = _d(r), _d(v), _d(m), _d(u)
        self.N = r.shape[0]
        self._s = lib().orc_sim_create(C.byref(p), self.N, _p(r), _p(v), _p(m), _p(u))

    def __del__(self):
        try:
            lib().orc_sim_destroy(self._s)
        except Exception:
            pass

    def setup(self):
        if lib().orc_sim_setup(self._s) != 0:
            raise RuntimeError("oracle setup failed")

    def step(self):
        if lib().orc_sim_step(self._s) != 0:
            raise RuntimeError("oracle step failed")

    def set_stars(self, r, v, m, h, softening=1, conservative=1, nbody_mult=0.1):
        r, v, m, h = _d(r).reshape(-1, 3), _d(v).reshape(-1, 3), _d(m), _d(h)
        self.Nstar = r.shape[0]
        lib().orc_sim_set_stars(self._s, self.Nstar, _p(r), _p(v), _p(m), _p(h), int(softening),
                                int(conservative), float(nbody_mult))

    def get(self, name):
        n = getattr(self, "Nstar", 0) if name.startswith("star_") else self.N
        w = 3 if name in self.FIELDS3 or name in ("star_r", "star_v", "star_a", "star_adot") \
            else 1
        out = np.zeros(max(n * w, 1))[: n * w]
        if lib().orc_sim_get(self._s, name.encode(), _p(out)) != 0:
            raise KeyError(name)
        return out.reshape(n, 3) if w == 3 else out

    @property
    def t(self):
        return lib().orc_sim_time(self._s)

    @property
    def timestep(self):
        return lib().orc_sim_timestep(self._s)

    def block_state(self):
        """(n, level_max, level_step, nresync, dt_max, particles that ended a step last)"""
        iv = np.zeros(4, dtype=np.int32)
        dv = np.zeros(2)
        lib().orc_sim_block_state(self._s, _p(iv), _p(dv))
        return int(iv[0]), int(iv[1]), int(iv[2]), int(iv[3]), float(dv[0]), int(dv[1])
