"""Seeded initial-condition generators for the particle hot path workloads.

Each generator returns a dict of numpy arrays (``r`` (N,3), ``v`` (N,3), ``m`` (N,), ``u`` (N,))
plus the simulation parameters that go with it (``params``), using the reference's own
parameter names (src/Common/Parameters.cpp:160-575) so the same dict can be written to a
reference ``.dat`` file or handed to the GPU simulation.

The generators restate the semantics of the reference's IC helpers rather than their random
streams: positions come from numpy's PCG64 with an explicit seed, and the same arrays are fed
to both the reference and the GPU path, so parity never depends on the random stream.

* ``jittered_lattice_box``  - cubic lattice at (i+1/2)/n (src/Ic/Ic.cpp:629-700) with a
  uniform +-jitter, periodic unit box, adiabatic gas, v_x = A sin(2 pi y).
* ``uniform_sphere``        - rejection sampling in the unit sphere (src/Ic/Ic.cpp:465-500).
* ``cold_sphere``           - self-gravitating isothermal sphere, U/|W| set by ``thermal_ratio``.
* ``rotating_m2_cloud``     - uniform sphere with an azimuthal density perturbation
  rho = rho0 (1 + A cos(m phi)) (cf. src/Ic/Ic.cpp:851-925) and solid-body rotation
  (src/Ic/Ic.cpp:974-1030), barotropic gas.
* ``plummer_sphere``        - Plummer model with Aarseth-Henon-Wielen sampling
  (cf. src/Ic/PlummerSphereIc.cpp:53-190), virialised isotropic velocities.
"""
from __future__ import annotations

import numpy as np

GAMMA_5_3 = 5.0 / 3.0


def _base_params(**kw):
    p = {
        "ndim": 3, "sim": "gradhsph", "dimensionless": 1, "ic": "file",
        "in_file_form": "column", "out_file_form": "column",
        "kernel": "m4", "tabulated_kernel": 0, "h_fac": 1.2, "h_converge": 0.01,
        "avisc": "mon97", "alpha_visc": 1.0, "beta_visc": 2.0, "acond": "none",
        "time_dependent_avisc": "none", "hydro_forces": 1, "self_gravity": 0,
        "gamma_eos": GAMMA_5_3, "mu_bar": 1.0, "temp0": 1.0, "rho_bary": 1.0e-14,
        "neib_search": "kdtree", "Nleafmax": 6, "thetamaxsqd": 0.1,
        "multipole": "quadrupole", "gravity_mac": "geometric",
        "courant_mult": 0.15, "accel_mult": 0.3, "energy_mult": 0.4,
        "Nlevels": 1, "sph_integration": "lfkdk",
        "tend": 1.0e30, "dt_snap": 1.0e30, "tsnapfirst": 1.0e30, "noutputstep": 1 << 30,
        "ndiagstep": 1 << 30, "nrestartstep": 1 << 30, "Nstepsmax": 1 << 30,
        "boundary_lhs[0]": "open", "boundary_rhs[0]": "open",
        "boundary_lhs[1]": "open", "boundary_rhs[1]": "open",
        "boundary_lhs[2]": "open", "boundary_rhs[2]": "open",
    }
    p.update(kw)
    return p


def jittered_lattice_box(n_side=32, seed=1, jitter=0.1, rho=1.0, u=1.0, vamp=0.01,
                         gamma=GAMMA_5_3):
    """Periodic unit box of n_side^3 equal-mass particles; workload of the hydro-only case."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_side
    g = (np.arange(n_side) + 0.5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    r = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    r = r + rng.uniform(-jitter * dx, jitter * dx, size=r.shape)
    r = np.mod(r, 1.0)
    N = r.shape[0]
    v = np.zeros_like(r)
    v[:, 0] = vamp * np.sin(2.0 * np.pi * r[:, 1])
    m = np.full(N, rho / N)
    uu = np.full(N, u)
    params = _base_params(
        gas_eos="energy_eqn", gamma_eos=gamma, self_gravity=0, courant_mult=0.2,
        **{"boundary_lhs[0]": "periodic", "boundary_rhs[0]": "periodic",
           "boundary_lhs[1]": "periodic", "boundary_rhs[1]": "periodic",
           "boundary_lhs[2]": "periodic", "boundary_rhs[2]": "periodic",
           "boxmin[0]": 0.0, "boxmin[1]": 0.0, "boxmin[2]": 0.0,
           "boxmax[0]": 1.0, "boxmax[1]": 1.0, "boxmax[2]": 1.0})
    return dict(r=r, v=v, m=m, u=uu, params=params)


def hot_blob_box(n_side=16, seed=1, blob_u=40.0, sigma=0.1, nlevels=4, level_diff_max=1,
                 jitter=0.1):
    """Periodic jittered-lattice box at rest (rho = 1, u = 1) with a hot Gaussian blob at the
    centre, u = 1 + blob_u exp(-|r - c|^2 / (2 sigma^2)): the sound speed, and with it the
    Courant timestep, varies by ~sqrt(1 + blob_u) across the box, so block timesteps spread
    the particles over several levels and the expanding blob moves them between levels."""
    g = jittered_lattice_box(n_side=n_side, seed=seed, jitter=jitter, vamp=0.0)
    d2 = ((g["r"] - 0.5) ** 2).sum(1)
    g["u"] = 1.0 + blob_u * np.exp(-d2 / (2.0 * sigma * sigma))
    g["params"].update(Nlevels=nlevels, level_diff_max=level_diff_max)
    return g


def uniform_sphere(N, seed, radius=1.0):
    """N points uniform in a sphere by seeded rejection sampling from the enclosing cube."""
    rng = np.random.default_rng(seed)
    out = np.empty((0, 3))
    while out.shape[0] < N:
        need = N - out.shape[0]
        c = rng.uniform(-radius, radius, size=(int(need * 2.0) + 64, 3))
        c = c[np.einsum("ij,ij->i", c, c) <= radius * radius]
        out = np.concatenate([out, c[:need]])
    return out


def cold_sphere(N=262144, seed=2, thermal_ratio=0.05, gamma=GAMMA_5_3, theta=0.5):
    """Self-gravitating isothermal uniform sphere (M=1, R=1, G=1), U/|W| = thermal_ratio."""
    r = uniform_sphere(N, seed)
    u = thermal_ratio * 0.6                      # |W| = 3/5 G M^2 / R
    temp0 = u * (gamma - 1.0)                    # isothermal u = temp0/(gamma-1)/mu
    params = _base_params(gas_eos="isothermal", gamma_eos=gamma, temp0=temp0,
                          self_gravity=1, thetamaxsqd=theta * theta, kernel="m4")
    return dict(r=r, v=np.zeros_like(r), m=np.full(N, 1.0 / N), u=np.full(N, u),
                params=params)


def _azimuthal_resample(r, amp, mpert):
    """Map each azimuth phi to phi' with phi' + (amp/m) sin(m phi') = phi, which turns a
    uniform azimuthal distribution into density (1 + amp cos(m phi'))."""
    R = np.hypot(r[:, 0], r[:, 1])
    phi = np.mod(np.arctan2(r[:, 1], r[:, 0]), 2.0 * np.pi)
    p = phi.copy()
    for _ in range(60):
        f = p + (amp / mpert) * np.sin(mpert * p) - phi
        p = p - f / (1.0 + amp * np.cos(mpert * p))
    out = r.copy()
    out[:, 0] = R * np.cos(p)
    out[:, 1] = R * np.sin(p)
    return out


def rotating_m2_cloud(N=1048576, seed=3, amp=0.1, mpert=2, rot_ratio=0.16,
                      thermal_ratio=0.26, gamma=GAMMA_5_3, rho_bary=10.0, theta=0.5,
                      nlevels=1):
    """Rotating, m=2-perturbed uniform cloud (M=1, R=1, G=1) with a barotropic EOS."""
    r = _azimuthal_resample(uniform_sphere(N, seed), amp, mpert)
    omega = np.sqrt(rot_ratio * 0.6 / 0.2)       # E_rot = (1/5) M R^2 Omega^2
    v = np.zeros_like(r)
    v[:, 0] = -omega * r[:, 1]
    v[:, 1] = omega * r[:, 0]
    u = thermal_ratio * 0.6
    temp0 = u * (gamma - 1.0)                    # barotropic u -> temp0/(gamma-1) at low rho
    params = _base_params(gas_eos="barotropic", gamma_eos=gamma, temp0=temp0,
                          rho_bary=rho_bary, self_gravity=1, kernel="quintic",
                          thetamaxsqd=theta * theta, Nlevels=nlevels)
    return dict(r=r, v=v, m=np.full(N, 1.0 / N), u=np.full(N, u), params=params)


def plummer_sphere(N=131072, seed=4, a=1.0, mass=1.0, rmax_frac=0.999):
    """Plummer sphere (G=1) by Aarseth-Henon-Wielen sampling: radius from the cumulative
    mass profile, speed by von Neumann rejection on g(q) = q^2 (1-q^2)^(7/2)."""
    rng = np.random.default_rng(seed)
    X1 = rng.uniform(0.0, rmax_frac, N)
    rad = a / np.sqrt(X1 ** (-2.0 / 3.0) - 1.0)
    cth = rng.uniform(-1.0, 1.0, N)
    phi = rng.uniform(0.0, 2.0 * np.pi, N)
    sth = np.sqrt(1.0 - cth * cth)
    r = np.stack([rad * sth * np.cos(phi), rad * sth * np.sin(phi), rad * cth], axis=1)
    q = np.empty(N)
    todo = np.arange(N)
    while todo.size:
        x4 = rng.uniform(0.0, 1.0, todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0 * mass) * (rad * rad + a * a) ** -0.25
    vmag = q * vesc
    cth = rng.uniform(-1.0, 1.0, N)
    phi = rng.uniform(0.0, 2.0 * np.pi, N)
    sth = np.sqrt(1.0 - cth * cth)
    v = np.stack([vmag * sth * np.cos(phi), vmag * sth * np.sin(phi), vmag * cth], axis=1)
    m = np.full(N, mass / N)
    r -= np.sum(m[:, None] * r, axis=0) / mass
    v -= np.sum(m[:, None] * v, axis=0) / mass
    params = _base_params(sim="nbody", hydro_forces=0, self_gravity=1, thetamaxsqd=0.16,
                          nbody_eps=0.01, fixed_dt=1.0e-3)
    return dict(r=r, v=v, m=m, u=np.zeros(N), params=params)


def write_column_file(path, ic, stars=False, h=None, star_ic=None):
    """Reference column snapshot (src/Common/SimulationIO.hpp:193-270): header Nhydro, Nstar,
    ndim, t; then x y z vx vy vz m h rho u per particle, at full precision.  With ``stars``
    every particle is written as a star (Nhydro = 0) and ``h`` gives the stars' smoothing
    lengths, which the reference's N-body sums and timesteps use.  ``star_ic`` (r, v, m, h)
    adds stars after the gas."""
    r, v, m, u = ic["r"], ic["v"], ic["m"], ic["u"]
    N = r.shape[0]
    hh = np.zeros(N) if h is None else np.asarray(h, dtype=np.float64)
    cols = np.column_stack([r, v, m, hh, np.zeros(N), u])
    ns = 0 if star_ic is None else star_ic["r"].shape[0]
    with open(path, "w") as f:
        f.write(f"0\n{N}\n3\n0.0\n" if stars else f"{N}\n{ns}\n3\n0.0\n")
        np.savetxt(f, cols, fmt="%.17g")
        if ns:
            sc = np.column_stack([star_ic["r"], star_ic["v"], star_ic["m"], star_ic["h"],
                                  np.zeros(ns), np.zeros(ns)])
            np.savetxt(f, sc, fmt="%.17g")


def write_param_file(path, params, in_file, run_id="REF"):
    """Reference parameter file, one ``comment : key = value`` line each
    (src/Common/Parameters.cpp:126-151)."""
    p = dict(params)
    p["in_file"] = in_file
    p["run_id"] = run_id
    with open(path, "w") as f:
        for k, val in p.items():
            if isinstance(val, float):
                val = repr(val)
            f.write(f"{k} : {k} = {val}\n")
