"""Calibrate the CPU oracle against the compiled reference (test infrastructure; runs in the
build container, where oracle/_ref is built from the reference sources by Makefile.ref).

The benchmark's CPU baseline is the oracle (oracle/gdf_oracle.c, OpenMP) timed on the GPU
host.  The reference itself cannot travel there, so its speed is related to the oracle's here:
both run the same seeded inputs on one thread (the reference's serial STANDARD build; its
OpenMP build does not compile under GCC >= 9 unmodified), timed over the step loop only
(setup excluded: a 0-step run is subtracted for the reference).  The ratio
oracle time / reference time is written to profiles/cpu_calibration.json.

    python oracle/calibrate_cpu.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
REF = os.path.join(HERE, "_ref", "ref_dump")

CASES = {
    "box32_m4_hydro": ("jittered periodic lattice box, 32^3 = 32768 particles, hydro only, "
                       "M4, 10 steps", lambda ic: ic.jittered_lattice_box(n_side=32, seed=1), 10),
    "cloud32k_quintic_grav": ("rotating m=2 cloud, 32768 particles, quintic, self-gravity "
                              "theta = 0.5, 4 steps",
                              lambda ic: ic.rotating_m2_cloud(N=32768, seed=3), 4),
}


def time_reference(g, nsteps, td):
    from gandalf_amd import ic
    col, dat = os.path.join(td, "in.col"), os.path.join(td, "in.dat")
    ic.write_column_file(col, g)
    ic.write_param_file(dat, g["params"], col, "CAL")
    wall = {}
    for n in (0, nsteps):
        t0 = time.perf_counter()
        subprocess.run([REF, dat, str(n), os.path.join(td, "out"), str(max(n, 1))], cwd=td,
                       check=True, stdout=subprocess.DEVNULL)
        wall[n] = time.perf_counter() - t0
    return wall[nsteps] - wall[0]


def oracle_script():
    """The oracle's timing run, in a child process so that OMP_NUM_THREADS = 1 applies before
    the OpenMP runtime starts."""
    return ("import sys, time; sys.path[:0] = [%r, %r]\n"
            "from gandalf_amd import ic; import oracle as O\n"
            "import calibrate_cpu as K\n"
            "g = K.CASES[%r][1](ic)\n"
            "s = O.Sim(O.params_from_dict(g['params']), g['r'], g['v'], g['m'], g['u'])\n"
            "s.setup(); t0 = time.perf_counter()\n"
            "for _ in range(%d): s.step()\n"
            "print(time.perf_counter() - t0)\n")


def main():
    if not os.path.exists(REF):
        sys.exit("reference not built: make -f oracle/Makefile.ref -j8")
    from gandalf_amd import ic
    out = {"threads": 1, "reference_build": "serial, COMPILER_MODE=STANDARD, PRECISION=DOUBLE "
           "(oracle/Makefile.ref)", "timed": "step loop only (setup excluded)", "cases": {}}
    for name, (desc, gen, nsteps) in CASES.items():
        g = gen(ic)
        with tempfile.TemporaryDirectory() as td:
            t_ref = time_reference(g, nsteps, td)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, "-c", oracle_script() % (ROOT, HERE, name, nsteps)],
                           env=env, check=True, capture_output=True, text=True)
        t_orc = float(r.stdout.strip().splitlines()[-1])
        n = g["r"].shape[0]
        out["cases"][name] = {
            "workload": desc, "particle_steps": n * nsteps,
            "reference_s": round(t_ref, 3), "oracle_s": round(t_orc, 3),
            "oracle_over_reference": round(t_orc / t_ref, 3),
            "reference_particle_steps_per_s": round(n * nsteps / t_ref, 1),
            "oracle_particle_steps_per_s": round(n * nsteps / t_orc, 1)}
        print(name, out["cases"][name], flush=True)
    with open(os.path.join(ROOT, "profiles", "cpu_calibration.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
