"""Benchmark: SPH + self-gravity particle-steps/s on one MI355X.

Workload: the rotating m=2-perturbed cloud of 1,048,576 SPH particles (uniform radius-1 sphere,
rho = rho0 (1 + 0.1 cos 2 phi), solid-body rotation E_rot/|W| = 0.16, thermal ratio 0.26,
barotropic EOS, quintic kernel, self-gravity with quadrupole cells, opening angle 0.5), seeded
synthetic initial conditions, block timesteps with up to 8 levels.  A "step" is one full
MainLoop at the base timestep: drift, tree build, h-rho iteration and forces of the active
particles (repeated for particles whose step CheckTimesteps ends early), block-timestep
update, end-of-step kick.  Particle-steps count the particles that complete a step (the
active ones), so steps where few particles are active count little.

Both precision paths are timed: the FP32 mixed path (the headline value) and the plain FP64
path; the same workload with global timesteps is timed beside them for reference.  Inputs are
resident on the device before the timed region; the timed region brackets exactly --steps
steps with device synchronisation on both sides.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12}   # MI355X dense vector peaks (FLOP/s)
# algorithmic work per interaction (frozen accounting): density pair 45 (converged evaluation
# only), hydro pair 75, hydro pair with the kernel-softened gravity of neighbours 135,
# direct/monopole gravity term 22, quadrupole particle-cell term 70, candidate-target distance
# test of the mixed hydro kernel 8 (three differences and a squared norm with two fused
# multiply-adds; the kernel decides the scatter-gather sets from the density pass's lists)
W_DENS, W_HYDRO, W_HYDRO_GRAV, W_PP, W_PC, W_TEST = 45, 75, 135, 22, 70, 8
NLEVELS = 8          # block timestep levels of the workload


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_path(ic, precision, steps, warmup):
    from gandalf_amd.gpu import GpuSim
    s = GpuSim(ic["params"], precision=precision)
    s.set_particles(ic["r"], ic["v"], ic["m"], ic["u"])
    t0 = time.perf_counter()
    s.setup()
    log(f"[{precision}] setup {time.perf_counter() - t0:.2f} s")
    for w in range(warmup):
        s.step(1)
    s.set_timing(True)
    passes = {k: 0.0 for k in ("tree", "density", "hydro", "gravity", "integrate",
                                "k_gravity", "k_hydro")}
    counts = None
    per_step = []
    finished = []
    launches = 0
    hlaunches = 0
    t_start = time.perf_counter()
    for k in range(steps):
        ts = time.perf_counter()
        s.step(1)                      # returns after the step's final device sync
        per_step.append(time.perf_counter() - ts)
        pt = s.pass_times()
        for key in passes:
            passes[key] += pt[key]
        launches += int(pt["k_gravity_launches"])
        hlaunches += int(pt["k_hydro_launches"])
        finished.append(s.block_state()[5])
        c = s.interaction_counts()
        counts = c if counts is None else {k2: counts[k2] + c[k2] for k2 in c}
    wall = time.perf_counter() - t_start
    t, dt, n = s.time()
    diag = s.diagnostics()
    s.close()
    N = ic["r"].shape[0]
    return dict(N=N, wall=wall, steps=steps, per_step=per_step, passes=passes, counts=counts,
                t=t, dt=dt, E=float(diag[0]), precision=precision, finished=finished,
                launches=launches, hlaunches=hlaunches, particle_steps=int(sum(finished)))


def kernel_rooflines(res, dtype):
    """The two pair kernels that dominate the step, each as algorithmic FLOPs over its own
    event-timed duration on the stream it runs on, summed over the timed steps' launches:
    k_hydro (135 FLOP per scatter-gather pair: hydro terms plus the kernel-softened gravity of
    the pair, and 8 per candidate-target distance test deciding the set, both counted on the
    device) and k_gravity (22 FLOP per direct term -- walk particle pairs minus the
    scatter-gather pairs -- and 70 per quadrupole particle-cell term).  The gravity walk runs
    on a second stream beside the hydro kernel, so both durations include the other's share of
    the device.  Returns (dominant, other)."""
    c, p = res["counts"], res["passes"]
    peak = PEAK[dtype]
    out = {}
    for name, flops, ms, nl in (
            ("k_hydro", W_HYDRO_GRAV * c["sg_pairs"] + W_TEST * c["sg_tests"], p["k_hydro"],
             res["hlaunches"]),
            ("k_gravity", W_PP * (c["grav_pp"] - c["sg_pairs"]) + W_PC * c["grav_pc"],
             p["k_gravity"], res["launches"])):
        nl = max(nl, 1)
        achieved = flops / (ms * 1e-3) if ms > 0 else 0.0
        out[name] = {"kernel": name, "bound": "valu", "achieved": achieved / 1e12,
                     "peak": peak / 1e12, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": measured_traffic(name) if dtype == "fp32" else None,
                     "flop_per_launch": flops / nl, "launch_ms": ms / nl, "launches": nl,
                     "total_ms": ms}
    dom = max(out.values(), key=lambda r: r["total_ms"])
    other = [r for r in out.values() if r is not dom][0]
    loop_flops = (W_DENS * c["gather_pairs"] + W_HYDRO_GRAV * c["sg_pairs"] +
                  W_TEST * c["sg_tests"] +
                  W_PP * (c["grav_pp"] - c["sg_pairs"]) + W_PC * c["grav_pc"])
    loop_rate = loop_flops / res["wall"]
    dom = dict(dom)
    dom["loop"] = {"achieved": loop_rate / 1e12, "frac": loop_rate / peak,
                   "flop_per_step": loop_flops / res["steps"]}
    return dom, other


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel (mixed path, timed steps of this benchmark) from the
    committed rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs, scaled by the same counters on a
    1 GiB copy; None if the summary is absent."""
    path = os.path.join(ROOT, "profiles", "pair_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return float(d[kernel]["bytes_per_launch"]) if kernel in d else None


def copy_calibration(mib):
    """A plain device-to-device copy of known size (torch's copy kernel), for scaling the
    FETCH_SIZE / WRITE_SIZE counters to bytes: run under rocprofv3 --pmc."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}), flush=True)


def cpu_baseline(nsample=131072, steps=4):
    """The CPU oracle (OpenMP restatement of the reference path) on a bounded sample of the
    same workload: the same generator and block timesteps at nsample particles, `steps` base
    steps after setup, counting the particles that complete a step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from gandalf_amd import ic as ICS
    g = ICS.rotating_m2_cloud(N=nsample, seed=3, rho_bary=10.0, nlevels=NLEVELS)
    p = O.params_from_dict(g["params"])
    s = O.Sim(p, g["r"], g["v"], g["m"], g["u"])
    s.setup()
    t0 = time.perf_counter()
    done = 0
    for _ in range(steps):
        s.step()
        done += s.block_state()[5]
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": done / dt, "unit": "particle-steps/s", "cores": cores,
           "kind": "port",
           "sample": f"rotating m=2 cloud at N={nsample}, block timesteps ({NLEVELS} levels), "
                     f"{steps} base steps after setup ({done} particle-steps), "
                     f"oracle/gdf_oracle.c OpenMP"}
    # the oracle's speed relative to the compiled reference on one thread, measured in the
    # build container by oracle/calibrate_cpu.py (the reference cannot run on the GPU host)
    cal = os.path.join(ROOT, "profiles", "cpu_calibration.json")
    if os.path.exists(cal):
        with open(cal) as f:
            c = json.load(f)
        out["oracle_over_reference_time"] = {
            k: v["oracle_over_reference"] for k, v in c["cases"].items()}
        out["calibration"] = "profiles/cpu_calibration.json (1 thread, step loop only)"
    return out


def ensemble(nreq=256):
    """The serving workload: 256 requests arriving as a Poisson process (20/s), sizes
    log-uniform in 4096-262144, the three generators in rotation, 20 mixed-precision steps
    each.  The stream is served in real time (aggregate particle-steps/s over the run, per-step
    latency, queueing before admission); the same requests are then served with every request
    queued at once (the batches' capacity) and one after another alone on the device (the
    sequential baseline, end to end: set-up and steps)."""
    from gandalf_amd import serving as S
    reqs = S.request_stream(nreq=nreq)
    srv = S.Server()
    live = srv.run(reqs)
    srv.close()
    log(f"[ensemble] poisson stream: {live}")
    sat_reqs = S.request_stream(nreq=nreq)
    for q in sat_reqs:
        q.arrival = 0.0
    srv = S.Server()
    sat = srv.run(sat_reqs)
    srv.close()
    log(f"[ensemble] all queued at once: {sat}")
    seq = S.run_sequential(S.request_stream(nreq=nreq))
    log(f"[ensemble] sequential: {seq}")
    return {"workload": f"{nreq} requests, Poisson 20/s, N log-uniform 4096-262144, box/sphere/"
                        "plummer in rotation, 20 mixed-precision steps each",
            "value": live["value"], "unit": "particle-steps/s",
            "step_latency_ms": live["step_latency_ms"],
            "admission_wait_ms": live["admission_wait_ms"], "wall_s": live["wall_s"],
            "particle_steps": live["particle_steps"],
            "saturated": {"value": sat["value"], "wall_s": sat["wall_s"],
                          "step_latency_ms": sat["step_latency_ms"]},
            "sequential": {"value": seq["value"], "wall_s": seq["wall_s"],
                           "steps_only_value": seq["steps_only_value"],
                           "step_latency_ms": seq["step_latency_ms"]},
            "saturated_over_sequential": sat["value"] / seq["value"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-global", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU path is a design note only)", file=sys.stderr)
        sys.exit(2)
    from gandalf_amd import ic as ICS
    log(f"generating rotating m=2 cloud N={args.n}")
    g = ICS.rotating_m2_cloud(N=args.n, seed=3, rho_bary=10.0, nlevels=NLEVELS)
    mixed = run_path(g, "mixed", args.steps, args.warmup)
    log(f"[mixed] {mixed['wall']:.3f} s for {args.steps} steps, particle-steps per step "
        f"{mixed['finished']}; passes(ms) {mixed['passes']}")
    fp64 = None
    if not args.no_fp64:
        fp64 = run_path(g, "fp64", args.steps, args.warmup)
        log(f"[fp64] {fp64['wall']:.3f} s for {args.steps} steps; passes(ms) {fp64['passes']}")
    glob = None
    if not args.no_global:
        gg = ICS.rotating_m2_cloud(N=args.n, seed=3, rho_bary=10.0, nlevels=1)
        glob = run_path(gg, "mixed", 3, 1)
        log(f"[mixed, global timesteps] {glob['wall']:.3f} s for 3 steps")
    N = args.n
    value = mixed["particle_steps"] / mixed["wall"]
    out = {
        "metric": "SPH+grav particle-steps/s at 1M (FP32 vs FP64); ensemble p50/p99 step latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * mixed["wall"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "rotating m=2 perturbed cloud, quintic kernel, barotropic EOS, "
                               "self-gravity theta=0.5 quadrupole, block timesteps up to "
                               f"{NLEVELS} levels",
                   "n_particles": N, "precision": "fp32 pair arithmetic + fp64 integration",
                   "particle_steps_per_step": mixed["finished"]},
        "roofline": kernel_rooflines(mixed, "fp32")[0],
        "roofline_other_kernel": kernel_rooflines(mixed, "fp32")[1],
        "passes_ms_per_step": {k: v / args.steps for k, v in mixed["passes"].items()},
        "interactions_per_step": {k: v // args.steps for k, v in mixed["counts"].items()},
    }
    if fp64 is not None:
        out["fp64"] = {"value": fp64["particle_steps"] / fp64["wall"],
                       "ms_per_step": 1e3 * fp64["wall"] / args.steps,
                       "roofline": kernel_rooflines(fp64, "fp64")[0],
                       "roofline_other_kernel": kernel_rooflines(fp64, "fp64")[1],
                       "passes_ms_per_step": {k: v / args.steps for k, v in fp64["passes"].items()}}
    if glob is not None:
        out["global_timesteps"] = {"value": glob["particle_steps"] / glob["wall"],
                                   "ms_per_step": 1e3 * glob["wall"] / glob["steps"],
                                   "precision": "mixed"}
    if not args.no_ensemble:
        out["ensemble"] = ensemble()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
