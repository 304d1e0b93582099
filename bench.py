#!/usr/bin/env python3
"""Benchmark of the ARTSS solver hot path on one MI355X.

Workload: the buoyant-plume LES (NSTempTurbSolver: SL advection, colored
Gauss-Seidel diffusion, ConstSmagorinsky, Boussinesq buoyancy, Gaussian heat
source, geometric multigrid pressure) on 256x128x128 cells over 2x1x1 m with
one box obstacle, from artss_amd/configs/plume_256x128x128.xml (1000 steps of
dt = 5e-3).  A step is one full time step over all inner cells; the pressure
equation is solved until ||rhs - lap p||_2 <= 1e-8 every step.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `value` is cell-updates/s = fluid inner cells x K /
wall time of K timed steps (after W untimed ones, from rest) of the default
path, fields resident in HBM: FP64 GCR preconditioned by a mixed-precision
V-cycle (FP32 smoothing, restriction and prolongation; an exact FP64 coarse
solve; FP64 residuals and norms), started from the pressure extrapolated in
time (--start projection starts from the combination of the last --depth
increments of the pressure closest to the solution in the energy norm).  The defaults (W = 10, K = 990) time steps 11..1000, the whole input.

`legs` times every path over the same steps W+1..W+K, each from rest:
  mixed               the default path (`value`)
  mixed_projection    the same from the projection start: fewer iterations,
                      but the residual target then leaves larger smooth
                      pressure errors (compare its field difference)
  mixed_learned       the same with the learned warm start: a BF16 U-Net
                      trained during the warm-up steps on the (input, target)
                      pairs the solver records, then used on every timed step
  gcr_fp64            FP64 GCR with the FP64 V-cycle (the reference's),
                      the default path's start
  plain_fp64          the reference's V-cycles on their own, to the same residual
  reference_schedule  the plain FP64 path with the reference's own V-cycle
                      schedule (n_cycle per step, no residual target): the
                      work the reference's CPU build does per step
Each leg reports its throughput, iterations, the steps whose GCR solve was
finished by V-cycles (`fallbacks`), and the largest relative field
difference to plain_fp64 after step W+K.

`roofline` is the GCR update's algorithmic bytes per launch over its mean
launch time (HIP events on the solver's stream around every launch in a
separate timing region) against the 8 TB/s HBM3E peak; `traffic` and
`step` come from the rocprofv3 summaries under profiles/.  `cpu_baseline`
runs the reference's own serial build (oracle/_ref, compiled from the
reference sources) for a bounded sample continuing from the GPU state.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s, 256x128x128 plume LES, learned+mixed vs plain FP64 MG at res 1e-8"
HBM_PEAK_GBS = 8000.0
FIELDS = ("u", "v", "w", "p", "T")

# algorithmic bytes per cell for the kernels the live timer can name (each
# distinct array element read or written once; artss_amd/roofline.py has the
# model of every kernel)
BYTES_PER_CELL = {
    "jacobi_mg": 24, "jacobi_mg32": 16, "cgs_mg": 16, "cgs_diffuse": 20, "jacobi_diffuse": 32, "residuum": 24,
    "residual": 16, "sl_advect": 40, "restrict": 10, "prolongate": 17,
    # GCR update of the mixed path, per cell of the whole array: x, r (f64) read+write,
    # z and the kept direction (f32) read, the new kept direction (f32) written
    "gcr_update": 44,
}
WHOLE_ARRAY = {"gcr_update"}

PATHS = {"mixed": "GCR in FP64 preconditioned by an FP32 V-cycle with an exact FP64 coarse solve, %s start",
         "mixed_projection": "GCR in FP64 preconditioned by an FP32 V-cycle with an exact FP64 coarse solve, "
                             "projection start",
         "mixed_learned": "the same, started from the extrapolation plus the trained U-Net's correction",
         "gcr_fp64": "FP64 GCR preconditioned by the FP64 V-cycle, %s start",
         "plain_fp64": "FP64, V-cycles alone (the reference's iteration) to the residual target",
         "reference_schedule": "FP64, the reference's V-cycle schedule (n_cycle per step, no residual target)"}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=990)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default=os.path.join(ROOT, "artss_amd", "configs", "plume_256x128x128.xml"))
    ap.add_argument("--tolerance", type=float, default=1e-8, help="pressure residual target per step")
    ap.add_argument("--path", choices=("mixed", "gcr_fp64", "plain_fp64"), default="mixed",
                    help="path of the timed region (`value`); the others are legs")
    ap.add_argument("--legs", default="mixed_projection,mixed_learned,gcr_fp64,plain_fp64,reference_schedule",
                    help="paths timed beside the default one over the same steps (comma list; '' for none)")
    ap.add_argument("--profile-kernel", default="gcr_update@0",
                    help="kernel@level timed for the roofline (the largest HBM-bound kernel of the path)")
    ap.add_argument("--timing-steps", type=int, default=3, help="steps of the kernel timing region")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the reference CPU sample (0: skip)")
    ap.add_argument("--mg-schedule", choices=("graph", "fused", "plain"), default="graph",
                    help="V-cycle launch schedule (bit-identical results)")
    ap.add_argument("--sweeps", default="1,1", help="pre,post smoothing sweeps of the FP32 levels")
    ap.add_argument("--weight", type=float, default=0.9, help="Jacobi weight of those sweeps")
    ap.add_argument("--coarse", choices=("direct", "tail"), default="direct",
                    help="exact dense solve of the coarse levels where they have no obstacle, or their V-cycle")
    ap.add_argument("--coarse-sweeps", type=int, default=5,
                    help="sweeps of the GCR preconditioner's coarse Jacobi solve (a fixed count; "
                         "the effective count follows the reference's pointer-swap rule)")
    ap.add_argument("--coarse-tol", type=float, default=-1.0,
                    help="stopping test of that coarse solve: <0 fixed sweeps, 0 the XML's absolute "
                         "tolerance, >0 relative reduction")
    ap.add_argument("--tail-sweeps", type=int, default=3,
                    help="n_relax of the preconditioner's coarse levels (single-workgroup kernel; 0: the XML's)")
    ap.add_argument("--tail-cells", type=int, default=4096,
                    help="levels of at most this many cells run in one single-workgroup kernel")
    ap.add_argument("--start", choices=("extrapolation", "projection"), default="extrapolation",
                    help="start of the GCR iteration: linear extrapolation in time, or the combination of the "
                         "last --depth increments closest to the solution in the energy norm")
    ap.add_argument("--depth", type=int, default=4, help="increments kept by the projection start")
    ap.add_argument("--train-steps", type=int, default=150, help="Adam steps of the learned leg's network")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the ensemble measurement (uneven plume boxes; 0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=50, help="timed steps of the ensemble measurement")
    ap.add_argument("--ensemble-threads", type=int, default=8, help="host threads issuing the members' steps")
    ap.add_argument("--ensemble-graph", type=int, default=1,
                    help="also time the members stepping through captured graphs with this margin of "
                         "iterations per loop (`concurrent_graph`; a negative value: not timed)")
    ap.add_argument("--device", type=int, default=0)
    return ap.parse_args()


def host_cpu():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count()
    return {"model": model, "nproc": allowed, "logical_cpus": os.cpu_count()}


def cpu_baseline(solver, config, first_step, steps, inner):
    """Reference serial build on this host, continuing from the GPU state."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if steps <= 0 or not os.path.exists(harness):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        for f in FIELDS:
            solver.field(f).tofile(os.path.join(tmp, f + ".f64"))
        r = subprocess.run([harness, "state", config, tmp, str(first_step), str(steps)], capture_output=True,
                           text=True, timeout=900)
    cpu = host_cpu()
    if r.returncode != 0:
        return {"value": None, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": "failed: " + r.stderr.strip()[-200:], "host": cpu}
    secs = float(r.stdout.split("seconds=")[1].split()[0])
    return {"value": inner * steps / secs, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": "%d steps (%d..%d) of the same workload continuing from the GPU state, reference serial "
                      "build (one thread) with its own V-cycle schedule (n_cycle per step), %.1f s; compare "
                      "legs.reference_schedule, the GPU doing the same work" % (
                          steps, first_step, first_step + steps - 1, secs),
            "host": cpu}


def path_text(args, leg):
    text = PATHS[leg]
    return text % args.start if "%s" in text else text


def make_solver(artss_amd, args, leg, net=None):
    solver = artss_amd.Solver(args.config, device=args.device)
    solver.set_mg_schedule({"plain": 0, "fused": 1, "graph": 2}[args.mg_schedule], args.tail_cells)
    if leg == "reference_schedule":
        return solver
    solver.set_pressure_mode(True, args.tolerance, 500)
    method = {"mixed": 2, "mixed_projection": 2, "mixed_learned": 2, "gcr_fp64": 1, "plain_fp64": 0}[leg]
    solver.set_pressure_method(method, args.coarse_sweeps)
    if method != 0:
        solver.set_coarse_tolerance(args.coarse_tol)
    if method == 2:
        pre, post = (int(v) for v in args.sweeps.split(","))
        solver.set_mixed_sweeps(pre, post, args.weight)
        solver.set_coarse_solve(args.coarse == "direct")
        solver.set_mixed_tail(args.tail_sweeps)
    if method != 0:
        if leg == "mixed_projection" or (args.start == "projection" and leg in ("mixed", "gcr_fp64")):
            solver.set_warm_start(3, depth=args.depth)
        else:
            solver.set_warm_start(1, record=leg == "mixed_learned")
    return solver


def profile_value(prefix, key):
    """`key = value` lines that the rocprofv3 summaries under profiles/ state"""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        if not name.startswith(prefix):
            continue
        with open(os.path.join(d, name)) as fh:
            for line in fh:
                line = line.strip().lstrip("- ").strip("`")
                if line.startswith(key + " = "):
                    v = line.split("=", 1)[1].strip().split()[0]
                    try:
                        return float(v) if "." in v or "e" in v else int(v)
                    except ValueError:
                        return None
    return None


class Run:
    """Steps one solver from rest; timed(k) times k steps between syncs."""

    def __init__(self, solver, dt):
        self.solver, self.dt, self.step = solver, dt, 1

    def timed(self, k, on_step=None):
        dom = self.solver.domain
        cyc, dif, res = [], [], []
        dom.sync()
        t0 = time.perf_counter()
        for _ in range(k):
            self.solver.step(self.step * self.dt)
            c = self.solver.counters()
            cyc.append(c["vcycles"])
            dif.append(c["diffusion"])
            res.append(c["residual"])
            if on_step:
                on_step(self.step)
            self.step += 1
        dom.sync()
        return time.perf_counter() - t0, cyc, dif, res


def rel_errors(fields, ref):
    return {f: float(np.abs(fields[f] - ref[f]).max() / max(np.abs(ref[f]).max(), 1e-300)) for f in FIELDS}


def learned_warmup(artss_amd, run, args):
    """warm-up of the learned leg: record the solver's (input, target) pairs,
    train the U-Net on them (forward, backward and Adam on the GPU), then
    switch the start to extrapolation plus the network's correction"""
    s = run.solver
    dom = s.domain
    Nx, Ny, Nz = dom.info(0)["N"]
    box = (Nz - 2, Ny - 2, Nx - 2)
    xs, ts = [], []

    def keep(step):
        if step < 3:
            return
        x = artss_amd.Buffer(dom, np.zeros(box + (2,), np.float32))
        t = artss_amd.Buffer(dom, np.zeros(box + (1,), np.float32))
        if s.warm_sample(x, t):
            xs.append(x)
            ts.append(t)

    run.timed(args.warmup, keep)
    net = artss_amd.UNet(dom, seed=5)
    rng = np.random.default_rng(0)
    t0 = time.perf_counter()
    losses = []
    for it in range(args.train_steps if xs else 0):
        j = int(rng.integers(len(xs)))
        loss, _ = net.gradients(xs[j], ts[j], want_grads=False)
        losses.append(loss)
        net.adam_step(1e-3 if it < 0.75 * args.train_steps else 3e-4)
    dom.sync()
    train_s = time.perf_counter() - t0
    for b in xs + ts:
        b._release()
    s.set_warm_start(2, net)
    return net, {"pairs": len(xs), "adam_steps": len(losses), "train_s": train_s,
                 "loss_first": losses[0] if losses else None,
                 "loss_last10": float(np.mean(losses[-10:])) if losses else None}


def leg_run(artss_amd, args, leg, dt, inner):
    s = make_solver(artss_amd, args, leg)
    run = Run(s, dt)
    extra = {}
    net = None
    if leg == "mixed_learned":
        net, extra["training"] = learned_warmup(artss_amd, run, args)
    else:
        run.timed(args.warmup)
    first = run.step
    secs, cyc, dif, res = run.timed(args.steps)
    fields = {f: s.field(f) for f in FIELDS}
    out = {"value": inner * args.steps / secs, "unit": "cell-updates/s", "ms_per_step": secs * 1e3 / args.steps,
           "steps": "%d..%d" % (first, run.step - 1), "path": path_text(args, leg),
           "mean_iterations": float(np.mean(cyc)), "max_iterations": int(np.max(cyc)),
           "max_final_residual": None if leg == "reference_schedule" else float(np.max(res)),
           "fallbacks": s.fallbacks()}
    out.update(extra)
    if net is not None:
        net.close()
    s.close()
    return out, fields


def ensemble_run(artss_amd, args):
    """Members of uneven size (64..160 x 64..96 x 32..64 cells, 0-3 obstacles,
    20-80 kW/m^3), stepped concurrently and, for comparison, one after
    another, from the same state."""
    from artss_amd import scenarios
    xmls = scenarios.ensemble_members(args.ensemble_members, seed=2026)
    out = {"members": args.ensemble_members, "threads": args.ensemble_threads, "steps": args.ensemble_steps,
           "step_graph_margin": args.ensemble_graph}
    runs = [("concurrent", args.ensemble_threads, None), ("sequential", 1, None)]
    if args.ensemble_graph is not None:
        runs.append(("concurrent_graph", args.ensemble_threads, args.ensemble_graph))
    for label, threads, graph in runs:
        members = []
        for x in xmls:
            m = artss_amd.Solver(x, device=args.device)
            m.set_mg_schedule(2, args.tail_cells)
            m.set_pressure_mode(True, args.tolerance, 500)
            m.set_pressure_method(2, args.coarse_sweeps)
            m.set_coarse_tolerance(args.coarse_tol)
            if args.start == "projection":
                m.set_warm_start(3, depth=args.depth)
            else:
                m.set_warm_start(1)
            if graph is not None:
                m.set_step_graph(True, graph)
            members.append(m)
        cells = sum(m.domain.info(0)["inner_cells"] for m in members)
        ens = artss_amd.Ensemble(members, threads=threads)
        for n in range(1, 4):
            ens.step(n)
        t0 = time.perf_counter()
        for n in range(4, 4 + args.ensemble_steps):
            ens.step(n)  # every member's stream is synchronised when step returns
        secs = time.perf_counter() - t0
        out["cells"] = cells
        out[label] = {"value": cells * args.ensemble_steps / secs, "unit": "cell-updates/s",
                      "ms_per_step": secs * 1e3 / args.ensemble_steps}
        if graph is not None:
            st = [m.step_graph_stats() for m in members]
            out[label]["step_graph"] = {"replays": sum(x["replays"] for x in st),
                                        "rollbacks": sum(x["rollbacks"] for x in st),
                                        "captures": sum(x["captures"] for x in st),
                                        "notes": sorted({x["note"] for x in st if x["note"]})}
        ens.close()
        for m in members:
            m.close()
    out["concurrency_gain"] = out["concurrent"]["value"] / out["sequential"]["value"]
    return out


def main():
    args = parse()
    if args.ensemble_graph is not None and args.ensemble_graph < 0:
        args.ensemble_graph = None
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (the step would shard "
              "as z-slabs with halo exchanges; not implemented)", file=sys.stderr)
        sys.exit(2)
    import artss_amd

    with open(args.config) as fh:
        dt = float(fh.read().split("<dt>")[1].split("</dt>")[0])
    solver = make_solver(artss_amd, args, args.path)
    dom = solver.domain
    info = dom.info(0)
    inner = info["inner_cells"]
    run = Run(solver, dt)
    run.timed(args.warmup)
    first = run.step

    # timed region: K steps as they run in production (V-cycles replayed from HIP graphs)
    secs, cycles, diff_it, residuals = run.timed(args.steps)
    value = inner * args.steps / secs
    fields = {f: solver.field(f) for f in FIELDS}
    fallbacks = solver.fallbacks()

    # kernel timing region: HIP events around every launch of the profiled
    # kernel on the solver's stream; the V-cycle then runs its launches
    # eagerly (events cannot bracket launches inside a replayed graph)
    timing_steps = max(1, min(args.steps, args.timing_steps))
    dom.profile_arm(args.profile_kernel)
    t_secs, _, _, _ = run.timed(timing_steps)
    launches, kernel_ms = dom.profile_read()
    tuned = [artss_amd.tuned_cells_per_thread(dom.info(l)["size"]) for l in range(4)]
    kname, klevel = args.profile_kernel.split("@")
    kinfo = dom.info(int(klevel))
    per_launch_bytes = BYTES_PER_CELL.get(kname, 0) * (kinfo["size"] if kname in WHOLE_ARRAY else kinfo["inner_cells"])
    mean_ms = kernel_ms / launches if launches else float("nan")
    achieved = per_launch_bytes / (mean_ms * 1e-3) / 1e9 if launches else None
    roofline = {"bound": "hbm", "kernel": args.profile_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                "traffic": profile_value("pmc_", "traffic[%s]" % args.profile_kernel),
                "bytes_per_launch": per_launch_bytes, "launches": launches, "mean_launch_us": mean_ms * 1e3,
                "timing_steps": timing_steps, "timing_region_ms_per_step": t_secs * 1e3 / timing_steps,
                "share_of_timing_region": kernel_ms / (t_secs * 1e3),
                "rocprof_mean_launch_us": profile_value("roofline_", "mean_us[gcr_update]"),
                "step": {"kernel_frac": profile_value("roofline_", "kernel_frac"),
                         "step_frac": profile_value("roofline_", "step_frac"),
                         "source": "rocprofv3 kernel trace of this command with --steps 3 --warmup 2, summarised "
                                   "by artss_amd/roofline.py (profiles/roofline_*.md)"}}

    # the CPU sample continues from the state after the GPU regions (same
    # workload, later steps)
    cpu = cpu_baseline(solver, args.config, run.step, args.cpu_steps, inner)
    solver.close()

    legs = {args.path: {"value": value, "unit": "cell-updates/s", "ms_per_step": secs * 1e3 / args.steps,
                        "steps": "%d..%d" % (first, first + args.steps - 1),
                        "path": path_text(args, args.path), "mean_iterations": float(np.mean(cycles)),
                        "max_iterations": int(np.max(cycles)), "max_final_residual": float(np.max(residuals)),
                        "fallbacks": fallbacks}}
    leg_fields = {args.path: fields}
    for leg in [x for x in args.legs.split(",") if x and x != args.path]:
        legs[leg], leg_fields[leg] = leg_run(artss_amd, args, leg, dt, inner)
    if "plain_fp64" in legs:
        ref = leg_fields["plain_fp64"]
        for leg, f in leg_fields.items():
            legs[leg]["speedup_over_plain_fp64"] = legs[leg]["value"] / legs["plain_fp64"]["value"]
            legs[leg]["max_rel_field_difference_to_plain_fp64"] = rel_errors(f, ref)

    ensemble = ensemble_run(artss_amd, args) if args.ensemble_members > 0 else None

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (pressure preconditioner f32)" if args.path == "mixed" else "f64",
        "data": "synthetic: plume from rest, timed steps %d..%d" % (first, first + args.steps - 1),
        "config": {"workload": "plume_256x128x128", "grid": list(info["n"]), "inner_cells": inner,
                   "path": path_text(args, args.path), "mg_schedule": args.mg_schedule,
                   "start": args.start, "depth": args.depth if args.start == "projection" else None,
                   "tail_cells": args.tail_cells, "sweeps": args.sweeps, "weight": args.weight,
                   "coarse": args.coarse, "tail_sweeps": args.tail_sweeps, "coarse_sweeps": args.coarse_sweeps,
                   "coarse_tol": args.coarse_tol, "tuned_cells_per_thread": tuned,
                   "pressure": "iterate until ||rhs - lap p||_2 <= %g every step" % args.tolerance,
                   "mean_vcycles": float(np.mean(cycles)),
                   "mean_diffusion_iterations": np.mean(diff_it, axis=0).tolist(),
                   "max_final_residual": float(np.max(residuals))},
        "legs": legs, "ensemble": ensemble,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
