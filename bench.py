#!/usr/bin/env python3
"""Benchmark of the ARTSS solver hot path on one MI355X.

Workload: the buoyant-plume LES (NSTempTurbSolver: SL advection, colored
Gauss-Seidel diffusion, ConstSmagorinsky, Boussinesq buoyancy, Gaussian heat
source, geometric multigrid pressure) on 256x128x128 cells over 2x1x1 m with
one box obstacle, from artss_amd/configs/plume_256x128x128.xml.  A step is
one full time step over all inner cells; the pressure equation is solved
until ||rhs - lap p||_2 <= 1e-8 every step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--pressure mixed|gcr|vcycle]

Prints one JSON line.  `value` is cell-updates/s = inner cells x K / wall
time of the K timed steps of the accelerated path (default: GCR with a
mixed-precision V-cycle), fields resident in HBM throughout.  `plain_fp64`
runs the plain FP64 path (the reference's V-cycles on their own, to the same
residual) from rest over the same steps, with its throughput, the speedup of
`value` over it and the largest relative difference of every field between
the two paths.  `roofline` is the dominant kernel's algorithmic bytes per
launch over its mean launch time (HIP events on the solver's stream around
every launch in a timing region) against the 8 TB/s HBM3E peak.
`cpu_baseline` runs the reference's own serial build (oracle/_ref, compiled
from the reference sources) for a bounded sample continuing from the same
state on this host.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s, 256x128x128 plume LES, learned+mixed vs plain FP64 MG at res 1e-8"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per level-inner cell for one launch of a kernel (each
# distinct array element read or written once, 8-byte doubles)
BYTES_PER_CELL = {
    "jacobi_mg": 24,       # read x, b; write x'
    "jacobi_mg32": 16,     # level 0 of the mixed V-cycle: read x (f32), b (f64 residual); write x' (f32)
    "cgs_mg": 16,          # one color: read x (all), b (half); write half
    "cgs_diffuse": 20,     # one color, turbulent: + eddy viscosity (half)
    "jacobi_diffuse": 32,  # turbulent: read x, b, ev; write x'
    "residuum": 24,        # read x, b; write r
    "residual": 16,        # read x, b
    "sl_advect": 40,       # read in, u, v, w; write out
    "restrict": 10,        # per fine cell: read it (8); per coarse cell write residual and reset error (16/8)
    "prolongate": 17,      # out of place: read fine error (8) and coarse (8/8), write (8)
    # GCR update of the mixed path, per cell of the whole array: x, r (f64) read+write,
    # z (f32) read, kept direction z_prev, w_prev (f32) read+write
    "gcr_update": 52,
}
# kernels that run over the whole array (ghosts included) rather than the fluid cells
WHOLE_ARRAY = {"gcr_update"}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default=os.path.join(ROOT, "artss_amd", "configs", "plume_256x128x128.xml"))
    ap.add_argument("--tolerance", type=float, default=1e-8, help="pressure residual target per step")
    ap.add_argument("--reference-schedule", action="store_true",
                    help="use the reference's V-cycle schedule instead of the residual target")
    ap.add_argument("--profile-kernel", default="gcr_update@0",
                    help="kernel@level timed for the roofline (the largest HBM-bound kernel of the path)")
    ap.add_argument("--timing-steps", type=int, default=3, help="steps of the kernel timing region")
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="timed steps of the plain FP64 V-cycle path run beside the accelerated one (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the reference CPU sample (0: skip)")
    ap.add_argument("--mg-schedule", choices=("graph", "fused", "plain"), default="graph",
                    help="V-cycle launch schedule (bit-identical results)")
    ap.add_argument("--pressure", choices=("vcycle", "gcr", "mixed"), default="mixed",
                    help="residual-target iteration: V-cycles alone (reference) or V-cycle-preconditioned GCR")
    ap.add_argument("--coarse-iterations", type=int, default=0,
                    help="mixed: cap of the preconditioner's coarse Jacobi solve (0: the XML's max_iter)")
    ap.add_argument("--sweeps", default="1,1",
                    help="mixed: pre,post smoothing sweeps of the FP32 levels (0: the XML's effective count)")
    ap.add_argument("--weight", type=float, default=0.8,
                    help="mixed: Jacobi weight of those sweeps (0: the XML's)")
    ap.add_argument("--coarse", choices=("direct", "tail"), default="direct",
                    help="mixed: exact dense solve of the coarse levels, or their single-workgroup V-cycle")
    ap.add_argument("--tail-sweeps", type=int, default=3,
                    help="mixed: n_relax of the coarse levels' single-workgroup kernel (0: the XML's)")
    ap.add_argument("--tail-cells", type=int, default=4096,
                    help="levels of at most this many cells run in one single-workgroup kernel")
    ap.add_argument("--warm", choices=("none", "extrapolate"), default="extrapolate",
                    help="start of each step's pressure iteration: the previous p (the reference's start) or "
                         "p extrapolated linearly in time")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the ensemble measurement (uneven plume boxes; 0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=5, help="timed steps of the ensemble measurement")
    ap.add_argument("--ensemble-threads", type=int, default=8, help="host threads issuing the members' steps")
    ap.add_argument("--device", type=int, default=0)
    return ap.parse_args()


def cpu_baseline(solver, config, first_step, steps, inner):
    """Reference serial build on this host, continuing from the GPU state."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if steps <= 0 or not os.path.exists(harness):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        for f in ("u", "v", "w", "p", "T"):
            solver.field(f).tofile(os.path.join(tmp, f + ".f64"))
        r = subprocess.run([harness, "state", config, tmp, str(first_step), str(steps)], capture_output=True,
                           text=True, timeout=900)
    if r.returncode != 0:
        return {"value": None, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
                "sample": "failed: " + r.stderr.strip()[-200:]}
    secs = float(r.stdout.split("seconds=")[1].split()[0])
    return {"value": inner * steps / secs, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": "%d steps (%d..%d) of the same workload continuing from the GPU state, reference "
                      "serial build with its own V-cycle schedule, %.1f s" % (steps, first_step,
                                                                              first_step + steps - 1, secs)}


PATHS = {"vcycle": "FP64, V-cycles alone (the reference's iteration)",
         "gcr": "FP64, V-cycle-preconditioned GCR",
         "mixed": "GCR in FP64 preconditioned by an FP32 V-cycle with FP64 coarse levels"}


def make_solver(artss_amd, args, pressure, warm="none"):
    solver = artss_amd.Solver(args.config, device=args.device)
    solver.set_mg_schedule({"plain": 0, "fused": 1, "graph": 2}[args.mg_schedule], args.tail_cells)
    if not args.reference_schedule:
        solver.set_pressure_mode(True, args.tolerance, 500)
        solver.set_pressure_method({"vcycle": 0, "gcr": 1, "mixed": 2}[pressure], args.coarse_iterations)
        if pressure != "vcycle" and warm == "extrapolate":
            solver.set_warm_start(1)
        if pressure == "mixed":
            pre, post = (int(v) for v in args.sweeps.split(","))
            solver.set_mixed_sweeps(pre, post, args.weight)
            solver.set_coarse_solve(args.coarse == "direct")
            solver.set_mixed_tail(args.tail_sweeps)
    return solver


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 PMC summary in
    profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a copy kernel of known size), or None."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for name in sorted(os.listdir(d)):
        if not name.startswith("pmc_"):
            continue
        with open(os.path.join(d, name)) as fh:
            for line in fh:
                line = line.strip().strip("`")
                if line.startswith("traffic[%s] = " % kernel):
                    return int(line.split("=")[1])
    return None


def ensemble_run(artss_amd, args):
    """Members of uneven size (64..160 x 64..96 x 32..64 cells, 0-3 obstacles,
    20-80 kW/m^3), stepped concurrently and, for comparison, one after
    another, from the same state."""
    from artss_amd import scenarios
    xmls = scenarios.ensemble_members(args.ensemble_members, seed=2026)
    out = {"members": args.ensemble_members, "threads": args.ensemble_threads, "steps": args.ensemble_steps}
    for label, threads in (("concurrent", args.ensemble_threads), ("sequential", 1)):
        members = []
        for x in xmls:
            m = artss_amd.Solver(x, device=args.device)
            m.set_mg_schedule(2, args.tail_cells)
            m.set_pressure_mode(True, args.tolerance, 500)
            m.set_pressure_method(2)
            m.set_warm_start(1)
            members.append(m)
        cells = sum(m.domain.info(0)["inner_cells"] for m in members)
        ens = artss_amd.Ensemble(members, threads=threads)
        for n in range(1, 4):
            ens.step(n)
        for m in members:
            m.domain.sync()
        t0 = time.perf_counter()
        for n in range(4, 4 + args.ensemble_steps):
            ens.step(n)
        for m in members:
            m.domain.sync()
        secs = time.perf_counter() - t0
        out["cells"] = cells
        out[label] = {"value": cells * args.ensemble_steps / secs, "unit": "cell-updates/s",
                      "ms_per_step": secs * 1e3 / args.ensemble_steps}
        ens.close()
        for m in members:
            m.close()
    out["concurrency_gain"] = out["concurrent"]["value"] / out["sequential"]["value"]
    return out


class Run:
    """Steps one solver from rest; timed(k) times k steps between syncs."""

    def __init__(self, solver, dt):
        self.solver, self.dt, self.step = solver, dt, 1

    def timed(self, k):
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
            self.step += 1
        dom.sync()
        return time.perf_counter() - t0, cyc, dif, res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (the step would shard "
              "as z-slabs with halo exchanges; not implemented)", file=sys.stderr)
        sys.exit(2)
    import artss_amd

    with open(args.config) as fh:
        dt = float(fh.read().split("<dt>")[1].split("</dt>")[0])
    solver = make_solver(artss_amd, args, args.pressure, args.warm)
    dom = solver.domain
    info = dom.info(0)
    inner = info["inner_cells"]
    run = Run(solver, dt)
    run.timed(args.warmup)
    first = run.step

    # timed region: K steps as they run in production (V-cycles replayed from HIP graphs)
    secs, cycles, diff_it, residuals = run.timed(args.steps)
    value = inner * args.steps / secs
    after = run.step - 1
    fields = {f: solver.field(f) for f in ("u", "v", "w", "p", "T")} if args.plain_steps > 0 else None

    # kernel timing region: HIP events around every launch of the profiled
    # kernel on the solver's stream; the V-cycle then runs its launches
    # eagerly (events cannot bracket launches inside a replayed graph)
    timing_steps = max(1, min(args.steps, args.timing_steps))
    dom.profile_arm(args.profile_kernel)
    t_secs, _, _, _ = run.timed(timing_steps)
    launches, kernel_ms = dom.profile_read()
    # cells per thread along z the autotuner picked for the fused kernels of levels 0..3
    tuned = [artss_amd.tuned_cells_per_thread(dom.info(l)["size"]) for l in range(4)]
    kname, klevel = args.profile_kernel.split("@")
    kinfo = dom.info(int(klevel))
    per_launch_bytes = BYTES_PER_CELL.get(kname, 0) * (kinfo["size"] if kname in WHOLE_ARRAY else kinfo["inner_cells"])
    mean_ms = kernel_ms / launches if launches else float("nan")
    achieved = per_launch_bytes / (mean_ms * 1e-3) / 1e9 if launches else None
    roofline = {"bound": "hbm", "kernel": args.profile_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                "traffic": measured_traffic(args.profile_kernel),
                "bytes_per_launch": per_launch_bytes, "launches": launches, "mean_launch_us": mean_ms * 1e3,
                "timing_steps": timing_steps, "timing_region_ms_per_step": t_secs * 1e3 / timing_steps,
                "share_of_timing_region": kernel_ms / (t_secs * 1e3)}

    # the CPU sample continues from the state after the GPU regions (same
    # workload, later steps)
    cpu = cpu_baseline(solver, args.config, run.step, args.cpu_steps, inner)
    solver.close()

    # plain FP64 path (V-cycles alone to the same residual) from rest over the
    # same steps: its throughput, and the field error of the accelerated path
    plain = None
    if args.plain_steps > 0 and args.pressure != "vcycle" and not args.reference_schedule:
        ps = make_solver(artss_amd, args, "vcycle")
        prun = Run(ps, dt)
        prun.timed(args.warmup)
        k_plain = min(args.plain_steps, args.steps)
        p_secs, p_cyc, _, p_res = prun.timed(k_plain)
        prun.timed(after - prun.step + 1)  # to the step the accelerated fields were taken at
        errors = {}
        for f, a in fields.items():
            ref = ps.field(f)
            errors[f] = float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))
        ps.close()
        plain = {"value": inner * k_plain / p_secs, "unit": "cell-updates/s", "steps": k_plain,
                 "ms_per_step": p_secs * 1e3 / k_plain, "mean_vcycles": float(np.mean(p_cyc)),
                 "max_final_residual": float(np.max(p_res)), "path": PATHS["vcycle"],
                 "speedup_of_value": value / (inner * k_plain / p_secs),
                 "max_rel_field_error_at_step_%d" % after: errors}

    ensemble = ensemble_run(artss_amd, args) if args.ensemble_members > 0 else None

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.pressure != "mixed" else "f64 (pressure preconditioner f32)",
        "data": "synthetic: plume from rest, timed steps %d..%d" % (first, first + args.steps - 1),
        "config": {"workload": "plume_256x128x128", "grid": list(info["n"]), "inner_cells": inner,
                   "path": PATHS[args.pressure], "mg_schedule": args.mg_schedule, "tail_cells": args.tail_cells,
                   "warm_start": ("previous p" if args.warm == "none" or args.pressure == "vcycle" else
                                  "p + (p - p_prev), linear extrapolation in time"),
                   "coarse_iterations": args.coarse_iterations, "sweeps": args.sweeps, "weight": args.weight,
                   "coarse": args.coarse, "tail_sweeps": args.tail_sweeps,
                   "tuned_cells_per_thread": tuned,
                   "pressure": ("reference V-cycle schedule" if args.reference_schedule else
                                "iterate until ||rhs - lap p||_2 <= %g every step" % args.tolerance),
                   "mean_vcycles": float(np.mean(cycles)), "mean_diffusion_iterations": np.mean(
                       diff_it, axis=0).tolist(), "max_final_residual": float(np.max(residuals))},
        "plain_fp64": plain, "ensemble": ensemble,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
