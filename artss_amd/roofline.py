"""Algorithmic HBM bytes of every kernel of the solver path, and a summary of
a rocprofv3 kernel trace against the MI355X HBM roofline.

    python -m artss_amd.roofline TRACE.csv --grid 256 128 128 --levels 5 \\
        --steps S [--ms-per-step T] [--tail-cells 4096] [--laminar] [--title ...]

TRACE.csv is the `*_kernel_trace.csv` of `rocprofv3 --kernel-trace
--output-format csv`; S is the number of time steps the traced command ran
(all of them, warm-up included), T the wall time per step that command
measured (optional).  Prints a markdown table: per kernel and launch shape
the launches, durations, algorithmic bytes per launch and achieved GB/s,
then the totals over the trace: sum of bytes / sum of kernel time (the
kernels' fraction of the 8 TB/s peak) and, with T, sum of bytes per step /
T (the whole step's fraction).

Algorithmic bytes count each distinct array element a launch must read or
write once, at its element size (8 B FP64, 4 B FP32), over the cells the
launch covers: inner cells of its level for stencils, the whole array for
whole-array kernels, the fill's entries for boundary kernels.  Neighbour
re-reads, index tables and face tables are not counted, so traffic above
these figures is waste.  The level a launch ran on follows from its grid
(64 x 4 threads per workgroup over the inner box or the whole array of a
level).  Launches of a gated kernel (one that returns at once when its
loop's device-side test has stopped the loop) shorter than a quarter of
the median of their kernel and shape did no work and count no bytes.

Diffusion kernels take an eddy viscosity on the turbulent solvers (the
plume); --laminar counts them without it.
"""
import argparse
import csv
import math
import re
import statistics
import sys

HBM_PEAK_GBS = 8000.0
BX, BY = 64, 4

# kernels whose launches return at once when their loop's device test has stopped
GATED = {"k_copy", "k_cgs_color", "k_cgs_iter", "k_update_norm", "k_reduce_gate", "k_gcr_update", "k_gcr_coef", "k_sum_gate",
         "k_bc_faces", "k_bc_domain"}


class Levels:
    """level sizes n_l = floor(n_{l-1} / 2) (DomainData.cpp:51) with one ghost layer"""

    def __init__(self, grid, levels):
        self.n = [tuple(grid)]
        for _ in range(levels):
            self.n.append(tuple(max(1, v // 2) for v in self.n[-1]))
        self.N = [tuple(v + 2 for v in n) for n in self.n]

    def inner(self, l):
        return self.n[l][0] * self.n[l][1] * self.n[l][2]

    def size(self, l):
        return self.N[l][0] * self.N[l][1] * self.N[l][2]

    def ghosts(self, l):
        return self.size(l) - self.inner(l)

    def level_of(self, gx, gy, half_x=False):
        """level whose inner box or whole array a 64x4 launch covers (grid in threads)"""
        def up(v, b):
            return -(-v // b) * b
        for l, (n, N) in enumerate(zip(self.n, self.N)):
            nx = -(-n[0] // 2) if half_x else n[0]
            if gx == up(nx, BX) and gy == up(n[1], BY):
                return l, "inner"
            if not half_x and gx == up(N[0], BX) and gy == up(N[1], BY):
                return l, "full"
        return None, None

    def level_of_1d(self, threads):
        """whole-array element-wise launch: the level whose array fits the grid"""
        for l in range(len(self.n)):
            if threads == min(-(-self.size(l) // 256), 65536) * 256:
                return l
        return None


def parse_name(name):
    """('k_jacobi_bc', ['float', 'double', 'true', 'false']) from a kernel name as rocprofv3
    prints it (demangled, with or without its parameter list)"""
    s = name.replace("(anonymous namespace)::", "").strip()
    m = re.search(r"(\w+)\s*(?:<([^()]*)>)?\s*(?:\(|$)", s)
    if not m:
        return s, []
    args = [a.strip() for a in (m.group(2) or "").split(",") if a.strip()]
    return m.group(1), args


def size_of(t):
    return 4 if t == "float" else 8


def bytes_of(kernel, args, gx, gy, gz, L, turbulent=True, tail_cells=4096):
    """algorithmic bytes of one launch and a short note, or (None, reason)"""
    ev = 8 if turbulent else 0
    level, kind = L.level_of(gx, gy)
    if kernel in ("k_copy", "k_fill", "k_fill32", "k_scale", "k_add", "k_scale_to") or kernel == "k_copy_fan":
        l = L.level_of_1d(gx)
        if l is None:
            return None, "array size unknown"
        n = L.size(l)
        per = {"k_copy": 16, "k_fill": 8, "k_fill32": 4, "k_scale": 16, "k_add": 24, "k_scale_to": 16}.get(kernel)
        if kernel == "k_copy_fan":
            f = int(args[0]) if args else 1
            per = 24 * f
            return per * n, "%d field(s): read 1, write 2, whole array, level %d" % (f, l)
        return per * n, "%d B per array cell, level %d" % (per, l)
    if kernel == "k_cgs_color":
        level, kind = L.level_of(gx, gy, half_x=True)
        if level is None:
            return None, "level unknown"
        # one color: its cells read and written, their b (and nu_t) read, the
        # other color's cells read as neighbours
        per = 8 + 8 + 8 + ev + 8
        return per * L.inner(level) // 2, "one color: read+write it, read b%s and the other color, level %d" % (
            ", nu_t" if ev else "", level)
    if kernel == "k_cgs_iter":
        level = 0  # the diffusion runs on the finest level only
        # one whole iteration: src read and dst written on every cell, b (and
        # nu_t) read on the fluid cells
        return 16 * L.size(level) + (8 + ev) * L.inner(level), \
            "one iteration: read src, write dst on every cell, read b%s, level %d" % (", nu_t" if ev else "", level)
    if kernel == "k_swap":
        l = L.level_of_1d(gx)
        if l is None:
            return None, "array size unknown"
        return 32 * L.size(l), "two arrays read and written, level %d" % l
    if kernel == "k_mg_tail":
        first = next((l for l in range(len(L.n)) if L.size(l) <= tail_cells), len(L.n) - 1)
        return 16 * L.size(first), "coarse levels on chip: read rhs, write error of level %d" % first
    if kernel == "k_dense_apply":
        rows = gx // 64
        return 8 * rows * rows + 16 * rows, "dense inverse %dx%d" % (rows, rows)
    if kernel in ("k_gcr_coef", "k_sum", "k_sum_gate", "k_reduce_gate", "k_reduce_sum", "k_gate_open",
                  "k_gate_open_iter"):
        return 8 * 4 * L.inner(0) // (BX * BY * 4), "partial sums (latency bound)"
    if kernel == "k_bc_domain":
        l = next((q for q in range(len(L.n)) if gx == -(-L.ghosts(q) // 256) * 256), None)
        if l is None:
            return None, "level unknown"
        return 16 * L.ghosts(l), "ghost cells: read source, write ghost, level %d" % l
    if kernel in ("k_bc_faces", "k_fill_bc", "k_bc_merged"):
        return 16 * gx * max(1, gy) * max(1, gz), "at most one entry per thread: read neighbour, write cell"
    if level is None:
        return None, "level unknown"
    cells = L.inner(level)
    whole = L.size(level)
    if kernel == "k_gcr_update":
        t = size_of(args[0]) if args else 8
        return (32 + 3 * t) * whole, "x, r read+write, z and kept z read, new kept z written (%d B), whole array" % t
    if kernel == "k_lap_dots":
        tz = size_of(args[0]) if args else 8
        return (2 * tz + 8) * cells, "read z, kept z, r"
    if kernel == "k_jacobi_bc":
        t = size_of(args[0])
        tb = size_of(args[1])
        zero = len(args) > 3 and args[3] == "true"
        return (t + (0 if zero else t) + tb) * cells, "level %d sweep%s" % (level, " from zero" if zero else "")
    if kernel == "k_residuum_faces":
        t, tb = size_of(args[0]), size_of(args[1])
        return (2 * t + tb) * cells, "level %d: read x, b; write r" % level
    if kernel == "k_prolongate_add_bc":
        tf, tc = size_of(args[0]), size_of(args[1])
        return (2 * tf) * cells + tc * L.inner(level + 1), "to level %d: read fine, coarse; write" % level
    if kernel == "k_restrict_sweep0":
        # grid over the coarse array; 8 fine FP32 reads per coarse cell, write rhs and error
        return (8 * 4 + 4 + 4) * cells, "to level %d: read 8 fine, write rhs and first sweep" % level
    if kernel == "k_restrict_fill":
        tin, tout = size_of(args[0]), size_of(args[1])
        terr = size_of(args[2]) if len(args) > 2 else 8
        return (8 * tin + tout) * cells + terr * whole, "to level %d: read 8 fine, write coarse, reset error" % level
    if kernel == "k_restrict":
        return (8 * 8 + 8) * cells, "to level %d" % level
    if kernel in ("k_prolongate", "k_prolongate_add"):
        extra = 8 if kernel == "k_prolongate_add" else 0
        return (8 + 8 * (8 + extra)) * cells, "from level %d: 8 fine cells per coarse cell" % level
    if kernel == "k_copy_bc":
        return 16 * (cells if kind == "inner" else whole), "copy with fill, level %d" % level
    per = {
        "k_sl_advect": 40, "k_jacobi_diffuse": 24 + ev, "k_jacobi_mg": 24, "k_jacobi_mg_first": 24,
        "k_explicit_diffuse": 16 + ev, "k_dissipate": 40, "k_update_norm": 16 + ev, "k_residuum": 24,
        "k_residual_partials": 16, "k_residual_vec": 24, "k_divergence": 32, "k_projection": 56,
        "k_smagorinsky": 32, "k_add_source": 24, "k_buoyancy": 24, "k_copy_domain": 16, "k_warm_partials": 16,
        "k_warm_apply": 24, "k_warm_pack": 24 + 12, "k_warm_target": 32 + 4,
    }.get(kernel)
    if per is None:
        return None, "no byte model"
    return per * cells, "%d B per cell, level %d" % (per, level)


def read_trace(path):
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            def col(*names):
                for n in names:
                    if n in r and r[n] != "":
                        return r[n]
                raise KeyError(names[0])
            rows.append({"name": col("Kernel_Name", "KernelName"),
                         "gx": int(col("Grid_Size_X", "Grid_X")), "gy": int(col("Grid_Size_Y", "Grid_Y")),
                         "gz": int(col("Grid_Size_Z", "Grid_Z")),
                         "ns": int(col("End_Timestamp", "EndNs")) - int(col("Start_Timestamp", "BeginNs"))})
    return rows


def summarize(rows, L, steps, ms_per_step=None, turbulent=True, tail_cells=4096, title=""):
    groups = {}
    for r in rows:
        key = (r["name"], r["gx"], r["gy"], r["gz"])
        groups.setdefault(key, []).append(r["ns"])
    out = []
    total_ns = total_bytes = 0
    unmodelled_ns = 0
    lines = []
    for (name, gx, gy, gz), ns in groups.items():
        kernel, args = parse_name(name)
        b, note = bytes_of(kernel, args, gx, gy, gz, L, turbulent, tail_cells)
        med = statistics.median(ns)
        worked = [t for t in ns if not (kernel in GATED and t < 0.25 * med)]
        idle = len(ns) - len(worked)
        t_all = sum(ns)
        total_ns += t_all
        if b is None:
            unmodelled_ns += t_all
            gbs = None
        else:
            total_bytes += b * len(worked)
            gbs = b * len(worked) / max(sum(worked), 1) if worked else None
        lines.append((t_all, kernel + ("<%s>" % ",".join(args) if args else ""), "%dx%dx%d" % (gx, gy, gz), len(ns),
                      idle, t_all / len(ns) / 1e3, min(ns) / 1e3, b, gbs, note))
    lines.sort(key=lambda x: -x[0])
    # mean duration of the level-0 GCR update launches that did work (the
    # kernel the benchmark's roofline names)
    mean_us = {}
    upd = [r["ns"] for r in rows if parse_name(r["name"])[0] == "k_gcr_update"]
    if upd:
        med = statistics.median(upd)
        worked = [t for t in upd if t >= 0.25 * med]
        mean_us["gcr_update"] = sum(worked) / len(worked) / 1e3
    if title:
        out.append("# " + title)
        out.append("")
    out.append("| kernel | grid (threads) | launches | idle | mean µs | min µs | total ms | share | bytes/launch | GB/s "
               "| frac | bytes |")
    out.append("|---|---|---|---|---|---|---|---|---|---|---|---|")
    for t_all, k, g, n, idle, mean, mn, b, gbs, note in lines:
        out.append("| %s | %s | %d | %d | %.2f | %.2f | %.3f | %.1f%% | %s | %s | %s | %s |" % (
            k.replace("<", "&lt;").replace(">", "&gt;"), g, n, idle, mean, mn, t_all / 1e6, 100.0 * t_all / total_ns,
            "%.3g" % b if b is not None else "", "%.0f" % gbs if gbs else "",
            "%.3f" % (gbs / HBM_PEAK_GBS) if gbs else "", note))
    out.append("")
    frac_k = total_bytes / max(total_ns, 1) / HBM_PEAK_GBS
    out.append("Totals over the trace (%d launches, %d steps): kernel time %.3f ms, algorithmic bytes %.4g, "
               "kernels without a byte model %.3f ms (%.1f%%)." % (len(rows), steps, total_ns / 1e6, total_bytes,
                                                                   unmodelled_ns / 1e6, 100.0 * unmodelled_ns /
                                                                   max(total_ns, 1)))
    out.append("")
    out.append("- sum of bytes / sum of kernel time: %.0f GB/s = **%.3f of 8 TB/s**" % (
        total_bytes / max(total_ns, 1), frac_k))
    per_step_bytes = total_bytes / steps
    per_step_kernel_ms = total_ns / 1e6 / steps
    out.append("- per step: %.4g bytes, %.3f ms of kernels" % (per_step_bytes, per_step_kernel_ms))
    frac_s = None
    if ms_per_step:
        frac_s = per_step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS
        out.append("- per step over the measured wall time %.3f ms: %.0f GB/s = **%.3f of 8 TB/s**" % (
            ms_per_step, per_step_bytes / (ms_per_step * 1e-3) / 1e9, frac_s))
    out.append("")
    out.append("Figures read by bench.py:")
    out.append("")
    out.append("- `kernel_frac = %.4f`" % frac_k)
    if frac_s is not None:
        out.append("- `step_frac = %.4f`" % frac_s)
    for name, mean in sorted(mean_us.items()):
        out.append("- `mean_us[%s] = %.3f`" % (name, mean))
    return "\n".join(out), {"kernel_frac": frac_k, "step_frac": frac_s, "bytes_per_step": per_step_bytes,
                            "kernel_ms_per_step": per_step_kernel_ms}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("trace")
    ap.add_argument("--grid", type=int, nargs=3, required=True)
    ap.add_argument("--levels", type=int, required=True)
    ap.add_argument("--steps", type=int, required=True)
    ap.add_argument("--ms-per-step", type=float, default=None)
    ap.add_argument("--tail-cells", type=int, default=4096)
    ap.add_argument("--laminar", action="store_true")
    ap.add_argument("--title", default="")
    a = ap.parse_args(argv)
    text, _ = summarize(read_trace(a.trace), Levels(a.grid, a.levels), a.steps, a.ms_per_step, not a.laminar,
                        a.tail_cells, a.title)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
