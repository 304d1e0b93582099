"""Writes the golden-vector fixtures under tests/golden/ from the real reference.

Runs oracle/_ref/ref_harness (the reference's own sources compiled by
`make -C oracle ref`) on small inputs and packs its raw dumps into .npz
files: the inputs it generated and the outputs the reference computed.
Only run in a container where /root/reference exists; the fixtures it writes
are data and are committed with this script.

    python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# obstacle boxes in physical coordinates on the unit-cube test grids
OBST_ONE = [(0.25, 0.5, 0.0, 0.375, 0.5, 0.75)]                      # stands on the floor
OBST_ADJ = [(0.125, 0.375, 0.25, 0.625, 0.25, 0.75),                 # A
            (0.375, 0.625, 0.375, 0.75, 0.125, 0.5),                 # B touches A on x = 0.375
            (0.625, 0.875, 0.0, 0.5, 0.25, 0.875)]                   # C touches B on x = 0.625

CASES = {
    # operator-level cases: seeded random fields through single reference calls
    "ops_jacobi": dict(xml=lambda: cases.plume(n=(16, 8, 8), X=(0.0, 2.0), levels=2, steps=2, dt=0.01,
                                               diffusion="Jacobi", smoother="Jacobi", obstacles=OBST_ONE,
                                               hrr=3.0, source_x=1.0),
                       ops=[("bc", 1.0), ("advect", 1.7), ("jacobi", 1e-2), ("jacobi_turb", 1e-2),
                            ("divergence", 1.0), ("projection", 1.0), ("vcycle1", 1.0), ("vcycle2", 1.0),
                            ("smagorinsky", 1.0), ("sources", 1.0)], lists=True),
    "ops_cgs": dict(xml=lambda: cases.plume(n=(16, 8, 8), X=(0.0, 2.0), levels=2, steps=2, dt=0.01,
                                            diffusion="ColoredGaussSeidel", smoother="ColoredGaussSeidel",
                                            obstacles=OBST_ONE, hrr=3.0, source_x=1.0),
                    ops=[("cgs", 1e-2), ("cgs_turb", 1e-2), ("vcycle1", 1.0), ("vcycle2", 1.0)], lists=False),
    "ops_explicit": dict(xml=lambda: cases.plume(n=(16, 8, 8), X=(0.0, 2.0), levels=2, steps=2, dt=0.01,
                                                 diffusion="Jacobi", smoother="Jacobi", obstacles=OBST_ONE,
                                                 hrr=3.0, source_x=1.0),
                         ops=[("explicit", 1e-2), ("explicit_turb", 1e-2)], lists=False),
    "geo_adjacent": dict(xml=lambda: _bc_mixed(OBST_ADJ), ops=[("bc", 1.0)], lists=True),
    # whole-solver runs: state after steps 1, 2 and the last one
    "run_ns_tg": dict(xml=lambda: cases.ns_taylor_green(n=16, levels=2, steps=5, dt=2e-3, reference=True),
                      run=("1,2", "taylor_green")),
    "run_plume_jacobi": dict(xml=lambda: cases.plume(n=(16, 16, 16), levels=2, steps=5, dt=0.02,
                                                     diffusion="Jacobi", obstacles=OBST_ONE, hrr=5.0),
                             run=("1,2", "")),
    "run_plume_explicit": dict(xml=lambda: cases.plume(n=(16, 16, 16), levels=2, steps=5, dt=0.02,
                                                       diffusion="Explicit", obstacles=OBST_ONE, hrr=5.0),
                               run=("1,2", "")),
    "run_nstemp_dissipation": dict(xml=lambda: _nstemp_dissipation(), run=("1,2", "")),
    "run_nsturb_tg": dict(xml=lambda: _tg_variant("NSTurbSolver", turbulence=True), run=("1,2", "taylor_green")),
    "run_diffusion_tg": dict(xml=lambda: _tg_variant("DiffusionSolver"), run=("1,2", "taylor_green")),
    "run_diffusion_turb_tg": dict(xml=lambda: _tg_variant("DiffusionTurbSolver", turbulence=True),
                                  run=("1,2", "taylor_green")),
    "run_advection_diffusion_tg": dict(xml=lambda: _tg_variant("AdvectionDiffusionSolver",
                                                               diffusion="ColoredGaussSeidel"),
                                       run=("1,2", "taylor_green")),
    "run_plume_cgs": dict(xml=lambda: cases.plume(n=(16, 16, 16), levels=2, steps=5, dt=0.02,
                                                  diffusion="ColoredGaussSeidel", smoother="ColoredGaussSeidel",
                                                  obstacles=OBST_ONE, hrr=5.0),
                          run=("1,2", "")),
}


def _tg_variant(solver, diffusion="Jacobi", turbulence=False):
    """the Taylor-Green NSSolver input with another solver description (the
    harness sets the Taylor-Green velocity, usr_fct stays "Zero")"""
    x = cases.ns_taylor_green(n=16, levels=2, steps=5, dt=2e-3, reference=True, diffusion=diffusion)
    x = x.replace('<solver description="NSSolver">', '<solver description="%s">' % solver)
    if turbulence:
        x = x.replace('<source type="ExplicitEuler"',
                      '<turbulence type="ConstSmagorinsky"><Cs> 0.2 </Cs></turbulence><source type="ExplicitEuler"', 1)
    return x


def _nstemp_dissipation():
    x = cases.plume(n=(16, 16, 16), levels=2, steps=5, dt=0.02, diffusion="Jacobi", obstacles=OBST_ONE, hrr=5.0)
    x = x.replace('<solver description="NSTempTurbSolver">', '<solver description="NSTempSolver">')
    x = x.replace('dissipation="No"', 'dissipation="Yes"')
    return x


def _bc_mixed(obstacles):
    """Unit-cube-like box with unequal spacing and every boundary type:
    periodic in x, non-zero Neumann in y, non-zero Dirichlet in z, Neumann on
    left/right for T (DomainBoundary.cpp:120 scales by dz there)."""
    x = cases.plume(n=(16, 8, 8), X=(0.0, 1.0), levels=2, steps=2, dt=0.01, diffusion="Jacobi",
                    obstacles=(), hrr=3.0)
    bnd = ('<boundary field="u,v,w,p" patch="left,right" type="periodic" />'
           '<boundary field="u,v,w" patch="bottom,top" type="neumann" value="0.7" />'
           '<boundary field="u,v,w" patch="front,back" type="dirichlet" value="1.3" />'
           '<boundary field="p" patch="bottom,top,front,back" type="neumann" value="-0.4" />'
           '<boundary field="T" patch="left,right,bottom" type="neumann" value="2.5" />'
           '<boundary field="T" patch="top,front,back" type="dirichlet" value="300.5" />'
           '<boundary field="nu" patch="left,right,bottom,top,front,back" type="neumann" value="0.0" />')
    start = x.index("<boundaries>") + len("<boundaries>")
    end = x.index("</boundaries>")
    x = x[:start] + bnd + x[end:]
    obs = cases.obstacles_xml(obstacles, vel_bc="dirichlet", scal_bc="neumann", vel_value=0.25, scal_value=1.5)
    x = x.replace('<obstacles enabled="No"/>', obs)
    return x


def _read_dir(d):
    out = {}
    for f in sorted(os.listdir(d)):
        p = os.path.join(d, f)
        if f.endswith(".f64"):
            out[f[:-4]] = np.fromfile(p, np.float64)
        elif f.endswith(".u64"):
            out[f[:-4]] = np.fromfile(p, np.uint64)
    return out


def main(only=None):
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"], check=True)
    for name, spec in CASES.items():
        if only and name not in only:
            continue
        with tempfile.TemporaryDirectory() as tmp:
            xml = spec["xml"]()
            xp = os.path.join(tmp, "case.xml")
            with open(xp, "w") as fh:
                fh.write(xml)
            arrays = {"xml": np.frombuffer(xml.encode(), np.uint8)}
            if spec.get("lists"):
                d = os.path.join(tmp, "lists")
                os.mkdir(d)
                subprocess.run([HARNESS, "lists", xp, d], check=True, capture_output=True)
                arrays.update({"list_" + k: v for k, v in _read_dir(d).items()})
            for op, amp in spec.get("ops", []):
                d = os.path.join(tmp, op)
                os.mkdir(d)
                subprocess.run([HARNESS, "op", xp, op, d, "20261202", repr(amp)], check=True, capture_output=True)
                arrays.update({op + "/" + k: v for k, v in _read_dir(d).items()})
            if "run" in spec:
                steps, ic = spec["run"]
                d = os.path.join(tmp, "run")
                os.mkdir(d)
                args = [HARNESS, "run", xp, d, steps] + ([ic] if ic else [])
                subprocess.run(args, check=True, capture_output=True)
                arrays.update({"run/" + k: v for k, v in _read_dir(d).items()})
            path = os.path.join(HERE, name + ".npz")
            np.savez_compressed(path, **arrays)
            print(name, len(arrays), "arrays", os.path.getsize(path) // 1024, "KiB")


if __name__ == "__main__":
    main(sys.argv[1:] or None)
