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
    "ops_dyn_smagorinsky": dict(xml=lambda: _dyn(cases.plume(n=(16, 8, 8), X=(0.0, 2.0), levels=2, steps=2, dt=0.01,
                                                            diffusion="Jacobi", obstacles=OBST_ONE, hrr=3.0,
                                                            source_x=1.0)),
                                ops=[("dyn_smagorinsky", 1.0)], lists=False),
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
    "run_nsturb_dyn_tg": dict(xml=lambda: _dyn(_tg_variant("NSTurbSolver", turbulence=True)),
                              run=("1,2", "taylor_green")),
    "run_diffusion_tg": dict(xml=lambda: _tg_variant("DiffusionSolver"), run=("1,2", "taylor_green")),
    "run_diffusion_turb_tg": dict(xml=lambda: _tg_variant("DiffusionTurbSolver", turbulence=True),
                                  run=("1,2", "taylor_green")),
    "run_advection_diffusion_tg": dict(xml=lambda: _tg_variant("AdvectionDiffusionSolver",
                                                               diffusion="ColoredGaussSeidel"),
                                       run=("1,2", "taylor_green")),
    # runs from the reference's own initial conditions (SolverController.cpp:94-196)
    "run_advection_gauss": dict(xml=lambda: _ic_case(
        "AdvectionSolver", ("GaussBubble", "<u_lin> 0.5 </u_lin><v_lin> 0.25 </v_lin><w_lin> -0.5 </w_lin>"
                                           "<x_shift> 0.4 </x_shift><y_shift> 0.5 </y_shift><z_shift> 0.55 </z_shift>"
                                           "<l> 0.125 </l>"), _PERIODIC, dt=0.01), run=("1,2", "")),
    "run_diffusion_expsinusprod": dict(xml=lambda: _ic_case(
        "DiffusionSolver", ("ExpSinusProd", "<l> 2.0 </l>"), _DIRICHLET0, nu=0.05), run=("1,2", "")),
    "run_diffusion_hat": dict(xml=lambda: _ic_case(
        "DiffusionSolver", ("Hat", "<x1> 0.25 </x1><x2> 0.625 </x2><y1> 0.3 </y1><y2> 0.7 </y2><z1> 0.125 </z1>"
                                   "<z2> 0.75 </z2><val_in> 2.0 </val_in><val_out> 1.0 </val_out>"),
        _DIRICHLET0, diffusion="ColoredGaussSeidel"), run=("1,2", "")),
    "run_burgers_expsinussum": dict(xml=lambda: _ic_case(
        "AdvectionDiffusionSolver", ("ExpSinusSum", ""), _PERIODIC, X=(0.0, 6.283185307179586),
        Y=(0.0, 6.283185307179586), Z=(0.0, 6.283185307179586)), run=("1,2", "")),
    "run_ns_mcdermott": dict(xml=lambda: _ic_case(
        "NSSolver", ("McDermott", "<A> 2.0 </A>"), _PERIODIC, X=(-3.141592653589793, 3.141592653589793),
        Y=(-3.141592653589793, 3.141592653589793), Z=(0.0, 1.0), n=(16, 16, 8)), run=("1,2", "")),
    "run_ns_vortex": dict(xml=lambda: _ic_case(
        "NSSolver", ("Vortex", "<u_lin> 0.1 </u_lin><v_lin> 0.0 </v_lin><w_lin> 0.05 </w_lin><pa> 0.0 </pa>"
                               "<rhoa> 1.0 </rhoa>"), _PERIODIC, X=(-0.1556, 0.1556), Y=(-0.1556, 0.1556),
        Z=(0.0, 0.0389), n=(16, 16, 8), nu=1e-4, dt=0.005), run=("1,2", "")),
    "run_ns_drift": dict(xml=lambda: _ic_case(
        "NSSolver", ("Drift", "<u_lin> 0.2 </u_lin><v_lin> -0.1 </v_lin><w_lin> 0.05 </w_lin><pa> 0.3 </pa>"),
        _PERIODIC, obstacles=OBST_ONE), run=("1,2", "")),
    "run_nstemp_layers": dict(xml=lambda: _nstemp_layers(), run=("1,2", "")),
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


def _dyn(x):
    """the same input with the dynamic Smagorinsky model"""
    y = x.replace('<turbulence type="ConstSmagorinsky"><Cs> 0.2 </Cs></turbulence>',
                  '<turbulence type="DynamicSmagorinsky"></turbulence>')
    assert y != x
    return y


def _nstemp_dissipation():
    x = cases.plume(n=(16, 16, 16), levels=2, steps=5, dt=0.02, diffusion="Jacobi", obstacles=OBST_ONE, hrr=5.0)
    x = x.replace('<solver description="NSTempTurbSolver">', '<solver description="NSTempSolver">')
    x = x.replace('dissipation="No"', 'dissipation="Yes"')
    return x


def _ic_case(solver, ic, bnd, X=(0.0, 1.0), Y=(0.0, 1.0), Z=(0.0, 1.0), n=(16, 16, 16), diffusion="Jacobi",
             nu=0.01, dt=2e-3, steps=5, levels=2, obstacles=(), extra=""):
    """a run from one of the reference's own initial conditions (Functions.cpp)"""
    phys = '<t_end> {!r} </t_end><dt> {!r} </dt><nu> {!r} </nu>'.format(steps * dt, dt, nu)
    body = '<advection type="SemiLagrangian" field="u,v,w"></advection>'
    if solver != "AdvectionSolver":
        body += cases.diffusion_xml(diffusion, "u,v,w", 100, 1e-7, 1.0)
    if solver in ("NSSolver", "NSTurbSolver"):
        body += ('<source type="ExplicitEuler" force_fct="Zero" dir="x"></source>' +
                 cases.pressure_xml(levels))
    body += extra
    return cases._wrap(phys, (solver, body), cases.domain_xml(X, Y, Z, n), bnd, cases.obstacles_xml(list(obstacles)),
                       '<initial_conditions usr_fct="{}" random="No">{}</initial_conditions>'.format(*ic))


_PERIODIC = '<boundary field="u,v,w,p" patch="front,back,left,right,bottom,top" type="periodic" />'
_DIRICHLET0 = '<boundary field="u,v,w,p" patch="front,back,left,right,bottom,top" type="dirichlet" value="0.0" />'


def _nstemp_layers():
    x = cases.plume(n=(16, 16, 16), levels=2, steps=5, dt=0.02, diffusion="Jacobi", obstacles=OBST_ONE, hrr=5.0)
    x = x.replace('<solver description="NSTempTurbSolver">', '<solver description="NSTempSolver">')
    start = x.index("<initial_conditions")
    end = x.index("</initial_conditions>") + len("</initial_conditions>")
    ic = ('<initial_conditions usr_fct="LayersT" random="No"><n_layers> 3 </n_layers><dir>y</dir>'
          '<border_1> 0.3 </border_1><border_2> 0.59375 </border_2><value_1> 293.15 </value_1>'
          '<value_2> 296.0 </value_2><value_3> 300.5 </value_3></initial_conditions>')
    return x[:start] + ic + x[end:]


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
