"""The CPU oracle against fixtures of the reference's own tests and against
results the mathematics fixes.

The golden vectors (test_oracle_golden.py) pin the oracle bit for bit to the
real reference on seeded inputs.  These checks are independent of both: a
dropped term, a wrong sign, a transposed axis or a wrong weight in the
oracle fails at least one of them.

- index lists: the sizes the reference's gtests expect
  (gtests/domain/Multigrid.cpp:30-57, 8x12x4 cells, two coarse levels), for
  the oracle and for the product's host-side lists;
- SL advection of a linear field with constant velocity: trilinear
  back-tracing is exact for linear data, so every cell whose departure point
  stays inside the domain holds f(x - C h);
- divergence of a linear field and projection with a quadratic pressure:
  central differences are exact on both;
- ConstSmagorinsky on a uniform shear u = a y: |S| = |a|, so
  nu_t = (Cs delta)^2 |a| with delta = (dx dy dz)^(1/3);
- implicit diffusion of a discrete eigenfunction on a periodic box: the
  converged iterate is b / (1 - D dt lambda_h);
- the V-cycle on sin(pi x) sin(pi y) sin(pi z) with homogeneous Dirichlet
  pressure: an exact eigenvector of the 7-point Laplacian with the
  antisymmetric ghost fill, so p -> rhs / lambda_h;
- the Gaussian heat source: the discrete integral of the spatial part over
  the domain is HRR / cp (GaussFunction.cpp:75-90 normalises by it).
"""
import math

import numpy as np
import pytest

import artss_amd
import artss_xml
import cases
import oracle_py as O

RHO, U, V, W, P, T = 0, 1, 2, 3, 4, 5  # reference FieldType numbering


def _xml(n, X=(0.0, 1.0), Y=(0.0, 1.0), Z=(0.0, 1.0), levels=2, bnd=None, smoother="Jacobi", nu=0.01, dt=1e-3,
         diffusion="Jacobi", obstacles=()):
    phys = '<t_end> {} </t_end><dt> {} </dt><nu> {} </nu>'.format(repr(10 * dt), repr(dt), repr(nu))
    solver = ('<advection type="SemiLagrangian" field="u,v,w"></advection>' +
              cases.diffusion_xml(diffusion, "u,v,w", 100, 1e-7, 1.0) +
              '<source type="ExplicitEuler" force_fct="Zero" dir="x"></source>' +
              cases.pressure_xml(levels, smoother=smoother))
    if bnd is None:
        bnd = '<boundary field="u,v,w,p" patch="front,back,left,right,bottom,top" type="periodic" />'
    ic = '<initial_conditions usr_fct="Zero" random="No"></initial_conditions>'
    return cases._wrap(phys, ("NSSolver", solver), cases.domain_xml(X, Y, Z, n), bnd,
                       cases.obstacles_xml(list(obstacles)), ic)


def _grid(dom, level=0):
    """cell-centre coordinates broadcast to the (k, j, i) array shape"""
    info = dom.info(level)
    c = dom.cfg["domain"]
    Nx, Ny, Nz = info["N"]
    hx, hy, hz = info["h"]
    x = c["X1"] + (np.arange(Nx) - 0.5) * hx
    y = c["Y1"] + (np.arange(Ny) - 0.5) * hy
    z = c["Z1"] + (np.arange(Nz) - 0.5) * hz
    return np.meshgrid(z, y, x, indexing="ij")[::-1]  # X, Y, Z each (Nz, Ny, Nx)


def _inner(dom, level=0, margin=0):
    info = dom.info(level)
    Nx, Ny, Nz = info["N"]
    m = np.zeros((Nz, Ny, Nx), bool)
    m[1 + margin:Nz - 1 - margin, 1 + margin:Ny - 1 - margin, 1 + margin:Nx - 1 - margin] = True
    return m


def test_list_sizes_match_reference_gtest():
    xml = _xml((8, 12, 4), X=(0.0, 5.0), Y=(0.0, 10.0), Z=(0.0, 15.0), levels=2)
    cfg = artss_xml.parse(xml)
    odom = O.Domain(cfg)
    hdom = artss_amd.Domain(xml, device=-1)
    domain_cells, inner_cells = (840, 192, 60), (384, 48, 6)
    for level in range(3):
        assert len(odom.list(1, level)) == domain_cells[level]
        assert len(odom.list(0, level)) == inner_cells[level]
        assert len(hdom.list(1, level)) == domain_cells[level]
        assert len(hdom.list(0, level)) == inner_cells[level]


def test_sl_advect_linear_field_constant_velocity():
    n = (24, 20, 16)
    xml = _xml(n, X=(0.0, 1.2), Y=(0.0, 1.0), Z=(0.0, 0.8))
    cfg = artss_xml.parse(xml)
    dom = O.Domain(cfg)
    X, Y, Z = _grid(dom)
    h = dom.info(0)["h"]
    dt = 0.01
    cfl = (0.3, -0.6, 1.7)  # fractional, negative and above one
    f = 1.5 + 2.0 * X - 3.0 * Y + 0.5 * Z
    u, v, w = (np.full(X.shape, cfl[a] * h[a] / dt) for a in range(3))
    out = np.zeros(X.size)
    dom.sl_advect(out, f.ravel(), u.ravel(), v.ravel(), w.ravel(), dt, U)
    exact = 1.5 + 2.0 * (X - cfl[0] * h[0]) - 3.0 * (Y - cfl[1] * h[1]) + 0.5 * (Z - cfl[2] * h[2])
    m = _inner(dom, margin=3)
    err = np.abs(out.reshape(X.shape) - exact)[m].max()
    assert err < 1e-12, err


def test_divergence_linear_and_projection_quadratic():
    n = (12, 10, 14)
    cfg = artss_xml.parse(_xml(n, X=(0.0, 1.2), Y=(-0.5, 0.5), Z=(0.0, 1.4)))
    dom = O.Domain(cfg)
    X, Y, Z = _grid(dom)
    u, v, w = X.ravel().copy(), (2 * Y).ravel().copy(), (3.0 - Z).ravel().copy()
    div = np.zeros(X.size)
    dom.divergence(div, u, v, w)
    m = _inner(dom)
    assert np.abs(div.reshape(X.shape)[m] - 2.0).max() < 1e-12
    # the ghost shell of the divergence is zero (IPressure.cpp:58-104)
    assert np.all(div.reshape(X.shape)[~m] == 0)
    p = (0.5 * X * X + Y * Y - 0.5 * Z * Z).ravel()
    ou, ov, ow = np.zeros(X.size), np.zeros(X.size), np.zeros(X.size)
    dom.projection(ou, ov, ow, u, v, w, p)
    for got, want in ((ou, 0.0), (ov, 0.0), (ow, 3.0)):
        assert np.abs(got.reshape(X.shape)[m] - want).max() < 1e-12


def test_smagorinsky_uniform_shear():
    n = (10, 12, 8)
    cfg = artss_xml.parse(_xml(n, X=(0.0, 1.0), Y=(0.0, 1.5), Z=(0.0, 0.6)))
    dom = O.Domain(cfg)
    X, Y, Z = _grid(dom)
    a, cs = 0.7, 0.2
    hx, hy, hz = dom.info(0)["h"]
    u, v, w = (a * Y).ravel().copy(), np.zeros(X.size), np.zeros(X.size)
    ev = np.zeros(X.size)
    dom.smagorinsky(ev, u, v, w, cs)
    delta = (hx * hy * hz) ** (1.0 / 3.0)
    m = _inner(dom)
    assert np.allclose(ev.reshape(X.shape)[m], cs * cs * delta * delta * abs(a), rtol=1e-12, atol=0)


@pytest.mark.parametrize("kind", ["Jacobi", "ColoredGaussSeidel"])
def test_implicit_diffusion_eigenfunction(kind):
    n = (16, 16, 16)
    cfg = artss_xml.parse(_xml(n))
    dom = O.Domain(cfg)
    X, Y, Z = _grid(dom)
    h = dom.info(0)["h"][0]
    D, dt = 0.05, 0.02
    e = np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) * np.sin(4 * np.pi * Z)
    lam = -(4 / h ** 2) * (np.sin(np.pi * h) ** 2 * 2 + np.sin(2 * np.pi * h) ** 2)
    b = e.ravel().copy()
    x0 = np.zeros(X.size)
    dom.apply_bc(b, U)
    out, inp = x0.copy(), x0.copy()
    it, res = dom.diffuse(kind, out, inp, b, D, np.zeros(X.size), dt, 5000, 1e-13, 1.0, U)
    exact = e / (1 - D * dt * lam)
    m = _inner(dom)
    err = np.abs(out.reshape(X.shape) - exact)[m].max()
    assert err < 1e-10, (it, res, err)


def test_vcycle_dirichlet_eigenvector():
    n = (32, 32, 32)
    bnd = ('<boundary field="u,v,w" patch="front,back,left,right,bottom,top" type="dirichlet" value="0.0" />'
           '<boundary field="p" patch="front,back,left,right,bottom,top" type="dirichlet" value="0.0" />')
    cfg = artss_xml.parse(_xml(n, levels=3, bnd=bnd))
    dom = O.Domain(cfg)
    X, Y, Z = _grid(dom)
    h = dom.info(0)["h"][0]
    rhs = (np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)).ravel()
    lam = -3 * (4 / h ** 2) * np.sin(np.pi * h / 2) ** 2
    p = np.zeros(X.size)
    mg = O.MG(dom, cfg)
    dt = cfg["dt"]
    m = _inner(dom).ravel()
    for cycle in range(60):
        mg.pressure(p, rhs, 2 * dt, dt)  # n_cycle V-cycles
    err = np.abs(p - rhs / lam)[m].max() / np.abs(rhs / lam)[m].max()
    assert err < 1e-9, err
    assert dom.residual_norm(p, rhs) < 1e-6


def test_gauss_source_integral_is_hrr_over_cp():
    n = (20, 24, 16)
    cfg = artss_xml.parse(_xml(n, X=(0.0, 1.0), Y=(0.0, 1.2), Z=(0.0, 0.8)))
    dom = O.Domain(cfg)
    hx, hy, hz = dom.info(0)["h"]
    hrr, cp = 3.5, 1.2
    out = np.zeros(dom.size(0))
    dom.gauss_spatial(out, (hrr, cp, 0.45, 0.1, 0.4, 0.1, 0.15, 0.12))
    total = out.sum() * hx * hy * hz
    assert math.isclose(total, hrr / cp, rel_tol=1e-12), total


def test_vcycle_reproduces_reference_pressure_fixture():
    """The reference's own pressure test (tests/pressure: 64^3 over [0, 2]^3,
    five levels, SinSinSin right-hand side with l = 2, Dirichlet 0, one step
    with the residual-driven cycle loop of VCycleMG::pressure) against its
    fixture p_ref.dat, to the fixture's six digits (the file lists the inner
    cells first, then the ghost cells)."""
    import artss_xml
    import oracle_py
    import refxml

    z = refxml.load("pressure")
    cfg = artss_xml.parse(bytes(z["xml"]).decode())
    dom = oracle_py.Domain(cfg)
    info = dom.info(0)
    N, h = info["N"], info["h"]
    l = float(cfg["ic"]["values"]["l"])
    # Functions::sin_sin_sin (Functions.cpp:781-811) on the inner cells
    lo = [cfg["domain"][k] for k in ("X1", "Y1", "Z1")]
    rhs = dom.field()
    inner = dom.list(0, 0).astype(np.int64)
    k, rem = np.divmod(inner, N[0] * N[1])
    j, i = np.divmod(rem, N[0])
    x = [lo[0] + (i - 0.5) * h[0], lo[1] + (j - 0.5) * h[1], lo[2] + (k - 0.5) * h[2]]
    rhs[inner] = np.sin(l * np.pi * x[0]) * np.sin(l * np.pi * x[1]) * np.sin(l * np.pi * x[2])
    p = dom.field()
    dom.apply_bc(p, 4)
    mg = oracle_py.MG(dom, cfg)
    mg.pressure(p, rhs, cfg["dt"], cfg["dt"])
    perm = refxml.inner_first(z["dims"])
    assert refxml.same6(p[perm], z["refdat/p"]) == 0


@pytest.mark.parametrize("solver, turbulent, adds", [("NSTempConSolver", False, True),
                                                     ("NSTempTurbConSolver", True, True),
                                                     ("NSTempTurbConSolver", False, False)])
def test_concentration_cube_source_at_rest(solver, turbulent, adds):
    """A fluid at rest (no heat release, uniform temperature, no thermal
    diffusivity) with a Cube
    concentration source and no molecular diffusivity: the concentration
    grows by dt * value per step on exactly the cells whose centres lie in
    the box and stays zero elsewhere, velocities stay zero.
    NSTempTurbConSolver adds the source only for a turbulent concentration
    (NSTempTurbConSolver.cpp:277-285)."""
    import artss_xml
    import cases
    import oracle_py

    x = cases.concentration_plume(solver, "Cube", turbulent, gamma=0.0, hrr=0.0, kappa=0.0, n=(16, 16, 16))
    cfg = artss_xml.parse(x)
    dom = oracle_py.Domain(cfg)
    sim = oracle_py.Sim(dom, cfg, None, None, None, cases.uniform_ic(dom, float(cfg["ic"]["values"]["val"])))
    info = dom.info(0)
    Nx, Ny, Nz = info["N"]
    h = info["h"]
    start, end, val = (0.25, 0.0, 0.3), (0.6, 0.25, 0.7), 3.5
    box = np.zeros((Nz, Ny, Nx), bool)
    for k in range(1, Nz - 1):
        for j in range(1, Ny - 1):
            for i in range(1, Nx - 1):
                c = ((i - 0.5) * h[0], (j - 0.5) * h[1], (k - 0.5) * h[2])
                box[k, j, i] = all(start[a] <= c[a] <= end[a] for a in range(3))
    assert box.sum() == 6 * 4 * 6
    expect = 0.0
    for step in range(1, 5):
        sim.step(step * cfg["dt"])
        expect = expect + cfg["dt"] * val if adds else 0.0
        C = sim.field("C").reshape(Nz, Ny, Nx)
        assert np.all(C[box] == expect), step
        inner = np.zeros_like(box)
        inner[1:-1, 1:-1, 1:-1] = True
        assert np.all(C[inner & ~box] == 0.0), step
        for f in "uvw":
            assert np.all(sim.field(f) == 0.0), (f, step)
