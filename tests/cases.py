"""XML inputs for the parity tests, in the reference's own input format.

Each case is written twice when it needs an initial condition the reference
lacks (Taylor-Green velocity): the product reads usr_fct="TaylorGreen", the
reference harness reads usr_fct="Zero" and sets the same field on the inner
list itself (oracle/ref_harness.cpp, `taylor_green`).
"""
import math

import numpy as np

_JACOBI = '<diffusion type="Jacobi" field="{f}"><max_iter> {mi} </max_iter><tol_res> {tol} </tol_res><w> {w} </w></diffusion>'
_CGS = ('<diffusion type="ColoredGaussSeidel" field="{f}"><max_iter> {mi} </max_iter><tol_res> {tol} </tol_res>'
        '<w> {w} </w></diffusion>')


def diffusion_xml(kind, field, max_iter=100, tol=1e-7, w=1.0):
    if kind == "Explicit":
        return '<diffusion type="Explicit" field="{}"></diffusion>'.format(field)
    t = _JACOBI if kind == "Jacobi" else _CGS
    return t.format(f=field, mi=max_iter, tol=repr(tol), w=repr(w))


def pressure_xml(levels, n_cycle=2, max_cycle=100, tol=1e-8, n_relax=4, smoother="Jacobi", s_max_iter=100,
                 s_tol=1e-8, s_w=0.6666666667):
    return ('<pressure type="VCycleMG" field="p"><n_level> {} </n_level><n_cycle> {} </n_cycle>'
            '<max_cycle> {} </max_cycle><tol_res> {} </tol_res><n_relax> {} </n_relax>{}</pressure>').format(
        levels, n_cycle, max_cycle, repr(tol), n_relax, diffusion_xml(smoother, "p", s_max_iter, s_tol, s_w))


def domain_xml(X, Y, Z, n):
    return ('<domain_parameters enable_computational_domain="No"><X1> {} </X1><X2> {} </X2><Y1> {} </Y1>'
            '<Y2> {} </Y2><Z1> {} </Z1><Z2> {} </Z2><nx> {} </nx><ny> {} </ny><nz> {} </nz></domain_parameters>'
            ).format(repr(X[0]), repr(X[1]), repr(Y[0]), repr(Y[1]), repr(Z[0]), repr(Z[1]), n[0], n[1], n[2])


def obstacles_xml(boxes, vel_bc="dirichlet", scal_bc="neumann", vel_value=0.0, scal_value=0.0):
    if not boxes:
        return '<obstacles enabled="No"/>'
    s = '<obstacles enabled="Yes">'
    for i, b in enumerate(boxes):
        s += ('<obstacle name="o{}"><geometry ox1="{}" ox2="{}" oy1="{}" oy2="{}" oz1="{}" oz2="{}"/>'
              '<boundary field="u,v,w" patch="front,back,left,right,top,bottom" type="{}" value="{}"/>'
              '<boundary field="p,T" patch="front,back,left,right,top,bottom" type="{}" value="{}"/>'
              '</obstacle>').format(i, *[repr(x) for x in b], vel_bc, repr(vel_value), scal_bc, repr(scal_value))
    return s + "</obstacles>"


def _wrap(physical, solver, domain, boundaries, obstacles, ic):
    return ('<?xml version="1.0" encoding="UTF-8" ?>\n<ARTSS>\n'
            '<physical_parameters>{}</physical_parameters>\n'
            '<solver description="{}">{}<solution available="No"></solution></solver>\n{}\n'
            '<adaption dynamic="No" data_extraction="No"> </adaption>\n<data_assimilation enabled="No" />\n'
            '<boundaries>{}</boundaries>\n{}\n<surfaces enabled="No"/>\n{}\n'
            '<visualisation save_vtk="No" save_csv="No"></visualisation>\n'
            '<logging file="artss.log" level="off"></logging>\n</ARTSS>\n').format(
        physical, solver[0], solver[1], domain, boundaries, obstacles, ic)


def ns_taylor_green(n=64, levels=4, steps=100, dt=1e-3, nu=0.01, reference=False, diffusion="Jacobi",
                    smoother="Jacobi", n_cycle=2):
    """NSSolver, unit cube, periodic u,v,w,p, Taylor-Green initial velocity."""
    phys = '<t_end> {} </t_end><dt> {} </dt><nu> {} </nu>'.format(repr(steps * dt), repr(dt), repr(nu))
    solver = ('<advection type="SemiLagrangian" field="u,v,w"></advection>' +
              diffusion_xml(diffusion, "u,v,w", 100, 1e-7, 1.0) +
              '<source type="ExplicitEuler" force_fct="Zero" dir="x"></source>' +
              pressure_xml(levels, n_cycle=n_cycle, smoother=smoother))
    bnd = '<boundary field="u,v,w,p" patch="front,back,left,right,bottom,top" type="periodic" />'
    ic = '<initial_conditions usr_fct="{}" random="No"></initial_conditions>'.format(
        "Zero" if reference else "TaylorGreen")
    return _wrap(phys, ("NSSolver", solver), domain_xml((0.0, 1.0), (0.0, 1.0), (0.0, 1.0), (n, n, n)), bnd,
                 obstacles_xml([]), ic)


def plume(n=(128, 128, 128), X=(0.0, 1.0), levels=5, steps=1000, dt=5e-3, diffusion="ColoredGaussSeidel",
          smoother="Jacobi", obstacles=(), hrr=None, source_x=None, n_cycle=2, T_a=293.15, nu=3.1e-5,
          kappa=4.2e-5, sigma=0.1, tau=5.0, diff_tol=1e-7, open_top=False):
    """NSTempTurbSolver buoyant plume: Boussinesq buoyancy along y, Gaussian
    volumetric heat source at the floor centre, ConstSmagorinsky Cs=0.2.

    open_top=False: closed box (u,v,w Dirichlet 0, p Neumann 0 everywhere).
    open_top=True: the top patch is an outflow (u,v,w and T Neumann 0,
    p Dirichlet 0), which makes the pressure Poisson problem non-singular."""
    if source_x is None:
        source_x = 0.5 * (X[0] + X[1])
    if hrr is None:
        hrr = 50.0 * gauss_volume(n, X, (source_x, 0.0, 0.5), sigma)
    phys = ('<t_end> {} </t_end><dt> {} </dt><nu> {} </nu><beta> 3.34e-3 </beta><g> -9.81 </g>'
            '<rhoa> 1 </rhoa><kappa> {} </kappa>').format(repr(steps * dt), repr(dt), repr(nu), repr(kappa))
    solver = ('<advection type="SemiLagrangian" field="u,v,w"></advection>' +
              diffusion_xml(diffusion, "u,v,w", 100, diff_tol, 1.0) +
              '<turbulence type="ConstSmagorinsky"><Cs> 0.2 </Cs></turbulence>'
              '<source type="ExplicitEuler" force_fct="Buoyancy" dir="y" use_init_values="No">'
              '<ambient_temperature_value> {} </ambient_temperature_value></source>'.format(repr(T_a)) +
              pressure_xml(levels, n_cycle=n_cycle, smoother=smoother) +
              '<temperature><advection type="SemiLagrangian" field="T"></advection>' +
              diffusion_xml(diffusion, "T", 100, diff_tol, 1.0) +
              '<turbulence include="Yes"><Pr_T> 0.5 </Pr_T></turbulence>'
              '<source type="ExplicitEuler" dir="y" temp_fct="Gauss" dissipation="No" random="No">'
              '<HRR> {} </HRR><cp> 1. </cp><x0> {} </x0><y0> 0. </y0><z0> 0.5 </z0><sigma_x> {} </sigma_x>'
              '<sigma_y> {} </sigma_y><sigma_z> {} </sigma_z><tau> {} </tau></source></temperature>'.format(
                  repr(hrr), repr(source_x), repr(sigma), repr(sigma), repr(sigma), repr(tau)))
    if open_top:
        bnd = ('<boundary field="u,v,w" patch="front,back,bottom,left,right" type="dirichlet" value="0.0" />'
               '<boundary field="u,v,w" patch="top" type="neumann" value="0.0" />'
               '<boundary field="p" patch="front,back,bottom,left,right" type="neumann" value="0.0" />'
               '<boundary field="p" patch="top" type="dirichlet" value="0.0" />'
               '<boundary field="T" patch="front,back,left,right" type="dirichlet" value="{}" />'
               '<boundary field="T" patch="bottom,top" type="neumann" value="0.0" />').format(repr(T_a))
    else:
        bnd = ('<boundary field="u,v,w" patch="front,back,bottom,top,left,right" type="dirichlet" value="0.0" />'
               '<boundary field="p" patch="front,back,bottom,top,left,right" type="neumann" value="0.0" />'
               '<boundary field="T" patch="front,back,top,left,right" type="dirichlet" value="{}" />'
               '<boundary field="T" patch="bottom" type="neumann" value="0.0" />').format(repr(T_a))
    ic = '<initial_conditions usr_fct="Uniform" random="No"><val> {} </val></initial_conditions>'.format(repr(T_a))
    return _wrap(phys, ("NSTempTurbSolver", solver), domain_xml(X, (0.0, 1.0), (0.0, 1.0), n), bnd,
                 obstacles_xml(list(obstacles)), ic)


def gauss_volume(n, X, pos, sigma):
    """Discrete integral of the Gaussian over the inner cells (GaussFunction.cpp:75-90, no obstacles)."""
    dx = (X[1] - X[0]) / n[0]
    dy, dz = 1.0 / n[1], 1.0 / n[2]
    x = X[0] + (np.arange(1, n[0] + 1) - 0.5) * dx
    y = (np.arange(1, n[1] + 1) - 0.5) * dy
    z = (np.arange(1, n[2] + 1) - 0.5) * dz
    r = 1.0 / (2 * sigma * sigma)
    gx = np.exp(-((x - pos[0]) ** 2) * r).sum()
    gy = np.exp(-((y - pos[1]) ** 2) * r).sum()
    gz = np.exp(-((z - pos[2]) ** 2) * r).sum()
    return float(round(gx * gy * gz * dx * dy * dz, 12))


def taylor_green_ic(dom):
    """u = sin2πx cos2πy cos2πz, v = -cos2πx sin2πy cos2πz, w = 0 on the inner list,
    evaluated with libm like the product and the harness."""
    info = dom.info(0)
    Nx, Ny, Nz = info["N"]
    hx, hy, hz = info["h"]
    c = dom.cfg["domain"]
    u = np.zeros(info["size"])
    v = np.zeros(info["size"])
    w = np.zeros(info["size"])
    sx = [(math.sin(2. * math.pi * (c["X1"] + (i - 0.5) * hx)), math.cos(2. * math.pi * (c["X1"] + (i - 0.5) * hx)))
          for i in range(Nx)]
    sy = [(math.sin(2. * math.pi * (c["Y1"] + (j - 0.5) * hy)), math.cos(2. * math.pi * (c["Y1"] + (j - 0.5) * hy)))
          for j in range(Ny)]
    sz = [(math.sin(2. * math.pi * (c["Z1"] + (k - 0.5) * hz)), math.cos(2. * math.pi * (c["Z1"] + (k - 0.5) * hz)))
          for k in range(Nz)]
    for idx in dom.list(0, 0):
        idx = int(idx)
        k, rem = divmod(idx, Nx * Ny)
        j, i = divmod(rem, Nx)
        u[idx] = sx[i][0] * sy[j][1] * sz[k][1]
        v[idx] = -sx[i][1] * sy[j][0] * sz[k][1]
    return u, v, w


def uniform_ic(dom, value):
    T = np.zeros(dom.size(0))
    T[dom.list(0, 0).astype(np.int64)] = value
    return T


def concentration_plume(solver="NSTempTurbConSolver", con_fct="Gauss", turbulent=True, gamma=1e-4, n=(16, 16, 16),
                        levels=2, steps=5, dt=0.02, obstacles=(), diffusion="Jacobi", dissipation=False, **kw):
    """The plume with a transported concentration (NSTempConSolver or
    NSTempTurbConSolver): a Gauss source at the floor or a Cube of constant
    value; concentration Dirichlet 0 on the left and right, Neumann 0
    elsewhere."""
    x = plume(n=n, levels=levels, steps=steps, dt=dt, obstacles=obstacles, diffusion=diffusion, **kw)
    x = x.replace('<solver description="NSTempTurbSolver">', '<solver description="%s">' % solver)
    x = x.replace('</physical_parameters>', '<gamma> {} </gamma></physical_parameters>'.format(repr(gamma)))
    if dissipation:
        x = x.replace('dissipation="No"', 'dissipation="Yes"')
    if con_fct == "Gauss":
        src = ('<source type="ExplicitEuler" dir="y" random="No" con_fct="Gauss"><HRR> 2.5 </HRR><cp> 1. </cp>'
               '<x0> 0.7 </x0><y0> 0. </y0><z0> 0.3 </z0><sigma_x> 0.15 </sigma_x><sigma_y> 0.1 </sigma_y>'
               '<sigma_z> 0.2 </sigma_z><tau> 2.0 </tau></source>')
    else:
        src = ('<source type="ExplicitEuler" dir="y" random="No" con_fct="Cube"><x_start> 0.25 </x_start>'
               '<y_start> 0.0 </y_start><z_start> 0.3 </z_start><x_end> 0.6 </x_end><y_end> 0.25 </y_end>'
               '<z_end> 0.7 </z_end><val> 3.5 </val></source>')
    turb = '<turbulence include="Yes"><Sc_T> 0.7 </Sc_T></turbulence>' if turbulent else \
        '<turbulence include="No"></turbulence>'
    con = ('<concentration><advection type="SemiLagrangian" field="rho"></advection>' +
           diffusion_xml(diffusion, "rho", 100, 1e-7, 1.0) + (turb if "Turb" in solver else "") + src +
           '</concentration>')
    x = x.replace('<solution available="No"></solution></solver>', con + '<solution available="No"></solution></solver>')
    x = x.replace('</boundaries>', '<boundary field="rho" patch="left,right" type="dirichlet" value="0.0" />'
                  '<boundary field="rho" patch="front,back,bottom,top" type="neumann" value="0.0" /></boundaries>')
    return x
