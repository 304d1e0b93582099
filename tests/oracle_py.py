"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

Test infrastructure: the oracle checks results, it is never the thing measured or
shipped.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

MAX_OBST = 32
NFT = 7


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("value", C.c_double)]


class Desc(C.Structure):
    _fields_ = [
        ("X1", C.c_double), ("X2", C.c_double), ("Y1", C.c_double), ("Y2", C.c_double),
        ("Z1", C.c_double), ("Z2", C.c_double),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("levels", C.c_int),
        ("n_obst", C.c_int), ("obst", (C.c_double * 6) * MAX_OBST),
        ("dom_has", C.c_int * NFT), ("dom", (BC * 6) * NFT),
        ("obst_has", (C.c_int * NFT) * MAX_OBST), ("obc", ((BC * 6) * NFT) * MAX_OBST),
    ]


class MGParams(C.Structure):
    _fields_ = [("n_level", C.c_int), ("n_cycle", C.c_int), ("max_cycle", C.c_int), ("n_relax", C.c_int),
                ("tol_res", C.c_double), ("smoother", C.c_int), ("max_iter", C.c_int),
                ("s_tol_res", C.c_double), ("w", C.c_double)]


class SimParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("dt", C.c_double), ("nu", C.c_double), ("kappa", C.c_double),
                ("beta", C.c_double), ("g", C.c_double), ("diff_type", C.c_int), ("diff_max_iter", C.c_int),
                ("diff_tol", C.c_double), ("diff_w", C.c_double), ("cs", C.c_double), ("prandtl", C.c_double),
                ("has_force", C.c_int), ("force_dir", C.c_int * 3), ("T_ambient", C.c_double),
                ("has_temp_source", C.c_int), ("temp_dir", C.c_int * 3), ("gauss", C.c_double * 8),
                ("tau", C.c_double), ("mg", MGParams), ("dissipation", C.c_int), ("u_lin", C.c_double * 3),
                ("turb_dynamic", C.c_int), ("gamma", C.c_double), ("con_turb", C.c_int), ("schmidt", C.c_double),
                ("con_dir", C.c_int * 3), ("con_fct", C.c_int), ("con_gauss", C.c_double * 8), ("con_tau", C.c_double),
                ("con_cube", C.c_double * 7)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.or_domain_create.restype = C.c_void_p
        L.or_domain_create.argtypes = [C.POINTER(Desc), C.c_char_p, C.c_size_t]
        L.or_domain_free.argtypes = [C.c_void_p]
        L.or_level_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_size_t), dp]
        L.or_list.restype = C.c_size_t
        L.or_list.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_uint64))]
        L.or_obstacle_box.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.or_apply_bc.argtypes = [C.c_void_p, dp, C.c_int, C.c_int]
        L.or_sl_advect.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, C.c_double, C.c_int]
        for fn in (L.or_jacobi_diffuse, L.or_cgs_diffuse):
            fn.restype = C.c_int
            fn.argtypes = [C.c_void_p, dp, dp, dp, C.c_double, dp, C.c_double, C.c_int, C.c_double, C.c_double,
                           C.c_int, dp]
        L.or_explicit_diffuse.argtypes = [C.c_void_p, dp, dp, C.c_double, dp, C.c_double, C.c_int]
        L.or_divergence.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.or_projection.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp, dp]
        L.or_smagorinsky.argtypes = [C.c_void_p, dp, dp, dp, dp, C.c_double, C.c_int]
        L.or_dyn_smagorinsky.argtypes = [C.c_void_p, dp, dp, dp, dp, C.c_double]
        L.or_add_source.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_int]
        L.or_buoyancy.argtypes = [C.c_void_p, dp, dp, dp, C.c_double, C.c_double]
        L.or_gauss_spatial.argtypes = [C.c_void_p, dp, dp]
        L.or_residual_norm.restype = C.c_double
        L.or_residual_norm.argtypes = [C.c_void_p, dp, dp, C.c_int]
        L.or_mg_create.restype = C.c_void_p
        L.or_mg_create.argtypes = [C.c_void_p, C.POINTER(MGParams)]
        L.or_mg_free.argtypes = [C.c_void_p]
        L.or_mg_pressure.restype = C.c_int
        L.or_mg_pressure.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double]
        L.or_sim_create.restype = C.c_void_p
        L.or_sim_create.argtypes = [C.c_void_p, C.POINTER(SimParams), dp, dp, dp, dp, dp]
        L.or_sim_free.argtypes = [C.c_void_p]
        L.or_sim_step.argtypes = [C.c_void_p, C.c_double]
        L.or_sim_field.restype = dp
        L.or_sim_field.argtypes = [C.c_void_p, C.c_char_p]
        L.or_sim_counters.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_desc(cfg):
    d = Desc()
    dom = cfg["domain"]
    for k in ("X1", "X2", "Y1", "Y2", "Z1", "Z2"):
        setattr(d, k, dom[k])
    d.nx, d.ny, d.nz = dom["nx"], dom["ny"], dom["nz"]
    d.levels = cfg["levels"]
    obst = cfg["obstacles"]
    if len(obst) > MAX_OBST:
        raise ValueError("too many obstacles")
    d.n_obst = len(obst)
    for ft in range(NFT):
        for p in range(6):
            d.dom[ft][p].type = -1
    for fields, patches, bc, val in cfg["boundaries"]:
        for ft in fields:
            d.dom_has[ft] = 1
            for p in patches:
                d.dom[ft][p].type = bc
                d.dom[ft][p].value = val
    for i, o in enumerate(obst):
        for a in range(6):
            d.obst[i][a] = o["box"][a]
        for ft in range(NFT):
            for p in range(6):
                d.obc[i][ft][p].type = -1
        for fields, patches, bc, val in o["boundaries"]:
            for ft in fields:
                d.obst_has[i][ft] = 1
                for p in patches:
                    d.obc[i][ft][p].type = bc
                    d.obc[i][ft][p].value = val
    return d


class Domain:
    """Oracle geometry for one configuration (level sizes + index lists)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self._desc = make_desc(cfg)
        err = C.create_string_buffer(256)
        self.h = self.L.or_domain_create(C.byref(self._desc), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())
        self.levels = cfg["levels"]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_domain_free(self.h)
            self.h = None

    def info(self, level=0):
        o = (C.c_size_t * 7)()
        h = (C.c_double * 3)()
        self.L.or_level_info(self.h, level, o, h)
        return dict(N=(o[0], o[1], o[2]), n=(o[3], o[4], o[5]), size=o[6], h=(h[0], h[1], h[2]))

    def size(self, level=0):
        return self.info(level)["size"]

    def shape(self, level=0):
        N = self.info(level)["N"]
        return (N[2], N[1], N[0])

    def list(self, which, level=0, oid=0):
        ptr = C.POINTER(C.c_uint64)()
        n = self.L.or_list(self.h, which, level, oid, C.byref(ptr))
        if n == 0:
            return np.zeros(0, np.uint64)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def obstacle_box(self, level, oid):
        s = (C.c_size_t * 3)()
        e = (C.c_size_t * 3)()
        self.L.or_obstacle_box(self.h, level, oid, s, e)
        return tuple(s), tuple(e)

    def field(self, level=0, fill=0.0):
        return np.full(self.size(level), fill, np.float64)

    # ---- operators ----
    def apply_bc(self, f, ft, level=0):
        self.L.or_apply_bc(self.h, _p(f), ft, level)

    def sl_advect(self, out, inp, u, v, w, dt, ft):
        self.L.or_sl_advect(self.h, _p(out), _p(inp), _p(u), _p(v), _p(w), dt, ft)

    def diffuse(self, kind, out, inp, b, D, ev, dt, max_iter, tol, w, ft):
        if kind == "Explicit":
            self.L.or_explicit_diffuse(self.h, _p(out), _p(inp), D, _p(ev), dt, ft)
            return 1, 0.0
        res = C.c_double()
        fn = self.L.or_jacobi_diffuse if kind == "Jacobi" else self.L.or_cgs_diffuse
        it = fn(self.h, _p(out), _p(inp), _p(b), D, _p(ev), dt, max_iter, tol, w, ft, C.byref(res))
        return it, res.value

    def divergence(self, out, x, y, z):
        self.L.or_divergence(self.h, _p(out), _p(x), _p(y), _p(z))

    def projection(self, ou, ov, ow, iu, iv, iw, ip):
        self.L.or_projection(self.h, _p(ou), _p(ov), _p(ow), _p(iu), _p(iv), _p(iw), _p(ip))

    def smagorinsky(self, ev, u, v, w, cs, ft=6):
        self.L.or_smagorinsky(self.h, _p(ev), _p(u), _p(v), _p(w), cs, ft)

    def dyn_smagorinsky(self, ev, u, v, w, nu):
        self.L.or_dyn_smagorinsky(self.h, _p(ev), _p(u), _p(v), _p(w), nu)

    def add_source(self, out, s, dt, ft):
        self.L.or_add_source(self.h, _p(out), _p(s), dt, ft)

    def buoyancy(self, out, T, Ta, beta, g):
        self.L.or_buoyancy(self.h, _p(out), _p(T), _p(Ta), beta, g)

    def gauss_spatial(self, out, par8):
        a = np.asarray(par8, np.float64)
        self.L.or_gauss_spatial(self.h, _p(out), _p(a))

    def residual_norm(self, x, b, level=0):
        return self.L.or_residual_norm(self.h, _p(x), _p(b), level)


def mg_params(cfg):
    pr = cfg["pressure"]
    sm = pr["smoother"]
    return MGParams(pr["n_level"], pr["n_cycle"], pr["max_cycle"], pr["n_relax"], pr["tol_res"],
                    0 if sm["type"] == "Jacobi" else 1, sm["max_iter"], sm["tol_res"], sm["w"])


class MG:
    def __init__(self, dom, cfg):
        self.dom = dom
        self._p = mg_params(cfg)
        self.h = dom.L.or_mg_create(dom.h, C.byref(self._p))

    def __del__(self):
        if getattr(self, "h", None):
            self.dom.L.or_mg_free(self.h)
            self.h = None

    def pressure(self, p, rhs, t, dt):
        return self.dom.L.or_mg_pressure(self.h, _p(p), _p(rhs), t, dt)


def sim_params(cfg):
    sp = SimParams()
    desc = cfg["solver"]
    kinds = {"NSSolver": 0, "NSTempTurbSolver": 1, "NSTempSolver": 3, "NSTurbSolver": 4, "DiffusionSolver": 5,
             "AdvectionDiffusionSolver": 6, "DiffusionTurbSolver": 7, "AdvectionSolver": 8, "NSTempConSolver": 9,
             "NSTempTurbConSolver": 10}
    if desc not in kinds:
        raise ValueError("the oracle has no driver for " + desc)
    sp.solver = kinds[desc]
    if desc == "AdvectionSolver":
        v = cfg["ic"]["values"]
        for a, k in enumerate(("u_lin", "v_lin", "w_lin")):
            sp.u_lin[a] = float(v[k])
    sp.dt, sp.nu, sp.kappa, sp.beta, sp.g = cfg["dt"], cfg["nu"], cfg["kappa"], cfg["beta"], cfg["g"]
    dif = cfg.get("diffusion", {"type": "Jacobi"})
    sp.diff_type = {"Jacobi": 0, "ColoredGaussSeidel": 1, "Explicit": 2}[dif["type"]]
    sp.diff_max_iter, sp.diff_tol, sp.diff_w = dif.get("max_iter", 0), dif.get("tol_res", 0.0), dif.get("w", 0.0)
    if "turbulence" in cfg:
        sp.cs = cfg["turbulence"]["cs"]
        sp.turb_dynamic = 1 if cfg["turbulence"]["type"] == "DynamicSmagorinsky" else 0
    src = cfg.get("source", {"force_fct": "Zero", "dir": []})
    sp.has_force = 1 if src["force_fct"] == "Buoyancy" else 0
    for a in src["dir"]:
        sp.force_dir[a] = 1
    if sp.has_force:
        sp.T_ambient = src.get("ambient", 0.0)
    t = cfg.get("temperature")
    if t is not None:
        sp.prandtl = t["prandtl"] if t["turbulence"] else 0.0
        ts = t["source"]
        sp.has_temp_source = 1 if ts["temp_fct"] == "Gauss" else 0
        for a in ts["dir"]:
            sp.temp_dir[a] = 1
        if sp.has_temp_source:
            for i, x in enumerate(ts["gauss"]):
                sp.gauss[i] = x
            sp.tau = ts["tau"]
        sp.dissipation = 1 if ts.get("dissipation") else 0
    if "pressure" in cfg:
        sp.mg = mg_params(cfg)
    c = cfg.get("concentration")
    if c is not None:
        sp.gamma = cfg["gamma"]
        sp.con_turb = 1 if c["turbulence"] else 0
        sp.schmidt = c.get("schmidt", 0.0)
        for a in c["source"]["dir"]:
            sp.con_dir[a] = 1
        if c["source"]["con_fct"] == "Gauss":
            sp.con_fct = 1
            for i, x in enumerate(c["source"]["gauss"]):
                sp.con_gauss[i] = x
            sp.con_tau = c["source"]["tau"]
        else:
            sp.con_fct = 2
            for i, x in enumerate(c["source"]["cube"]):
                sp.con_cube[i] = x
    return sp


class Sim:
    """Oracle NSSolver / NSTempTurbSolver with TimeIntegration's per-step tail."""

    def __init__(self, dom, cfg, u=None, v=None, w=None, T=None, p=None):
        self.dom = dom
        self._p = sim_params(cfg)
        self.h = dom.L.or_sim_create(dom.h, C.byref(self._p), _p(u), _p(v), _p(w), _p(T), _p(p))
        self.n = dom.size(0)

    def __del__(self):
        if getattr(self, "h", None):
            self.dom.L.or_sim_free(self.h)
            self.h = None

    def step(self, t):
        self.dom.L.or_sim_step(self.h, t)

    def field(self, name):
        ptr = self.dom.L.or_sim_field(self.h, name.encode())
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).copy()

    def counters(self):
        c = (C.c_int * 5)()
        self.dom.L.or_sim_counters(self.h, c)
        return list(c)
