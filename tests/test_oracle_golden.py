"""Pins the CPU oracle (oracle/*.c) to the real reference.

The fixtures under tests/golden/ were written by tests/golden/make_golden.py
from the reference's own sources (oracle/_ref/ref_harness): seeded random
inputs through single reference calls, index lists, and whole solver steps.
Every oracle function is compared bit for bit (np.array_equal on float64):

  geometry lists (a16)          test_lists, test_obstacle_faces
  boundary fill (a14, a15)      test_bc (all field types, all levels, every BC type)
  SL advection (a1)             test_advect
  Jacobi / turbulent (a2, a3)   test_diffusion[jacobi, jacobi_turb]
  CGS / turbulent (a5)          test_diffusion[cgs, cgs_turb]
  explicit diffusion (8(f) 3)   test_explicit_diffusion[explicit, explicit_turb]
  V-cycle MG (a4, a6-a11)       test_vcycle (Jacobi and CGS smoothers, Nt == 1 and Nt > 1)
  divergence, projection (a12)  test_divergence, test_projection
  ConstSmagorinsky (a13)        test_smagorinsky
  DynamicSmagorinsky (8(f) 3)   test_dyn_smagorinsky, test_run_drivers[run_nsturb_dyn_tg]
  sources (a17)                 test_sources (add_source scalar/vector, buoyancy, Gauss)
  solver drivers (a18, a19)     test_run (NSSolver Taylor-Green, NSTempTurb plume Jacobi/CGS/explicit),
                                test_run_drivers (the other solvers and the reference's initial conditions)

The reference's own fixture files (tests/*/*_ref.dat of the reference) list
the inner cells before the ghost cells; in that order the current reference
reproduces 11 of its 16 inputs to their six digits
(tests/golden/make_reference_inputs.py).  The oracle's V-cycle is pinned to
tests/pressure/p_ref.dat in test_oracle_closed_forms.py; the GPU path runs
all of those inputs in tests/test_gpu_reference_inputs.py.
"""
import math
import os

import numpy as np
import pytest

import artss_xml
import oracle_py as O
from cases import taylor_green_ic, uniform_ic

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    cfg = artss_xml.parse(bytes(z["xml"]).decode())
    return z, cfg


@pytest.fixture(scope="module")
def ops_jacobi():
    z, cfg = load("ops_jacobi")
    return z, cfg, O.Domain(cfg)


@pytest.fixture(scope="module")
def ops_cgs():
    z, cfg = load("ops_cgs")
    return z, cfg, O.Domain(cfg)


@pytest.fixture(scope="module")
def ops_explicit():
    z, cfg = load("ops_explicit")
    return z, cfg, O.Domain(cfg)


@pytest.fixture(scope="module")
def geo_adjacent():
    z, cfg = load("geo_adjacent")
    return z, cfg, O.Domain(cfg)


def same(a, b):
    assert a.shape == b.shape
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError("mismatch: %d cells, max |d| %.3g at %d" % ((d != 0).sum(), d.max(), d.argmax()))


@pytest.mark.parametrize("case", ["ops_jacobi", "geo_adjacent"])
def test_lists(case, request):
    z, cfg, dom = request.getfixturevalue(case)
    for level in range(cfg["levels"] + 1):
        same(dom.list(0, level), z["list_inner_%d" % level])
        same(dom.list(1, level), z["list_domain_%d" % level])
        same(dom.list(2, level), z["list_obstacle_%d" % level])


@pytest.mark.parametrize("case", ["ops_jacobi", "geo_adjacent"])
def test_obstacle_faces(case, request):
    z, cfg, dom = request.getfixturevalue(case)
    n = 0
    for level in range(cfg["levels"] + 1):
        for oid in range(len(cfg["obstacles"])):
            for p in range(6):
                same(dom.list(10 + p, level, oid), z["list_oface_%d_%d_%d" % (level, oid, p)])
                n += 1
    assert n > 0


def test_circular_constraints_removed_cells(geo_adjacent):
    """The adjacent obstacles really lose face cells (Obstacle.cpp:562-866)."""
    z, cfg, dom = geo_adjacent
    full = dom.obstacle_box(0, 0)
    sy = full[1][1] - full[0][1] + 1
    sz = full[1][2] - full[0][2] + 1
    assert len(dom.list(10 + 1, 0, 0)) < sy * sz  # A's RIGHT face lost the part facing B


@pytest.mark.parametrize("case", ["ops_jacobi", "geo_adjacent"])
def test_bc(case, request):
    z, cfg, dom = request.getfixturevalue(case)
    checked = 0
    for ft in range(7):
        for level in range(cfg["levels"] + 1):
            key = "%d_%d" % (ft, level)
            f = z["bc/bc_in_" + key].copy()
            dom.apply_bc(f, ft, level)
            same(f, z["bc/bc_out_" + key])
            checked += int(not np.array_equal(z["bc/bc_in_" + key], z["bc/bc_out_" + key]))
    assert checked >= 5


def test_advect(ops_jacobi):
    z, cfg, dom = ops_jacobi
    for ft in (1, 5):
        out = z["advect/adv_out0_%d" % ft].copy()
        dom.sl_advect(out, z["advect/adv_in_%d" % ft], z["advect/adv_u"], z["advect/adv_v"], z["advect/adv_w"],
                      cfg["dt"], ft)
        same(out, z["advect/adv_out_%d" % ft])


@pytest.mark.parametrize("case,op", [("ops_jacobi", "jacobi"), ("ops_jacobi", "jacobi_turb"),
                                     ("ops_cgs", "cgs"), ("ops_cgs", "cgs_turb")])
def test_diffusion(case, op, request):
    z, cfg, dom = request.getfixturevalue(case)
    d = cfg["diffusion"]
    out = z[op + "/dif_out0"].copy()
    inp = z[op + "/dif_in0"].copy()
    ev = z[op + "/dif_ev"] if op.endswith("turb") else None
    it, res = dom.diffuse(d["type"], out, inp, z[op + "/dif_b"], cfg["nu"], ev, cfg["dt"], d["max_iter"],
                          d["tol_res"], d["w"], 1)
    assert it >= 2
    same(out, z[op + "/dif_out"])
    same(inp, z[op + "/dif_in"])


@pytest.mark.parametrize("op", ["explicit", "explicit_turb"])
def test_explicit_diffusion(ops_explicit, op):
    z, cfg, dom = ops_explicit
    out = z[op + "/dif_out0"].copy()
    inp = z[op + "/dif_in0"].copy()
    ev = z[op + "/dif_ev"] if op.endswith("turb") else None
    dom.diffuse("Explicit", out, inp, z[op + "/dif_b"], cfg["nu"], ev, cfg["dt"], 0, 0.0, 0.0, 1)
    same(out, z[op + "/dif_out"])
    same(inp, z[op + "/dif_in"])  # the explicit step leaves its input alone
    assert not np.array_equal(out, z[op + "/dif_out0"])


@pytest.mark.parametrize("case", ["ops_jacobi", "ops_cgs"])
@pytest.mark.parametrize("op", ["vcycle1", "vcycle2"])
def test_vcycle(case, op, request):
    z, cfg, dom = request.getfixturevalue(case)
    mg = O.MG(dom, cfg)
    p = z[op + "/mg_p0"].copy()
    t = cfg["dt"] if op == "vcycle1" else 2 * cfg["dt"]
    cycles = mg.pressure(p, z[op + "/mg_rhs"], t, cfg["dt"])
    assert cycles >= cfg["pressure"]["n_cycle"]
    same(p, z[op + "/mg_p"])


def test_divergence(ops_jacobi):
    z, cfg, dom = ops_jacobi
    out = z["divergence/div_out0"].copy()
    dom.divergence(out, z["divergence/div_x"], z["divergence/div_y"], z["divergence/div_z"])
    same(out, z["divergence/div_out"])


def test_projection(ops_jacobi):
    z, cfg, dom = ops_jacobi
    o = [z["projection/prj_o%s0" % c].copy() for c in "uvw"]
    dom.projection(*o, z["projection/prj_iu"], z["projection/prj_iv"], z["projection/prj_iw"],
                   z["projection/prj_ip"])
    for c, a in zip("uvw", o):
        same(a, z["projection/prj_o" + c])


def test_smagorinsky(ops_jacobi):
    z, cfg, dom = ops_jacobi
    ev = z["smagorinsky/smg_ev0"].copy()
    dom.smagorinsky(ev, z["smagorinsky/smg_u"], z["smagorinsky/smg_v"], z["smagorinsky/smg_w"],
                    cfg["turbulence"]["cs"])
    same(ev, z["smagorinsky/smg_ev"])


def test_sources(ops_jacobi):
    z, cfg, dom = ops_jacobi
    out = z["sources/src_out0"].copy()
    dom.add_source(out, z["sources/src_s"], cfg["dt"], 5)
    same(out, z["sources/src_out"])
    f = {c: z["sources/srcv_%s0" % c].copy() for c in "xyz"}
    for a, c in enumerate("xyz"):
        if a in cfg["source"]["dir"]:
            dom.add_source(f[c], z["sources/srcv_s" + c], cfg["dt"], 1 + a)
        same(f[c], z["sources/srcv_" + c])
    bo = z["sources/buo_out0"].copy()
    dom.buoyancy(bo, z["sources/buo_T"], z["sources/buo_Ta"], cfg["beta"], cfg["g"])
    same(bo, z["sources/buo_out"])
    st = dom.field()
    dom.gauss_spatial(st, cfg["temperature"]["source"]["gauss"])
    st *= math.tanh(0.37 / cfg["temperature"]["source"]["tau"])  # libm, as GaussFunction.cpp:122
    same(st, z["sources/gauss_st"])


def test_dyn_smagorinsky():
    """two calls on one model object (its work fields persist between calls)"""
    z, cfg = load("ops_dyn_smagorinsky")
    dom = O.Domain(cfg)
    u, v, w = (z["dyn_smagorinsky/dyn_%s" % c].copy() for c in "uvw")
    ev = dom.field()
    for call in range(2):
        dom.dyn_smagorinsky(ev, u, v, w, cfg["nu"])
        same(ev, z["dyn_smagorinsky/dyn_ev%d" % call])
        u *= 0.5
        v *= 0.5
        w *= 0.5


@pytest.mark.parametrize("case", ["run_nstemp_dissipation", "run_nsturb_tg", "run_nsturb_dyn_tg", "run_diffusion_tg",
                                  "run_diffusion_turb_tg", "run_advection_diffusion_tg", "run_advection_gauss",
                                  "run_diffusion_expsinusprod", "run_diffusion_hat", "run_burgers_expsinussum",
                                  "run_ns_mcdermott", "run_ns_vortex", "run_ns_drift", "run_nstemp_layers"])
def test_run_drivers(case):
    """the other solver drivers, started from the reference's own initial
    fields (the golden run's step 0; the product computes those itself and
    is checked on them in test_gpu_parity.py)"""
    z, cfg = load(case)
    dom = O.Domain(cfg)
    ic = {f: z["run/%s_0" % f] for f in "uvwTp"}
    sim = O.Sim(dom, cfg, ic["u"], ic["v"], ic["w"], ic["T"], ic["p"])
    fields = ["u", "v", "w", "p", "T"]
    for f in fields:
        same(sim.field(f), z["run/%s_0" % f])
    nsteps = int(round(cfg["t_end"] / cfg["dt"]))
    for step in range(1, nsteps + 1):
        sim.step(step * cfg["dt"])
        if step in (1, 2):
            for f in fields + ["rhs"]:
                same(sim.field(f), z["run/%s_%d" % (f, step)])
    for f in fields:
        same(sim.field(f), z["run/%s_-1" % f])


@pytest.mark.parametrize("case", ["run_ns_tg", "run_plume_jacobi", "run_plume_cgs", "run_plume_explicit"])
def test_run(case):
    z, cfg = load(case)
    dom = O.Domain(cfg)
    if cfg["solver"] == "NSSolver":
        u, v, w = taylor_green_ic(dom)
        sim = O.Sim(dom, cfg, u, v, w, None)
    else:
        sim = O.Sim(dom, cfg, None, None, None, uniform_ic(dom, float(cfg["ic"]["values"]["val"])))
    nsteps = int(round(cfg["t_end"] / cfg["dt"]))
    fields = ["u", "v", "w", "p", "T"]
    for f in fields:
        same(sim.field(f), z["run/%s_0" % f])
    for step in range(1, nsteps + 1):
        sim.step(step * cfg["dt"])
        if step in (1, 2):
            for f in fields + ["rhs"]:
                same(sim.field(f), z["run/%s_%d" % (f, step)])
    for f in fields:
        same(sim.field(f), z["run/%s_-1" % f])
