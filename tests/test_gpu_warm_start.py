"""Warm start of the residual-target GCR pressure iteration.

Every start (the previous p, linear extrapolation in time, extrapolation plus
the U-Net's correction) must reach the same residual target and give the same
fields to within what the target allows.  The recorded training sample is
checked against the fields it is defined on, recomputed here in numpy, and a
network trained on the first steps of a plume must start later steps of the
same plume closer to their solution, with fewer GCR iterations, than the
reference's start from the previous p.
"""
import numpy as np
import pytest

import artss_amd
import cases

pytestmark = pytest.mark.gpu

OBST = [(0.375, 0.625, 0.0, 0.375, 0.625, 0.8125)]
N = (64, 32, 32)
TOL = 1e-9


def _xml(steps):
    return cases.plume(n=N, X=(0.0, 2.0), levels=4, steps=steps, dt=0.01, obstacles=OBST, hrr=3.0,
                       source_x=1.0, open_top=True)


def _solver(xml):
    s = artss_amd.Solver(xml)
    s.set_mg_schedule(2, 4096)
    s.set_pressure_mode(True, TOL, 400)
    s.set_pressure_method(2)
    return s


def _inner_box(dom, a):
    """(Z, Y, X) view of the inner cells of a full level-0 field"""
    info = dom.info(0)
    Nx, Ny, Nz = info["N"]
    return a.reshape(Nz, Ny, Nx)[1:-1, 1:-1, 1:-1]


def test_recorded_sample_matches_fields():
    s = _solver(_xml(5))
    dom = s.domain
    info = dom.info(0)
    Nx, Ny, Nz = info["N"]
    hx, hy, hz = info["h"]
    box = (Nz - 2, Ny - 2, Nx - 2)
    x_buf = artss_amd.Buffer(dom, np.zeros(box + (2,), np.float32))
    t_buf = artss_amd.Buffer(dom, np.zeros(box + (1,), np.float32))
    s.set_warm_start(0, record=True)
    s.step(0.01)
    assert not s.warm_sample(x_buf, t_buf)  # no previous solution at the first solve
    # the plume starts from rest: p stays zero until the buoyancy force has
    # moved the fluid, so the sample is taken a few steps in
    s.step(0.02)
    s.step(0.03)
    p1 = s.field("p")
    s.step(0.04)
    p2 = s.field("p")
    s.step(0.05)
    p3, rhs3 = s.field("p"), s.field("rhs")
    assert s.warm_sample(x_buf, t_buf)
    x, t = x_buf.download(), t_buf.download()[..., 0]

    fluid = np.zeros(Nx * Ny * Nz, bool)
    fluid[dom.list(0)] = True
    fluid = _inner_box(dom, fluid)
    assert fluid.sum() == info["inner_cells"] and not fluid.all()
    d = p2 - p1
    srms = np.sqrt((_inner_box(dom, d)[fluid] ** 2).sum() / fluid.sum())
    assert srms > 0
    # laplacian of the extrapolated start 2 p2 - p1 (both with their fill) on the inner box
    P = (2 * p2 - p1).reshape(Nz, Ny, Nx)
    lap = ((P[1:-1, 1:-1, 2:] - 2 * P[1:-1, 1:-1, 1:-1] + P[1:-1, 1:-1, :-2]) / hx ** 2 +
           (P[1:-1, 2:, 1:-1] - 2 * P[1:-1, 1:-1, 1:-1] + P[1:-1, :-2, 1:-1]) / hy ** 2 +
           (P[2:, 1:-1, 1:-1] - 2 * P[1:-1, 1:-1, 1:-1] + P[:-2, 1:-1, 1:-1]) / hz ** 2)
    c = 1.0 / (2 / hx ** 2 + 2 / hy ** 2 + 2 / hz ** 2)
    r_exp = np.where(fluid, c * (_inner_box(dom, rhs3) - lap) / srms, 0.0)
    d_exp = np.where(fluid, _inner_box(dom, d) / srms, 0.0)
    t_exp = np.where(fluid, _inner_box(dom, (p3 - p2) - d) / srms, 0.0)
    for got, exp, name in ((x[..., 0], r_exp, "residual"), (x[..., 1], d_exp, "increment"), (t, t_exp, "target")):
        scale = max(np.abs(exp).max(), 1e-30)
        err = np.abs(got - exp).max() / scale
        print(name, "max rel err", err, "scale", scale)
        assert err < 1e-5, name  # FP32 storage of FP64 values
    assert np.all(x[~fluid] == 0) and np.all(t[~fluid] == 0)


def _run(s, steps, first, modes):
    out = []
    for k in range(first, first + steps):
        s.step(k * 0.01)
        out.append(s.counters())
    return out


@pytest.mark.parametrize("mode", [1, 2, 3], ids=["extrapolate", "learned_untrained", "projection"])
def test_every_start_meets_the_target(mode):
    xml = _xml(6)
    ref = _solver(xml)
    c_ref = _run(ref, 6, 1, None)
    s = _solver(xml)
    net = artss_amd.UNet(s.domain, seed=3) if mode == 2 else None
    s.set_warm_start(mode, net)
    c = _run(s, 6, 1, None)
    print("previous p:", [q["vcycles"] for q in c_ref], "mode", mode, [q["vcycles"] for q in c])
    for q in c:
        assert 0 <= q["residual"] <= TOL
    for f in ("u", "v", "w", "p", "T"):
        a, b = s.field(f), ref.field(f)
        err = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
        assert err < 1e-5, (f, err)


def test_projection_start_needs_fewer_iterations():
    """The projection start is the combination of the last increments of the
    solution closest to the solution in the energy norm; the extrapolation
    is one of the combinations, so the projection starts at most as far from
    the solution, and on a plume that accelerates it needs fewer GCR
    iterations to the same target, with the same fields to within what the
    target allows."""
    xml = _xml(12)
    runs = {}
    for mode in (1, 3):
        s = _solver(xml)
        s.set_warm_start(mode)
        res, its = [], []
        for k in range(1, 13):
            s.step(0.01 * k)
            c = s.counters()
            res.append(c["initial_residual"])
            its.append(c["vcycles"])
            assert 0 <= c["residual"] <= TOL
        runs[mode] = (res, its, {f: s.field(f) for f in ("u", "v", "w", "p", "T")})
        s.close()
    (r1, i1, f1), (r3, i3, f3) = runs[1], runs[3]
    print("initial residuals, extrapolation", ["%.2e" % x for x in r1])
    print("initial residuals, projection   ", ["%.2e" % x for x in r3])
    print("iterations", i1, i3)
    assert sum(i3) < sum(i1)
    for f in f1:
        err = float(np.abs(f3[f] - f1[f]).max() / max(np.abs(f1[f]).max(), 1e-30))
        assert err < 1e-5, (f, err)


def test_projection_depth_is_checked():
    s = _solver(_xml(2))
    for depth in (0, 9):
        with pytest.raises(artss_amd.ArtssError):
            s.set_warm_start(3, depth=depth)
    s.set_warm_start(3, depth=8)
    s.step(0.01)
    s.close()


def test_learned_warm_start_reduces_iterations():
    """Train on steps 2..K of a plume from rest, then compare the starts on
    steps K+1..K+M, all three solvers sharing the trajectory up to step K."""
    K, M = 24, 8
    xml = _xml(K + M)
    solvers = [_solver(xml) for _ in range(3)]  # previous p, extrapolation, learned
    dom = solvers[2].domain
    info = dom.info(0)
    Nx, Ny, Nz = info["N"]
    box = (Nz - 2, Ny - 2, Nx - 2)
    solvers[1].set_warm_start(0, record=True)  # keeps p_prev for the switch
    solvers[2].set_warm_start(0, record=True)
    xs, ts = [], []
    for k in range(1, K + 1):
        for s in solvers:
            s.step(k * 0.01)
        x_buf = artss_amd.Buffer(dom, np.zeros(box + (2,), np.float32))
        t_buf = artss_amd.Buffer(dom, np.zeros(box + (1,), np.float32))
        if solvers[2].warm_sample(x_buf, t_buf):
            xs.append(x_buf)
            ts.append(t_buf)
    assert len(xs) == K - 1
    net = artss_amd.UNet(dom, seed=7)
    rng = np.random.default_rng(0)
    first = last = None
    for it in range(400):
        j = int(rng.integers(len(xs)))
        loss, _ = net.gradients(xs[j], ts[j], want_grads=False)
        first = loss if first is None else first
        last = loss
        net.adam_step(1e-3 if it < 300 else 3e-4)
    print("training loss", first, "->", last)
    solvers[1].set_warm_start(1)
    solvers[2].set_warm_start(2, net)
    res = [_run(s, M, K + 1, None) for s in solvers]
    names = ("previous p", "extrapolation", "learned")
    for name, c in zip(names, res):
        print(name, "iterations", [q["vcycles"] for q in c], "initial residual",
              ["%.2e" % q["initial_residual"] for q in c])
        for q in c:
            assert 0 <= q["residual"] <= TOL
    it0 = sum(q["vcycles"] for q in res[0])
    it2 = sum(q["vcycles"] for q in res[2])
    r0 = np.mean([np.log(q["initial_residual"]) for q in res[0]])
    r2 = np.mean([np.log(q["initial_residual"]) for q in res[2]])
    assert r2 < r0 - np.log(3.0)  # starts at least 3x closer on average
    assert it2 < it0


def test_solver_close_releases_network_and_buffers():
    """A network and buffers made on a solver's domain are released before the
    solver (whose stream they use) is destroyed, whatever order Python drops
    them in afterwards."""
    s = _solver(_xml(2))
    dom = s.domain
    info = dom.info(0)
    Nx, Ny, Nz = info["N"]
    net = artss_amd.UNet(dom, seed=1)
    buf = artss_amd.Buffer(dom, np.zeros((Nz - 2, Ny - 2, Nx - 2, 1), np.float32))
    s.set_warm_start(2, net)
    s.step(0.01)
    s.close()
    assert net.h is None and buf.ptr is None
    del net, buf
    s2 = _solver(_xml(2))
    net2 = artss_amd.UNet(s2.domain, seed=1)
    s2.set_warm_start(2, net2)
    net2.close()  # detaches itself from the solver first
    s2.step(0.01)
    assert s2.counters()["residual"] <= TOL
