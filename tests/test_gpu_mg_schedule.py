"""The fused V-cycle schedule against the one-launch-per-operator schedule.

VCycleMG can run a V-cycle as the reference issues it (one launch per
operator, host-driven coarse solve) or with launches folded and the coarse
levels in one single-workgroup kernel.  Both must give bit-identical
pressure fields and cycle counts; the plain schedule itself is pinned to the
reference by the golden and oracle tests in test_gpu_parity.py.

Cases cover both smoothers, odd and even n_relax (the Jacobi smoother's
buffer swaps), periodic / Neumann / Dirichlet pressure boundaries with a
non-zero Dirichlet value, no obstacles, one obstacle and three obstacles of
which two touch (a boundary fill that is not idempotent), two-dimensional
inputs (one cell along z), and every split between multi-workgroup and
single-workgroup levels.
"""
import zlib

import numpy as np
import pytest

import artss_amd
import cases
import refxml

pytestmark = pytest.mark.gpu

OBST3 = [(0.25, 0.5, 0.0, 0.375, 0.5, 0.75),
         (0.5, 0.75, 0.25, 0.5, 0.25, 0.5),
         (1.25, 1.5, 0.5, 0.75, 0.25, 0.75)]


def _ref_xml(name):
    """one of the reference's own inputs (kept in its fixture)"""
    return bytes(refxml.load(name)["xml"]).decode()


def _with(xml, n_relax=None, top_p_value=None):
    if n_relax is not None:
        xml = xml.replace("<n_relax> 4 </n_relax>", "<n_relax> %d </n_relax>" % n_relax)
    if top_p_value is not None:
        xml = xml.replace('<boundary field="p" patch="top" type="dirichlet" value="0.0" />',
                          '<boundary field="p" patch="top" type="dirichlet" value="%r" />' % top_p_value)
    return xml


CASES = {
    "plume_obst3_jacobi": lambda: cases.plume(n=(64, 32, 32), X=(0.0, 2.0), levels=3, steps=2, dt=0.01,
                                              obstacles=OBST3, hrr=3.0, source_x=1.0),
    "plume_obst3_cgs": lambda: cases.plume(n=(64, 32, 32), X=(0.0, 2.0), levels=3, steps=2, dt=0.01,
                                           smoother="ColoredGaussSeidel", obstacles=OBST3, hrr=3.0, source_x=1.0),
    "plume_open_relax3": lambda: _with(cases.plume(n=(48, 40, 24), X=(0.0, 1.5), levels=3, steps=2, dt=0.01,
                                                   obstacles=OBST3[:1], hrr=3.0, open_top=True),
                                       n_relax=3, top_p_value=0.25),
    "plume_open_relax2_cgs": lambda: _with(cases.plume(n=(40, 24, 32), levels=2, steps=2, dt=0.01,
                                                       smoother="ColoredGaussSeidel", obstacles=OBST3[:1],
                                                       hrr=3.0, open_top=True), n_relax=2),
    "tg_periodic": lambda: cases.ns_taylor_green(n=32, levels=3, steps=2),
    # level 0 above 2^20 cells: the kernels that give each thread four cells
    # along z (and the prolongation that shares its coarse planes among them)
    "plume_obst3_large": lambda: cases.plume(n=(128, 64, 128), X=(0.0, 2.0), levels=4, steps=2, dt=0.01,
                                             obstacles=OBST3, hrr=3.0, source_x=1.0, open_top=True),
    "tg_periodic_relax1": lambda: _with(cases.ns_taylor_green(n=32, levels=2, steps=2), n_relax=1),
    # two-dimensional inputs (nz = 1): the restriction reads the fine ghost
    # layer along z, so the residuum's fill must not be dropped; 40 -> 5 -> 2
    # along x and y (odd level sizes)
    "cavity_2d": lambda: _ref_xml("ns_cavity"),
    "vortex_2d_periodic": lambda: _ref_xml("ns_vortex"),
}


def _run(xml, mode, tail_cells, p0, rhs, dt):
    dom = artss_amd.Domain(xml)
    mg = dom.multigrid()
    mg.set_schedule(mode, tail_cells)
    gp, grhs = dom.field(p0), dom.field(rhs)
    counts, states = [], []
    for t in (dt, 2 * dt, 3 * dt):
        counts.append(mg.pressure(gp, grhs, t))
        states.append(gp.download())
    for _ in range(3):
        mg.vcycle(gp, grhs)
    states.append(gp.download())
    res = mg.residual(gp, grhs)
    tail, lds = mg.tail_info()
    print("first tail level", tail, "LDS bytes", lds)
    return counts, states, res, tail


@pytest.mark.parametrize("name", sorted(CASES))
def test_fused_schedule_matches_plain(name):
    xml = CASES[name]()
    size = artss_amd.Domain(xml, device=-1).info(0)["size"]
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    p0, rhs = 0.01 * rng.standard_normal(size), rng.standard_normal(size)
    dt = float(xml.split("<dt>")[1].split("</dt>")[0])
    ref_counts, ref_states, ref_res, _ = _run(xml, 0, 0, p0, rhs, dt)
    tails = set()
    for mode, tail_cells in ((1, 0), (1, 3000), (1, 20000), (1, 10 ** 9), (2, 3000), (2, 20000), (2, 10 ** 9)):
        counts, states, res, tail = _run(xml, mode, tail_cells, p0, rhs, dt)
        tails.add(tail)
        print(name, "mode", mode, "tail_cells", tail_cells, "first tail level", tail, "cycles", counts)
        assert counts == ref_counts, (mode, tail_cells, counts, ref_counts)
        for i, (a, b) in enumerate(zip(states, ref_states)):
            if not np.array_equal(a, b):
                d = np.abs(a - b)
                raise AssertionError("%s mode=%d tail_cells=%d state %d: %d cells differ, max %.3g" %
                                     (name, mode, tail_cells, i, (d != 0).sum(), d.max()))
        assert res == ref_res
    # the sweep over tail_cells ran without the single-workgroup kernel and with it covering every level
    assert 0 in tails and len(tails) >= 2, tails


def test_solver_tolerance_mode_graph_replay_matches_plain():
    """Whole plume steps with V-cycles to a residual target: every schedule
    gives the same fields and counts, and the live kernel timer (which runs
    the V-cycle eagerly while armed) sees the same launches in every mode."""
    xml = cases.plume(n=(64, 32, 32), X=(0.0, 2.0), levels=4, steps=3, dt=0.01, obstacles=OBST3[:1], hrr=3.0,
                      source_x=1.0, open_top=True)
    runs = {}
    for mode in (0, 1, 2):
        for timed in (False, True):
            s = artss_amd.Solver(xml)
            s.set_mg_schedule(mode, 20000)
            s.set_pressure_mode(True, 1e-9, 300)
            if timed:
                s.domain.profile_arm("jacobi_mg@0")
            counts = []
            for step in (1, 2, 3):
                s.step(step * 0.01)
                counts.append(s.counters())
            launches = s.domain.profile_read()[0] if timed else None
            runs[mode, timed] = ({f: s.field(f) for f in ("u", "v", "w", "p", "T")}, counts, launches)
            print("mode", mode, "timed", timed, [c["vcycles"] for c in counts], "jacobi_mg@0 launches", launches)
    fields0, counts0, _ = runs[0, False]
    assert sum(c["vcycles"] for c in counts0) > 0
    for key, (fields, counts, launches) in runs.items():
        assert counts == counts0, key
        for f in fields0:
            assert np.array_equal(fields[f], fields0[f]), (key, f)
        if key[1]:
            assert launches == runs[0, True][2] > 0, (key, launches)
