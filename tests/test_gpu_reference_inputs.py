"""The reference's own test inputs (its tests/*/*.xml) run unmodified through
the C ABI on the GPU, against the real reference's serial build.

Every field a solver computes must equal the reference's bit for bit
(SHA-256 of the whole float64 array, ghosts included) after steps 1, 2, 10,
100, 1000 and the last one (tests/golden/make_reference_inputs.py).  At
t_end the `.dat` files the library writes (Analysis::write_file's format,
analysis/Analysis.cpp:381-390) must also give the values of the reference's
own fixture files (<field>_ref.dat) to their sixth digit wherever the
current reference itself does; those files list the inner cells before the
ghost cells.  The Steckler input (160x128x128, seven obstacles) runs its
first three steps.

Covered this way: two-dimensional grids (nz = 1) with odd multigrid sizes,
the 64^3 five-level pressure solve, periodic, Dirichlet and Neumann
patches, the reference's own initial conditions (GaussBubble, ExpSinusSum,
ExpSinusProd, Hat, Drift, McDermott, Vortex, SinSinSin, Uniform), the
manufactured buoyancy source (temp_fct "Buoyancy"), the uniform momentum
source, an unknown force axis (dir="xyz") and viscous dissipation.
"""
import os
import tempfile

import numpy as np
import pytest

import artss_amd
import refxml

pytestmark = pytest.mark.gpu

CASES = refxml.cases()


@pytest.mark.parametrize("name", CASES)
def test_reference_input_bit_exact(name):
    z = refxml.load(name)
    xml = bytes(z["xml"]).decode()
    fields = refxml.SOLVER_FIELDS[refxml.solver_of(xml)]
    steps = [int(s) for s in z["steps"]]
    print("case", name, "solver", refxml.solver_of(xml), "steps", steps, flush=True)
    s = artss_amd.Solver(xml)
    done = 0
    for cp in steps:
        s.run(done + 1, cp - done)
        done = cp
        for f in fields:
            a = s.field(f)
            want = str(z["sha256/%s/%d" % (f, cp)])
            assert refxml.sha(a) == want, "%s: field %s differs after step %d (max |f| %.17g, reference %.17g)" % (
                name, f, cp, np.abs(a).max(), float(z["absmax/%s/%d" % (f, cp)]))
    refdat = [k[len("refdat/"):] for k in z if k.startswith("refdat/")]
    if not refdat:
        return
    perm = refxml.inner_first(z["dims"])
    with tempfile.TemporaryDirectory() as tmp:
        s.write_dat(tmp)
        for f in refdat:
            if not int(z["refdat_ok/" + f]):
                continue  # the current reference does not reproduce this fixture either
            written = np.loadtxt(os.path.join(tmp, f + ".dat"))
            bad = refxml.same6(written[perm], z["refdat/" + f])
            assert bad == 0, "%s: %s.dat differs from the reference's %s_ref.dat at %d cells" % (name, f, f, bad)
