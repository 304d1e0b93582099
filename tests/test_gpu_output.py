"""Output parity: the reference's final `.dat` files (Analysis::write_file and
Analysis::write_obstacles, analysis/Analysis.cpp:381-408) written from GPU
fields, compared with the same format applied to the oracle's fields.

The reference streams each double with the default ostream format (six
significant digits, the %g rules) one per line, and the obstacle file as
"index;value" over the obstacle cells of level 0, for every field type that
has domain boundary data (BoundaryDataController::get_used_fields).
"""
import os

import numpy as np
import pytest

import artss_amd
import artss_xml
import cases
import oracle_py

pytestmark = pytest.mark.gpu


def test_dat_files_match_reference_format(tmp_path):
    xml = cases.plume(n=(16, 16, 16), levels=2, steps=2, dt=0.02, diffusion="ColoredGaussSeidel",
                      obstacles=[(0.25, 0.5, 0.0, 0.375, 0.5, 0.75)], hrr=5.0)
    cfg = artss_xml.parse(xml)
    gpu = artss_amd.Solver(xml, device=0)
    dom = oracle_py.Domain(cfg)
    ref = oracle_py.Sim(dom, cfg, None, None, None, cases.uniform_ic(dom, float(cfg["ic"]["values"]["val"])))
    for step in (1, 2):
        gpu.step(step * cfg["dt"])
        ref.step(step * cfg["dt"])
    n = gpu.write_dat(tmp_path)
    names = ["u", "v", "w", "p", "T"]
    assert n == 2 * len(names)
    obst = dom.list(2, 0)  # obstacle cells of level 0
    assert len(obst) > 0
    for f in names:
        want = ref.field(f)
        with open(os.path.join(tmp_path, f + ".dat")) as fh:
            lines = fh.read().split("\n")
        assert lines[-1] == ""
        assert lines[:-1] == ["%g" % x for x in want], f
        with open(os.path.join(tmp_path, f + "_obstacle.dat")) as fh:
            olines = fh.read().split("\n")[:-1]
        assert olines == ["%d;%g" % (int(c), want[int(c)]) for c in obst], f
        assert np.array_equal(gpu.field(f), want)
