"""Ensemble member inputs (host only): every generated XML parses into a
domain whose sizes, obstacles and source follow the drawn parameters, and
the draw is a pure function of the seed."""
import re

import numpy as np

import artss_amd
from artss_amd import scenarios


def test_members_parse_and_follow_the_ranges():
    xmls = scenarios.ensemble_members(12, seed=5)
    assert xmls == scenarios.ensemble_members(12, seed=5)
    assert xmls != scenarios.ensemble_members(12, seed=6)
    n_obst = []
    for x in xmls:
        nx, ny, nz = [int(v) for v in re.search(r"<nx> (\d+) </nx><ny> (\d+) </ny><nz> (\d+)", x).groups()]
        assert 64 <= nx <= 160 and 64 <= ny <= 96 and 32 <= nz <= 64
        assert nx % 8 == 0 and ny % 8 == 0 and nz % 8 == 0
        peak = float(re.search(r"<HRR> ([^<]+) </HRR>", x).group(1)) / scenarios.HRR_PER_PEAK
        assert 20.0 <= peak <= 80.0
        dom = artss_amd.Domain(x, device=-1)
        info = dom.info(0)
        assert info["n"] == (nx, ny, nz)
        k = x.count("<obstacle ")
        n_obst.append(k)
        assert info["n_obstacles"] == k
        if k == 0:
            assert info["inner_cells"] == nx * ny * nz
        else:
            assert info["inner_cells"] < nx * ny * nz
        dom.close()
    assert min(n_obst) == 0 and max(n_obst) >= 2


def test_hrr_per_peak_matches_the_discrete_source_integral():
    # the half-Gaussian integral over the floor, summed on a fine grid
    h = 1.0 / 256
    x = (np.arange(-128, 128) + 0.5) * h
    y = (np.arange(0, 128) + 0.5) * h
    g = np.exp(-x ** 2 / (2 * scenarios.SIGMA ** 2))
    gy = np.exp(-y ** 2 / (2 * scenarios.SIGMA ** 2))
    integral = g.sum() * h * gy.sum() * h * g.sum() * h
    assert abs(integral - scenarios.HRR_PER_PEAK) / integral < 5e-3
