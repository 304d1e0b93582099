"""artss_mg_residual_static (no device allocation) against artss_mg_residual.

The static entry point computes the pressure residual norm with its partial
sums in the library's fixed device scratch array.  It must give the value of
the allocating entry point bit for bit, on two calls in a row (the second on
a smaller domain, so the shared array still holds the first call's larger
partial-sum set), and again after the allocating entry point has run.
"""
import numpy as np
import pytest

import artss_amd
import cases

pytestmark = pytest.mark.gpu

OBST = [(0.25, 0.5, 0.0, 0.375, 0.5, 0.75)]


def _setup(n, levels, seed):
    xml = cases.plume(n=n, X=(0.0, 2.0), levels=levels, steps=1, dt=0.01, obstacles=OBST)
    dom = artss_amd.Domain(xml)
    mg = dom.multigrid()
    size = dom.info(0)["size"]
    rng = np.random.default_rng(seed)
    p, rhs = dom.field(rng.standard_normal(size)), dom.field(rng.standard_normal(size))
    return dom, mg, p, rhs


def test_static_residual_matches_allocating_entry_point():
    big = _setup((128, 64, 64), 3, 1)
    small = _setup((32, 16, 16), 2, 2)
    r_big = big[1].residual_static(big[2], big[3])
    r_small = small[1].residual_static(small[2], small[3])
    assert r_big > 0 and r_small > 0
    assert r_big == big[1].residual(big[2], big[3])
    assert r_small == small[1].residual(small[2], small[3])
    # again after the allocating entry point ran, smaller input second
    assert big[1].residual_static(big[2], big[3]) == r_big
    assert small[1].residual_static(small[2], small[3]) == r_small
    print("static residual", r_big, r_small)


def test_static_residual_of_solved_pressure():
    _, mg, p, rhs = _setup((64, 32, 32), 3, 3)
    dt = 0.01
    mg.pressure(p, rhs, dt)
    r = mg.residual_static(p, rhs)
    assert r == mg.residual(p, rhs)
    assert np.isfinite(r)
