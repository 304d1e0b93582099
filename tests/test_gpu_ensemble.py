"""Ensembles: members of uneven size advanced together on host threads, each
on its own stream, must give every member exactly the fields it gets when
it runs alone (the members share nothing but the device)."""
import numpy as np
import pytest

import artss_amd
from artss_amd import scenarios

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "T")


def _member(xml):
    s = artss_amd.Solver(xml)
    s.set_pressure_mode(True, 1e-8, 400)
    s.set_pressure_method(2)
    s.set_warm_start(1)
    return s


@pytest.mark.parametrize("threads", [1, 4])
def test_ensemble_matches_members_run_alone(threads):
    xmls = scenarios.ensemble_members(5, seed=11, nx=(32, 64), ny=(32, 48), nz=(32, 48), steps=3)
    alone = []
    for x in xmls:
        s = _member(x)
        for n in range(1, 4):
            s.step(n * 0.005)
        alone.append(({f: s.field(f) for f in FIELDS}, s.counters()))
        s.close()
    members = [_member(x) for x in xmls]
    ens = artss_amd.Ensemble(members, threads=threads)
    for n in range(1, 4):
        ens.step(n)
    for m, (fields, counters) in zip(members, alone):
        assert m.counters() == counters
        for f in FIELDS:
            assert np.array_equal(m.field(f), fields[f]), f
    sizes = [m.domain.info(0)["n"] for m in members]
    assert len(set(sizes)) > 1  # uneven members
    ens.close()
