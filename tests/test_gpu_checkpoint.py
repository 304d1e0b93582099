"""Checkpoint and rollback: a run continued from an HBM snapshot or from a
checkpoint file gives bit for bit the fields of the run that was never
interrupted (the reference's FieldIO / DataAssimilation rollback)."""
import numpy as np
import pytest

import artss_amd
import cases

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "T")
OBST = [(0.375, 0.625, 0.0, 0.375, 0.625, 0.8125)]


def _xml():
    return cases.plume(n=(64, 32, 32), X=(0.0, 2.0), levels=4, steps=12, dt=0.01, obstacles=OBST, hrr=3.0,
                       source_x=1.0, open_top=True)


def _solver(warm):
    s = artss_amd.Solver(_xml())
    s.set_pressure_mode(True, 1e-9, 400)
    s.set_pressure_method(2)
    s.set_warm_start(warm)
    return s


def _steps(s, first, last):
    for n in range(first, last + 1):
        s.step(n * 0.01)


@pytest.mark.parametrize("warm", [0, 1], ids=["previous_p", "extrapolated"])
def test_rollback_from_snapshot_repeats_the_run(warm):
    s = _solver(warm)
    _steps(s, 1, 4)
    slot = s.snapshot(0.04)
    _steps(s, 5, 8)
    first = {f: s.field(f) for f in FIELDS}
    assert s.restore(slot) == 0.04
    _steps(s, 5, 8)
    for f in FIELDS:
        assert np.array_equal(s.field(f), first[f]), f
    s.drop_snapshot(slot)
    with pytest.raises(artss_amd.ArtssError):
        s.restore(slot)


def test_checkpoint_file_continues_in_another_solver(tmp_path):
    a = _solver(1)
    _steps(a, 1, 5)
    path = tmp_path / "plume.ck"
    a.save(path, 0.05)
    _steps(a, 6, 9)
    b = _solver(1)
    assert b.load(path) == 0.05
    _steps(b, 6, 9)
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
    other = artss_amd.Solver(cases.plume(n=(32, 32, 32), levels=3, steps=2, dt=0.01, hrr=3.0))
    with pytest.raises(artss_amd.ArtssError):
        other.load(path)
