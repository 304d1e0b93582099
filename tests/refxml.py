"""Helpers for the reference's own test inputs (tests/golden/refxml_*.npz,
written by tests/golden/make_reference_inputs.py)."""
import glob
import hashlib
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# fields each solver computes (the harness dumps every FieldController field;
# those a solver never writes keep the reference's uninitialised contents)
SOLVER_FIELDS = {
    "AdvectionSolver": ("u", "v", "w"),
    "DiffusionSolver": ("u", "v", "w"),
    "DiffusionTurbSolver": ("u", "v", "w"),
    "AdvectionDiffusionSolver": ("u", "v", "w"),
    "NSSolver": ("u", "v", "w", "p", "rhs"),
    "NSTurbSolver": ("u", "v", "w", "p", "rhs"),
    "NSTempSolver": ("u", "v", "w", "p", "rhs", "T"),
    "NSTempTurbSolver": ("u", "v", "w", "p", "rhs", "T"),
    "PressureSolver": ("p",),
}


def cases():
    return sorted(os.path.basename(p)[len("refxml_"):-len(".npz")]
                  for p in glob.glob(os.path.join(GOLDEN, "refxml_*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, "refxml_%s.npz" % name), allow_pickle=False)
    return {k: z[k] for k in z.files}


def solver_of(xml):
    start = xml.index('description="', xml.index("<solver")) + len('description="')
    return xml[start:xml.index('"', start)]


def inner_first(dims):
    """permutation from index order to the order of the reference's *_ref.dat
    files: the inner cells, then the ghost cells, each in index order"""
    nx, ny, nz = (int(d) for d in dims)
    ghost = np.ones((nz, ny, nx), bool)
    ghost[1:-1, 1:-1, 1:-1] = False
    g = ghost.ravel()
    return np.concatenate([np.flatnonzero(~g), np.flatnonzero(g)])


def same6(values, ref):
    """number of entries where `values`, printed with 6 significant digits
    (the default ostream format of Analysis::write_file), differ from the
    fixture's values by more than one unit in their sixth digit, ignoring
    differences below 1e-5 of the field's largest magnitude (the fixtures were
    written by a build that rounds differently in the last bits, which moves
    some printed values by one unit, and near-zero values by more)"""
    printed = np.array([float("%g" % v) for v in values])
    mag = np.abs(ref)
    unit = np.where(mag > 0, 10.0 ** (np.floor(np.log10(np.where(mag > 0, mag, 1.0))) - 5), 0.0)
    floor = 1e-5 * max(float(mag.max()), 1e-300)
    bad = np.abs(printed - ref) > np.maximum(unit * 1.000001, floor)
    return int(bad.sum())


def sha(a):
    """SHA-256 of the raw little-endian float64 bytes"""
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()
