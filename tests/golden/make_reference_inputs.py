"""Writes tests/golden/refxml_<case>.npz: the reference's own test inputs
(tests/*/*.xml of the reference) run by the real reference.

For every input the reference's serial build (oracle/_ref/ref_harness, the
reference's own sources compiled by `make -C oracle ref`) runs the time loop
of TimeIntegration.cpp:86-149 and the fixture keeps, as data:

  xml                 the input text
  steps               the checkpoints (time steps) below
  sha256/<f>/<step>   SHA-256 of the raw float64 bytes of field f after that
                      step (u, v, w, p, T, rhs), the whole array with ghosts
  absmax/<f>/<step>   max |f| there (for readable failure messages)
  dims                Nx, Ny, Nz
  refdat/<f>          the values of the reference's own fixture file
                      <f>_ref.dat, as written (6 significant digits), where
                      the reference keeps one for this input
  refdat_order        "inner_first": those files list the inner cells and
                      then the ghost cells, each in index order (the
                      current Analysis::write_file writes index order)
  refdat_ok/<f>       1 when the current reference's own field at t_end,
                      printed with 6 significant digits, gives the fixture's
                      values to their last digit (see refxml.same6), else 0

Only run where /root/reference exists; the fixtures are data and are
committed with this script.

    python tests/golden/make_reference_inputs.py [case ...]
"""
import glob
import hashlib
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
from refxml import inner_first, same6, sha  # noqa: E402
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_TESTS = "/root/reference/tests"
FIELDS = ("u", "v", "w", "p", "T", "rhs")

# case name -> (reference input relative to tests/, steps to run or None for
# the whole run to t_end).  Left out: the adaption input
# (navierStokesTempTurb/dynamicBoundaries), whose computational sub-domain
# and dynamic domain are outside the hot path.
CASES = {
    "advection": ("advection/Test_Advection.xml", None),
    "burgers": ("burgers/Test_Burgers.xml", None),
    "diffusion": ("diffusion/Test_Diffusion.xml", None),
    "diffusion_hat": ("diffusion/hat/Test_Diffusion_Hat.xml", None),
    "diffusion_turb": ("diffusionTurb/Test_DiffusionTurb.xml", None),
    "dissipation": ("dissipation/Test_Dissipation.xml", None),
    "ns_cavity": ("navierStokes/cavityFlow/Test_NavierStokes_Cavity.xml", None),
    "ns_channel": ("navierStokes/channelFlow/Test_NavierStokes_Channel.xml", None),
    "ns_mcdermott": ("navierStokes/mcDermott/Test_NavierStokes_McDermott.xml", None),
    "ns_vortex": ("navierStokes/vortex/Test_NavierStokes_Vortex.xml", None),
    "nstemp_mms": ("navierStokesTemp/mms/Test_NavierStokesTemp_BuoyancyMMS.xml", None),
    "nstempturb_mms": ("navierStokesTempTurb/mms/Test_NavierStokesTempTurb_BuoyancyMMS.xml", None),
    "nstempturb_steckler": ("navierStokesTempTurb/steckler/Test_NavierStokesTempTurb_Steckler.xml", 3),
    "nsturb_mcdermott": ("navierStokesTurb/mcDermott/Test_NavierStokesTurb_McDermott.xml", None),
    "nsturb_vortex": ("navierStokesTurb/vortex/Test_NavierStokesTurb_Vortex.xml", None),
    "pressure": ("pressure/Test_Pressure.xml", None),
}


def total_steps(xml):
    """the iteration count of TimeIntegration's loop (t = n*dt while t < t_end + dt/2)"""
    t_end = float(re.search(r"<t_end>([^<]*)</t_end>", xml).group(1))
    dt = float(re.search(r"<dt>([^<]*)</dt>", xml).group(1))
    n = 1
    while n * dt < t_end + dt / 2:
        n += 1
    return n - 1


def checkpoints(n):
    return sorted({s for s in (1, 2, 10, 100, 1000, n) if s <= n})


def main(only=None):
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref", "-j8"], check=True)
    for name, (rel, limit) in CASES.items():
        if only and name not in only:
            continue
        path = os.path.join(REF_TESTS, rel)
        xml = open(path).read()
        n = limit if limit is not None else total_steps(xml)
        steps = checkpoints(n)
        with tempfile.TemporaryDirectory() as tmp:
            r = subprocess.run([HARNESS, "run", path, tmp, ",".join(map(str, steps)), "-", str(n)], check=True,
                               capture_output=True, text=True)
            line = [ln for ln in r.stdout.splitlines() if "Nx=" in ln][-1]
            kv = dict(t.split("=") for t in line.split())
            assert int(kv["steps"]) == n, (name, kv["steps"], n)
            dims = tuple(int(kv[k]) for k in ("Nx", "Ny", "Nz"))
            arrays = {"xml": np.frombuffer(xml.encode(), np.uint8), "steps": np.array(steps, np.int64),
                      "dims": np.array(dims, np.int64), "refdat_order": np.array("inner_first")}
            final = {}
            for f in FIELDS:
                for s in steps:
                    a = np.fromfile(os.path.join(tmp, "%s_%d.f64" % (f, s)), np.float64)
                    arrays["sha256/%s/%d" % (f, s)] = np.array(sha(a))
                    arrays["absmax/%s/%d" % (f, s)] = np.array(float(np.abs(a).max()))
                    if s == n:
                        final[f] = a
            perm = inner_first(dims)
            notes = []
            for rf in sorted(glob.glob(os.path.join(os.path.dirname(path), "*_ref.dat"))):
                f = os.path.basename(rf)[:-len("_ref.dat")]
                ref = np.loadtxt(rf)
                arrays["refdat/" + f] = ref
                ok = limit is None and ref.size == final[f].size and same6(final[f][perm], ref) == 0
                arrays["refdat_ok/" + f] = np.array(int(ok))
                notes.append("%s_ref.dat %s" % (f, "reproduced" if ok else "differs"))
        out = os.path.join(HERE, "refxml_%s.npz" % name)
        np.savez_compressed(out, **arrays)
        print(name, n, "steps", dims, os.path.getsize(out) // 1024, "KiB;", ", ".join(notes), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
