"""Write the golden fixtures: each case of tests/golden/cases.py is run
through the reference program itself (oracle/_ref/FSILBM3D, compiled from the
reference sources by oracle/build_ref.sh), from the seeded initial state
injected through its own restart file, and the state it writes after N steps
is stored as tests/golden/<case>.npz:

  f_final  (19, xDim, yDim, zDim) float64  post-stream + BC state after N steps
  forces   per body group, the DatInfo Group<g>_forces.dat rows (10 digits)

Run:  python tests/golden/make_golden.py [case ...]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.refcase import run_reference  # noqa: E402
from tests.golden.cases import CASES, fixture_path  # noqa: E402


def make(pc):
    case = pc.full_case()
    case.npsize = 1
    if pc.plates:
        case.info_every = 1
    r = run_reference(case, pc.initial_state(), pc.nsteps, keep=True)
    if isinstance(r.f, list):
        out = {"f_final": r.f[0], "step": np.array(r.step)}
        for k, fs in enumerate(r.f[1:]):
            out[f"f_final_son{k + 1}"] = fs
    else:
        out = {"f_final": r.f, "step": np.array(r.step)}
    for g in range(len(pc.plates)):
        p = os.path.join(r.workdir, "DatInfo", f"Group{g + 1:03d}_forces.dat")
        rows = [l.split() for l in open(p) if l.strip() and not l.lstrip().startswith(("ZONE", "VARIABLES", "TITLE", "variables"))]
        rows = [[float(v) for v in row] for row in rows if all(_isnum(v) for v in row)]
        out[f"forces_g{g + 1}"] = np.array(rows)
    import shutil
    shutil.rmtree(r.workdir, ignore_errors=True)
    np.savez_compressed(fixture_path(pc.name), **out)
    return r.step


def _isnum(s):
    try:
        float(s)
        return True
    except ValueError:
        return False


if __name__ == "__main__":
    names = sys.argv[1:]
    for pc in CASES:
        if names and pc.name not in names:
            continue
        step = make(pc)
        print(f"{pc.name}: {step} steps -> {os.path.basename(fixture_path(pc.name))}", flush=True)
