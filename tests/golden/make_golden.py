"""Write the golden fixtures: each case of tests/golden/cases.py is run
through the reference program itself (oracle/_ref/FSILBM3D, compiled from the
reference sources by oracle/build_ref.sh), from the seeded initial state
injected through its own restart file, and the state it writes after N steps
is stored as tests/golden/<case>.npz:

  f_final        (19, xDim, yDim, zDim) float64  post-stream + BC state after N steps
  datflow, continue_bytes  (cases that write flow output) the reference's last
                 DatFlow/Flow*_b001 file and its DatContinue/continue* file, raw
  <file>_g<g>    per body group, the rows of DatInfo/Group<g>_<file>.dat
                 (file: forces, firstNode, lastNode, centerNode, nodeAverage,
                 power, energy; 10 significant digits, one row per step)

The full-size cases (LONG_CASES: the 128x64x64 heaving plate for 2000 steps,
the 384x192x192 flexible plate for 20 steps) keep every DatInfo row and, in
place of f_final, small windows of the final state (window<k>, with the
bounds in windows).

Run:  python tests/golden/make_golden.py [case ...]
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.refcase import run_reference  # noqa: E402
from tests.golden.cases import CASES, INFO_FILES, LONG_CASES, fixture_path  # noqa: E402


def _isnum(s):
    try:
        float(s)
        return True
    except ValueError:
        return False


def read_info(workdir, ngroups):
    """Numeric rows of every DatInfo group file (ZONE/title lines skipped)."""
    out = {}
    for g in range(ngroups):
        for name in INFO_FILES:
            p = os.path.join(workdir, "DatInfo", f"Group{g + 1:03d}_{name}.dat")
            if not os.path.exists(p):
                continue
            rows = [l.split() for l in open(p) if l.strip()]
            rows = [[float(v) for v in row] for row in rows if all(_isnum(v) for v in row)]
            out[f"{name}_g{g + 1}"] = np.array(rows)
    return out


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
    out.update(read_info(r.workdir, len(pc.plates)))
    if case.flow_every:
        # the last DatFlow file and the continue file, byte for byte
        flows = sorted(x for x in os.listdir(os.path.join(r.workdir, "DatFlow")) if x.startswith("Flow"))
        out["datflow"] = np.frombuffer(open(os.path.join(r.workdir, "DatFlow", flows[-1]), "rb").read(), np.uint8)
        out["datflow_name"] = np.array(flows[-1])
        conts = sorted(x for x in os.listdir(os.path.join(r.workdir, "DatContinue")) if x != "continue")
        out["continue_bytes"] = np.frombuffer(open(os.path.join(r.workdir, "DatContinue", conts[-1]), "rb").read(),
                                              np.uint8)
        out["continue_name"] = np.array(conts[-1])
    shutil.rmtree(r.workdir, ignore_errors=True)
    np.savez_compressed(fixture_path(pc.name), **out)
    return r.step


def make_long(lc):
    pc = lc.pc
    case = pc.full_case()
    case.npsize = lc.npsize
    case.info_every = 1
    r = run_reference(case, lc.uniform_state(), pc.nsteps, keep=True, timeout=3600.0)
    out = {"step": np.array(r.step), "windows": np.array(lc.windows, dtype=np.int32),
           "step_seconds": np.array(r.step_seconds)}
    for k in range(len(lc.windows)):
        out[f"window{k}"] = lc.window(r.f, k)
    out.update(read_info(r.workdir, len(pc.plates)))
    shutil.rmtree(r.workdir, ignore_errors=True)
    np.savez_compressed(fixture_path(pc.name), **out)
    return r.step


if __name__ == "__main__":
    names = sys.argv[1:]
    for pc in CASES:
        if names and pc.name not in names:
            continue
        step = make(pc)
        print(f"{pc.name}: {step} steps -> {os.path.basename(fixture_path(pc.name))}", flush=True)
    for lc in LONG_CASES:
        if names and lc.name not in names:
            continue
        step = make_long(lc)
        print(f"{lc.name}: {step} steps -> {os.path.basename(fixture_path(lc.name))}", flush=True)
