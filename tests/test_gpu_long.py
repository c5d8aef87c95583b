"""The full-size workloads on the device against the reference program's own
runs of them (tests/golden/make_golden.py LONG_CASES): every DatInfo row the
reference wrote (Cd/Cl/Cz, first/last/centre node, node average, power,
energy) at every step, and windows of the final state.

  * 128x64x64 rigid heaving plate, 2000 steps, FP64 plain form: every row at
    the reference's print precision (10 digits), the windows bit for bit.
  * 384x192x192 flexible plate, 20 steps, FP64 plain form: the rows at print
    precision within the elastic beam's device-libm tolerance (1e-12
    relative), the windows within 1e-12 relative; the FP32-deviation form's
    drag and lift within 1e-3 of the reference's.
"""
import numpy as np
import pytest

from tests.golden.cases import INFO_FILES, LONG_BY_NAME, load_fixture
from tests.helpers import print_equal, run_device_long

pytestmark = pytest.mark.gpu


def _check_rows(name, got, fx, rtol=0.0, atol_frac=0.0):
    bad = []
    for fname in INFO_FILES:
        ref = fx[f"{fname}_g1"]
        dev = got[fname]
        assert dev.shape == ref.shape, (fname, dev.shape, ref.shape)
        atol = atol_frac * np.abs(ref).max(axis=0, keepdims=True) if atol_frac else 0.0
        ok = print_equal(dev, ref, rtol=rtol, atol=atol)
        if not ok.all():
            i, j = np.argwhere(~ok)[0]
            bad.append(f"{fname}: {int((~ok).sum())} of {ok.size} differ, first step {i + 1} col {j + 1}: "
                       f"{dev[i, j]!r} vs {ref[i, j]!r}")
    assert not bad, f"{name}: " + "; ".join(bad)


def test_heaving_plate_2000_steps_matches_reference():
    lc = LONG_BY_NAME["c1_heaving_plate"]
    print("[gpu] 128x64x64 rigid heaving plate, 2000 steps, FP64, vs the reference's DatInfo", flush=True)
    fx = load_fixture(lc.name)
    out = run_device_long(lc)
    assert all(i == 3 for i in out["iters"])
    _check_rows(lc.name, out["rows"][0], fx)
    for k in range(len(lc.windows)):
        w, ref = out["windows"][k], fx[f"window{k}"]
        n = int((w != ref).sum())
        assert n == 0, f"window {k}: {n} values differ, max {np.abs(w - ref).max():.3e}"


def test_flexible_plate_full_size_matches_reference():
    lc = LONG_BY_NAME["c3_flexible_plate"]
    print("[gpu] 384x192x192 flexible plate, 20 steps, FP64, vs the reference's DatInfo", flush=True)
    fx = load_fixture(lc.name)
    out = run_device_long(lc)
    # the elastic beam runs sin/cos/acos on the device (last-place differences from the host libm)
    _check_rows(lc.name, out["rows"][0], fx, rtol=1e-12, atol_frac=1e-12)
    for k in range(len(lc.windows)):
        w, ref = out["windows"][k], fx[f"window{k}"]
        err = np.abs(w - ref).max()
        print(f"[gpu] window {k}: {int((w != ref).sum())} values differ, max {err:.3e}", flush=True)
        assert err <= 1e-12 * np.abs(ref).max()


def test_flexible_plate_fp32_forces_within_1e3_of_reference():
    lc = LONG_BY_NAME["c3_flexible_plate"]
    print("[gpu] 384x192x192 flexible plate, 20 steps, FP32-deviation form vs the reference", flush=True)
    fx = load_fixture(lc.name)
    out = run_device_long(lc, precision=1)
    F = out["rows"][0]["forces"][:, 3:5]
    R = fx["forces_g1"][:, 3:5]
    rel = np.abs(F - R).max(axis=0) / np.abs(R).max(axis=0)
    tip = np.abs(out["rows"][0]["lastNode"][:, 3:6] - fx["lastNode_g1"][:, 3:6]).max()
    print(f"[gpu] FP32 vs reference: drag {rel[0]:.2e} lift {rel[1]:.2e} (relative to max |Cd|, |Cl|); "
          f"tip displacement {tip:.2e} L", flush=True)
    assert rel.max() < 1e-3
