"""Device beam solver (plate bodies on the GPU) against the reference's
fixtures and the oracle's beam restatement."""
import numpy as np
import pytest

from tests.golden.cases import BY_NAME, load_fixture
from tests.helpers import mismatch, run_device, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["ibm_static_plate", "ibm_heaving_plate", "ibm_pitching_two_plates"])
def test_device_rigid_bodies_match_reference_fixture(name):
    """Rigid bodies through the device body pipeline (markers and kinematics
    on the device, placement from host libm) reproduce the reference bit for
    bit: the rigid branch uses no device transcendental function."""
    pc = BY_NAME[name]
    print(f"[gpu] device-FEM rigid {name}", flush=True)
    f = run_device(pc, device_fem=True)
    n, m = mismatch(f, load_fixture(name)["f_final"])
    assert n == 0, f"{n} differ, max {m:.3e}"


@pytest.mark.parametrize("name", ["fem_flexible_plate", "fem_clamped_heaving_plate"])
def test_device_elastic_plate_vs_reference_and_oracle(name):
    """Elastic plate (Newmark + Newton + PCG on the device).  The device libm
    (sin, cos, acos) may differ from the host C library in the last place, so
    agreement is stated as the plain form's tolerance, 1e-12 relative: fluid
    state against the reference's fixture, node positions against the oracle;
    Newton/CG iteration counts equal on the well-conditioned case."""
    pc = BY_NAME[name]
    print(f"[gpu] device-FEM elastic {name}", flush=True)
    ro, rd = [], []
    fo = run_oracle(pc, record=ro)
    fd = run_device(pc, record=rd, device_fem=True)
    fx = load_fixture(name)["f_final"]
    n_ref, m_ref = mismatch(fd, fx)
    n_orc, m_orc = mismatch(fd, fo)
    print(f"[gpu] {name}: vs reference {n_ref} differ (max {m_ref:.3e}); vs oracle {n_orc} (max {m_orc:.3e})", flush=True)
    assert m_ref <= 1e-12 * np.abs(fx).max()
    for so, sd in zip(ro, rd):
        po, pd_ = so["beam"][0][0], sd["beam"][0][0]
        assert np.abs(po - pd_).max() <= 1e-12 * np.abs(po).max(), np.abs(po - pd_).max()
        if name == "fem_flexible_plate":
            assert so["fem_info"][0][:2] == sd["fem_info"][0][:2]


@pytest.mark.parametrize("name", ["fem_flexible_plate", "ibm_heaving_plate"])
def test_coupling_without_read_backs_matches(name):
    """Coupling calls that ask nothing back (no host wait per step) give the
    same state bit for bit as the waiting calls."""
    pc = BY_NAME[name]
    print(f"[gpu] no-wait coupling {name}", flush=True)
    fa = run_device(pc, device_fem=True)
    fb = run_device(pc, device_fem=True, wait=False)
    n, m = mismatch(fb, fa)
    assert n == 0, f"{n} differ, max {m:.3e}"


def test_queued_substeps_with_report_match_waiting_call():
    """The step queues the beam sub-steps behind its coupling call on the device
    and then collects the call's report (fsil_ibm_report); against the waiting
    coupling call followed by the sub-steps, every step's forcing-loop
    iteration count and residual, the beam loads and the final state agree bit
    for bit."""
    pc = BY_NAME["fem_flexible_plate"]
    print("[gpu] queued sub-steps + report vs waiting coupling call", flush=True)
    ra, rb = [], []
    fa = run_device(pc, device_fem=True, record=ra, wait_first=True)
    fb = run_device(pc, device_fem=True, record=rb)
    n, m = mismatch(fb, fa)
    assert n == 0, f"{n} differ, max {m:.3e}"
    assert len(ra) == len(rb) > 0
    for a, b in zip(ra, rb):
        assert a["iters"] == b["iters"] and a["dmax"] == b["dmax"]
        for la, lb in zip(a["lod"], b["lod"]):
            assert np.array_equal(la, lb)

