"""Learned warm start of the multi-direct-forcing loop on the device: the
features/targets recording changes nothing, a guess that does not lower the
residual falls back to the plain loop's exact result, and a network trained
on the solver's own trajectories cuts the iterations needed to reach the
same tolerance."""
import numpy as np
import pytest

from tests.helpers import run_plate_fsi

pytestmark = pytest.mark.gpu


def _w(body_model=2):
    from fsilbm3d_amd.workloads import FlexiblePlate
    return FlexiblePlate(dims=(96, 64, 64), L=16.0, Re=100.0, le=(24.0, 32.0, 32.0), body_model=body_model)


def test_recording_changes_nothing():
    print("[gpu] warm start: recording", flush=True)
    w = _w()
    a = run_plate_fsi(w, 0, 8)
    b = run_plate_fsi(w, 0, 8, collect=8)
    assert np.array_equal(a["forces"], b["forces"]) and a["iters"] == b["iters"]
    assert b["X"].shape == (8 * 16 * 16, 12) and np.isfinite(b["X"]).all() and np.isfinite(b["Y"]).all()


def test_bad_guess_falls_back_to_the_plain_result():
    """A network whose guess is a large constant force raises the residual:
    every call after the first (which has no previous force to start from)
    restarts from zero and reproduces the plain run bit for bit."""
    from fsilbm3d_amd.mlp import NPARAM, WarmStartNet
    print("[gpu] warm start: fallback", flush=True)
    w = _w()
    net = WarmStartNet(1)
    th = np.zeros(NPARAM, np.float32)
    th[-3:] = 50.0                      # b4: a constant guess of 50 Uref |Ea|
    net.params = th
    a = run_plate_fsi(w, 0, 6)
    b = run_plate_fsi(w, 0, 6, net=net, warm_from=0)
    assert not b["fell"][0] and all(b["fell"][1:]) and not any(b["warm"])
    assert np.array_equal(a["forces"], b["forces"]) and a["iters"] == b["iters"]
    assert np.array_equal(a["pos"], b["pos"])


def test_trained_network_cuts_iterations():
    """A heaving rigid plate (prescribed motion, so converging the forcing
    further cannot destabilise the body).  The loop's residual falls slowly
    (about 1e-2 after 16-32 iterations on this plate), so the tolerance is
    set where the plain loop needs several iterations: dtolLBM 3e-2.  Train
    on the first 40 steps, then warm-start steps 40-59: fewer direct-forcing
    iterations on average than the plain loop and the same tolerance met.
    Both runs stop at a residual of 3e-2 of Uref from different starting
    forces, so their forces differ at that level, not at rounding: the drag
    is asserted within 5e-2 relative, the lift difference is printed (the
    heave force is the most sensitive to where an unconverged loop stops)."""
    from fsilbm3d_amd.mlp import WarmStartNet
    print("[gpu] warm start: trained", flush=True)
    w = _w(body_model=1)
    solver = dict(dtolLBM=3e-2, ntolLBM=20)
    plain = run_plate_fsi(w, 0, 60, collect=40, solver=solver)
    net = WarmStartNet(5)
    before, after = net.train(plain["X"], plain["Y"], epochs=40, batch=512, lr=2e-3, seed=3)
    print(f"[gpu] warm start: training loss {before:.3e} -> {after:.3e}", flush=True)
    assert after < 0.1 * before
    warm = run_plate_fsi(w, 0, 60, net=net, warm_from=40, solver=solver)
    ip, iw = np.mean(plain["iters"][40:]), np.mean(warm["iters"][40:])
    print(f"[gpu] warm start: mean iterations plain {ip:.2f} warm {iw:.2f}; guess used {np.mean(warm['warm']):.2f}",
          flush=True)
    assert iw < ip
    print(f"[gpu] warm start: residuals plain {np.round(plain['dmax'][40:], 4)} warm {np.round(warm['dmax'][40:], 4)}",
          flush=True)
    assert max(warm["dmax"][40:]) <= 3e-2 and max(warm["iters"][40:]) < 20
    F0, F1 = plain["forces"][40:, :2], warm["forces"][40:, :2]
    rel = np.abs(F1 - F0).max(axis=0) / np.abs(plain["forces"][:, :2]).max(axis=0)
    print(f"[gpu] warm start: force difference {rel}", flush=True)
    assert rel[0] < 5e-2
