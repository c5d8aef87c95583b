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


def test_learned_form_at_full_size():
    """The flexible-plate case at full size (384x192x192, 2304 markers) in the
    FP32-deviation form with the learned warm start standing for the first 2
    of the 10 forcing rounds: trained on two seeded runs of the case (heave
    amplitude U(0.2, 0.4) L, frequency U(0.15, 0.3) U/L, 30 steps each, every
    marker of every step a sample), then 30 steps warm-started from step 2.
    The guess is kept on most steps (no fallback to the zero start), and each
    such call runs 9 forcing rounds (the guess's correction, then rounds 3-10)
    instead of 10.  The loop stops at its
    cap unconverged (residual about 1e-2 of Uref), so its end point moves with
    the starting force: the drag and lift differences against the plain FP32
    run are printed and bounded at the level measured (over 250 steps of the
    bench: drag 1.9e-2, lift 0.55 of their largest magnitudes); 1e-3 is reached
    by none of the skips measured (2, 5, 8)."""
    import dataclasses
    from fsilbm3d_amd.mlp import WarmStartNet
    from fsilbm3d_amd.workloads import flexible_plate
    print("[gpu] warm start: full size", flush=True)
    w = flexible_plate()
    skip, steps = 2, 30
    u = np.random.default_rng(20261202).random(4)
    XS, YS = [], []
    for r in range(2):
        wr = dataclasses.replace(w, amp=0.2 + 0.2 * u[2 * r], freq=0.15 + 0.15 * u[2 * r + 1])
        d = run_plate_fsi(wr, 1, steps, collect=steps, skip=skip, fill=True)
        XS.append(d["X"])
        YS.append(d["Y"])
    X, Y = np.concatenate(XS), np.concatenate(YS)
    assert X.shape == (2 * steps * 2304, 12)
    net = WarmStartNet(11)
    before, after = net.train(X, Y, epochs=10, batch=1024, lr=2e-3, seed=11)
    print(f"[gpu] warm start full size: {X.shape[0]} samples, loss {before:.3e} -> {after:.3e}", flush=True)
    assert after < 0.1 * before
    plain = run_plate_fsi(w, 1, steps, fill=True)
    warm = run_plate_fsi(w, 1, steps, net=net, warm_from=1, skip=skip, fill=True)
    net.close()
    kept = [k for k in range(1, steps) if warm["warm"][k - 1] and not warm["fell"][k - 1]]
    print(f"[gpu] warm start full size: guess kept on {len(kept)} of {steps - 1} steps; rounds plain "
          f"{np.mean(plain['rounds']):.2f} learned {np.mean(warm['rounds'][1:]):.2f}", flush=True)
    assert len(kept) >= 0.8 * (steps - 1)
    assert all(warm["rounds"][k] == plain["rounds"][k] - 1 for k in kept)
    F0, F1 = plain["forces"][:, :2], warm["forces"][:, :2]
    rel = np.abs(F1 - F0).max(axis=0) / np.abs(F0).max(axis=0)
    print(f"[gpu] warm start full size: drag/lift difference against plain FP32 {rel}", flush=True)
    assert np.isfinite(warm["forces"]).all()
    assert rel[0] < 0.1 and rel[1] < 1.0
