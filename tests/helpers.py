"""Shared drivers: run a parity case on the CPU oracle or on the device."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from fsilbm3d_amd import params as P  # noqa: E402
from fsilbm3d_amd.solid import RigidPlate  # noqa: E402


def derived(case):
    """nu, tau, Omega, Omega2, MRT matrices, Uref, Tref of a refcase.Case."""
    if case.UrefType == 0:
        uref = abs(case.uvwIn[0])
    else:
        uref = case.Uref
    lref = case.Lref
    tref = lref / uref if case.TrefType == 0 else case.Tref
    nu = P.viscosity(uref, lref, case.Re)
    tau, om = P.srt(nu, case.dh)
    om2 = P.trt_omega2(om, case.params[0]) if case.collide == 2 else om
    Mc = Mf = None
    if case.collide == 3:
        a, b = P.mrt_matrices(om)
        Mc, Mf = P.fortran_order(a), P.fortran_order(b)
    return dict(nu=nu, tau=tau, Omega=om, Omega2=om2, Mc=Mc, Mf=Mf, Uref=uref, Tref=tref)


def info_scales(case, d, asfacs):
    """Lref, Uref, Aref, Fref, Pref, Eref of a case whose reference area is
    the largest body's (allocate_solid_memory, Solidbody.f90:302-305)."""
    from oracle.oracle import reference_scales
    asfac = asfacs[int(np.argmax(asfacs))] if len(asfacs) else 0.0
    return reference_scales(case.UrefType, case.uvwIn, case.Uref, case.Lref, case.TrefType, case.Tref, case.denIn,
                            asfac)


def plate_asfac(p):
    """Sum of the element lengths areaElem00 in element order (SolidSolver.f90:83-85, 1375)."""
    x = p.nodes()[:, :3]
    s = 0.0
    for e in range(len(x) - 1):
        d = x[e + 1] - x[e]
        s = s + float(np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
    return s


def make_bodies(pc, denIn):
    from fsilbm3d_amd.solid import SurfaceBody
    from tests.golden.cases import SurfacePlate
    out = []
    for p in pc.plates:
        if isinstance(p, SurfacePlate):
            out.append(SurfaceBody(p.mesh(), XYZo=p.XYZo, XYZAmpl=p.XYZAmpl, XYZPhi=p.XYZPhi, AoAo=p.AoAo,
                                   AoAAmpl=p.AoAAmpl, AoAPhi=p.AoAPhi, Freq=p.Freq, initXYZVel=p.initXYZVel,
                                   denIn=denIn, IBPenaltyAlpha=pc.case.IBPenaltyAlpha))
            continue
        out.append(RigidPlate(p.nodes(), p.elements(), XYZo=p.XYZo, XYZAmpl=p.XYZAmpl, XYZPhi=p.XYZPhi,
                              AoAo=p.AoAo, AoAAmpl=p.AoAAmpl, AoAPhi=p.AoAPhi, Freq=p.Freq,
                              initXYZVel=p.initXYZVel, denIn=denIn, IBPenaltyAlpha=pc.case.IBPenaltyAlpha))
    return out


def flow_kwargs(case):
    return dict(denIn=case.denIn, uvwIn=case.uvwIn, shearRateIn=case.shearRateIn, velocityKind=case.velocityKind,
                volumeForceIn=case.volumeForceIn, volumeForceAmp=case.volumeForceAmp,
                volumeForceFreq=case.volumeForceFreq, volumeForcePhi=case.volumeForcePhi)


def oracle_bodies(pc, d):
    """Plate bodies on the oracle's beam restatement (rigid or elastic)."""
    from oracle.oracle import OracleBeam, OracleSurface
    from tests.golden.cases import SurfacePlate, icosphere
    case = pc.case
    out = []
    for p in pc.plates:
        cons = np.zeros((p.n_nodes, 6), np.int32)
        cons[0] = 1
        els = np.array(p.elements(), np.int32)
        nodes = p.nodes()
        if isinstance(p, SurfacePlate):
            # the beam the reference writes of the mesh's first two nodes (F20.10)
            xyz, tri = icosphere(p.radius)
            nodes = np.array([[float("%.10f" % v) for v in xyz[i]] + [1.0, 1.0, 0.0, 0.0, 1.0] for i in range(2)])
            cons = np.zeros((2, 6), np.int32)
            cons[0] = 1
            els = np.array([(1, 2, 1)], np.int32)
        beam = OracleBeam(nodes, els, cons, iBodyModel=p.iBodyModel, isMotionGiven=p.isMotionGiven,
                              denR=p.denR, KB=p.KB, KS=p.KS, psR=p.psR, Freq=p.Freq, XYZo=p.XYZo, XYZAmpl=p.XYZAmpl,
                              XYZPhi=p.XYZPhi, AoAo=p.AoAo, AoAAmpl=p.AoAAmpl, AoAPhi=p.AoAPhi,
                              initXYZVel=p.initXYZVel, Lref=case.Lref, Uref=d["Uref"], denIn=case.denIn,
                              dampK=case.dampK, dampM=case.dampM, GeoGamma=case.GeoGamma,
                              NewmarkGamma=case.NewmarkGamma, NewmarkBeta=case.NewmarkBeta, dtolFEM=case.dtolFEM,
                              ntolFEM=case.ntolFEM, IBPenaltyAlpha=case.IBPenaltyAlpha)
        if isinstance(p, SurfacePlate):
            beam = OracleSurface(beam, tri, xyz, case.IBPenaltyAlpha, case.denIn)
        out.append(beam)
    return out


def run_oracle(pc, f0=None, nsteps=None, record=None):
    """Run a ParityCase on the C oracle; returns the final state (and the
    per-step IBM outputs when record is a list)."""
    from oracle.oracle import OracleBlock, lod_flow
    case = pc.case
    d = derived(case)
    ob = OracleBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["Omega"], d["Omega2"], d["Mc"], d["Mf"],
                     tau=d["tau"], nu=d["nu"], **flow_kwargs(case))
    f = np.array(pc.initial_state() if f0 is None else f0, dtype=np.float64, copy=True)
    ob.init_boundary(f, 0.0)
    bodies = oracle_bodies(pc, d)
    ibm = dict(ntol=case.ntolLBM, dtol=case.dtolLBM, Uref=d["Uref"], denIn=case.denIn)
    n = pc.nsteps if nsteps is None else nsteps
    time = 0.0
    for _ in range(n):
        time = time + case.dh
        bl = None
        if bodies:
            for b in bodies:
                if b.v_move == 1 or b.iBodyModel == 2:
                    b.update_markers()
            bl = [dict(Exyz=b.Exyz.copy(), Evel=b.Evel.copy(), Ea=b.Ea.copy()) for b in bodies]
            xcs = [b.element_centres() for b in bodies]
        out = ob.step(f, time, bl, ibm)
        if bodies:
            lods = [lod_flow(E, bb["Exyz"], b.marker_elem, b.elem_nodes, xc, b.nND)
                    for E, bb, b, xc in zip(out["Eforce"], bl, bodies, xcs)]
            for b, lod in zip(bodies, lods):
                b.set_lodflow(lod)
            if record is not None:
                record.append(dict(iters=out["iters"], Eforce=out["Eforce"], lod=lods,
                                   Exyz=[bb["Exyz"] for bb in bl]))
            dts = case.dh / float(case.numsubstep)
            for isub in range(1, case.numsubstep + 1):
                for b in bodies:
                    b.structure(time, isub, case.dh, dts)
            if record is not None:
                record[-1]["beam"] = [b.state() for b in bodies]
                record[-1]["fem_info"] = [b.info() for b in bodies]
                refs = info_scales(case, d, [b.Asfac for b in bodies])
                record[-1]["solid_info"] = [b.solid_info(p.XYZo, refs) for b, p in zip(bodies, pc.plates)]
    return f


def run_oracle_block(pc):
    """A body-free ParityCase on the oracle; returns (state, tau_all as (x, y, z))."""
    from oracle.oracle import OracleBlock
    case = pc.case
    d = derived(case)
    ob = OracleBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["Omega"], d["Omega2"], d["Mc"], d["Mf"],
                     tau=d["tau"], nu=d["nu"], **flow_kwargs(case))
    f = np.array(pc.initial_state(), dtype=np.float64, copy=True)
    ob.init_boundary(f, 0.0)
    time = 0.0
    for _ in range(pc.nsteps):
        time = time + case.dh
        ob.step(f, time)
    tau = ob.tau_all.reshape(case.dims) if ob.tau_all is not None else np.full(case.dims, d["tau"])
    return f, tau


def run_device_block(pc):
    """A body-free ParityCase through the C-ABI; returns (state, tau_all)."""
    from fsilbm3d_amd.block import LBMBlock
    case = pc.case
    d = derived(case)
    blk = LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params)
    blk.set_flow(**flow_kwargs(case))
    blk.upload(np.array(pc.initial_state(), dtype=np.float64))
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    time = 0.0
    for _ in range(pc.nsteps):
        time = time + case.dh
        blk.update_volume_force(time)
        blk.collision_streaming_bc()
    out = blk.download(), blk.tau_all()
    blk.close()
    return out


def son_derived(case, sb):
    """Omega / Omega2 / MRT matrices of a son block (viscosity of the root)."""
    nu = derived(case)["nu"]
    tau, om = P.srt(nu, sb.dh)
    om2 = P.trt_omega2(om, sb.params[0]) if sb.collide == 2 else om
    Mc = Mf = None
    if sb.collide == 3:
        a, b = P.mrt_matrices(om)
        Mc, Mf = P.fortran_order(a), P.fortran_order(b)
    return dict(nu=nu, tau=tau, Omega=om, Omega2=om2, Mc=Mc, Mf=Mf)


def run_oracle_tree(pc, fs0=None, nsteps=None):
    """Run a father/son ParityCase on the oracle tree; returns [father, sons...]."""
    from oracle.oracle import OracleBlock
    from oracle.oracle_tree import OracleTree, comm_pair
    case = pc.case
    d = derived(case)
    root = OracleBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["Omega"], d["Omega2"], d["Mc"],
                       d["Mf"], **flow_kwargs(case))
    sons, pairs, taus = [], [], [d["tau"]]
    for sb in case.sons:
        sd = son_derived(case, sb)
        sons.append(OracleBlock(sb.dims, sb.dh, sb.origin, sb.bc, sb.collide, sd["Omega"], sd["Omega2"], sd["Mc"],
                                sd["Mf"], **flow_kwargs(case)))
        per = [int(sb.bc[2 * a] == 301) for a in range(3)]
        pairs.append(comm_pair((case.dims, case.dh, case.origin), (sb.dims, sb.dh, sb.origin), sb.bc, per))
        taus.append(sd["tau"])
    fs = [np.array(a, dtype=np.float64, copy=True) for a in (pc.initial_state() if fs0 is None else fs0)]
    root.init_boundary(fs[0], 0.0)
    for k, sb in enumerate(sons):
        sb.init_boundary(fs[1 + k], 0.0)
    tree = OracleTree(root, sons, pairs, taus, case.interpolateScheme)
    n = pc.nsteps if nsteps is None else nsteps
    time = 0.0
    for _ in range(n):
        time = time + case.dh
        tree.step(fs, time)
    return fs


def plate_specs(pc):
    from fsilbm3d_amd.fsi import PlateSpec
    out = []
    for p in pc.plates:
        cons = np.zeros((p.n_nodes, 6), np.int32)
        cons[0] = 1
        out.append(PlateSpec(nodes8=p.nodes(), elements=np.array(p.elements(), np.int32), constraints=cons,
                             iBodyModel=p.iBodyModel, isMotionGiven=p.isMotionGiven, denR=p.denR, KB=p.KB, KS=p.KS,
                             psR=p.psR, Freq=p.Freq, XYZo=p.XYZo, XYZAmpl=p.XYZAmpl, XYZPhi=p.XYZPhi, AoAo=p.AoAo,
                             AoAAmpl=p.AoAAmpl, AoAPhi=p.AoAPhi, initXYZVel=p.initXYZVel))
    return out


def _lib_check(rc):
    from fsilbm3d_amd import _lib
    _lib.check(rc)


def run_device(pc, f0=None, nsteps=None, record=None, device_fem=None, wait=True, wait_first=False):
    """Run a ParityCase through the C-ABI on the GPU; returns the final state.
    Bodies run on the device beam solver when device_fem (default: when any
    plate is elastic), else with the host rigid kinematics.  wait=False: the
    device-body coupling calls read nothing back (no host wait per step).
    wait_first: each step waits for its coupling call before queuing the beam
    sub-steps (instead of queuing them first and then collecting the report)."""
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock, FSIBlock
    case = pc.case
    d = derived(case)
    blk = LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params)
    blk.set_flow(**flow_kwargs(case))
    f = np.array(pc.initial_state() if f0 is None else f0, dtype=np.float64, copy=True)
    blk.upload(f)
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    if device_fem is None:
        device_fem = any(p.iBodyModel == 2 for p in pc.plates)
    if device_fem and pc.plates:
        bodies = pc.plates
        fsi = DeviceBodyBlock(blk, plate_specs(pc), ntolLBM=case.ntolLBM, dtolLBM=case.dtolLBM, Uref=d["Uref"],
                              denIn=case.denIn, Lref=case.Lref, numsubstep=case.numsubstep,
                              IBPenaltyAlpha=case.IBPenaltyAlpha, dampK=case.dampK, dampM=case.dampM,
                              GeoGamma=case.GeoGamma, NewmarkGamma=case.NewmarkGamma, NewmarkBeta=case.NewmarkBeta,
                              dtolFEM=case.dtolFEM, ntolFEM=case.ntolFEM)
    else:
        bodies = make_bodies(pc, case.denIn)
        fsi = FSIBlock(blk, bodies, ntolLBM=case.ntolLBM, dtolLBM=case.dtolLBM, Uref=d["Uref"], denIn=case.denIn,
                       numsubstep=case.numsubstep)
    n = pc.nsteps if nsteps is None else nsteps
    time = 0.0
    for _ in range(n):
        time = time + case.dh
        if isinstance(fsi, DeviceBodyBlock) and wait_first:
            blk.update_volume_force(time)
            fsi.interaction_force(blk.dh, True)
            _lib_check(fsi.lib.fsil_fem_substeps(blk.handle, float(time), fsi.numsubstep, blk.dh))
            blk.collision_streaming_bc()
        elif isinstance(fsi, DeviceBodyBlock):
            fsi.step(time, wait=wait)
        else:
            fsi.step(time)
        if record is not None and bodies:
            rec = dict(iters=fsi.iters, dmax=getattr(fsi, "dmax", None), Eforce=fsi.Eforce[:fsi.M].copy(),
                       lod=[fsi.body_loads(i).copy() for i in range(len(bodies))])
            if isinstance(fsi, DeviceBodyBlock):
                rec["beam"] = [fsi.state(i) for i in range(len(bodies))]
                rec["fem_info"] = [fsi.info(i) for i in range(len(bodies))]
            record.append(rec)
    out = blk.download()
    blk.close()
    return out


def run_device_tree(pc, fs0=None, nsteps=None, per_call=False):
    """Run a father/son ParityCase through the C-ABI; returns [father, sons...]
    and the CommPair of every son.
    per_call steps the recursion from the host mirror (one library call per
    exchange); otherwise the whole tree step is one library call."""
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.tree import BlockTree
    case = pc.case
    d = derived(case)
    blocks = [LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params, block_id=1)]
    for k, sb in enumerate(case.sons):
        blocks.append(LBMBlock(sb.dims, sb.dh, sb.origin, sb.bc, sb.collide, d["nu"], sb.params, block_id=2 + k))
    for b in blocks:
        b.set_flow(**flow_kwargs(case))
    fs = pc.initial_state() if fs0 is None else fs0
    for b, f in zip(blocks, fs):
        b.upload(np.asarray(f, dtype=np.float64))
    tree = BlockTree(blocks, case.interpolateScheme)
    tree.set_boundary_conditions()
    n = pc.nsteps if nsteps is None else nsteps
    if per_call:
        time = 0.0
        for _ in range(n):
            time = time + case.dh
            tree.step(time)
    else:
        tree.steps(n, 0.0, case.dh)
    out = [b.download() for b in blocks]
    pairs = {s: tree.comm_pair(s) for s in range(len(blocks)) if tree.father[s] >= 0}
    for b in blocks[::-1]:
        b.close()
    return out, pairs


def mismatch(a, b):
    d = np.abs(a - b)
    return int((a != b).sum()), float(d.max()) if d.size else 0.0


def run_plate_fsi(w, precision, nsteps, record_every=1, net=None, warm_from=None, collect=0, solver=None, skip=0,
                  fill=False):
    """A flexible-plate workload (fsilbm3d_amd.workloads.FlexiblePlate) on the
    device in the given storage form; returns per-step total fluid force on
    the body (sum of lodFlow x, y, z), MDF iterations, forcing rounds and
    residuals, and the final node positions.  net/warm_from: learned warm
    start from that step on, standing for `skip` forcing rounds; collect:
    record the network's features/targets (after `skip` rounds) of the first
    steps.  fill: start from the uniform equilibrium written on the device
    (large blocks) instead of an uploaded state."""
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    blk = LBMBlock(w.dims, 1.0, (0.0, 0.0, 0.0), w.bc, 1, w.nu, precision=precision)
    blk.set_flow(**w.flow_kwargs())
    if fill:
        blk.fill(w.equilibrium())
    else:
        blk.upload(w.initial_state())
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    kw = w.solver_kwargs()
    kw.update(solver or {})
    fsi = DeviceBodyBlock(blk, [w.plate()], **kw)
    forces, iters, dmax, warm, fell, rounds = [], [], [], [], [], []
    XS, YS = [], []
    t = 0.0
    for k in range(nsteps):
        use = net is not None and warm_from is not None and k >= warm_from
        fsi.set_warm_start(net if use else None, record=(k < collect) or use, skip=skip)
        t = t + 1.0
        fsi.step(t)
        if (k + 1) % record_every == 0:
            forces.append(fsi.body_loads(0).reshape(-1, 6).sum(0)[:3].copy())
        iters.append(fsi.iters)
        dmax.append(fsi.dmax)
        rounds.append(fsi.rounds)
        if k < collect or use:
            X, Y, used, fb = fsi.warm_start_data()
            if k < collect:
                XS.append(X)
                YS.append(Y)
            warm.append(used)
            fell.append(fb)
    pos = fsi.state(0)[0].copy()
    den, uuu = blk.calculate_macro_quantities()
    blk.close()
    out = dict(forces=np.array(forces), iters=iters, dmax=dmax, pos=pos, u=uuu, warm=warm, fell=fell, rounds=rounds)
    if collect:
        out["X"] = np.concatenate(XS)
        out["Y"] = np.concatenate(YS)
    return out


def print_equal(dev, ref, rtol=0.0, atol=0.0):
    """True where a value equals the reference's DatInfo entry (written with
    10 significant digits, E20.10) after rounding to 10 digits, or, with rtol
    or atol, lies within half a unit of the 10th digit plus that tolerance."""
    dev = np.asarray(dev, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    fmt = np.vectorize(lambda v: float("%.9e" % v))(dev) if dev.size else dev
    same = fmt == ref
    if rtol or atol:
        mag = np.where(ref != 0.0, np.floor(np.log10(np.abs(np.where(ref != 0.0, ref, 1.0)))), 0.0)
        half = 0.5 * 10.0 ** (mag - 9.0)
        same = same | (np.abs(dev - ref) <= half * (1.0 + 1e-9) + rtol * np.abs(ref) + atol)
    return same


def run_device_long(lc, nsteps=None, precision=0, info=True, net=None, warm_from=None, skip=0, collect=0):
    """A full-size workload (tests/golden/cases.py LongCase) on the device
    from the uniform inflow state (filled on the device): per-step DatInfo rows
    of every body (dict file -> array [steps, width]), iteration counts, the
    forcing rounds run and the final state's windows (FP64 form only)."""
    from fsilbm3d_amd.block import LBMBlock
    from fsilbm3d_amd.params import equilibrium_point
    from fsilbm3d_amd.fsi import DeviceBodyBlock
    pc = lc.pc
    case = pc.case
    d = derived(case)
    blk = LBMBlock(case.dims, case.dh, case.origin, case.bc, case.collide, d["nu"], case.params, precision=precision)
    blk.set_flow(**flow_kwargs(case))
    blk.fill(equilibrium_point(1.0, pc.ubase))
    blk.update_volume_force(0.0)
    blk.set_boundary_conditions()
    fsi = DeviceBodyBlock(blk, plate_specs(pc), ntolLBM=case.ntolLBM, dtolLBM=case.dtolLBM, Uref=d["Uref"],
                          denIn=case.denIn, Lref=case.Lref, numsubstep=case.numsubstep,
                          IBPenaltyAlpha=case.IBPenaltyAlpha, dampK=case.dampK, dampM=case.dampM,
                          GeoGamma=case.GeoGamma, NewmarkGamma=case.NewmarkGamma, NewmarkBeta=case.NewmarkBeta,
                          dtolFEM=case.dtolFEM, ntolFEM=case.ntolFEM)
    refs = info_scales(case, d, [plate_asfac(p) for p in pc.plates])
    n = pc.nsteps if nsteps is None else nsteps
    rows = [dict() for _ in pc.plates]
    iters, rounds, XS, YS = [], [], [], []
    time = 0.0
    for k in range(n):
        use = net is not None and warm_from is not None and k >= warm_from
        if net is not None or collect or skip:
            fsi.set_warm_start(net if use else None, record=k < collect, skip=skip)
        time = time + case.dh
        fsi.step(time)
        iters.append(fsi.iters)
        rounds.append(fsi.rounds)
        if k < collect:
            X, Y, _, _ = fsi.warm_start_data()
            XS.append(X)
            YS.append(Y)
        if info:
            for i in range(len(pc.plates)):
                for name, v in fsi.solid_info(i, refs).items():
                    rows[i].setdefault(name, []).append(v)
    out = dict(rows=[{k: np.array(v) for k, v in r.items()} for r in rows], iters=iters, rounds=rounds,
               pos=fsi.state(0)[0].copy())
    if precision == 0:
        f = blk.download()
        out["windows"] = [lc.window(f, k) for k in range(len(lc.windows))]
        del f
    if collect:
        out["X"] = np.concatenate(XS)
        out["Y"] = np.concatenate(YS)
    blk.close()
    return out
