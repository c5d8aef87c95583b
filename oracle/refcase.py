"""Drive the reference program (oracle/_ref/FSILBM3D) on a described case.

Test infrastructure only: the parity tests, the golden-fixture script and the
CPU-baseline leg of bench.py use it to run the reference's own Fortran path.

A case is written as the reference's inputs:
  * ``inFlow.dat`` keyword sections, in the read order of
    FlowCondition.f90:37-71 (FlowCondition), Solidbody.f90:100-161 (SolidBody),
    FluidDomain.f90:69-85 (FluidBlocks) and FlowCondition.f90:87-111 (probes);
    comment lines start with ``#`` (Util.f90:89-104 skips them).
  * ``plate.dat`` in the Frame3D layout read by SolidSolver.f90:1282-1349.
  * ``DatContinue/continue``: the stream-unformatted restart file read by
    FluidDomain.f90:139-151 (header int32 nblocks, int32 step, real8 time;
    per block real8 xmin,ymin,zmin,dh; int32 xDim,yDim,zDim; real8
    fIn(z,y,x,0:18), FluidDomain.f90:1770-1789).  On identical grids the
    restart's trilinear re-interpolation (FluidDomain.f90:172-223) has weights
    1 and 0 and reproduces the state bit for bit.

After N steps the reference writes ``DatContinue/continue<10 digits>``
(FluidDomain.f90:268-285, cadence main.f90:117); that file is the oracle's
output state: post-stream + boundary-condition (pre-collision) fIn.
"""
from __future__ import annotations

import dataclasses
import os
import re
import shutil
import subprocess
import tempfile
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BIN = os.path.join(HERE, "_ref", "FSILBM3D")


@dataclasses.dataclass
class Group:
    """One body group of the SolidBody section (Solidbody.f90:121-155)."""
    plate: str                      # plate.dat text (Frame3D), or the gmsh mesh text of a surface body
    iBodyModel: int = 1             # 1 rigid prescribed, 2 elastic beam
    iBodyType: int = 1              # 1 plate given by a line mesh, -1 surface given by a gmsh mesh
    isMotionGiven: Sequence[int] = (1, 1, 1, 1, 1, 1)
    denR: float = 1.0
    psR: float = 0.3
    KB: float = 0.1
    KS: float = 1000.0
    freq: float = 0.0
    St: float = 0.0
    firstXYZ: Sequence[float] = (0.0, 0.0, 0.0)
    deltaXYZ: Sequence[float] = (0.0, 0.0, 0.0)
    initXYZVel: Sequence[float] = (0.0, 0.0, 0.0)
    XYZAmpl: Sequence[float] = (0.0, 0.0, 0.0)
    XYZPhi: Sequence[float] = (0.0, 0.0, 0.0)
    AoAo: Sequence[float] = (0.0, 0.0, 0.0)
    AoAAmpl: Sequence[float] = (0.0, 0.0, 0.0)
    AoAPhi: Sequence[float] = (0.0, 0.0, 0.0)
    fishNum: int = 1
    numXYZ: Sequence[int] = (1, 1, 1)


@dataclasses.dataclass
class SubBlock:
    """A further fluid block (FluidBlocks section, FluidDomain.f90:69-85):
    a refined son block (dh = father dh / 2) with BCfluid (0) faces."""
    dims: Sequence[int]
    dh: float
    origin: Sequence[float]
    bc: Sequence[int] = (0, 0, 0, 0, 0, 0)
    collide: int = 1
    params: Sequence[float] = (0.0,) * 10


@dataclasses.dataclass
class Case:
    dims: Sequence[int] = (16, 8, 8)            # xDim, yDim, zDim
    dh: float = 1.0
    origin: Sequence[float] = (0.0, 0.0, 0.0)
    bc: Sequence[int] = (301, 301, 301, 301, 301, 301)
    collide: int = 1                              # iCollidModel
    params: Sequence[float] = (0.0,) * 10
    Re: float = 3.0
    denIn: float = 1.0
    uvwIn: Sequence[float] = (0.0, 0.0, 0.0)
    shearRateIn: Sequence[float] = (0.0, 0.0, 0.0)
    velocityKind: int = 0
    volumeForceIn: Sequence[float] = (0.0, 0.0, 0.0)
    volumeForceAmp: float = 0.0
    volumeForceFreq: float = 0.0
    volumeForcePhi: float = 0.0
    LrefType: int = 1
    Lref: float = 1.0
    TrefType: int = 9
    Tref: float = 20.0
    UrefType: int = 9
    Uref: float = 0.05
    ntolLBM: int = 3
    dtolLBM: float = 1e-14
    IBPenaltyAlpha: float = 1.0
    GeoGamma: float = 1.0
    NewmarkGamma: float = 0.5
    NewmarkBeta: float = 0.25
    dampK: float = 0.0
    dampM: float = 0.0
    dtolFEM: float = 1e-10
    ntolFEM: int = 20
    isKB: int = 1
    numsubstep: int = 1
    groups: List[Group] = dataclasses.field(default_factory=list)
    npsize: int = 1
    info_every: Optional[int] = None    # write DatInfo every this many steps
    flow_every: Optional[int] = None    # write DatFlow every this many steps (and at start-up)
    offsetOutput: int = 0               # cells left out of DatFlow on every face
    outputtype: int = 0                 # 1: instantaneous DatFlow files
    sons: List[SubBlock] = dataclasses.field(default_factory=list)
    interpolateScheme: int = 1

    def blocks(self):
        """(dims, dh, origin) of every block, the root first."""
        return [(tuple(self.dims), self.dh, tuple(self.origin))] + [(tuple(b.dims), b.dh, tuple(b.origin)) for b in self.sons]


def _f(x: float) -> str:
    # repr() round-trips every float64 exactly; Fortran list-directed input
    # accepts the 'e' exponent form.
    return repr(float(x))


def _v(xs) -> str:
    return " ".join(_f(x) for x in xs)


def steps_to_times(case: Case, nsteps: int):
    """timeSimTotal / timeContiDelta that make the loop (main.f90:93) run
    exactly ``nsteps`` steps and write the continue file at the last one."""
    tref = reference_tref(case)
    total = (nsteps - 0.5) * case.dh / tref
    conti = nsteps * case.dh / tref
    return total, conti, tref


def reference_tref(case: Case) -> float:
    """Tref as calculate_reference_params (Solidbody.f90:219-284) derives it,
    for the reference-type choices these cases use."""
    if case.LrefType == 0:
        raise ValueError("cases here give Lref explicitly (LrefType != 0)")
    lref = case.Lref
    if case.UrefType == 0:
        uref = abs(case.uvwIn[0])
    elif case.UrefType >= 8:
        uref = case.Uref
    else:
        raise ValueError("UrefType not supported by this writer")
    if case.TrefType == 0:
        return lref / uref
    if case.TrefType >= 2:
        return case.Tref
    raise ValueError("TrefType not supported by this writer")


def inflow_text(case: Case, nsteps: int, restart: bool = True) -> str:
    total, conti, tref = steps_to_times(case, nsteps)
    big = 1.0e9
    if not restart:
        conti = big
    info = big if case.info_every is None else case.info_every * case.dh / tref
    flow = big if case.flow_every is None else case.flow_every * case.dh / tref
    wbegin = big if case.flow_every is None else 0.0
    L = []
    a = L.append
    a("# generated case")
    a("Parallel")
    a(str(case.npsize))
    a("FlowCondition")
    a("# isConCmpt numsubstep")
    a(f"{1 if restart else 0} {case.numsubstep}")
    a("# timeSimTotal timeContiDelta")
    a(f"{_f(total)} {_f(conti)}")
    a("# timeWriteBegin timeWriteEnd")
    a(f"{_f(wbegin)} {_f(big + 1.0)}")
    a("# timeFlowDelta timeBodyDelta timeInfoDelta")
    a(f"{_f(flow)} {_f(big)} {_f(info)}")
    a("# Re denIn")
    a(f"{_f(case.Re)} {_f(case.denIn)}")
    a("# uvwIn")
    a(_v(case.uvwIn))
    a("# shearRateIn velocityKind")
    a(f"{_v(case.shearRateIn)} {case.velocityKind}")
    a("# volumeForceIn")
    a(_v(case.volumeForceIn))
    a("# volumeForceAmp Freq Phi")
    a(f"{_f(case.volumeForceAmp)} {_f(case.volumeForceFreq)} {_f(case.volumeForcePhi)}")
    a("# LrefType Lref")
    a(f"{case.LrefType} {_f(case.Lref)}")
    a("# TrefType Tref")
    a(f"{case.TrefType} {_f(case.Tref)}")
    a("# UrefType Uref")
    a(f"{case.UrefType} {_f(case.Uref)}")
    a("# ntolLBM dtolLBM")
    a(f"{case.ntolLBM} {_f(case.dtolLBM)}")
    a("# interpolateScheme")
    a(str(case.interpolateScheme))
    a("SolidBody")
    a(f"{_f(case.IBPenaltyAlpha)} {_f(case.GeoGamma)}")
    a(f"{_f(case.NewmarkGamma)} {_f(case.NewmarkBeta)}")
    a(f"{_f(case.dampK)} {_f(case.dampM)}")
    a(f"{_f(case.dtolFEM)} {case.ntolFEM}")
    nfish = sum(g.fishNum for g in case.groups)
    a(f"{nfish} {len(case.groups)} {case.isKB}")
    for gi, g in enumerate(case.groups):
        a(f"{g.fishNum} {g.numXYZ[0]} {g.numXYZ[1]} {g.numXYZ[2]}")
        a(f"plate{gi + 1}.dat")
        a(f"{g.iBodyModel} {g.iBodyType}")
        a(" ".join(str(int(v)) for v in g.isMotionGiven[:3]))
        a(" ".join(str(int(v)) for v in g.isMotionGiven[3:]))
        a(f"{_f(g.denR)} {_f(g.psR)}")
        a(f"{_f(g.KB)} {_f(g.KS)}")
        a(f"{_f(g.freq)} {_f(g.St)}")
        a(_v(g.firstXYZ))
        a(_v(g.deltaXYZ))
        a(_v(g.initXYZVel))
        a(_v(g.XYZAmpl))
        a(_v(g.XYZPhi))
        a(_v(g.AoAo))
        a(_v(g.AoAAmpl))
        a(_v(g.AoAPhi))
        if gi < len(case.groups) - 1:
            a("=")
    a("FluidBlocks")
    a(str(1 + len(case.sons)))
    blocks = [(1, case.collide, case.dims, case.dh, case.origin, case.bc, case.params)]
    for i, b in enumerate(case.sons):
        blocks.append((2 + i, b.collide, b.dims, b.dh, b.origin, b.bc, b.params))
    for k, (bid, col, dims, dh, org, bc, prm) in enumerate(blocks):
        a("# ID iCollideModel offsetOutput outputtype")
        a(f"{bid} {col} {case.offsetOutput if k == 0 else 0} {case.outputtype if k == 0 else 0}")
        a("# xDim yDim zDim")
        a(" ".join(str(int(d)) for d in dims))
        a("# dh xmin ymin zmin")
        a(f"{_f(dh)} {_v(org)}")
        a("# BndConds")
        a(" ".join(str(int(v)) for v in bc))
        a("# params")
        a(_v(prm))
        if k < len(blocks) - 1:
            a("=")
    a("ProbingFluid")
    a("0 1")
    a("ProbingSolid")
    a("0")
    return "\n".join(L) + "\n"


def plate_text(nodes_xyz, lspan, rspan, direction, elements_nspan, constraint_first=True) -> str:
    """Frame3D plate file: nodes on a line, 2-node beam elements
    (SolidSolver.f90:1305-1369 read order; element line
    ``id i j k type mat Nspan``)."""
    n = len(nodes_xyz)
    ne = n - 1
    L = ["plate", f"{n} {ne} 1", "POINT", f"{n}"]
    for i, p in enumerate(nodes_xyz):
        L.append(f"{i + 1} {_v(p)} {_f(lspan)} {_f(rspan)} {_v(direction)}")
    L += ["ELEMENT", f"{ne}"]
    for e in range(ne):
        L.append(f"{e + 1} {e + 1} {e + 2} {e + 2} 2 1 {int(elements_nspan)}")
    L += ["CONSTRAINT", f"{n}"]
    for i in range(n):
        c = "1 1 1 1 1 1" if (i == 0 and constraint_first) else "0 0 0 0 0 0"
        L.append(f"{i + 1} {c}")
    L += ["MATERIAL", "1", "1 1 1 1 1 0 1 1 1", "END"]
    return "\n".join(L) + "\n"


def write_continue(path: str, fs, case: Case, step: int = 0, time_over_tref: float = 0.0):
    """fs: one state per block (root first), each of shape (19, xDim, yDim,
    zDim) in C order = the Fortran fIn(z,y,x,q) memory order."""
    if isinstance(fs, np.ndarray):
        fs = [fs]
    blocks = case.blocks()
    assert len(fs) == len(blocks)
    with open(path, "wb") as fh:
        fh.write(np.array([len(blocks), step], dtype="<i4").tobytes())
        fh.write(np.array([time_over_tref], dtype="<f8").tobytes())
        for f, (dims, dh, org) in zip(fs, blocks):
            xd, yd, zd = dims
            assert f.shape == (19, xd, yd, zd) and f.dtype == np.float64
            fh.write(np.array([org[0], org[1], org[2], dh], dtype="<f8").tobytes())
            fh.write(np.array([xd, yd, zd], dtype="<i4").tobytes())
            fh.write(np.ascontiguousarray(f).astype("<f8").tobytes())


def read_continue(path: str, dims):
    """Returns (step, time, f) for a single-block file; for several blocks f
    is the list of states (dims: list of block dims)."""
    raw = open(path, "rb").read()
    nb, step = np.frombuffer(raw[:8], dtype="<i4")
    time = np.frombuffer(raw[8:16], dtype="<f8")[0]
    multi = isinstance(dims[0], (tuple, list))
    dl = list(dims) if multi else [dims]
    off = 16
    out = []
    for d in dl:
        xd, yd, zd = d
        off += 32 + 12
        n = 19 * xd * yd * zd
        out.append(np.frombuffer(raw[off:off + n * 8], dtype="<f8").reshape(19, xd, yd, zd).copy())
        off += n * 8
    return int(step), float(time), (out if multi else out[0])


@dataclasses.dataclass
class RefResult:
    f: np.ndarray
    step: int
    time: float
    stdout: str
    workdir: str
    step_seconds: List[float]

    def info_file(self, name: str) -> np.ndarray:
        return np.loadtxt(os.path.join(self.workdir, "DatInfo", name), ndmin=2)


def run_reference(case: Case, f0: np.ndarray, nsteps: int, keep: bool = False,
                  plates: Optional[List[str]] = None, timeout: float = 600.0,
                  env_extra: Optional[dict] = None, start_time_over_tref: float = 0.0,
                  binary: Optional[str] = None) -> RefResult:
    """Run the reference binary (or another build of the reference program
    that reads the same inputs, ``binary``) for ``nsteps`` steps from state
    ``f0``."""
    exe = binary or REF_BIN
    if not os.path.exists(exe):
        raise FileNotFoundError(f"{exe} missing: run oracle/build_ref.sh (or the program's own build)")
    wd = tempfile.mkdtemp(prefix="fsilref_")
    try:
        for d in ("DatContinue", "DatInfo", "DatBody", "DatBodySpan", "DatFlow"):
            os.makedirs(os.path.join(wd, d))
        with open(os.path.join(wd, "inFlow.dat"), "w") as fh:
            fh.write(inflow_text(case, nsteps))
        for gi, g in enumerate(case.groups):
            # a surface body's mesh is plate<k>.msh; the reference writes the .dat itself
            ext = "msh" if g.iBodyType == -1 else "dat"
            with open(os.path.join(wd, f"plate{gi + 1}.{ext}"), "w") as fh:
                fh.write(g.plate)
        write_continue(os.path.join(wd, "DatContinue", "continue"), f0, case, 0, start_time_over_tref)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(case.npsize)
        if env_extra:
            env.update(env_extra)
        p = subprocess.run([exe], cwd=wd, capture_output=True, text=True, timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError(f"reference failed rc={p.returncode}\n{p.stdout[-3000:]}\n{p.stderr[-3000:]}")
        outs = sorted(x for x in os.listdir(os.path.join(wd, "DatContinue")) if x != "continue")
        if not outs:
            raise RuntimeError("reference wrote no continue file\n" + p.stdout[-3000:])
        dims = [b[0] for b in case.blocks()] if case.sons else case.dims
        step, time, f = read_continue(os.path.join(wd, "DatContinue", outs[-1]), dims)
        secs = [float(m) for m in re.findall(r"Time for\s+one step:\s*([-+0-9.Ee]+)", p.stdout)]
        return RefResult(f=f, step=step, time=time, stdout=p.stdout, workdir=wd, step_seconds=secs)
    finally:
        if not keep:
            shutil.rmtree(wd, ignore_errors=True)


def time_reference_steps(case: Case, nsteps: int, threads: int, timeout: float = 900.0):
    """Run the reference from its own uniform initial state (no restart file,
    no outputs) for ``nsteps`` steps with ``threads`` OpenMP threads; returns
    the per-step wall times it logs ("Time for one step", main.f90:145-146)."""
    if not os.path.exists(REF_BIN):
        raise FileNotFoundError(f"{REF_BIN} missing: run oracle/build_ref.sh")
    c = dataclasses.replace(case, npsize=int(threads))
    wd = tempfile.mkdtemp(prefix="fsilrefb_")
    try:
        for d in ("DatContinue", "DatInfo", "DatBody", "DatBodySpan", "DatFlow"):
            os.makedirs(os.path.join(wd, d))
        with open(os.path.join(wd, "inFlow.dat"), "w") as fh:
            fh.write(inflow_text(c, nsteps, restart=False))
        for gi, g in enumerate(c.groups):
            ext = "msh" if g.iBodyType == -1 else "dat"
            with open(os.path.join(wd, f"plate{gi + 1}.{ext}"), "w") as fh:
                fh.write(g.plate)
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(threads)
        env.setdefault("OMP_PROC_BIND", "close")
        p = subprocess.run([REF_BIN], cwd=wd, capture_output=True, text=True, timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError(f"reference failed rc={p.returncode}\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
        secs = [float(m) for m in re.findall(r"Time for\s+one step:\s*([-+0-9.Ee]+)", p.stdout)]
        return secs
    finally:
        shutil.rmtree(wd, ignore_errors=True)
