"""Build the reference program as a GPU host of libfsil.so (the drop-in check).

The reference's own Fortran sources under /root/reference/src are compiled
as oracle/build_ref.sh compiles them, with main.f90 changed at four call sites
(fsil_host.f90 describes them) and two more objects: include/fsil_bridge.f90
(the ISO_C_BINDING layer) and tests/fortran_host/fsil_host.f90 (the glue).
The sources are copied into _build/src only for the compiler and removed
again; what stays is _build/obj and the program _build/FSILBM3D_fsil, linked
against fsilbm3d_amd/libfsil.so through a relative rpath.  Without
/root/reference nothing is built and a prebuilt program is kept.

Run:  python tests/fortran_host/build_host.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BUILD = os.path.join(HERE, "_build")
HOST_BIN = os.path.join(BUILD, "FSILBM3D_fsil")
MODULES = ["ConstParams", "FlowCondition", "SolidSolver", "Solidbody", "FluidDomain", "LBMBlockComm", "Util"]
FFLAGS = ["-O3", "-fopenmp", "-ffree-form", "-ffp-contract=off"]


def _once(text, anchor, new, what):
    n = text.count(anchor)
    if n != 1:
        raise RuntimeError(f"build_host: {what}: expected one '{anchor.strip()}' in main.f90, found {n}")
    return text.replace(anchor, new)


def patch_main(text: str) -> str:
    """The four call sites of main.f90 redirected to the fsil_host module."""
    text = _once(text, "    USE LBMBlockComm\n", "    USE LBMBlockComm\n    USE fsil_host\n", "module use")
    text = _once(text, "    call update_volume_force_blocks()\n",
                 "    call fsil_host_init()\n    call update_volume_force_blocks()\n", "start-up")
    step = "call tree_collision_streaming_IBM_FEM(blockTreeRoot,"
    text = _once(text, step, "call fsil_host_step(blockTreeRoot,", "time step")
    i = text.index("call fsil_host_step(blockTreeRoot,")
    macro = "call calculate_macro_quantities_blocks()"
    j = text.index(macro, i)
    text = text[:j] + "call fsil_host_macro()" + text[j + len(macro):]
    text = _once(text, "            call write_continue_blocks(step,",
                 "            call fsil_host_fin()\n            call write_continue_blocks(step,", "continue output")
    return text


def build(ref: str = None, verbose: bool = False) -> str:
    ref = ref or os.environ.get("FSIL_REFERENCE", "/root/reference")
    src_dir = os.path.join(ref, "src")
    if not os.path.isdir(src_dir):
        print(f"build_host: {src_dir} not found; keeping any prebuilt {HOST_BIN}", file=sys.stderr)
        return HOST_BIN
    fc = os.environ.get("FC", "amdflang")
    src, obj = os.path.join(BUILD, "src"), os.path.join(BUILD, "obj")
    os.makedirs(obj, exist_ok=True)
    os.makedirs(src, exist_ok=True)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("build_host failed:\n" + " ".join(cmd) + "\n" + p.stdout[-4000:] + p.stderr[-4000:])

    try:
        for m in MODULES:
            shutil.copyfile(os.path.join(src_dir, m + ".f90"), os.path.join(src, m + ".f90"))
        with open(os.path.join(src_dir, "main.f90")) as fh:
            main = patch_main(fh.read())
        with open(os.path.join(src, "main.f90"), "w") as fh:
            fh.write(main)
        units = [os.path.join(src, m + ".f90") for m in MODULES]
        units += [os.path.join(ROOT, "include", "fsil_bridge.f90"), os.path.join(HERE, "fsil_host.f90"),
                  os.path.join(src, "main.f90")]
        objs = []
        for u in units:
            o = os.path.join(obj, os.path.splitext(os.path.basename(u))[0] + ".o")
            run([fc] + FFLAGS + ["-module-dir", obj, "-I", obj, "-c", u, "-o", o])
            objs.append(o)
        ft = os.path.join(obj, "forkthread.o")
        run(["gcc", "-O3", "-c", os.path.join(src_dir, "forkthread.c"), "-o", ft])
        objs.append(ft)
        tmp = HOST_BIN + ".tmp"
        run([fc, "-fopenmp", "-o", tmp] + objs +
            ["-L", os.path.join(ROOT, "fsilbm3d_amd"), "-lfsil", "-Wl,-rpath,$ORIGIN/../../../fsilbm3d_amd"])
        os.replace(tmp, HOST_BIN)
    finally:
        shutil.rmtree(src, ignore_errors=True)
    return HOST_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
