"""The C-ABI library: it loads, exports exactly the entry points that
include/fsilbm3d_amd.h declares, and fails loudly when it is missing.  No
compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

from fsilbm3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fsilbm3d_amd.h")


def header_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsil_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_table():
    assert header_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fsil_[a-z0-9_]+)", out))
    assert exported == set(header_symbols())


def test_binding_loads_and_reports_version():
    lib = _lib.load()
    assert lib.fsil_version() == 1
    assert isinstance(lib.fsil_last_error(), bytes)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.FsilError):
        _lib.load()


def test_fortran_binding_compiles():
    """include/fsil_bridge.f90 (the ISO_C_BINDING module a Fortran host uses)
    compiles with the image's Fortran compiler."""
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not present")
    src = os.path.join(ROOT, "include", "fsil_bridge.f90")
    out = subprocess.run([fc, "-c", "-ffree-form", src, "-module-dir", "/tmp", "-o", "/tmp/fsil_bridge.o"],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr


def test_fortran_coupling_call_with_absent_arguments_compiles(tmp_path):
    """A host with its bodies on the device beam solver calls the coupling
    with the marker arrays and the unwanted outputs omitted (they arrive as
    null pointers); the call compiles against the module."""
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not present")
    src = os.path.join(ROOT, "include", "fsil_bridge.f90")
    mod = str(tmp_path)
    out = subprocess.run([fc, "-c", "-ffree-form", src, "-module-dir", mod, "-o", os.path.join(mod, "b.o")],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    prog = tmp_path / "host.f90"
    prog.write_text(
        "subroutine device_bodies_step(h, dt, t, nsub)\n"
        "    use iso_c_binding\n"
        "    use fsil_bridge\n"
        "    integer(c_int), intent(in) :: h, nsub\n"
        "    real(c_double), intent(in) :: dt, t\n"
        "    integer(c_int) :: it\n"
        "    real(c_double) :: dm\n"
        "    call fsil_check(fsil_fem_update_markers(h), 'markers')\n"
        "    call fsil_check(fsil_ibm_interaction_force(h, dt), 'ibm')\n"
        "    call fsil_check(fsil_fem_substeps(h, t, nsub, dt), 'fem')\n"
        "    call fsil_check(fsil_ibm_report(h, it, dm), 'report')\n"
        "    call fsil_check(fsil_ibm_interaction_force(h, dt, iters=it, dmax=dm), 'ibm with report')\n"
        "end subroutine\n")
    out = subprocess.run([fc, "-c", "-ffree-form", str(prog), "-module-dir", mod, "-I", mod,
                          "-o", os.path.join(mod, "host.o")], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr


def test_fortran_binding_covers_every_entry_point():
    """Every function the header declares has a bind(C) interface in the
    Fortran module, under the same name."""
    import re
    hdr = open(os.path.join(ROOT, "include", "fsilbm3d_amd.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(fsil_\w+)\s*\(", hdr, re.M))
    f90 = open(os.path.join(ROOT, "include", "fsil_bridge.f90")).read()
    bound = set(re.findall(r"name='(fsil_\w+)'", f90))
    assert declared and declared == bound, (declared - bound, bound - declared)
