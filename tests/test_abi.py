"""The C-ABI library loads and exports every entry point include/fvm_gpu.h
declares; argument checks that need no GPU return error codes, never abort."""
import ctypes
import os
import re
import subprocess

import pytest

import fvmgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fvm_gpu.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvm_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("fvm_ctx_create", "fvm_scheme_set", "fvm_block_create", "fvm_upload",
              "fvm_download", "fvm_step", "fvm_norms", "fvm_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = fvmgpu.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(fvmgpu.EXPORTS) <= set(declared())


def test_version_and_null_arguments():
    lib = fvmgpu.load_library()
    assert b"gfx950" in lib.fvm_version()
    assert lib.fvm_scheme_set(None, 201, 251, 300, 101, 0.3, 0.2, 1.4, 1) == -1
    assert b"null context" in lib.fvm_last_error()
    assert lib.fvm_block_set_bc(None, 0, 0) == -1
    assert lib.fvm_upload(None, None, 0) == -1
    assert lib.fvm_step(None, 1, 1e-3, 0.0, None) == -1
    assert lib.fvm_norms(None, None, None) == -1
    assert lib.fvm_ctx_destroy(None) == 0
    out = ctypes.c_void_p()
    assert lib.fvm_ctx_create(0, None) == -1


def test_scheme_codes_match_reference_and_oracle():
    import _oracle as O
    for name in ("RK_SSP3", "RK4", "SIMPLE", "MUSCL", "WENO", "LIM_NONE", "LIM_MINMOD",
                 "LIM_VANLEER", "LIM_VANALBADA", "ROE", "VANLEER", "AUSM", "AUSMPLUS"):
        assert getattr(fvmgpu, name) == getattr(O, name)
    # spellings of the case file (src/defs/type_defs_scheme.f90:41-84)
    assert fvmgpu.SCHEME_INDEX["minmod_"] == 251 and fvmgpu.SCHEME_INDEX["Roe"] == 300
    hdr = open(HEADER).read()
    for macro, val in (("FVM_MUSCL", 201), ("FVM_WENO", 202), ("FVM_ROE", 300),
                       ("FVM_AUSM", 303), ("FVM_LIM_VANALBADA", 253), ("FVM_RK_SSP3", 101)):
        assert re.search(rf"#define {macro}\s+{val}\b", hdr)


def test_request_struct_layout_matches_header(tmp_path):
    """fvmgpu.Request (ctypes) has the size and field offsets of the C
    compiler's fvm_request."""
    fields = [f[0] for f in fvmgpu.Request._fields_]
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "fvm_gpu.h"\n'
        "int main(void) {\n"
        '  printf("size %zu\\n", sizeof(fvm_request));\n'
        + "".join(f'  printf("{f} %zu\\n", offsetof(fvm_request, {f}));\n' for f in fields)
        + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True,
                                                       capture_output=True, text=True)
               .stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(fvmgpu.Request)
    for f in fields:
        assert int(out[f]) == getattr(fvmgpu.Request, f).offset, f


def test_fortran_request_layout(tmp_path):
    """The Fortran module's bind(C) fvm_request has the C struct's size and
    field offsets (compared with the ctypes structure, itself checked against
    the C compiler above)."""
    import shutil
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler")
    mod = os.path.join(ROOT, "fvm-3d-f90_amd", "fortran", "fvm_gpu_mod.f90")
    fields = ["n", "spacing", "recon", "kappa", "ref_compat", "bc", "t_end", "q0", "borrow",
              "ic_kind", "ic_nparams", "ic_params", "ic_origin"]
    lines = ["program layout", "  use iso_c_binding", "  use fvm_gpu", "  implicit none",
             "  type(fvm_request), target :: r",
             "  integer(c_intptr_t) :: b",
             "  b = transfer(c_loc(r), b)",
             "  print '(a, i0)', 'size ', c_sizeof(r)"]
    for f in fields:
        lines.append(f"  print '(a, i0)', '{f} ', transfer(c_loc(r%{f}), b) - b")
    lines.append("end program layout")
    src = tmp_path / "layout.f90"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([fc, "-O0", "-module-dir", str(tmp_path), "-c", mod, "-o", str(tmp_path / "m.o")],
                   check=True, capture_output=True)
    lib = os.path.join(ROOT, "fvm-3d-f90_amd", "lib")   # the module's helpers name the library
    subprocess.run([fc, "-O0", "-I", str(tmp_path), str(src), str(tmp_path / "m.o"), "-o", str(exe),
                    "-L" + lib, "-lfvmgpu", "-Wl,-rpath," + lib], check=True, capture_output=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.strip().splitlines())
    assert int(out["size"]) == ctypes.sizeof(fvmgpu.Request)
    for f in fields:
        assert int(out[f]) == getattr(fvmgpu.Request, f).offset, f


def test_bc_codes_agree_across_header_fortran_python_and_oracle():
    """Every FVM_BC_* code of the header has the same value in the Fortran
    module, the Python package and, for the kinds the oracle applies, the
    oracle (the reference's codes, src/defs/type_defs_bc.f90:11-30)."""
    import _oracle as O
    hdr = dict((k, int(v)) for k, v in re.findall(r"#define (FVM_BC_\w+)\s+(\d+)", open(HEADER).read()))
    fmod = open(os.path.join(ROOT, "fvm-3d-f90_amd", "fortran", "fvm_gpu_mod.f90")).read()
    fort = dict((k, int(v)) for k, v in re.findall(r"(FVM_BC_\w+)\s*=\s*(\d+)", fmod))
    assert len(hdr) >= 15
    assert hdr == fort
    for name, val in hdr.items():
        assert getattr(fvmgpu, name[4:]) == val, name
    assert hdr["FVM_BC_WALL"] == O.BC_WALL and hdr["FVM_BC_WALL_VISCOUS"] == O.BC_NOSLIP
    assert hdr["FVM_BC_SYMMETRY"] == O.BC_SYMMETRY and hdr["FVM_BC_PERIODIC"] == O.BC_PERIODIC
