"""The C-ABI library loads and exports every symbol include/dcsr.h declares
(no compute calls: no GPU is needed)."""
import ctypes
import os
import subprocess

import pytest

import dcsr_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_entry_points():
    names = dcsr_amd.header_functions()
    for must in ("dcsr_create", "dcsr_destroy", "dcsr_load_coo", "dcsr_load_csr", "dcsr_export_csr", "dcsr_update",
                 "dcsr_compact", "dcsr_rebin", "dcsr_spmv", "dcsr_query", "dcsr_csr_spmv"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(dcsr_amd.LIB_PATH):
        pytest.fail("libdcsr.so not built (run python -m dcsr_amd.build)")
    L = ctypes.CDLL(dcsr_amd.LIB_PATH)
    missing = [n for n in dcsr_amd.header_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", dcsr_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T dcsr_" in ln}
    undeclared = exported - set(dcsr_amd.header_functions())
    assert not undeclared, f"exported but not declared: {undeclared}"


def test_status_strings():
    L = dcsr_amd.lib()
    assert L.dcsr_status_string(0) == b"ok"
    assert b"pool" in L.dcsr_status_string(3)


def test_invalid_arguments_rejected_without_gpu():
    L = dcsr_amd.lib()
    assert L.dcsr_destroy(None) == 1
    assert L.dcsr_create(None, 1, 1, None) == 1


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "dcsr.h"\nint main(void){ dcsr_info i; (void)i; return (int)DCSR_OK; }\n')
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", inc, "-c", str(src), "-o", str(tmp_path / "t.o")],
                   check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-I", inc, "-c", str(src), "-o",
                    str(tmp_path / "t2.o")], check=True)


def test_reference_style_driver_compiles_against_the_shim(tmp_path):
    """A caller written like the reference's drivers (every device:: entry
    point of include/dcsr_matrix.hpp: DCSR, CSR and HYB overloads) compiles
    and links against libdcsr.so."""
    from dcsr_amd import build as B
    out = str(tmp_path / "driver")
    r = subprocess.run(B.driver_cmd(out), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.getsize(out) > 0
