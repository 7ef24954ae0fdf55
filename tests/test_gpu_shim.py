"""GPU parity through the C++ drop-in shim (include/dcsr_matrix.hpp): the
reference-style driver tests/shim/ref_style_driver (built by
__graft_entry__.build()) runs the banded 256 x 256 flow of the reference's
drivers -- load, four insert-or-add batches, SpMV, CSR SpMV, SpGEMM,
defragmentation, DCSR -> HYB, HYB SpMV, HYB append update, HYB -> DCSR --
and every output is compared with the golden fixtures and the oracle:
structure and update values bit-exact, SpMV within 1e-5 of FP64, SpGEMM and
the HYB store bit-exact against the host path's order."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from tests.helpers import assert_same_structure, fp64_spmv_check

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "shim", "ref_style_driver")


def rd(d, name, dt):
    return np.fromfile(os.path.join(d, name), dtype=dt)


def csr_out(d, name, R, C):
    return O.Csr(R, C, rd(d, f"{name}_off.i32", np.int32), rd(d, f"{name}_col.i32", np.int32),
                 rd(d, f"{name}_val.f32", np.float32))


def hyb_out(d, name, R, C, width):
    pitch = (R + 31) // 32 * 32
    return O.Hyb(R, C, width, pitch, rd(d, f"{name}_rs.i32", np.int32), rd(d, f"{name}_ell_col.i32", np.int32),
                 rd(d, f"{name}_ell_val.f32", np.float32), rd(d, f"{name}_coo_row.i32", np.int32),
                 rd(d, f"{name}_coo_col.i32", np.int32), rd(d, f"{name}_coo_val.f32", np.float32))


def same_hyb(a, b):
    for f in ("row_sizes", "ell_col", "coo_row", "coo_col"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
    for f in ("ell_val", "coo_val"):
        np.testing.assert_array_equal(getattr(a, f).view(np.uint32), getattr(b, f).view(np.uint32), err_msg=f)


def test_reference_style_driver_banded256(tmp_path):
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"
    g = np.load(os.path.join(HERE, "golden", "banded256.npz"))
    R = int(g["R"])
    width = int(g["hyb_width"])
    d = str(tmp_path)
    np.array([R, 4, width], np.int32).tofile(os.path.join(d, "meta.i32"))
    for k in ("in", "b0", "b1", "b2", "b3"):
        g[f"{k}_row"].astype(np.int32).tofile(os.path.join(d, f"{k}_row.i32"))
        g[f"{k}_col"].astype(np.int32).tofile(os.path.join(d, f"{k}_col.i32"))
        g[f"{k}_val"].astype(np.float32).tofile(os.path.join(d, f"{k}_val.f32"))
    x = g["x"].astype(np.float32)
    x.tofile(os.path.join(d, "x.f32"))
    print("[gpu-test] running the reference-style driver", flush=True)
    sys.stdout.flush()
    r = subprocess.run([DRIVER, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    gold = lambda p: O.Csr(R, R, g[f"{p}_off"], g[f"{p}_col"], g[f"{p}_val"])  # noqa: E731
    assert_same_structure(csr_out(d, "load", R, R), gold("load"))
    for k in range(4):
        assert_same_structure(csr_out(d, f"upd{k}", R, R), gold(f"upd{k}"))
    A = gold("upd3")
    fp64_spmv_check(A, x, rd(d, "y.f32", np.float32))
    fp64_spmv_check(A, x, rd(d, "ycsr.f32", np.float32))
    Ac = A.canonical()
    assert_same_structure(csr_out(d, "spgemm", R, R), gold("spgemm"), values=None)
    assert_same_structure(csr_out(d, "spgemm", R, R), O.spgemm(Ac, Ac))
    assert_same_structure(csr_out(d, "compact", R, R), A)
    H = O.csr_to_hyb_full(Ac, width)
    same_hyb(hyb_out(d, "hyb", R, R, width), H)
    np.testing.assert_array_equal(rd(d, "yhyb.f32", np.float32).view(np.uint32), O.hyb_spmv(H, x).view(np.uint32))
    H2 = O.hyb_update(H, g["b0_row"], g["b0_col"], g["b0_val"])
    same_hyb(hyb_out(d, "hupd", R, R, width), H2)
    np.testing.assert_array_equal(rd(d, "yhupd.f32", np.float32).view(np.uint32), O.hyb_spmv(H2, x).view(np.uint32))
    hr, hc, hv = H2.to_coo()
    assert_same_structure(csr_out(d, "hback", R, R), O.coo_to_csr(R, R, hr, hc, hv))
