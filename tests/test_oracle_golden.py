"""The oracle restatement reproduces the reference's own outputs (committed
fixtures written by tests/golden/make_golden.py from the cusp host path),
bit for bit and in stored order."""
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def csr(d, prefix, R, C):
    return O.Csr(R, C, d[f"{prefix}_off"], d[f"{prefix}_col"], d[f"{prefix}_val"])


def same(a: O.Csr, b: O.Csr):
    np.testing.assert_array_equal(a.off, b.off)
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.val.view(np.uint32), b.val.view(np.uint32))


@pytest.fixture(scope="module")
def banded():
    return load("banded256.npz")


@pytest.fixture(scope="module")
def edge():
    return load("edge.npz")


def test_banded_load(banded):
    R = int(banded["R"])
    A = O.coo_to_csr(R, R, banded["in_row"], banded["in_col"], banded["in_val"])
    same(A, csr(banded, "load", R, R))


def test_banded_updates(banded):
    R = int(banded["R"])
    A = csr(banded, "load", R, R)
    for i in range(4):
        A = O.update(A, banded[f"b{i}_row"], banded[f"b{i}_col"], banded[f"b{i}_val"])
        same(A, csr(banded, f"upd{i}", R, R))


def test_banded_spmv(banded):
    R = int(banded["R"])
    A = csr(banded, "upd3", R, R)
    y = O.spmv(A, banded["x"])
    np.testing.assert_array_equal(y.view(np.uint32), banded["y"].view(np.uint32))


def test_banded_spgemm(banded):
    R = int(banded["R"])
    A = csr(banded, "upd3", R, R)
    same(O.spgemm(A, A), csr(banded, "spgemm", R, R))


def test_banded_hyb(banded):
    R = int(banded["R"])
    A = csr(banded, "upd3", R, R)
    ec, ev, cr, cc, cv, pitch = O.csr_to_hyb(A, int(banded["hyb_width"]))
    assert pitch == int(banded["hyb_pitch"])
    np.testing.assert_array_equal(ec, banded["hyb_ell_col"])
    np.testing.assert_array_equal(ev.view(np.uint32), banded["hyb_ell_val"].view(np.uint32))
    np.testing.assert_array_equal(cr, banded["hyb_coo_row"])
    np.testing.assert_array_equal(cc, banded["hyb_coo_col"])
    np.testing.assert_array_equal(cv.view(np.uint32), banded["hyb_coo_val"].view(np.uint32))


def test_edge_load_keeps_duplicates_and_zeros(edge):
    R, C = int(edge["R"]), int(edge["C"])
    A = O.coo_to_csr(R, C, edge["in_row"], edge["in_col"], edge["in_val"])
    same(A, csr(edge, "load", R, C))
    assert A.nnz == edge["in_row"].size
    assert (A.val == 0).any()


def test_edge_canonicalize_and_update(edge):
    R, C = int(edge["R"]), int(edge["C"])
    A = csr(edge, "load", R, C)
    E = O.Csr(R, C, np.zeros(R + 1, np.int32), [], [])
    A1 = O.csr_add(A, E)
    same(A1, csr(edge, "canon", R, C))
    assert not (A1.val == 0).any()
    A2 = O.update(A1, edge["b0_row"], edge["b0_col"], edge["b0_val"])
    same(A2, csr(edge, "upd0", R, C))


def test_edge_spgemm_and_spmv(edge):
    R, C = int(edge["R"]), int(edge["C"])
    A = csr(edge, "upd0", R, C)
    same(O.spgemm(A, A), csr(edge, "spgemm", R, C))
    y = O.spmv(A, edge["x"])
    np.testing.assert_array_equal(y.view(np.uint32), edge["y"].view(np.uint32))
