"""The oracle restatement against the cusp host path compiled from the
reference checkout (oracle/_ref/libdcsr_ref.so) on randomized inputs:
bit-exact, stored order included.  Skipped where _ref was not built."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import random_coo

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference checkout)")


def same(a, b):
    np.testing.assert_array_equal(a.off, b.off)
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.val.view(np.uint32), b.val.view(np.uint32))


CASES = [
    # R, C, n, dup_frac, zero_frac, values
    (1, 1, 5, 0.5, 0.2, "small_int"),
    (17, 9, 60, 0.3, 0.1, "sym"),
    (64, 64, 800, 0.2, 0.05, "small_int"),
    (300, 200, 5000, 0.1, 0.0, "sym"),
    (1000, 1000, 20000, 0.05, 0.01, "pos"),
]


@pytest.mark.parametrize("R,C,n,dup,zero,vals", CASES)
def test_sort_and_load(R, C, n, dup, zero, vals):
    rng = np.random.default_rng(R * 7 + n)
    r, c, v = random_coo(rng, R, C, n, dup, zero, vals)
    a = O.sort_by_row(r, c, v)
    b = O.sort_by_row(r, c, v, use_ref=True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32))
    same(O.coo_to_csr(R, C, r, c, v), O.coo_to_csr(R, C, r, c, v, use_ref=True))


@pytest.mark.parametrize("R,C,n,dup,zero,vals", CASES)
def test_update_chain(R, C, n, dup, zero, vals):
    rng = np.random.default_rng(R + n)
    r, c, v = random_coo(rng, R, C, n, dup, zero, vals)
    A = O.coo_to_csr(R, C, r, c, v)
    for k in range(3):
        br, bc, bv = random_coo(rng, R, C, max(1, n // 3), dup, zero, vals)
        if k == 1 and A.nnz:  # deletions by -current
            rows_of = np.repeat(np.arange(R), np.diff(A.off))
            pick = rng.integers(0, A.nnz, max(1, A.nnz // 10))
            br = np.concatenate([br, rows_of[pick]]).astype(np.int32)
            bc = np.concatenate([bc, A.col[pick]]).astype(np.int32)
            bv = np.concatenate([bv, -A.val[pick]]).astype(np.float32)
        mine = O.update(A, br, bc, bv)
        ref = O.update(A, br, bc, bv, use_ref=True)
        same(mine, ref)
        A = mine


@pytest.mark.parametrize("R,C,n,dup,zero,vals", CASES)
def test_spmv_spgemm_hyb(R, C, n, dup, zero, vals):
    rng = np.random.default_rng(3 * R + n)
    r, c, v = random_coo(rng, R, R, n, dup, zero, vals)  # square for A*A
    A = O.coo_to_csr(R, R, r, c, v)
    x = rng.uniform(-1, 1, R).astype(np.float32)
    np.testing.assert_array_equal(O.spmv(A, x).view(np.uint32), O.spmv(A, x, use_ref=True).view(np.uint32))
    same(O.spgemm(A, A), O.spgemm(A, A, use_ref=True))
    for width in (1, 3, 8):
        got = O.csr_to_hyb(A, width)
        ref = O.csr_to_hyb(A, width, use_ref=True)
        assert got[5] == ref[5]
        for g, w in zip(got[:5], ref[:5]):
            np.testing.assert_array_equal(np.asarray(g).view(np.uint32), np.asarray(w).view(np.uint32))


@pytest.mark.parametrize("R,C,n,dup,zero,vals", CASES)
def test_hyb_spmv_after_appends(R, C, n, dup, zero, vals):
    """The oracle's HYB SpMV equals cusp::multiply on a host hyb_matrix bit
    for bit, also after append updates that leave duplicates in both parts."""
    rng = np.random.default_rng(5 * R + n)
    r, c, v = random_coo(rng, R, C, n, dup, zero, vals)
    for width in (1, 4, 9):
        H = O.coo_to_hyb(R, C, r, c, v, width)
        Hr = O.coo_to_hyb(R, C, r, c, v, width, use_ref=True)
        for a, b in zip(H.to_coo(), Hr.to_coo()):
            np.testing.assert_array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
        for _ in range(2):
            H = O.hyb_update(H, *random_coo(rng, R, C, max(1, n // 3), dup, zero, vals))
        x = rng.uniform(-1, 1, C).astype(np.float32)
        np.testing.assert_array_equal(O.hyb_spmv(H, x).view(np.uint32), O.hyb_spmv(H, x, use_ref=True).view(np.uint32))
