"""Shared test helpers: comparisons between the HIP engine and the oracle."""
from __future__ import annotations

import numpy as np

import oracle as O


def to_host_csr(mat, canonical=False):
    """Export a DcsrMatrix to an oracle.Csr (host)."""
    off, col, val = mat.to_csr(canonical=canonical)
    return O.Csr(mat.num_rows, mat.num_cols, off.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy())


def assert_same_structure(got: O.Csr, want: O.Csr, values="exact"):
    """Row-by-row comparison after a stable column sort of each row.

    values: "exact" (bitwise float32), or a float relative tolerance, or None."""
    g = got.canonical()
    w = want.canonical()
    assert g.rows == w.rows and g.cols == w.cols
    np.testing.assert_array_equal(g.off, w.off, err_msg="row lengths differ")
    np.testing.assert_array_equal(g.col, w.col, err_msg="column indices differ")
    if values == "exact":
        bad = np.nonzero(g.val.view(np.uint32) != w.val.view(np.uint32))[0]
        assert bad.size == 0, f"{bad.size} values differ, first at {bad[:5]}: {g.val[bad[:5]]} vs {w.val[bad[:5]]}"
    elif values is not None:
        np.testing.assert_allclose(g.val, w.val, rtol=values, atol=0)


def fp64_spmv_check(A: O.Csr, x, y, rtol=1e-5):
    """|y - y64| <= rtol * (|A| |x|)_i  (conditioned bound; FP64 via scipy)."""
    S = A.to_scipy(np.float64)
    x64 = np.asarray(x, np.float64)
    y64 = S @ x64
    scale = abs(S) @ np.abs(x64)
    err = np.abs(np.asarray(y, np.float64) - y64)
    tol = rtol * scale + 1e-30
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, f"{bad.size} rows out of tolerance, e.g. row {bad[0]}: {y[bad[0]]} vs {y64[bad[0]]}"


def random_coo(rng, R, C, n, dup_frac=0.0, zero_frac=0.0, values="sym"):
    r = rng.integers(0, R, n).astype(np.int32)
    c = rng.integers(0, C, n).astype(np.int32)
    if dup_frac > 0 and n > 1:
        k = int(n * dup_frac)
        src = rng.integers(0, n, k)
        dst = rng.integers(0, n, k)
        r[dst] = r[src]
        c[dst] = c[src]
    if values == "sym":
        v = rng.uniform(-1, 1, n).astype(np.float32)
    elif values == "small_int":
        v = rng.integers(-3, 4, n).astype(np.float32)
    else:
        v = rng.uniform(0, 1, n).astype(np.float32)
    if zero_frac > 0:
        z = rng.random(n) < zero_frac
        v[z] = 0.0
    return r, c, v


def check_layout(mat):
    """Allocator invariants of a DcsrMatrix (SURVEY-style pins of the pool):
    every row's chunk [start, start + cap) lies below the bump pointer and
    chunks are pairwise disjoint, len <= cap, bump <= pool capacity, the
    live lengths sum to nnz, and bump = sum of capacities + garbage."""
    st, ln, cp = (t.cpu().numpy().astype(np.int64) for t in mat.layout())
    info = mat.info()
    bump, pool_cap, garbage = info["pool_used"], info["pool_cap"], info["garbage"]
    assert 0 <= bump <= pool_cap, (bump, pool_cap)
    assert (ln >= 0).all() and (ln <= cp).all(), "a row is longer than its chunk"
    assert int(ln.sum()) == info["nnz"], (int(ln.sum()), info["nnz"])
    assert int(cp.sum()) + garbage == bump, (int(cp.sum()), garbage, bump)
    live = cp > 0
    s, c = st[live], cp[live]
    order = np.argsort(s, kind="stable")
    s, c = s[order], c[order]
    assert (s + c <= bump).all(), "a chunk lies past the bump pointer"
    assert (s[1:] >= s[:-1] + c[:-1]).all(), "two chunks overlap"
    return info


def check_healthy(mat):
    """No sticky error and no scan ever took its look-back fallback."""
    assert mat.status() == 0
    assert mat.info()["scan_fallback"] == 0
