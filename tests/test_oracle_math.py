"""Pins of the oracle that do not depend on the reference build: brute force
on small inputs, closed forms and an independent library (scipy, FP64).
Each check is chosen so that a plausible mistake in the restatement (wrong
accumulation order, missing zero drop, unstable sort, transposed operand)
fails it."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import fp64_spmv_check, random_coo


def brute_sort_by_row(r, c, v):
    order = sorted(range(len(r)), key=lambda i: (int(r[i]), i))  # stable by row
    return r[order], c[order], v[order]


def test_sort_by_row_is_stable_counting_sort():
    rng = np.random.default_rng(1)
    r, c, v = random_coo(rng, 20, 20, 400, 0.3, 0.0, "sym")
    got = O.sort_by_row(r, c, v)
    want = brute_sort_by_row(r, c, v)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


def test_indices_to_offsets_closed_form():
    rng = np.random.default_rng(2)
    R = 50
    r = np.sort(rng.integers(0, R, 1000)).astype(np.int32)
    A = O.coo_to_csr(R, 10, r, np.zeros_like(r), np.ones(r.size, np.float32))
    want = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=R))])
    np.testing.assert_array_equal(A.off, want)


def brute_add(A, B):
    """Dictionary-of-keys A + B with float32 left-to-right sums per
    coordinate, A part and B part accumulated separately, exact zero drop."""
    out = []
    for i in range(A.rows):
        sa, sb, order = {}, {}, []
        for M, acc in ((A, sa), (B, sb)):
            cols, vals = M.row(i)
            for j, x in zip(cols.tolist(), vals.tolist()):
                if j not in sa and j not in sb:
                    order.append(j)
                acc[j] = np.float32(acc.get(j, np.float32(0)) + np.float32(x))
        row = {}
        for j in order:
            s = np.float32(sa.get(j, np.float32(0)) + sb.get(j, np.float32(0)))
            if s != 0:
                row[j] = s
        out.append(row)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_add_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    R, C = 30, 40
    A = O.coo_to_csr(R, C, *random_coo(rng, R, C, 300, 0.3, 0.1, "small_int"))
    B = O.coo_to_csr(R, C, *random_coo(rng, R, C, 200, 0.3, 0.1, "small_int"))
    got = O.csr_add(A, B)
    want = brute_add(A, B)
    for i in range(R):
        cols, vals = got.row(i)
        assert dict(zip(cols.tolist(), vals.tolist())) == {k: float(v) for k, v in want[i].items()}
        assert len(set(cols.tolist())) == cols.size  # unique columns
    # exact zeros are dropped and the stored order is reverse first touch
    assert not (got.val == 0).any()


def test_add_sums_batch_duplicates_before_adding():
    # a + (b1 + b2) differs from (a + b1) + b2 in float32 for these values
    a, b1, b2 = np.float32(1.0), np.float32(1e8), np.float32(-1e8)
    A = O.Csr(1, 4, [0, 1], [2], [a])
    B = O.Csr(1, 4, [0, 2], [2, 2], [b1, b2])
    got = O.csr_add(A, B)
    assert got.val[0] == a + (b1 + b2) == np.float32(1.0)
    assert got.val[0] != (a + b1) + b2


def test_update_order_of_rows_is_reverse_first_touch():
    A = O.Csr(1, 8, [0, 3], [5, 1, 3], [1, 1, 1])
    got = O.update(A, np.array([0, 0]), np.array([7, 0]), np.array([2, 2], np.float32))
    np.testing.assert_array_equal(got.col, [0, 7, 3, 1, 5])


@pytest.mark.parametrize("seed", [0, 1])
def test_spmv_against_fp64(seed):
    rng = np.random.default_rng(seed)
    R, C = 500, 400
    A = O.coo_to_csr(R, C, *random_coo(rng, R, C, 20000, 0.0, 0.0, "sym"))
    x = rng.uniform(-1, 1, C).astype(np.float32)
    y = O.spmv(A, x)
    fp64_spmv_check(A, x, y, rtol=1e-5)
    # the check is sharp enough to reject a transposed operand
    Rs = 300
    B = O.coo_to_csr(Rs, Rs, *random_coo(rng, Rs, Rs, 5000, 0.0, 0.0, "sym"))
    xs = rng.uniform(-1, 1, Rs).astype(np.float32)
    yt = (B.to_scipy().T @ xs.astype(np.float64)).astype(np.float32)
    with pytest.raises(AssertionError):
        fp64_spmv_check(B, xs, yt, rtol=1e-5)


def test_spmv_sequential_order():
    # ((0 + 1e8*1) + (-1e8*1)) + 1*1 = 1 ; a tree order would give 0 or 1 too,
    # so use three terms where order matters
    A = O.Csr(1, 3, [0, 3], [0, 1, 2], [1e8, 1.0, -1e8])
    y = O.spmv(A, np.ones(3, np.float32))
    assert y[0] == np.float32(np.float32(np.float32(1e8) + 1.0) - np.float32(1e8))


def test_spmm_dense_is_columnwise_spmv():
    rng = np.random.default_rng(5)
    R, C, k = 60, 50, 4
    A = O.coo_to_csr(R, C, *random_coo(rng, R, C, 700, 0.1, 0.0, "sym"))
    X = rng.uniform(-1, 1, (C, k)).astype(np.float32)
    Y = O.spmm_dense(A, X)
    for j in range(k):
        np.testing.assert_array_equal(Y[:, j], O.spmv(A, X[:, j].copy()))


def test_spgemm_structure_matches_scipy_without_cancellation():
    rng = np.random.default_rng(6)
    R = 120
    A = O.coo_to_csr(R, R, *random_coo(rng, R, R, 900, 0.0, 0.0, "pos"))
    G = O.spgemm(A, A)
    S = (A.to_scipy() @ A.to_scipy()).tocsr()
    S.sort_indices()
    Gc = G.canonical()
    np.testing.assert_array_equal(Gc.off, S.indptr)
    np.testing.assert_array_equal(Gc.col, S.indices)
    np.testing.assert_allclose(Gc.val, S.data, rtol=1e-5)
    assert O.spgemm_count(A, A) == S.nnz


def test_spgemm_drops_exact_zeros():
    # row 0 of A*B: 1*1 + 1*(-1) = 0 at column 0 -> dropped
    A = O.Csr(1, 2, [0, 2], [0, 1], [1, 1])
    B = O.Csr(2, 2, [0, 1, 3], [0, 0, 1], [1, -1, 2])
    G = O.spgemm(A, B)
    np.testing.assert_array_equal(G.col, [1])
    np.testing.assert_array_equal(G.val, [2])
    assert O.spgemm_count(A, B) == 2


def test_hyb_brute_force():
    rng = np.random.default_rng(8)
    R = 40
    A = O.coo_to_csr(R, R, *random_coo(rng, R, R, 300, 0.0, 0.0, "sym"))
    w = 3
    ec, ev, cr, cc, cv, pitch = O.csr_to_hyb(A, w)
    assert pitch % 32 == 0 and pitch >= R
    coo = []
    for i in range(R):
        cols, vals = A.row(i)
        for k, (j, x) in enumerate(zip(cols, vals)):
            if k < w:
                assert ec[k * pitch + i] == j and ev[k * pitch + i] == x
            else:
                coo.append((i, j, x))
        for k in range(len(cols), w):
            assert ec[k * pitch + i] == -1
    assert [(int(a), int(b), float(c)) for a, b, c in zip(cr, cc, cv)] == [(a, int(b), float(c)) for a, b, c in coo]


@pytest.mark.parametrize("seed", [0, 1])
def test_canonical_row_sort_matches_stable_lexsort(seed):
    """The C per-row sort (short rows by insertion, long ones by merge) gives
    the same stable order as numpy's lexsort, duplicates in stored order."""
    rng = np.random.default_rng(seed)
    R, C = 300, 60
    r, c, v = random_coo(rng, R, C, 6000, 0.4, 0.0, "sym")
    r = np.concatenate([r, np.full(3000, 7, np.int32)])  # one long row with duplicates
    c = np.concatenate([c, rng.integers(0, C, 3000).astype(np.int32)])
    v = np.concatenate([v, rng.uniform(-1, 1, 3000).astype(np.float32)])
    A = O.coo_to_csr(R, C, r, c, v)  # rows keep the (random) input order inside
    got, want = A.canonical(), A.canonical_np()
    np.testing.assert_array_equal(got.col, want.col)
    np.testing.assert_array_equal(got.val.view(np.uint32), want.val.view(np.uint32))
    assert (np.diff(got.col.astype(np.int64))[np.diff(np.repeat(np.arange(R), np.diff(A.off))) == 0] >= 0).all()


def test_spgemm_fp64_checker_against_scipy_and_planted_errors():
    """orc_spgemm_check_f64 accepts the FP32 oracle product (whose error is
    far below 1e-5 of |A||B|), agrees with scipy's FP64 product, and counts
    a planted relative error of 1e-4 and a stored entry without partials."""
    rng = np.random.default_rng(5)
    R = 400
    r, c, v = random_coo(rng, R, R, 8000, 0.0, 0.0, "sym")
    A = O.coo_to_csr(R, R, r, c, v)
    Cm = O.spgemm(A, A).canonical()
    assert O.spgemm_check_f64(A, A, Cm) == (0, -1)
    S = A.to_scipy(np.float64)
    P = (S @ S).tocsr()
    P.sort_indices()
    dense_want = P.toarray()
    rows_of = np.repeat(np.arange(R), np.diff(Cm.off))
    np.testing.assert_allclose(Cm.val, dense_want[rows_of, Cm.col], rtol=1e-4, atol=1e-6)
    bad = O.Csr(R, R, Cm.off, Cm.col, Cm.val.copy())
    k = Cm.nnz // 2
    bad.val[k] = bad.val[k] + np.float32(abs(bad.val[k]) * 1e-3 + 1e-3)
    assert O.spgemm_check_f64(A, A, bad) == (1, k)
    # a stored coordinate with no partial product in its row
    absent = np.setdiff1d(np.arange(R), Cm.col[Cm.off[3]:Cm.off[4]])[0]
    col = Cm.col.copy()
    col[Cm.off[3]] = absent
    assert O.spgemm_check_f64(A, A, O.Csr(R, R, Cm.off, col, Cm.val))[0] >= 1


def brute_hyb_update(H, rows, cols, vals):
    """Per-row lists: the batch in stable row order, each entry into the next
    free ELL slot of its row, else onto the row's overflow list; the COO part
    is the rows' overflow lists in row order."""
    R, w = H.rows, H.width
    slots = {i: [] for i in range(R)}
    for k in range(w):
        for i in range(R):
            if H.ell_col[k * H.pitch + i] != -1:
                slots[i].append((int(H.ell_col[k * H.pitch + i]), np.float32(H.ell_val[k * H.pitch + i])))
    ovf = {i: [] for i in range(R)}
    for r_, c_, v_ in zip(H.coo_row, H.coo_col, H.coo_val):
        ovf[int(r_)].append((int(c_), np.float32(v_)))
    size = {i: int(H.row_sizes[i]) for i in range(R)}
    order = sorted(range(len(rows)), key=lambda e: (int(rows[e]), e))
    for e in order:
        i = int(rows[e])
        if size[i] < w:
            slots[i].append((int(cols[e]), np.float32(vals[e])))
        else:
            ovf[i].append((int(cols[e]), np.float32(vals[e])))
        size[i] += 1
    return slots, ovf, size


@pytest.mark.parametrize("width,seed", [(1, 0), (3, 1), (6, 2)])
def test_hyb_update_brute_force_and_operator_identity(width, seed):
    """orc_hyb_update equals the per-row list model, and the represented
    matrix (duplicates summed) of HYB(A) updated with B is A + B."""
    rng = np.random.default_rng(seed)
    R, C = 40, 30
    r, c, v = random_coo(rng, R, C, 200, 0.2, 0.0, "sym")
    H = O.coo_to_hyb(R, C, r, c, v, width)
    br, bc, bv = random_coo(rng, R, C, 150, 0.3, 0.1, "sym")
    H2 = O.hyb_update(H, br, bc, bv)
    slots, ovf, size = brute_hyb_update(H, br, bc, bv)
    for i in range(R):
        got_slots = [(int(H2.ell_col[k * H2.pitch + i]), H2.ell_val[k * H2.pitch + i]) for k in range(width)
                     if H2.ell_col[k * H2.pitch + i] != -1]
        assert [(a, np.float32(b).view(np.uint32)) for a, b in got_slots] == \
            [(a, np.float32(b).view(np.uint32)) for a, b in slots[i]]
        assert H2.row_sizes[i] == size[i]
    want_coo = [(i, cc, vv) for i in range(R) for cc, vv in ovf[i]]
    assert H2.row_sizes[R] == len(want_coo) == H2.ncoo
    assert [(int(a), int(b), np.float32(cv).view(np.uint32)) for a, b, cv in zip(H2.coo_row, H2.coo_col, H2.coo_val)] \
        == [(a, b, np.float32(cv).view(np.uint32)) for a, b, cv in want_coo]
    # operator identity in FP64: sum of stored entries per coordinate
    import scipy.sparse as sp
    hr, hc, hv = H2.to_coo()
    got = sp.coo_matrix((hv.astype(np.float64), (hr, hc)), shape=(R, C)).toarray()
    want = sp.coo_matrix((np.concatenate([v, bv]).astype(np.float64), (np.concatenate([r, br]),
                          np.concatenate([c, bc]))), shape=(R, C)).toarray()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_hyb_spmv_against_fp64():
    rng = np.random.default_rng(9)
    R, C = 500, 400
    r, c, v = random_coo(rng, R, C, 6000, 0.1, 0.0, "sym")
    H = O.hyb_update(O.coo_to_hyb(R, C, r, c, v, 8), *random_coo(rng, R, C, 3000, 0.2, 0.0, "sym"))
    x = rng.uniform(-1, 1, C).astype(np.float32)
    hr, hc, hv = H.to_coo()
    import scipy.sparse as sp
    S = sp.coo_matrix((hv.astype(np.float64), (hr, hc)), shape=(R, C)).tocsr()
    y64 = S @ x.astype(np.float64)
    scale = abs(S) @ np.abs(x.astype(np.float64))
    assert (np.abs(O.hyb_spmv(H, x) - y64) <= 1e-5 * scale + 1e-30).all()


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_update_mt_equals_sequential(threads):
    """The multi-threaded update (multi-core CPU baseline) gives the
    sequential update's CSR exactly: offsets, column order within rows and
    value bits, with duplicates, cancellations and empty rows."""
    rng = np.random.default_rng(90 + threads)
    R, C = 500, 300
    r = rng.integers(0, R, 4000).astype(np.int32)
    c = rng.integers(0, C, 4000).astype(np.int32)
    v = rng.uniform(-1, 1, 4000).astype(np.float32)
    A = O.coo_to_csr(R, C, r, c, v)
    for k in range(3):
        br = rng.integers(0, R // 2, 3000).astype(np.int32)  # half the rows untouched
        bc = rng.integers(0, C, 3000).astype(np.int32)
        bv = rng.integers(-2, 3, 3000).astype(np.float32)
        want = O.update(A, br, bc, bv)
        got = O.update_mt(A, br, bc, bv, threads)
        np.testing.assert_array_equal(got.off, want.off)
        np.testing.assert_array_equal(got.col, want.col)
        np.testing.assert_array_equal(got.val.view(np.uint32), want.val.view(np.uint32))
        A = want


def test_update_mt_rejects_bad_row():
    A = O.coo_to_csr(4, 4, np.array([0], np.int32), np.array([0], np.int32), np.array([1.0], np.float32))
    with pytest.raises(ValueError):
        O.update_mt(A, np.array([4], np.int32), np.array([0], np.int32), np.array([1.0], np.float32), 2)


def test_spmv_mt_equals_sequential():
    rng = np.random.default_rng(95)
    R, C = 700, 400
    A = O.coo_to_csr(R, C, rng.integers(0, R, 9000).astype(np.int32), rng.integers(0, C, 9000).astype(np.int32),
                     rng.uniform(-1, 1, 9000).astype(np.float32))
    x = rng.uniform(-1, 1, C).astype(np.float32)
    np.testing.assert_array_equal(O.spmv_mt(A, x, 5).view(np.uint32), O.spmv(A, x).view(np.uint32))
