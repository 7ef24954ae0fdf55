"""GPU parity: load / export / insert-or-add update / compaction / SpMV of the
HIP engine against the oracle (cusp host-path semantics) and the reference's
own fixtures.  Structure and update values are compared bit-exactly; SpMV
values within 1e-5 of FP64 (conditioned: |y - y64| <= 1e-5 (|A||x|)_i)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, fp64_spmv_check, random_coo, to_host_csr

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def gold_csr(d, prefix, R, C):
    return O.Csr(R, C, d[f"{prefix}_off"], d[f"{prefix}_col"], d[f"{prefix}_val"])


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    import dcsr_amd
    dcsr_amd.lib()
    return t


def spmv_host(mat, x):
    import torch
    y = mat.spmv(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    return y.cpu().numpy()


def test_banded_flow_matches_reference_fixtures(torch):
    from dcsr_amd import DcsrMatrix
    d = np.load(os.path.join(GOLD, "banded256.npz"))
    R = int(d["R"])
    say("banded256 load")
    m = DcsrMatrix(R, R)
    m.LoadMatrix(d["in_row"], d["in_col"], d["in_val"])
    assert_same_structure(to_host_csr(m), gold_csr(d, "load", R, R))
    for i in range(4):
        say(f"banded256 update {i}")
        m.UpdateMatrix(d[f"b{i}_row"], d[f"b{i}_col"], d[f"b{i}_val"])
        assert_same_structure(to_host_csr(m), gold_csr(d, f"upd{i}", R, R))
    say("banded256 spmv")
    A = gold_csr(d, "upd3", R, R)
    y = spmv_host(m, d["x"])
    fp64_spmv_check(A, d["x"], y)
    say("banded256 compact")
    m.SortMatrixRow()
    assert m.info()["garbage"] == 0
    assert_same_structure(to_host_csr(m), A)
    fp64_spmv_check(A, d["x"], spmv_host(m, d["x"]))
    m.close()


def test_edge_fixture_duplicates_zeros_deletes(torch):
    from dcsr_amd import DcsrMatrix
    d = np.load(os.path.join(GOLD, "edge.npz"))
    R, C = int(d["R"]), int(d["C"])
    say("edge load")
    m = DcsrMatrix(R, C)
    m.LoadMatrix(d["in_row"], d["in_col"], d["in_val"])
    info = m.info()
    assert info["canonical"] == 0 and info["nnz"] == d["in_row"].size
    assert_same_structure(to_host_csr(m), gold_csr(d, "load", R, C))
    say("edge canonical export")
    assert_same_structure(to_host_csr(m, canonical=True), gold_csr(d, "canon", R, C))
    say("edge update with deletes")
    m.UpdateMatrix(d["b0_row"], d["b0_col"], d["b0_val"])
    assert_same_structure(to_host_csr(m), gold_csr(d, "upd0", R, C))
    assert m.info()["canonical"] == 1
    y = spmv_host(m, d["x"])
    np.testing.assert_array_equal(y, d["y"])  # small integers: every order is exact
    m.close()


@pytest.mark.parametrize("R,C,n,seed", [(1, 1, 3, 0), (33, 65, 500, 1), (1000, 700, 30000, 2), (5000, 5000, 200000, 3)])
def test_random_update_chain(torch, R, C, n, seed):
    from dcsr_amd import DcsrMatrix
    rng = np.random.default_rng(seed)
    r, c, v = random_coo(rng, R, C, n, 0.0, 0.0, "sym")
    say(f"random chain R={R} C={C} n={n}")
    m = DcsrMatrix(R, C, pool_factor=1.2)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    for k in range(4):
        br, bc, bv = random_coo(rng, R, C, max(1, n // 2), 0.2, 0.02, "small_int" if k % 2 else "sym")
        if k == 2 and A.nnz:
            rows_of = np.repeat(np.arange(R), np.diff(A.off))
            pick = rng.integers(0, A.nnz, max(1, A.nnz // 5))
            br = np.concatenate([br, rows_of[pick]]).astype(np.int32)
            bc = np.concatenate([bc, A.col[pick]]).astype(np.int32)
            bv = np.concatenate([bv, -A.val[pick]]).astype(np.float32)
        say(f"  update {k} ({br.size} triples)")
        m.UpdateMatrix(br, bc, bv)
        A = O.update(A, br, bc, bv)
        assert_same_structure(to_host_csr(m), A)
        info = m.info()
        assert info["nnz"] == A.nnz
        x = rng.uniform(-1, 1, C).astype(np.float32)
        fp64_spmv_check(A, x, spmv_host(m, x))
    say("  compact")
    m.SortMatrixRow()
    assert_same_structure(to_host_csr(m), A)
    m.close()


@pytest.mark.parametrize("pool_factor", [20.0, 1.5])
def test_nosync_update_chain_and_spmv(torch, pool_factor):
    """NOSYNC updates (no state read-back) back to back, each followed by an
    SpMV sized from the host's bound on the live entries: results match the
    oracle and a handle updated synchronously.  Every batch touches nearly
    every row, so each rewrites the whole matrix into fresh chunks: the large
    pool holds all six rewrites, the small one makes the NOSYNC path compact
    in stream order and grow."""
    from dcsr_amd import DcsrMatrix, UPDATE_NOSYNC
    R, C, n = 3000, 2500, 20000
    rng = np.random.default_rng(21)
    r, c, v = random_coo(rng, R, C, 40000, 0.0, 0.0, "sym")
    say(f"nosync chain pool_factor={pool_factor}")
    sync = DcsrMatrix(R, C, pool_factor=pool_factor)
    m = DcsrMatrix(R, C, pool_factor=pool_factor)
    for h in (sync, m):
        h.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    x = rng.uniform(-1, 1, C).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    for k in range(6):
        br, bc, bv = random_coo(rng, R, C, n, 0.2, 0.02, "small_int" if k % 2 else "sym")
        # cancel some stored entries exactly (deletions)
        rows_of = np.repeat(np.arange(R), np.diff(A.off))
        pick = rng.choice(A.nnz, 500, replace=False)
        br[:500], bc[:500], bv[:500] = rows_of[pick], A.col[pick], -A.val[pick]
        say(f"  update {k}")
        sync.UpdateMatrix(br, bc, bv)
        m.UpdateMatrix(br, bc, bv, flags=UPDATE_NOSYNC)
        A = O.update(A, br, bc, bv)
        y = m.spmv(xd)  # no state read-back in between
        if k % 2 == 0:
            continue  # two NOSYNC updates back to back before the next check
        torch.cuda.synchronize()
        fp64_spmv_check(A, x, y.cpu().numpy())
        assert m.status() == 0
        got = to_host_csr(m)
        assert_same_structure(got, A)
        assert_same_structure(got, to_host_csr(sync))
        assert m.info()["nnz"] == A.nnz
    m.close()
    sync.close()


def test_nosync_snapshots_keep_the_bounds_tight(torch):
    """Each NOSYNC update's write kernel stores the live count and bump
    pointer it leaves into host memory; the next call folds the newest one
    into the host's bounds.  Batches that only add to stored coordinates claim
    nothing, but the host must assume each may claim the whole live matrix
    plus the batch: without landed snapshots the bump bound would pass the
    pool within a few calls and force an in-order compaction.  With the
    stream synchronized between calls every snapshot has landed, so no
    compaction may happen; the results still match the oracle."""
    from dcsr_amd import DcsrMatrix, UPDATE_NOSYNC
    R, C, nnz, n = 2000, 2000, 20000, 4000
    rng = np.random.default_rng(33)
    r, c, v = random_coo(rng, R, C, nnz, 0.0, 0.0, "pos")  # positive: no sum below cancels
    # distinct coordinates: the matrix is canonical at load, so the first
    # NOSYNC call does not take the synchronous canonicalizing path
    _, first = np.unique(r.astype(np.int64) * C + c, return_index=True)
    r, c, v = r[first], c[first], v[first]
    A = O.coo_to_csr(R, C, r, c, v)
    # pool 4x the entries: the host's bound grows by ~nnz + n + 7 * touched
    # (~50K slots) per call without a snapshot, past the 80K pool at call 2
    m = DcsrMatrix(R, C, pool_factor=4.0)
    m.LoadMatrix(r, c, v)
    rows_of = np.repeat(np.arange(R), np.diff(A.off)).astype(np.int32)
    c0 = m.info()["compactions"]
    for k in range(12):
        pick = rng.integers(0, A.nnz, n)
        br, bc = rows_of[pick], A.col[pick].astype(np.int32)
        bv = rng.integers(1, 4, n).astype(np.float32)
        m.UpdateMatrix(br, bc, bv, flags=UPDATE_NOSYNC)
        A = O.update(A, br, bc, bv)
        torch.cuda.synchronize()
    assert m.status() == 0
    assert m.info()["compactions"] == c0
    assert_same_structure(to_host_csr(m), A)
    m.close()


def test_update_tag_wraps_and_spmv_offsets_stay_right(torch, monkeypatch):
    """Each update tags the rows it touches (tag << 32 | touched index) so
    that its move launch can scan the new row lengths for the next SpMV.
    Started two updates before the 32-bit tag wraps, the chain crosses the
    wrap (the map is cleared, tags restart at 1); every SpMV after an update,
    and the structure, must match the oracle."""
    from dcsr_amd import DcsrMatrix, UPDATE_NOSYNC
    monkeypatch.setenv("DCSR_UPDATE_TAG_START", str(0xFFFFFFFE))
    R, C = 3000, 2000
    rng = np.random.default_rng(41)
    r, c, v = random_coo(rng, R, C, 30000, 0.0, 0.0, "sym")
    m = DcsrMatrix(R, C, pool_factor=4.0)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    x = rng.uniform(-1, 1, C).astype(np.float32)
    for k in range(5):
        # rows drawn from a shifting window: some rows touched by one update
        # and not the next
        br = rng.integers(k * 400, k * 400 + 1200, 3000).astype(np.int32)
        bc = rng.integers(0, C, 3000).astype(np.int32)
        bv = rng.uniform(-1, 1, 3000).astype(np.float32)
        say(f"tag wrap update {k}")
        m.UpdateMatrix(br, bc, bv, flags=UPDATE_NOSYNC if k % 2 else 0)
        A = O.update(A, br, bc, bv)
        fp64_spmv_check(A, x, spmv_host(m, x))
    assert m.status() == 0
    assert_same_structure(to_host_csr(m), A)
    m.close()


def test_powerlaw_with_hub_rows(torch):
    """Rows longer than the hub threshold go through the split SpMV path."""
    from dcsr_amd import DcsrMatrix
    say("powerlaw hubs")
    R = 20000
    r, c, v = W.powerlaw(R, 400000, max_degree=19000, seed=5, signed=True)
    m = DcsrMatrix(R, R)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, R, r, c, v)
    assert np.diff(A.off).max() > 4096
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, R).astype(np.float32)
    y = spmv_host(m, x)
    fp64_spmv_check(A, x, y)
    y2 = spmv_host(m, x)
    np.testing.assert_array_equal(y, y2)  # deterministic
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, 100000, signed=True)
    m.UpdateMatrix(br, bc, bv)
    A = O.update(A, br, bc, bv)
    assert_same_structure(to_host_csr(m), A)
    fp64_spmv_check(A, x, spmv_host(m, x))
    m.close()


@pytest.mark.parametrize("n", [2_090_000, 1024 * 2048, 2_300_000])
def test_large_batches(torch, n):
    """Batches around 1,024 tiles of 2,048 triples: up to that size each pass
    sums the tile counts before its tile itself (2,090,000: 1,021 tiles, every
    summing lane used); above it a last workgroup scans them (2,300,000).  At
    exactly 1024 * 2048 the heads (1,024 tiles of n) sum directly while the
    per-entry passes (tiles of n + 1 entries: 1,025) use the tables."""
    from dcsr_amd import DcsrMatrix
    R = 200_000
    r, c, v = W.powerlaw(R, 1_000_000, max_degree=3000, seed=21, signed=True)
    m = DcsrMatrix(R, R)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, R, r, c, v)
    rng = np.random.default_rng(23)
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, n, signed=True)
    say(f"large batch n={n}")
    m.UpdateMatrix(br, bc, bv)
    A = O.update(A, br, bc, bv)
    assert_same_structure(to_host_csr(m), A)
    x = rng.uniform(-1, 1, R).astype(np.float32)
    fp64_spmv_check(A, x, spmv_host(m, x))
    m.close()


def test_alpha_beta_and_csr_comparator(torch):
    from dcsr_amd import DcsrMatrix, csr_spmv
    rng = np.random.default_rng(11)
    R, C = 3000, 2000
    r, c, v = random_coo(rng, R, C, 60000, 0.0, 0.0, "sym")
    m = DcsrMatrix(R, C)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    x = rng.uniform(-1, 1, C).astype(np.float32)
    y0 = rng.uniform(-1, 1, R).astype(np.float32)
    say("alpha/beta spmv")
    y = torch.from_numpy(y0.copy()).cuda()
    m.spmv(torch.from_numpy(x).cuda(), y, alpha=2.0, beta=-0.5)
    torch.cuda.synchronize()
    want = 2.0 * (A.to_scipy() @ x.astype(np.float64)) - 0.5 * y0
    scale = 2.0 * (abs(A.to_scipy()) @ np.abs(x.astype(np.float64))) + 0.5 * np.abs(y0)
    assert (np.abs(y.cpu().numpy() - want) <= 1e-5 * scale + 1e-30).all()
    say("csr comparator")
    yc = csr_spmv(A.off, A.col, A.val, x)
    torch.cuda.synchronize()
    fp64_spmv_check(A, x, yc.cpu().numpy())
    m.close()


def test_errors_leave_matrix_unchanged(torch):
    from dcsr_amd import DcsrError, DcsrMatrix, UPDATE_NOGROW
    rng = np.random.default_rng(12)
    R = 200
    keys = rng.choice(R * R, 2000, replace=False)
    r, c = (keys // R).astype(np.int32), (keys % R).astype(np.int32)
    v = rng.uniform(-1, 1, keys.size).astype(np.float32)
    m = DcsrMatrix(R, R, pool_factor=1.0, min_pool=16, allow_grow=False)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, R, r, c, v)
    say("bad index update")
    with pytest.raises(DcsrError) as e:
        m.UpdateMatrix(np.array([1, R]), np.array([1, 1]), np.array([1.0, 1.0], np.float32))
    assert e.value.code == 2
    assert_same_structure(to_host_csr(m), A)
    say("pool exhaustion without growth")
    br, bc, bv = random_coo(rng, R, R, 4000, 0.0, 0.0, "sym")
    with pytest.raises(DcsrError) as e:
        m.UpdateMatrix(br, bc, bv, flags=UPDATE_NOGROW)
    assert e.value.code == 3
    assert_same_structure(to_host_csr(m), A)
    say("empty batch")
    m.UpdateMatrix(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    assert_same_structure(to_host_csr(m), A)
    say("bad load index")
    with pytest.raises(DcsrError):
        m.LoadMatrix(np.array([0, -1]), np.array([0, 0]), np.ones(2, np.float32))
    m.close()
    say("growth on demand")
    m = DcsrMatrix(R, R, pool_factor=1.0, min_pool=16)
    m.LoadMatrix(r, c, v)
    m.UpdateMatrix(br, bc, bv)
    assert_same_structure(to_host_csr(m), O.update(A, br, bc, bv))
    m.close()


def test_nosync_errors_are_sticky_and_leave_matrix_unchanged(torch):
    """NOSYNC calls return at once; a rejected batch is a no-op whose cause is
    reported by the next status read, which then clears it.  Later batches
    still apply."""
    from dcsr_amd import DcsrMatrix, UPDATE_NOGROW, UPDATE_NOSYNC
    rng = np.random.default_rng(13)
    R = 300
    keys = rng.choice(R * R, 3000, replace=False)
    r, c = (keys // R).astype(np.int32), (keys % R).astype(np.int32)
    v = rng.uniform(-1, 1, keys.size).astype(np.float32)
    m = DcsrMatrix(R, R, pool_factor=1.0, min_pool=16, allow_grow=False)
    m.LoadMatrix(r, c, v)
    # an empty synchronous update canonicalizes the loaded matrix, so the
    # NOSYNC calls below take the asynchronous path
    m.UpdateMatrix(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    A = O.coo_to_csr(R, R, r, c, v)
    say("nosync bad index")
    m.UpdateMatrix(np.array([3, R + 5], np.int32), np.array([1, 1], np.int32), np.array([1.0, 2.0], np.float32),
                   flags=UPDATE_NOSYNC)
    x = rng.uniform(-1, 1, R).astype(np.float32)
    # the rejected batch's launch still rebuilt the SpMV's row offsets, from
    # the unchanged lengths
    fp64_spmv_check(A, x, spmv_host(m, x))
    assert m.status() == 2  # DCSR_ERR_BAD_INDEX
    assert m.status() == 0  # read clears it
    assert_same_structure(to_host_csr(m), A)
    say("nosync pool exhaustion without growth")
    br, bc, bv = random_coo(rng, R, R, 6000, 0.0, 0.0, "sym")
    m.UpdateMatrix(br, bc, bv, flags=UPDATE_NOSYNC | UPDATE_NOGROW)
    fp64_spmv_check(A, x, spmv_host(m, x))
    assert m.status() == 3  # DCSR_ERR_POOL_EXHAUSTED
    assert_same_structure(to_host_csr(m), A)
    say("nosync batch that fits, after the rejected ones")
    sr, sc, sv = A.off, A.col, A.val
    rows_of = np.repeat(np.arange(R), np.diff(sr))
    pick = rng.choice(A.nnz, 100, replace=False)
    ur, uc, uv = rows_of[pick].astype(np.int32), sc[pick].astype(np.int32), (0.5 * sv[pick]).astype(np.float32)
    m.UpdateMatrix(ur, uc, uv, flags=UPDATE_NOSYNC)  # adds only: values in place, no claims
    A = O.update(A, ur, uc, uv)
    fp64_spmv_check(A, x, spmv_host(m, x))
    assert m.status() == 0
    assert_same_structure(to_host_csr(m), A)
    m.close()


def test_csr_load_and_initialize(torch):
    from dcsr_amd import DcsrMatrix
    rng = np.random.default_rng(13)
    R, C = 400, 300
    r, c, v = random_coo(rng, R, C, 5000, 0.1, 0.0, "sym")
    A = O.coo_to_csr(R, C, r, c, v)
    say("load_csr")
    m = DcsrMatrix(R, C)
    m.load_csr(A.off, A.col, A.val)
    assert_same_structure(to_host_csr(m), A)
    say("initialize + update")
    m2 = DcsrMatrix(R, C)
    m2.Initialize_Matrix(8)
    assert m2.info()["nnz"] == 0
    m2.UpdateMatrix(r, c, v)
    E = O.Csr(R, C, np.zeros(R + 1, np.int32), [], [])
    assert_same_structure(to_host_csr(m2), O.update(E, r, c, v))
    m.close()
    m2.close()


def test_spmv_empty_rows_and_long_rows(torch):
    """Flat SpMV edge cases: empty matrix, leading/trailing empty rows, a row
    spanning many waves (carry chain), beta applied to empty rows."""
    from dcsr_amd import DcsrMatrix
    rng = np.random.default_rng(21)
    say("empty matrix spmv")
    R, C = 5000, 3000
    m = DcsrMatrix(R, C)
    x = rng.uniform(-1, 1, C).astype(np.float32)
    y0 = rng.uniform(-1, 1, R).astype(np.float32)
    y = torch.from_numpy(y0.copy()).cuda()
    m.spmv(torch.from_numpy(x).cuda(), y, alpha=1.0, beta=0.5)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(y.cpu().numpy(), (0.5 * y0).astype(np.float32))
    say("sparse rows + one 50k-entry row")
    rows = [np.full(2900, 4000), rng.integers(100, 200, 500), np.full(2500, 4001)]
    cols = [rng.choice(C, 2900, replace=False), rng.integers(0, C, 500), rng.choice(C, 2500, replace=False)]
    r = np.concatenate(rows).astype(np.int32)
    c = np.concatenate(cols).astype(np.int32)
    v = rng.uniform(-1, 1, r.size).astype(np.float32)
    # a very long row: 50k entries with repeated columns (duplicates kept)
    r = np.concatenate([r, np.full(50000, 2500, np.int32)])
    c = np.concatenate([c, rng.integers(0, C, 50000).astype(np.int32)])
    v = np.concatenate([v, rng.uniform(-1, 1, 50000).astype(np.float32)])
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    yv = spmv_host(m, x)
    fp64_spmv_check(A, x, yv)
    assert (yv[np.diff(A.off) == 0] == 0).all()
    y = torch.from_numpy(y0.copy()).cuda()
    m.spmv(torch.from_numpy(x).cuda(), y, alpha=-1.5, beta=2.0)
    torch.cuda.synchronize()
    want = -1.5 * (A.to_scipy() @ x.astype(np.float64)) + 2.0 * y0
    scale = 1.5 * (abs(A.to_scipy()) @ np.abs(x.astype(np.float64))) + 2.0 * np.abs(y0)
    assert (np.abs(y.cpu().numpy() - want) <= 1e-5 * scale + 1e-30).all()
    np.testing.assert_array_equal(spmv_host(m, x), yv)  # reproducible
    m.close()


def test_scan_fallback_path_gives_identical_results(torch):
    """With the look-back wait limit at 0 every scan tile past the first takes
    the exact fallback (its prefix summed from the earlier tiles' inputs):
    load (row-start scan), SpMV (row-length scan + row map), export and
    compaction give bit-identical results, and the handle reports that the
    fallback ran.  With the default limit a fresh handle reports none."""
    import dcsr_amd
    from tests.helpers import check_layout
    R = 600_000
    r, c, v = W.powerlaw(R, 3_000_000, max_degree=3000, seed=61, signed=True)
    rng = np.random.default_rng(62)
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, 500_000, signed=True)
    x = rng.uniform(-1, 1, R).astype(np.float32)

    def run():
        m = dcsr_amd.DcsrMatrix(R, R)
        m.LoadMatrix(r, c, v)
        m.UpdateMatrix(br, bc, bv)
        y = spmv_host(m, x)
        m.SortMatrixRow()
        y2 = spmv_host(m, x)
        return m, to_host_csr(m), y, y2

    say("default scan wait limit")
    m0, C0, y0, y0b = run()
    assert m0.info()["scan_fallback"] == 0
    say("scan wait limit 0: every tile takes the fallback")
    dcsr_amd.set_scan_wait_limit(0)
    try:
        m1, C1, y1, y1b = run()
    finally:
        dcsr_amd.set_scan_wait_limit(-1)
    assert m1.info()["scan_fallback"] == 1
    assert m1.status() == 0  # a diagnostic, not an error
    assert_same_structure(C1, C0)
    np.testing.assert_array_equal(y1, y0)
    np.testing.assert_array_equal(y1b, y0b)
    check_layout(m1)
    m0.close()
    m1.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_allocator_invariants_through_updates_and_compaction(torch, seed):
    """Chunks stay disjoint and below the bump pointer, and bump = sum of
    capacities + garbage, across synchronous and asynchronous updates (with
    deletions), in-stream compaction, growth, SpGEMM output and HYB load."""
    import dcsr_amd
    from tests.helpers import check_layout
    rng = np.random.default_rng(70 + seed)
    R, C = 20000, 15000
    r, c, v = random_coo(rng, R, C, 200000, 0.0, 0.0, "sym")
    m = dcsr_amd.DcsrMatrix(R, C, pool_factor=1.3)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    check_layout(m)
    for k in range(5):
        br, bc, bv = random_coo(rng, R, C, 60000, 0.1, 0.0, "sym")
        rows_of = np.repeat(np.arange(R), np.diff(A.off))
        pick = rng.choice(A.nnz, 3000, replace=False)
        br[:3000], bc[:3000], bv[:3000] = rows_of[pick], A.col[pick], -A.val[pick]
        m.UpdateMatrix(br, bc, bv, flags=dcsr_amd.UPDATE_NOSYNC if k % 2 else 0)
        A = O.update(A, br, bc, bv)
        info = check_layout(m)
        assert info["nnz"] == A.nnz
    m.SortMatrixRow()
    assert check_layout(m)["garbage"] == 0
    assert_same_structure(to_host_csr(m), A)
    sq = dcsr_amd.DcsrMatrix(R, R)
    sq.LoadMatrix(r[r < R], np.minimum(c, R - 1)[r < R], v[r < R])
    Cm = dcsr_amd.spgemm(sq, sq)
    check_layout(Cm)
    ec, ev, cr, cc, cv, pitch = m.to_hyb(8)
    h2 = dcsr_amd.DcsrMatrix(R, C)
    h2.load_hyb(8, pitch, ec, ev, cr, cc, cv)
    check_layout(h2)
    for x in (m, sq, Cm, h2):
        x.close()
