"""GPU parity: batched persistent step (many small matrices, each with its own
insert-or-add batch + SpMV in one launch) against the oracle, per matrix."""
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, fp64_spmv_check

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def make_case(count, seed, batch=300, max_rows=512):
    mats = W.many_small(count=count, min_rows=16, max_rows=max_rows, lam=6.0, seed=seed)
    rng = np.random.default_rng(seed + 1)
    sizes = np.array([m[0] for m in mats], np.int32)
    A = [O.coo_to_csr(n, n, r, c, v) for n, r, c, v in mats]
    batches = [W.batch_mixed(rng, r, c, n, n, batch, signed=True) for n, r, c, v in mats]
    return mats, sizes, A, batches


def concat(parts):
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([p[0].size for p in parts])
    return off, np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]), \
        np.concatenate([p[2] for p in parts])


def split_csr(b, off, col, val):
    out = []
    for m in range(b.count):
        r0, r1 = int(b.row_base[m]), int(b.row_base[m + 1])
        o = off[r0:r1 + 1] - off[r0]
        out.append(O.Csr(r1 - r0, r1 - r0, o, col[off[r0]:off[r1]], val[off[r0]:off[r1]]))
    return out


@pytest.fixture(scope="module")
def dev():
    import torch
    import dcsr_amd
    dcsr_amd.lib()
    return torch, dcsr_amd


@pytest.mark.parametrize("graph", [False, True])
def test_batch_steps_match_per_matrix_oracle(dev, graph):
    torch, dcsr_amd = dev
    count = 96
    mats, sizes, A, batches = make_case(count, seed=5)
    say(f"batch create/load ({count} matrices, graph={graph})")
    b = dcsr_amd.DcsrBatch(sizes, sizes)
    off, r, c, v = concat([(m[1], m[2], m[3]) for m in mats])
    b.load_coo(off, r, c, v)
    rng = np.random.default_rng(77)
    tri = concat(batches)
    d_tri = [torch.from_numpy(a).cuda() for a in tri]
    x = rng.uniform(-1, 1, b.total_cols).astype(np.float32)
    d_x = torch.from_numpy(x).cuda()
    d_y = torch.empty(b.total_rows, dtype=torch.float32, device="cuda")
    for step in range(2):
        say(f"  step {step}")
        if step == 1:  # same device buffers, new contents (graph replay reads them)
            batches = [W.batch_mixed(rng, mm[1], mm[2], mm[0], mm[0], 300, signed=True) for mm in mats]
            tri = concat(batches)
            for dst, src in zip(d_tri, tri):
                dst.copy_(torch.from_numpy(src))
        b.step(*d_tri, x=d_x, y=d_y, graph=graph)
        torch.cuda.synchronize()
        code, per = b.status()
        assert code == 0 and not per.any()
        A = [O.update(Am, *bt) for Am, bt in zip(A, batches)]
        # the live count the step's workgroups summed up (one atomic per
        # workgroup, not per matrix)
        assert b.info()["nnz"] == sum(Am.nnz for Am in A)
        got = split_csr(b, *[t.cpu().numpy() for t in b.to_csr()])
        y = d_y.cpu().numpy()
        for m in range(count):
            assert_same_structure(got[m], A[m])
            r0, r1 = int(b.row_base[m]), int(b.row_base[m + 1])
            c0 = int(b.col_base[m])
            fp64_spmv_check(A[m], x[c0:c0 + A[m].cols], y[r0:r1])
    b.close()


def test_batch_autotune_and_errors(dev):
    torch, dcsr_amd = dev
    count = 40
    mats, sizes, A, batches = make_case(count, seed=9, max_rows=2048)
    b = dcsr_amd.DcsrBatch(sizes, sizes)
    off, r, c, v = concat([(m[1], m[2], m[3]) for m in mats])
    b.load_coo(off, r, c, v)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.uniform(-1, 1, b.total_cols).astype(np.float32)).cuda()
    y = torch.empty(b.total_rows, dtype=torch.float32, device="cuda")
    say("autotune")
    team, thresh = b.autotune(x, y)
    assert team in (4, 8, 16, 32) and thresh in (64, 256, 1024)
    for tm, th in ((4, 64), (32, 1024)):
        b.set_shape(tm, th)
        y2 = b.step(x=x)
        torch.cuda.synchronize()
        xs = x.cpu().numpy()
        for m in range(count):
            r0, r1 = int(b.row_base[m]), int(b.row_base[m + 1])
            c0 = int(b.col_base[m])
            fp64_spmv_check(A[m], xs[c0:c0 + A[m].cols], y2.cpu().numpy()[r0:r1])
    say("bad index in one matrix")
    bad = [list(bt) for bt in batches]
    bad[3][0] = bad[3][0].copy()
    bad[3][0][0] = sizes[3]  # row out of range for matrix 3
    tri = concat([tuple(t) for t in bad])
    b.step(*tri)
    torch.cuda.synchronize()
    code, per = b.status()
    assert code == 2 and per[3] == 1 and per.sum() == 1
    got = split_csr(b, *[t.cpu().numpy() for t in b.to_csr()])
    for m in range(count):
        want = A[m] if m == 3 else O.update(A[m], *batches[m])
        assert_same_structure(got[m], want)
    b.close()


def test_exhausted_matrix_leaves_room_for_the_others(dev):
    """A pool that may not grow: matrix 0's claim (its whole 2,000-row matrix
    rewritten plus 4,000 inserts) cannot fit, the small matrices' claims can.
    Matrix 0 fails alone with its exhausted bit and stays unchanged; every
    other matrix commits (a claim that does not fit never advances the bump
    pointer); the allocator's invariants hold afterwards."""
    torch, dcsr_amd = dev
    from tests.helpers import check_layout
    rng = np.random.default_rng(77)
    big = (2000, *W.powerlaw(2000, 32000, 2.1, 200, seed=78, signed=True)[:3])
    small = [(100, *W.powerlaw(100, 400, 2.1, 20, seed=80 + i, signed=True)[:3]) for i in range(15)]
    mats = [big] + small
    sizes = np.array([m[0] for m in mats], np.int32)
    nnz = sum(m[1].size for m in mats)
    free = 20000
    b = dcsr_amd.DcsrBatch(sizes, sizes, pool_factor=(nnz + free) / nnz, allow_grow=False, min_pool=16)
    b.load_coo(*concat([(m[1], m[2], m[3]) for m in mats]))
    A = [O.coo_to_csr(n, n, r, c, v) for n, r, c, v in mats]
    # matrix 0: 4,000 new coordinates over all its rows; the others: 50 triples each
    br0 = np.arange(4000) % 2000
    bc0 = (np.arange(4000) * 7 + 1) % 2000
    key = set(zip(mats[0][1].tolist(), mats[0][2].tolist()))
    sel = np.array([(r, c) not in key for r, c in zip(br0, bc0)])
    batches = [(br0[sel].astype(np.int32), bc0[sel].astype(np.int32), np.ones(int(sel.sum()), np.float32))]
    batches += [W.batch_mixed(rng, r, c, n, n, 50, signed=True) for n, r, c, v in small]
    say("batch step that exhausts the pool for matrix 0 only")
    b.step(*concat(batches))
    code, per = b.status()
    assert code == 3  # DCSR_ERR_POOL_EXHAUSTED
    assert per[0] != 0 and not per[1:].any()
    off, col, val = (t.cpu().numpy() for t in b.to_csr())
    got = split_csr(b, off, col, val)
    assert_same_structure(got[0], A[0])
    for i in range(1, len(mats)):
        assert_same_structure(got[i], O.update(A[i], *batches[i]))
    check_layout(b)
    b.close()
