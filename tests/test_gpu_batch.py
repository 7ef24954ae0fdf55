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
