"""GPU parity of the serving loop: a mixed stream of updates (with deletions
by -current), SpMV and SpMM queries submitted back to back (so queries get
batched into shared multi-column products) gives, for every query, the
product with exactly the matrix version it was submitted after -- the same
result as the plain per-call path on its own, within 1e-5 (conditioned)."""
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, to_host_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def within(A, X, Y, rtol=1e-5):
    S = A.to_scipy(np.float64)
    X = X.reshape(A.cols, -1).astype(np.float64)
    want = S @ X
    scale = abs(S) @ np.abs(X)
    return bool((np.abs(Y.reshape(want.shape) - want) <= rtol * scale + 1e-30).all())


def make_stream(rng, A, nreq):
    R = A.rows
    reqs = []
    for i in range(nreq):
        u = rng.random()
        if u < 0.25:
            n = int(np.exp(rng.uniform(np.log(100), np.log(5000))))
            rows_of = np.repeat(np.arange(R), np.diff(A.off))
            ex = rng.integers(0, A.nnz, n // 2)
            r = np.concatenate([rows_of[ex], rng.integers(0, R, n - n // 2)]).astype(np.int32)
            c = np.concatenate([A.col[ex], rng.integers(0, A.cols, n - n // 2)]).astype(np.int32)
            v = rng.uniform(-1, 1, n).astype(np.float32)
            nd = int(0.3 * (n // 2))
            v[:nd] = -A.val[ex[:nd]]  # deletions by -current
            reqs.append(("update", (r, c, v)))
            A = O.update(A, r, c, v)
        else:
            k = int(rng.choice([1, 1, 4, 16, 32]))
            X = rng.uniform(-1, 1, (A.cols, k)).astype(np.float32)
            reqs.append(("query", (k, X, A)))
    return reqs, A


def test_server_mixed_stream_matches_alone(torch_dev=None):
    import torch
    import dcsr_amd
    R = 20000
    r, c, v = W.powerlaw(R, 300000, max_degree=5000, seed=12, signed=True)
    A0 = O.coo_to_csr(R, R, r, c, v)
    rng = np.random.default_rng(4)
    reqs, A_final = make_stream(rng, A0, 80)
    say("server: submit mixed stream")
    m = dcsr_amd.DcsrMatrix(R, R, pool_factor=1.5)
    m.LoadMatrix(r, c, v)
    srv = dcsr_amd.DcsrServer(m, max_cols=64)
    tickets, outs = [], []
    for kind, p in reqs:
        if kind == "update":
            tickets.append(srv.update(*p))
            outs.append(None)
        else:
            k, X, _ = p
            if k == 1:
                t, y = srv.spmv(torch.from_numpy(X[:, 0].copy()).cuda())
            else:
                t, y = srv.spmm(torch.from_numpy(X).cuda())
            tickets.append(t)
            outs.append(y)
    lat = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["queries"] == sum(1 for k, _ in reqs if k == "query")
    assert st["query_groups"] < st["queries"], "queries were not batched"
    assert all(x >= 0 for x in lat)
    assert_same_structure(to_host_csr(m), A_final)
    say("plain per-call path on its own")
    alone = dcsr_amd.DcsrMatrix(R, R)
    alone.LoadMatrix(r, c, v)
    for (kind, p), y in zip(reqs, outs):
        if kind == "update":
            alone.UpdateMatrix(*p)
            continue
        k, X, Aver = p
        if k == 1:
            ya = alone.spmv(torch.from_numpy(X[:, 0].copy()).cuda())
        else:
            ya = alone.spmm(torch.from_numpy(X).cuda())
        torch.cuda.synchronize()
        mixed = y.cpu().numpy()
        single = ya.cpu().numpy()
        assert within(Aver, X, mixed), "mixed result outside tolerance of its version"
        assert within(Aver, X, single), "alone result outside tolerance"
    m.close()
    alone.close()
