"""GPU parity at the full sizes of the measured workloads, built with the same
generators and seeds as bench.py:

* power-law update chain: 325,557 x 325,557 with 3.2M nnz, eight chained
  1M-triple insert-or-add batches (50% existing coordinates, rows by
  degree), an SpMV after each; structure and update values bit-exact against
  the oracle after every batch, SpMV within 1e-5 of FP64 (conditioned),
  nnz and the pool allocator's invariants checked;
* batched: 4,096 matrices (rows log-uniform in 64..4096, Poisson(16) nnz per
  row), one autotuned graph step of 2,000 triples per matrix plus SpMV, then
  a second step replayed from the captured graph with new triples in the
  same buffers; every matrix against the oracle, SpMV against FP64;
* SpGEMM: 2M x 2M Poisson(8) with N(0,1) values, one 4M-triple update, then
  C = A*A: structure and values bit-exact against the oracle's product of
  the column-sorted rows, values within 1e-5 of the FP64 product;
* serving: the resident 10M x 10M / 160M-nnz power-law matrix and a mixed
  request stream; every query answered by the server (continuous batching)
  equals the plain per-call path on its own within 2e-5 relative (all
  values and right-hand sides are positive, so A x = |A||x| and a relative
  bound is the conditioned one), and the final matrices are identical.
"""
import sys
import time

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, check_healthy, check_layout, fp64_spmv_check, to_host_csr

pytestmark = pytest.mark.gpu

T0 = time.time()


def say(msg):
    print(f"[gpu-full {time.time() - T0:7.1f}s] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    import dcsr_amd
    dcsr_amd.lib()
    return t


def test_powerlaw_update_chain_325k(torch):
    import dcsr_amd
    R = 325_557
    say("powerlaw 325,557 x 3.2M: generate")
    r, c, v = W.powerlaw(R, 3_200_000, 2.1, 2700, seed=1)
    rng = np.random.default_rng(101)
    batches = [W.batch_mixed(rng, r, c, R, R, 1_000_000) for _ in range(8)]
    x = (1.0 - np.random.default_rng(7).random(R, dtype=np.float32)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    m = dcsr_amd.DcsrMatrix(R, R, pool_factor=8.0)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, R, r, c, v)
    for k, (br, bc, bv) in enumerate(batches):
        # odd batches through the asynchronous call, as the bench times them
        flags = dcsr_amd.UPDATE_NOSYNC if k % 2 else 0
        say(f"  batch {k} ({'nosync' if flags else 'sync'})")
        m.UpdateMatrix(br, bc, bv, flags=flags)
        y = m.spmv(xd)
        A = O.update(A, br, bc, bv)
        torch.cuda.synchronize()
        assert_same_structure(to_host_csr(m), A)  # structure and values bit-exact
        fp64_spmv_check(A, x, y.cpu().numpy())
        info = check_layout(m)
        assert info["nnz"] == A.nnz
    say(f"  final nnz {A.nnz}")
    check_healthy(m)
    m.close()


def _concat(parts):
    off = np.zeros(len(parts) + 1, np.int64)
    off[1:] = np.cumsum([p[0].size for p in parts])
    return off, np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]), \
        np.concatenate([p[2] for p in parts])


def test_batched_4096_graph_steps(torch):
    import dcsr_amd
    say("batched 4,096 matrices: generate")
    count, per = 4096, 2000
    mats = W.many_small(count=count, min_rows=64, max_rows=4096, lam=16.0, seed=4)
    sizes = np.array([mm[0] for mm in mats], np.int32)
    rng = np.random.default_rng(202)
    steps = [[W.batch_mixed(rng, mm[1], mm[2], mm[0], mm[0], per, signed=True) for mm in mats] for _ in range(2)]
    b = dcsr_amd.DcsrBatch(sizes, sizes, pool_factor=8.0)
    b.load_coo(*_concat([(mm[1], mm[2], mm[3]) for mm in mats]))
    rb, cb = b.row_base, b.col_base
    x = rng.uniform(-1, 1, b.total_cols).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty(b.total_rows, dtype=torch.float32, device="cuda")
    team, thresh = b.autotune(xd, yd)
    say(f"  autotuned team={team} thresh={thresh}")
    dev = [torch.from_numpy(a).cuda() for a in _concat(steps[0])]
    oracle = [O.coo_to_csr(n, n, rr, cc, vv) for n, rr, cc, vv in mats]
    for k in range(2):
        if k:
            for d, a in zip(dev, _concat(steps[k])):
                d.copy_(torch.from_numpy(a).cuda())  # same buffers: the captured graph is replayed
        say(f"  step {k}")
        b.step(*dev, x=xd, y=yd, graph=True)
        torch.cuda.synchronize()
        code, per_status = b.status()
        assert code == 0 and not per_status.any()
        oracle = [O.update(A, *steps[k][i]) for i, A in enumerate(oracle)]
        off, col, val = (t.cpu().numpy() for t in b.to_csr())
        for i, A in enumerate(oracle):
            lo, hi = off[rb[i]], off[rb[i + 1]]
            got = O.Csr(A.rows, A.cols, off[rb[i]:rb[i + 1] + 1] - lo, col[lo:hi], val[lo:hi])
            assert_same_structure(got, A)
        # SpMV of the whole block-diagonal batch against FP64
        import scipy.sparse as sp
        rows_of = np.repeat(np.arange(b.total_rows), np.diff(off))
        mat_of_row = np.repeat(np.arange(count), np.diff(rb))
        gcol = col.astype(np.int64) + cb[mat_of_row[rows_of]]
        S = sp.csr_matrix((val.astype(np.float64), gcol, off.astype(np.int64)), shape=(b.total_rows, b.total_cols))
        y64 = S @ x.astype(np.float64)
        scale = abs(S) @ np.abs(x.astype(np.float64))
        err = np.abs(yd.cpu().numpy().astype(np.float64) - y64)
        assert (err <= 1e-5 * scale + 1e-30).all()
    info = b.info()
    assert info["nnz"] == sum(A.nnz for A in oracle)
    b.close()


def test_spgemm_2m_after_4m_update(torch):
    import dcsr_amd
    R = 2_000_000
    say("spgemm 2M Poisson(8): generate")
    r, c, v = W.poisson(R, R, 8.0, seed=3, normal=True)
    rng = np.random.default_rng(303)
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, 4_000_000, signed=True)
    A = dcsr_amd.DcsrMatrix(R, R, pool_factor=2.0)
    A.LoadMatrix(r, c, v)
    say("  4M-triple update")
    A.UpdateMatrix(br, bc, bv)
    Ao = O.update(O.coo_to_csr(R, R, r, c, v), br, bc, bv)
    assert_same_structure(to_host_csr(A), Ao)
    say("  C = A*A")
    C = dcsr_amd.spgemm(A, A)
    Cg = to_host_csr(C)
    say(f"  C nnz {Cg.nnz}; oracle product")
    Ac = Ao.canonical()
    Co = O.spgemm(Ac, Ac)
    say("  compare")
    assert_same_structure(Cg, Co)  # column-sorted A: the same FP32 order, so values are bit-exact
    bad, first = O.spgemm_check_f64(Ac, Ac, Cg, 1e-5)
    assert bad == 0, f"{bad} entries outside 1e-5 of FP64, first at {first}"
    check_layout(C)
    check_healthy(A)
    check_healthy(C)
    A.close()
    C.close()


def test_serving_10m_160m_server_equals_per_call(torch):
    import dcsr_amd
    R, N = 10_000_000, 160_000_000
    say("serving 10M / 160M: generate")
    r, c, v = W.serving_matrix(R, N)
    rng = np.random.default_rng(55)
    reqs = W.serving_stream(rng, r, c, v, R, 40)
    nupd = sum(1 for q in reqs if q[0] == "update")
    say(f"  {len(reqs)} requests, {nupd} updates")
    d_r, d_c, d_v = (torch.from_numpy(a).cuda() for a in (r, c, v))
    del r, c, v
    X = {k: torch.rand((R, k), generator=torch.Generator(device="cuda").manual_seed(k), device="cuda")
         for k in (1, 4, 16, 32)}
    m = dcsr_amd.DcsrMatrix(R, R, pool_factor=2.0)
    m.LoadMatrix(d_r, d_c, d_v)
    srv = dcsr_amd.DcsrServer(m, max_cols=64)
    tickets, outs = [], []
    upd = {}
    for i, q in enumerate(reqs):
        if q[0] == "update":
            upd[i] = tuple(torch.from_numpy(a).cuda() for a in q[2])
            tickets.append(srv.update(*upd[i]))
            outs.append(None)
        elif q[1] == 1:
            t, y = srv.spmv(X[1][:, 0].contiguous())
            tickets.append(t)
            outs.append(y)
        else:
            t, y = srv.spmm(X[q[1]])
            tickets.append(t)
            outs.append(y)
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    srv.close()
    say(f"  server done: {st['query_groups']} groups for {st['queries']} queries")
    assert st["query_groups"] < st["queries"], "queries were not batched"
    alone = dcsr_amd.DcsrMatrix(R, R, pool_factor=2.0)
    alone.LoadMatrix(d_r, d_c, d_v)
    for i, q in enumerate(reqs):
        if q[0] == "update":
            alone.UpdateMatrix(*upd[i])
            continue
        ya = alone.spmv(X[1][:, 0].contiguous()) if q[1] == 1 else alone.spmm(X[q[1]])
        ys = outs[i]
        err = (ys - ya).abs()
        assert bool((err <= 2e-5 * ya.abs() + 1e-30).all()), f"request {i} (k={q[1]}) differs"
        outs[i] = None
    say("  final matrices")
    g = [t for t in m.to_csr()]
    w = [t for t in alone.to_csr()]
    assert torch.equal(g[0], w[0]) and torch.equal(g[1], w[1])
    assert torch.equal(g[2].view(torch.int32), w[2].view(torch.int32))
    check_healthy(m)
    check_healthy(alone)
    m.close()
    alone.close()
