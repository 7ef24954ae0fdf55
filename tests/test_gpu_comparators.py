"""GPU parity of the CSR-rebuild comparators (csrcmp.hip) against the oracle
and against the Dynamic-CSR engine itself.

dcsr_csr_add rebuilds C = A + B whole (cusp's generic add algorithm) and must
equal the oracle's cusp::add host path, which dcsr_update also follows, bit
for bit.  dcsr_csr_spgemm is a global-ESC product and must equal the oracle's
host product of the same column-sorted matrices, and dcsr_spgemm, bit for
bit (structure, exact-zero drops, values)."""
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, random_coo, to_host_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-cmp] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    import dcsr_amd
    dcsr_amd.lib()
    return torch, dcsr_amd


def merged(R, C, r, c, v):
    """A CSR matrix without duplicate coordinates: the host path's
    cusp::add of the triples into an empty matrix, columns then sorted."""
    e = np.zeros(0, np.int32)
    return O.update(O.coo_to_csr(R, C, e, e, np.zeros(0, np.float32)), r, c, v).canonical()


def host(R, C, t):
    o, c, v = t
    return O.Csr(R, C, o.cpu().numpy(), c.cpu().numpy(), v.cpu().numpy())


@pytest.mark.parametrize("R,C,n,seed", [(1, 1, 3, 0), (40, 70, 900, 1), (3000, 2000, 50_000, 2),
                                        (100_000, 100_000, 400_000, 3)])
def test_csr_add_chain_matches_oracle_and_update(dev, R, C, n, seed):
    torch, D = dev
    rng = np.random.default_rng(seed)
    r, c, v = random_coo(rng, R, C, n, 0.0, 0.0, "sym")
    # as loaded: random coordinates collide, so A holds duplicates at first
    A = O.coo_to_csr(R, C, r, c, v)
    m = D.DcsrMatrix(R, C)
    m.LoadMatrix(r, c, v)
    cur = (torch.from_numpy(A.off).cuda(), torch.from_numpy(A.col).cuda(), torch.from_numpy(A.val).cuda())
    for k in range(3):
        br, bc, bv = random_coo(rng, R, C, max(1, n // 3), 0.2, 0.02, "small_int" if k == 1 else "sym")
        if k == 2 and A.nnz:  # exact cancellations (deletions)
            rows_of = np.repeat(np.arange(R), np.diff(A.off))
            pick = rng.integers(0, A.nnz, max(1, A.nnz // 6))
            br = np.concatenate([br, rows_of[pick]]).astype(np.int32)
            bc = np.concatenate([bc, A.col[pick]]).astype(np.int32)
            bv = np.concatenate([bv, -A.val[pick]]).astype(np.float32)
        say(f"R={R} add {k}: {br.size} triples")
        cur = D.csr_add(R, C, *cur, br, bc, bv)
        A = O.update(A, br, bc, bv)
        got = host(R, C, cur)
        assert_same_structure(got, A)
        assert all((np.diff(got.col[got.off[i]:got.off[i + 1]]) > 0).all() for i in range(R))  # sorted, unique
        m.UpdateMatrix(br, bc, bv)
        assert_same_structure(got, to_host_csr(m))
    m.close()


def test_csr_add_power_law_batch(dev):
    torch, D = dev
    R = 50_000
    r, c, v = W.powerlaw(R, 500_000, max_degree=3000, seed=7)
    A = O.coo_to_csr(R, R, r, c, v)
    rng = np.random.default_rng(8)
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, 200_000)
    say("power-law add")
    C = host(R, R, D.csr_add(R, R, A.off, A.col, A.val, br, bc, bv))
    assert_same_structure(C, O.update(A, br, bc, bv))


def test_csr_add_rejects_bad_index(dev):
    torch, D = dev
    R = 100
    A = O.coo_to_csr(R, R, np.array([1], np.int32), np.array([2], np.int32), np.array([1.0], np.float32))
    with pytest.raises(D.DcsrError):
        D.csr_add(R, R, A.off, A.col, A.val, np.array([5, 100], np.int32), np.array([1, 1], np.int32),
                  np.array([1.0, 2.0], np.float32))


@pytest.mark.parametrize("R,K,C,lam,seed", [(1, 1, 1, 1.0, 0), (300, 200, 250, 6.0, 1), (20_000, 20_000, 20_000, 8.0, 2)])
def test_csr_spgemm_matches_oracle_and_dcsr(dev, R, K, C, lam, seed):
    torch, D = dev
    rng = np.random.default_rng(seed)
    ar, ac, av = random_coo(rng, R, K, int(R * lam), 0.0, 0.0, "sym")
    br, bc, bv = random_coo(rng, K, C, int(K * lam), 0.0, 0.0, "sym")
    A = merged(R, K, ar, ac, av)
    B = merged(K, C, br, bc, bv)
    say(f"spgemm {R}x{K}x{C}")
    got = host(R, C, D.csr_spgemm(R, K, C, A.off, A.col, A.val, B.off, B.col, B.val))
    want = O.spgemm(A, B)
    assert_same_structure(got, want)
    ma, mb = D.DcsrMatrix(R, K), D.DcsrMatrix(K, C)
    ma.load_csr(A.off, A.col, A.val)
    mb.load_csr(B.off, B.col, B.val)
    mc = D.spgemm(ma, mb)
    assert_same_structure(got, to_host_csr(mc))
    for h in (ma, mb, mc):
        h.close()


def test_csr_spgemm_order_dependent_cancellation(dev):
    """Row 0 of A has (1e8, 1, -1e8) on columns 0..2 and B's rows 0..2 all
    hit column 0: in A-row order (1e8 + 1) - 1e8 = 0 in FP32, so the entry
    is dropped, as in the host path; (0, 1) gets 1e8 and stays."""
    torch, D = dev
    A = O.coo_to_csr(1, 3, np.array([0, 0, 0], np.int32), np.array([0, 1, 2], np.int32),
                     np.array([1e8, 1.0, -1e8], np.float32))
    B = O.coo_to_csr(3, 2, np.array([0, 0, 1, 2], np.int32), np.array([0, 1, 0, 0], np.int32),
                     np.array([1.0, 1.0, 1.0, 1.0], np.float32))
    got = host(1, 2, D.csr_spgemm(1, 3, 2, A.off, A.col, A.val, B.off, B.col, B.val))
    want = O.spgemm(A, B)
    assert want.nnz == 1 and want.col[0] == 1
    assert_same_structure(got, want)
