"""GPU parity: dense multi-column SpMM (Y = alpha A X + beta Y) against FP64
(scipy) with the conditioned bound |Y - Y64| <= 1e-5 (|A||X| + |beta Y|)."""
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    import torch
    import dcsr_amd
    dcsr_amd.lib()
    R = 6000
    r, c, v = W.powerlaw(R, 120000, max_degree=5000, seed=31, signed=True)
    # empty rows at both ends and in the middle
    keep = (r > 50) & (r < R - 40) & ((r < 3000) | (r > 3010))
    r, c, v = r[keep], c[keep], v[keep]
    m = dcsr_amd.DcsrMatrix(R, R)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, R, r, c, v)
    yield torch, m, A
    m.close()


def check_spmm(A, X, Y0, Y, alpha, beta):
    S = A.to_scipy(np.float64)
    want = alpha * (S @ X.astype(np.float64)) + beta * Y0.astype(np.float64)
    scale = abs(alpha) * (abs(S) @ np.abs(X.astype(np.float64))) + abs(beta) * np.abs(Y0.astype(np.float64))
    err = np.abs(Y.astype(np.float64) - want)
    bad = np.argwhere(err > 1e-5 * scale + 1e-30)
    assert bad.size == 0, f"{len(bad)} entries out of tolerance, first {bad[:3]}"


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 8, 12, 16, 31, 32, 33, 40])
def test_spmm_columns(setup, k):
    torch, m, A = setup
    rng = np.random.default_rng(k)
    ldx = k + (3 if k % 2 else 0)
    Xfull = rng.uniform(-1, 1, (A.cols, ldx)).astype(np.float32)
    X = Xfull[:, :k]
    say(f"spmm k={k} ldx={ldx}")
    Xd = torch.from_numpy(Xfull).cuda()[:, :k]
    Y = m.spmm(Xd)
    torch.cuda.synchronize()
    check_spmm(A, X, np.zeros((A.rows, k), np.float32), Y.cpu().numpy(), 1.0, 0.0)
    Y2 = m.spmm(Xd)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Y.cpu().numpy(), Y2.cpu().numpy())  # reproducible


@pytest.mark.parametrize("k", [4, 16, 32])
def test_spmm_alpha_beta(setup, k):
    torch, m, A = setup
    rng = np.random.default_rng(100 + k)
    X = rng.uniform(-1, 1, (A.cols, k)).astype(np.float32)
    Y0 = rng.uniform(-1, 1, (A.rows, k)).astype(np.float32)
    say(f"spmm alpha/beta k={k}")
    Y = torch.from_numpy(Y0.copy()).cuda()
    m.spmm(torch.from_numpy(X).cuda(), Y, alpha=0.75, beta=-1.25)
    torch.cuda.synchronize()
    check_spmm(A, X, Y0, Y.cpu().numpy(), 0.75, -1.25)


def test_spmm_after_updates(setup):
    torch, m, A = setup
    rng = np.random.default_rng(5)
    R = A.rows
    br = rng.integers(0, R, 30000).astype(np.int32)
    bc = rng.integers(0, R, 30000).astype(np.int32)
    bv = rng.uniform(-1, 1, 30000).astype(np.float32)
    say("spmm after update")
    m.UpdateMatrix(br, bc, bv)
    A2 = O.update(A, br, bc, bv)
    X = rng.uniform(-1, 1, (R, 16)).astype(np.float32)
    Y = m.spmm(torch.from_numpy(X).cuda())
    torch.cuda.synchronize()
    check_spmm(A2, X, np.zeros((R, 16), np.float32), Y.cpu().numpy(), 1.0, 0.0)


@pytest.fixture(scope="module")
def wide_setup():
    """A matrix whose dense operand exceeds one XCD's 4 MB L2 even for one
    column (1.2M columns): SpMV and SpMM stream its entries with
    non-temporal loads.  Power-law rows, so carries and hub rows occur."""
    import torch
    import dcsr_amd
    dcsr_amd.lib()
    R, C = 3000, 1_200_000
    rng = np.random.default_rng(17)
    d = np.minimum(rng.zipf(1.7, R), 20000)
    rows = np.repeat(np.arange(R, dtype=np.int64), d)
    cols = rng.integers(0, C, rows.size)
    key = np.unique(rows * C + cols)
    r, c = (key // C).astype(np.int32), (key % C).astype(np.int32)
    v = rng.uniform(-1, 1, r.size).astype(np.float32)
    m = dcsr_amd.DcsrMatrix(R, C)
    m.LoadMatrix(r, c, v)
    yield torch, m, O.coo_to_csr(R, C, r, c, v)
    m.close()


@pytest.mark.parametrize("k", [1, 3, 4, 16, 32])
def test_spmm_streamed_entries(wide_setup, k):
    torch, m, A = wide_setup
    rng = np.random.default_rng(200 + k)
    ldx = k + (1 if k == 16 else 0)  # k = 16 with an unaligned row: the scalar wide kernel
    Xfull = rng.uniform(-1, 1, (A.cols, ldx)).astype(np.float32)
    X = Xfull[:, :k]
    say(f"streamed spmm k={k} ldx={ldx} cols={A.cols}")
    Xd = torch.from_numpy(Xfull).cuda()[:, :k]
    Y = m.spmv(Xd[:, 0]).unsqueeze(1) if k == 1 else m.spmm(Xd)
    torch.cuda.synchronize()
    check_spmm(A, X, np.zeros((A.rows, k), np.float32), Y.cpu().numpy(), 1.0, 0.0)
