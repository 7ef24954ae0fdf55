"""GPU parity: SpGEMM C = A*B against the oracle (cusp host-path semantics).

Structure is compared bit-exactly.  Values: bit-exact where every partial sum
is exactly representable (small-integer data, including exact cancellations
to zero); otherwise within |c - c64| <= 1e-5 (|A||B|)_ik of an FP64 product
(scipy), since the engine accumulates each column in A-column order while the
host path follows its rows' stored order."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, random_coo

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    import dcsr_amd
    dcsr_amd.lib()
    return torch, dcsr_amd


def host(m):
    from tests.helpers import to_host_csr
    return to_host_csr(m)


def fp64_check(A: O.Csr, B: O.Csr, C: O.Csr, rtol=1e-5):
    SA, SB = A.to_scipy(), B.to_scipy()
    want = (SA @ SB).tocsr()
    scale = (abs(SA) @ abs(SB)).tocsr()
    Cc = C.canonical()
    rows = np.repeat(np.arange(C.rows), np.diff(Cc.off))
    w = np.asarray(want[rows, Cc.col]).ravel()
    sc = np.asarray(scale[rows, Cc.col]).ravel()
    err = np.abs(Cc.val.astype(np.float64) - w)
    bad = np.nonzero(err > rtol * sc + 1e-30)[0]
    assert bad.size == 0, f"{bad.size} values out of tolerance"


def load(dcsr_amd, A: O.Csr):
    m = dcsr_amd.DcsrMatrix(A.rows, A.cols)
    m.load_csr(A.off, A.col, A.val)
    return m


def test_banded_square_vs_reference_fixture(dev):
    torch, dcsr_amd = dev
    d = np.load(os.path.join(GOLD, "banded256.npz"))
    R = int(d["R"])
    A = O.Csr(R, R, d["upd3_off"], d["upd3_col"], d["upd3_val"])
    G = O.Csr(R, R, d["spgemm_off"], d["spgemm_col"], d["spgemm_val"])
    say("banded256 A*A")
    mA = load(dcsr_amd, A)
    mC = dcsr_amd.spgemm(mA, mA)
    C = host(mC)
    assert_same_structure(C, G, values=None)
    fp64_check(A, A, C)
    mA.close()
    mC.close()


def test_integer_cancellation_fixture_exact(dev):
    torch, dcsr_amd = dev
    d = np.load(os.path.join(GOLD, "edge.npz"))
    R, Cn = int(d["R"]), int(d["C"])
    A = O.Csr(R, Cn, d["upd0_off"], d["upd0_col"], d["upd0_val"])
    G = O.Csr(R, Cn, d["spgemm_off"], d["spgemm_col"], d["spgemm_val"])
    say("edge A*A (integer, exact zeros)")
    mA = load(dcsr_amd, A)
    mC = dcsr_amd.spgemm(mA, mA)
    assert_same_structure(host(mC), G, values="exact")
    mA.close()
    mC.close()


# the last three: the widest B the 32-bit sort keys allow, for rows of <= 128
# partials (columns up to 2^25 - 2, keys just below the padding value) and of
# 129..1024 partials (up to 2^22 - 2), and one column past the first limit
# (small rows then take the 64-bit-key path)
@pytest.mark.parametrize("R,K,N,nnzA,nnzB,seed", [(1, 1, 1, 1, 1, 0), (50, 70, 30, 400, 500, 1),
                                                  (3000, 2000, 2500, 30000, 25000, 2),
                                                  (300, 400, (1 << 25) - 1, 3000, 4000, 3),
                                                  (200, 300, (1 << 22) - 1, 4000, 6000, 5),
                                                  (300, 400, 1 << 25, 3000, 4000, 4)])
def test_rectangular_small_int(dev, R, K, N, nnzA, nnzB, seed):
    torch, dcsr_amd = dev
    rng = np.random.default_rng(seed)
    A = O.coo_to_csr(R, K, *random_coo(rng, R, K, nnzA, 0.1, 0.0, "small_int"))
    B = O.coo_to_csr(K, N, *random_coo(rng, K, N, nnzB, 0.1, 0.0, "small_int"))
    say(f"rectangular {R}x{K} * {K}x{N}")
    mA, mB = load(dcsr_amd, A), load(dcsr_amd, B)
    mC = dcsr_amd.spgemm(mA, mB)
    # small integers: every accumulation order is exact, so values match too
    assert_same_structure(host(mC), O.spgemm(A, B), values="exact")
    for m in (mA, mB, mC):
        m.close()


def test_hub_rows_global_path(dev):
    """Power-law A*A: rows with > 4096 partial products (all three classes)."""
    torch, dcsr_amd = dev
    R = 20000
    r, c, v = W.powerlaw(R, 200000, max_degree=3000, seed=41, signed=True)
    A = O.coo_to_csr(R, R, r, c, v)
    P = np.array([np.diff(A.off)[A.col[A.off[i]:A.off[i + 1]]].sum() for i in range(R)])
    assert (P > 4096).any() and ((P > 1024) & (P <= 4096)).any() and ((P > 0) & (P <= 1024)).any()
    say("powerlaw A*A with hub rows")
    mA = load(dcsr_amd, A)
    mC = dcsr_amd.spgemm(mA, mA)
    C = host(mC)
    G = O.spgemm(A, A)
    assert_same_structure(C, G, values=None)
    fp64_check(A, A, C)
    mC2 = dcsr_amd.spgemm(mA, mA)
    np.testing.assert_array_equal(host(mC2).val, C.val)  # deterministic
    for m in (mA, mC, mC2):
        m.close()


def test_spgemm_after_update_and_use_of_result(dev):
    torch, dcsr_amd = dev
    rng = np.random.default_rng(9)
    R = 4000
    r, c, v = W.poisson(R, R, lam=8.0, seed=9)
    A = O.coo_to_csr(R, R, r, c, v)
    mA = load(dcsr_amd, A)
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, 20000, signed=True)
    say("update then A*A, then SpMV on C")
    mA.UpdateMatrix(br, bc, bv)
    A2 = O.update(A, br, bc, bv)
    mC = dcsr_amd.spgemm(mA, mA)
    C = host(mC)
    assert_same_structure(C, O.spgemm(A2, A2), values=None)
    fp64_check(A2, A2, C)
    x = rng.uniform(-1, 1, R).astype(np.float32)
    y = mC.spmv(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    from tests.helpers import fp64_spmv_check
    fp64_spmv_check(C, x, y.cpu().numpy())
    mA.close()
    mC.close()
