"""GPU parity: SpGEMM C = A*B against the oracle (cusp host-path semantics).

The engine sums each output column's partial products left to right in the
order of A's row entries (then B's), as the host path does
(cusp/system/detail/sequential/multiply/csr_spgemm.h:104-129), and its rows
are always stored column-sorted.  So C equals the host path's product of the
same matrices with column-sorted rows (``O.spgemm(A.canonical(),
B.canonical())``) bit for bit: structure, exact-zero drops and values.  The
tests compare exactly that, including FP32 data built to cancel in an
order-dependent way, and check values within |c - c64| <= 1e-5 (|A||B|)_ik
of an FP64 product.  (Where the host path holds a row in another order, as
cusp::add leaves it, its FP32 sums may differ in the last bits and an
order-dependent cancellation may drop an entry there that this engine
keeps, or the reverse.)"""
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
    Ac = A.canonical()
    G = O.spgemm(Ac, Ac)
    assert_same_structure(C, G, values="exact")
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
    A2c = A2.canonical()
    assert_same_structure(C, O.spgemm(A2c, A2c), values="exact")
    fp64_check(A2, A2, C)
    x = rng.uniform(-1, 1, R).astype(np.float32)
    y = mC.spmv(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    from tests.helpers import fp64_spmv_check
    fp64_spmv_check(C, x, y.cpu().numpy())
    mA.close()
    mC.close()


def cancellation_case(seed=17):
    """A (R x 4R) and B (4R x 2R).  Row i of A holds columns 3i, 3i+1, 3i+2
    and row j < 3R of B holds only column j // 3 with value 1, so C[i, i]
    is A's three values summed in stored order.  The triples are the six
    orders of {1e8, 1, -1e8} (times a power of two): in FP32 the 1 survives
    only in (1e8 - 1e8) + 1 and (-1e8 + 1e8) + 1 (patterns 1 and 3); in the
    four other orders it is absorbed and the sum cancels to 0.  A's other
    entries (columns >= 3R) meet B rows >= 3R whose N(0,1) entries lie in
    columns >= R, so they never touch C's diagonal."""
    rng = np.random.default_rng(seed)
    R = 3000
    base = np.array([[1e8, 1.0, -1e8], [1e8, -1e8, 1.0], [1.0, 1e8, -1e8], [-1e8, 1e8, 1.0],
                     [1.0, -1e8, 1e8], [-1e8, 1.0, 1e8]], np.float32)
    pick = rng.integers(0, len(base), R)
    scale = np.float32(2.0) ** rng.integers(-4, 5, R).astype(np.float32)
    ar = np.repeat(np.arange(R), 3)
    ac = (3 * ar + np.tile(np.arange(3), R)).astype(np.int64)
    av = (base[pick] * scale[:, None]).astype(np.float32).ravel()
    xr = rng.integers(0, R, 6 * R)
    xc = rng.integers(3 * R, 4 * R, 6 * R)
    ar = np.concatenate([ar, xr])
    ac = np.concatenate([ac, xc])
    av = np.concatenate([av, rng.standard_normal(xr.size).astype(np.float32)])
    key, first = np.unique(ar.astype(np.int64) * 4 * R + ac, return_index=True)
    A = O.coo_to_csr(R, 4 * R, (key // (4 * R)).astype(np.int32), (key % (4 * R)).astype(np.int32), av[first])
    br = np.arange(3 * R)
    bc = br // 3
    bv = np.ones(3 * R, np.float32)
    yr = rng.integers(3 * R, 4 * R, 9 * R)
    yc = rng.integers(R, 2 * R, 9 * R)
    br = np.concatenate([br, yr])
    bc = np.concatenate([bc, yc])
    bv = np.concatenate([bv, rng.standard_normal(yr.size).astype(np.float32)])
    key, first = np.unique(br.astype(np.int64) * 2 * R + bc, return_index=True)
    B = O.coo_to_csr(4 * R, 2 * R, (key // (2 * R)).astype(np.int32), (key % (2 * R)).astype(np.int32), bv[first])
    return A, B, pick


def test_float_cancellation_follows_column_sorted_order(dev):
    torch, dcsr_amd = dev
    A, B, pick = cancellation_case()
    R = A.rows
    Ac, Bc = A.canonical(), B.canonical()
    want = O.spgemm(Ac, Bc)
    W_ = want.canonical()
    diag = np.array([(W_.col[W_.off[i]:W_.off[i + 1]] == i).any() for i in range(R)])
    keeps = np.isin(pick, [1, 3])
    assert diag[keeps].all() and not diag[~keeps].any()
    # the same rows stored in reverse order give another structure on the
    # host path: the comparison below is sensitive to the summation order
    rev = O.Csr(Ac.rows, Ac.cols, Ac.off, np.concatenate([Ac.col[Ac.off[i]:Ac.off[i + 1]][::-1] for i in range(R)]),
                np.concatenate([Ac.val[Ac.off[i]:Ac.off[i + 1]][::-1] for i in range(R)]))
    assert O.spgemm(rev, Bc).nnz != want.nnz
    say("float cancellation A*B")
    mA, mB = load(dcsr_amd, A), load(dcsr_amd, B)
    mC = dcsr_amd.spgemm(mA, mB)
    assert_same_structure(host(mC), want, values="exact")
    for m in (mA, mB, mC):
        m.close()


@pytest.mark.parametrize("R,lam,seed", [(50_000, 8.0, 31), (6_000, 40.0, 32)])
def test_random_float_values_bit_exact(dev, R, lam, seed):
    """N(0,1) data: every class of rows (team widths, wave, block) sums in the
    host path's order for column-sorted rows, so values are bit-exact."""
    torch, dcsr_amd = dev
    r, c, v = W.poisson(R, R, lam=lam, seed=seed)
    A = O.coo_to_csr(R, R, r, c, v)
    say(f"random float A*A R={R} lam={lam}")
    mA = load(dcsr_amd, A)
    mC = dcsr_amd.spgemm(mA, mA)
    Ac = A.canonical()
    assert_same_structure(host(mC), O.spgemm(Ac, Ac), values="exact")
    mA.close()
    mC.close()
