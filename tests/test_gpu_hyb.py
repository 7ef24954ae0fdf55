"""GPU parity: DCSR <-> HYB conversions against the oracle's host CSR -> HYB
conversion (pinned to cusp's generic csr->hyb in test_oracle_ref.py)."""
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


@pytest.fixture(scope="module")
def dev():
    import torch
    import dcsr_amd
    dcsr_amd.lib()
    return torch, dcsr_amd


def cpu(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("width", [0, 1, 4, 19, 64])
def test_export_hyb_matches_host_conversion(dev, width):
    torch, dcsr_amd = dev
    d = np.load(os.path.join(GOLD, "banded256.npz"))
    R = int(d["R"])
    A = O.Csr(R, R, d["upd3_off"], d["upd3_col"], d["upd3_val"]).canonical()
    say(f"export_hyb width={width}")
    m = dcsr_amd.DcsrMatrix(R, R)
    m.load_csr(A.off, A.col, A.val)
    ec, ev, cr, cc, cv, pitch = m.to_hyb(width)
    torch.cuda.synchronize()
    wec, wev, wcr, wcc, wcv, wpitch = O.csr_to_hyb(A, width)
    assert pitch == wpitch
    np.testing.assert_array_equal(cpu(ec), wec)
    np.testing.assert_array_equal(cpu(ev).view(np.uint32), wev.view(np.uint32))
    np.testing.assert_array_equal(cpu(cr), wcr)
    np.testing.assert_array_equal(cpu(cc), wcc)
    np.testing.assert_array_equal(cpu(cv).view(np.uint32), wcv.view(np.uint32))
    say("load_hyb round trip")
    m2 = dcsr_amd.DcsrMatrix(R, R)
    m2.load_hyb(width, pitch, ec, ev, cr, cc, cv)
    assert_same_structure(to_host_csr(m2), A)
    m.close()
    m2.close()


def test_load_hyb_duplicates_and_padding(dev):
    torch, dcsr_amd = dev
    rng = np.random.default_rng(3)
    R, C, width = 70, 50, 3
    pitch = 96
    ec = np.full(width * pitch, -1, np.int32)
    ev = np.zeros(width * pitch, np.float32)
    trip = []
    for k in range(width):
        for r in range(R):
            if rng.random() < 0.6:
                c = int(rng.integers(0, 10))  # narrow columns: duplicates likely
                v = float(rng.integers(-2, 3))
                ec[k * pitch + r], ev[k * pitch + r] = c, v
                trip.append((r, c, v))
    # garbage in padding rows >= R must be ignored
    ec[R:pitch] = 7
    n = 120
    cr = rng.integers(0, R, n).astype(np.int32)
    cc = rng.integers(0, 10, n).astype(np.int32)
    cv = rng.integers(-2, 3, n).astype(np.float32)
    trip += list(zip(cr.tolist(), cc.tolist(), cv.tolist()))
    tr = np.array([t[0] for t in trip], np.int32)
    tc = np.array([t[1] for t in trip], np.int32)
    tv = np.array([t[2] for t in trip], np.float32)
    say("load_hyb with duplicates")
    m = dcsr_amd.DcsrMatrix(R, C)
    m.load_hyb(width, pitch, ec, ev, cr, cc, cv)
    want = O.coo_to_csr(R, C, tr, tc, tv)
    assert_same_structure(to_host_csr(m), want)
    # and the canonical form after the first update equals the host path
    m.UpdateMatrix(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))
    E = O.Csr(R, C, np.zeros(R + 1, np.int32), [], [])
    assert_same_structure(to_host_csr(m), O.csr_add(want, E))
    m.close()


def test_hyb_powerlaw_roundtrip(dev):
    torch, dcsr_amd = dev
    R = 30000
    r, c, v = W.powerlaw(R, 300000, max_degree=6000, seed=8, signed=True)
    A = O.coo_to_csr(R, R, r, c, v)
    say("powerlaw hyb round trip")
    m = dcsr_amd.DcsrMatrix(R, R)
    m.LoadMatrix(r, c, v)
    ec, ev, cr, cc, cv, pitch = m.to_hyb(10)
    m2 = dcsr_amd.DcsrMatrix(R, R)
    m2.load_hyb(10, pitch, ec, ev, cr, cc, cv)
    assert_same_structure(to_host_csr(m2), A)
    m.close()
    m2.close()


def hyb_host(H):
    ec, ev, rs, cr, cc, cv = (t.cpu().numpy() for t in H.export())
    return O.Hyb(H.num_rows, H.num_cols, H.width, H.pitch, rs, ec, ev, cr, cc, cv)


def assert_same_hyb(got, want):
    """Bit-exact: ELL slots (columns and values), row sizes, COO in order."""
    assert got.pitch == want.pitch and got.width == want.width
    np.testing.assert_array_equal(got.row_sizes, want.row_sizes)
    np.testing.assert_array_equal(got.ell_col, want.ell_col)
    np.testing.assert_array_equal(got.ell_val.view(np.uint32), want.ell_val.view(np.uint32))
    np.testing.assert_array_equal(got.coo_row, want.coo_row)
    np.testing.assert_array_equal(got.coo_col, want.coo_col)
    np.testing.assert_array_equal(got.coo_val.view(np.uint32), want.coo_val.view(np.uint32))


@pytest.mark.parametrize("R,C,n,width,seed", [(1, 1, 3, 1, 0), (300, 200, 4000, 4, 1), (5000, 4000, 80000, 16, 2),
                                              (20000, 20000, 300000, 8, 3)])
def test_hyb_store_load_update_spmv_match_oracle(R, C, n, width, seed):
    """The standalone HYB store: COO load, three append updates (duplicates,
    explicit zeros, rows pushed past the ELL width), CSR load; every state
    equals the oracle bit for bit and SpMV equals the host hyb multiply bit
    for bit (same left-to-right order), within 1e-5 of FP64 too."""
    import torch
    import dcsr_amd
    rng = np.random.default_rng(seed)
    r, c, v = random_coo(rng, R, C, n, 0.1, 0.02, "sym")
    say(f"hyb store R={R} n={n} width={width}")
    H = dcsr_amd.DcsrHyb(R, C, width, coo_cap=16)
    H.LoadMatrix(r, c, v)
    want = O.coo_to_hyb(R, C, r, c, v, width)
    assert_same_hyb(hyb_host(H), want)
    x = rng.uniform(-1, 1, C).astype(np.float32)
    for k in range(3):
        br, bc, bv = random_coo(rng, R, C, max(1, n // 3), 0.2, 0.05, "sym")
        if k == 1:  # many entries on a few rows: deep overflow
            br[: br.size // 2] = rng.integers(0, min(R, 7), br.size // 2)
        H.UpdateMatrix(br, bc, bv)
        want = O.hyb_update(want, br, bc, bv)
        assert_same_hyb(hyb_host(H), want)
        y = H.spmv(torch.from_numpy(x).cuda())
        torch.cuda.synchronize()
        yw = O.hyb_spmv(want, x)
        np.testing.assert_array_equal(y.cpu().numpy().view(np.uint32), yw.view(np.uint32))
        hr, hc, hv = want.to_coo()
        fp64_spmv_check(O.coo_to_csr(R, C, hr, hc, hv), x, y.cpu().numpy())
    say("hyb store CSR load")
    A = O.coo_to_csr(R, C, r, c, v)
    H.load_csr(A.off, A.col, A.val)
    assert_same_hyb(hyb_host(H), O.csr_to_hyb_full(A, width))
    H.close()


def test_hyb_store_dcsr_round_trip_and_bad_index():
    """HYB store -> DCSR keeps every stored entry (duplicates summed on the
    first update, as any loaded matrix); DCSR -> HYB store gives the
    column-sorted conversion; an out-of-range index rejects the update and
    leaves the store unchanged."""
    import torch
    import dcsr_amd
    from dcsr_amd import DcsrError
    rng = np.random.default_rng(7)
    R, C, width = 3000, 2500, 6
    r, c, v = random_coo(rng, R, C, 40000, 0.1, 0.0, "sym")
    H = dcsr_amd.DcsrHyb(R, C, width)
    H.LoadMatrix(r, c, v)
    br, bc, bv = random_coo(rng, R, C, 10000, 0.3, 0.0, "sym")
    H.UpdateMatrix(br, bc, bv)
    want = O.hyb_update(O.coo_to_hyb(R, C, r, c, v, width), br, bc, bv)
    say("hyb store -> dcsr")
    m = dcsr_amd.DcsrMatrix(R, C)
    H.to_dcsr(m)
    hr, hc, hv = want.to_coo()
    assert_same_structure(to_host_csr(m), O.coo_to_csr(R, C, hr, hc, hv))
    say("dcsr -> hyb store")
    m.UpdateMatrix(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))  # canonicalize
    H2 = dcsr_amd.DcsrHyb(R, C, width)
    H2.from_dcsr(m)
    assert_same_hyb(hyb_host(H2), O.csr_to_hyb_full(to_host_csr(m), width))
    say("bad index update")
    before = hyb_host(H)
    with pytest.raises(DcsrError) as e:
        H.UpdateMatrix(np.array([1, R], np.int32), np.array([0, 0], np.int32), np.ones(2, np.float32))
    assert e.value.code == 2
    assert_same_hyb(hyb_host(H), before)
    for x in (H, H2, m):
        x.close()
