"""GPU parity: DCSR <-> HYB conversions against the oracle's host CSR -> HYB
conversion (pinned to cusp's generic csr->hyb in test_oracle_ref.py)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, random_coo, to_host_csr

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
