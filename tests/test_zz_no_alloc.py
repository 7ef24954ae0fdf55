"""The allocation-free update entry point (dcsr_update_noalloc) against the
allocating one (dcsr_update): the same matrix, the same batches, two calls in
a row with the second one smaller, synchronous and NOSYNC; results compared
bit-exactly (structure and values) and with the oracle.  Sizes past the fixed
scratch are rejected with DCSR_ERR_TOO_LARGE and leave the matrix unchanged."""
import ctypes
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, check_healthy, to_host_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def D():
    import dcsr_amd
    dcsr_amd.lib()
    return dcsr_amd


def noalloc(D, m, br, bc, bv, flags=0):
    import torch
    r = torch.from_numpy(np.ascontiguousarray(br, np.int32)).cuda()
    c = torch.from_numpy(np.ascontiguousarray(bc, np.int32)).cuda()
    v = torch.from_numpy(np.ascontiguousarray(bv, np.float32)).cuda()
    L = D.lib()
    code = L.dcsr_update_noalloc(m.handle, r.data_ptr(), c.data_ptr(), v.data_ptr(), r.numel(), flags,
                                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return code


@pytest.mark.parametrize("flags", [0, 1])
def test_noalloc_matches_update_two_calls(D, flags):
    from dcsr_amd import DcsrMatrix, UPDATE_NOSYNC
    fl = UPDATE_NOSYNC if flags else 0
    R = 40_000
    r, c, v = W.powerlaw(R, 400_000, max_degree=3000, seed=71, signed=True)
    rng = np.random.default_rng(72)
    ref = DcsrMatrix(R, R, pool_factor=4.0)
    m = DcsrMatrix(R, R, pool_factor=4.0)
    for h in (ref, m):
        h.LoadMatrix(r, c, v)
        h.UpdateMatrix(np.zeros(1, np.int32), np.zeros(1, np.int32), np.ones(1, np.float32))  # canonical
    A = O.coo_to_csr(R, R, r, c, v)
    A = O.update(A, np.zeros(1, np.int32), np.zeros(1, np.int32), np.ones(1, np.float32))
    maxn = D.lib().dcsr_update_noalloc_max_triples()
    for k, n in enumerate([maxn, 5_000]):  # the second call smaller
        br, bc, bv = W.batch_mixed(rng, r, c, R, R, n, signed=True)
        say(f"flags={flags} call {k}: {n} triples")
        ref.UpdateMatrix(br, bc, bv, flags=fl)
        assert noalloc(D, m, br, bc, bv, fl) == 0
        A = O.update(A, br, bc, bv)
        got = to_host_csr(m)
        assert_same_structure(got, to_host_csr(ref))
        assert_same_structure(got, A)
        assert m.info()["nnz"] == A.nnz
    check_healthy(m)
    for h in (ref, m):
        h.close()


def test_noalloc_rejects_past_the_limit(D):
    from dcsr_amd import DcsrMatrix
    R = 1000
    rng = np.random.default_rng(73)
    m = DcsrMatrix(R, R)
    m.LoadMatrix(np.arange(R, dtype=np.int32), np.arange(R, dtype=np.int32), np.ones(R, np.float32))
    m.UpdateMatrix(np.zeros(1, np.int32), np.zeros(1, np.int32), np.ones(1, np.float32))
    before = to_host_csr(m)
    n = D.lib().dcsr_update_noalloc_max_triples() + 1
    br = rng.integers(0, R, n).astype(np.int32)
    code = noalloc(D, m, br, br, np.ones(n, np.float32))
    assert code == D.DCSR_ERR_TOO_LARGE
    assert_same_structure(to_host_csr(m), before)
    m.close()
