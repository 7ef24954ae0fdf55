"""GPU run of the real-matrix suite flow (dcsr_amd.suite, the reference's
FillTests per matrix): the reference's bundled MatrixMarket files and small
synthetic look-alikes of the SuiteSparse list, each through load, compaction,
SpMV and 50 update slices in DCSR, HYB and rebuilt CSR.  The runner's own
check compares the three formats; here the DCSR result is also compared with
the oracle's cusp::add chain over the same slices (bit-exact)."""
import os
import sys

import numpy as np
import pytest

import oracle as O
from tests.helpers import assert_same_structure

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def D():
    import dcsr_amd
    dcsr_amd.lib()
    return dcsr_amd


def oracle_final(R, C, r, c, v, trace):
    A = O.coo_to_csr(R, C, r, c, v)
    for br, bc, bv in trace["slices"]:
        A = O.update(A, br, bc, bv)
    return A


@pytest.mark.parametrize("name", ["can_256.mtx", "tols90.mtx"])
def test_bundled_matrix_flow(D, name):
    from dcsr_amd import suite
    print(f"[gpu-suite] {name}", flush=True)
    R, C, r, c, v = D.read_mtx(os.path.join(GOLD, name))
    trace = {}
    res = suite.run_matrix(name, R, C, r, c, v, check=True, trace=trace)
    assert res["check"] == "ok", res
    off, col, val = trace["final"]
    assert_same_structure(O.Csr(R, C, off, col, val), oracle_final(R, C, r, c, v, trace))


@pytest.mark.parametrize("name,scale", [("cnr-2000", 0.05), ("rail4284", 0.02), ("enron", 0.2)])
def test_lookalike_flow(D, name, scale):
    from dcsr_amd import suite
    print(f"[gpu-suite] {name} x{scale}", flush=True)
    (R0, C0, N0), k = next(((R, C, N), k) for k, (n, R, C, N) in enumerate(suite.SUITE) if n == name)
    R = max(1, int(R0 * scale))
    C = R if R0 == C0 else C0
    r, c, v = suite.lookalike(R, C, int(N0 * scale), 100 + k)
    trace = {}
    res = suite.run_matrix(name, R, C, r, c, v, check=True, trace=trace)
    assert res["check"] == "ok", res
    off, col, val = trace["final"]
    assert_same_structure(O.Csr(R, C, off, col, val), oracle_final(R, C, r, c, v, trace))
