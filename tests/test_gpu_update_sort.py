"""GPU parity of the update's batch sort paths against the oracle.

Batches of up to 2^21 triples are sorted by row group (rows split into groups
of 2^gshift consecutive rows: one scatter, then one workgroup sorts each
group; gsort.hip); larger batches use the LSD radix sort.  A group of up to
3,072 triples sorts in LDS: in one 32-bit word per triple when its local key
(gshift + cbits bits) has at most 20 bits beyond the first digit (28 with
8-bit digits, 29 with 9-bit ones), else with 32-bit local keys up to 32 bits
and 64-bit ones beyond; a larger group sorts out of global memory.  Each
case below is built so that one of these paths carries it; structure and
update values are compared bit-exactly with the oracle, SpMV within 1e-5 of
FP64 (conditioned)."""
import sys

import numpy as np
import pytest

import oracle as O
from dcsr_amd import workloads as W
from tests.helpers import assert_same_structure, check_healthy, check_layout, fp64_spmv_check, random_coo, \
    to_host_csr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-sort] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    import dcsr_amd
    dcsr_amd.lib()
    return t


def spmv_host(mat, x):
    import torch
    y = mat.spmv(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    return y.cpu().numpy()


def run_chain(R, C, base, batches, seed, pool_factor=2.0):
    from dcsr_amd import DcsrMatrix
    rng = np.random.default_rng(seed)
    r, c, v = base
    m = DcsrMatrix(R, C, pool_factor=pool_factor)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    for k, (br, bc, bv) in enumerate(batches):
        say(f"  batch {k}: {br.size} triples")
        m.UpdateMatrix(br, bc, bv)
        A = O.update(A, br, bc, bv)
        assert_same_structure(to_host_csr(m), A)
        info = check_layout(m)
        assert info["nnz"] == A.nnz
        x = rng.uniform(-1, 1, C).astype(np.float32)
        fp64_spmv_check(A, x, spmv_host(m, x))
    check_healthy(m)
    m.close()


def test_groups_in_lds_powerlaw(torch):
    """Power-law matrix, degree-drawn batches with in-batch duplicates: every
    group fits in LDS (32-bit local keys)."""
    R = 60_000
    say("power-law groups in LDS")
    r, c, v = W.powerlaw(R, 600_000, max_degree=2000, seed=11, signed=True)
    rng = np.random.default_rng(12)
    batches = []
    for k in range(3):
        br, bc, bv = W.batch_mixed(rng, r, c, R, R, 200_000, signed=True)
        dup = rng.integers(0, br.size, 20_000)
        batches.append((np.concatenate([br, br[dup]]), np.concatenate([bc, bc[dup]]),
                        np.concatenate([bv, bv[dup]])))
    run_chain(R, R, (r, c, v), batches, 13)


@pytest.mark.parametrize("R,n,lb", [(100_000, 300_000, 25), (1 << 20, 600_000, 30)])
def test_group_key_widths(torch, R, n, lb):
    """Groups in LDS at both key widths of the 32-bit sort: 100,000 rows and
    300,000 triples give groups of 2^8 rows and 25-bit local keys (9-bit
    digits, one word per triple); 2^20 rows and 600,000 triples give groups
    of 2^10 rows and 30-bit local keys (8-bit digits, key and position
    arrays)."""
    cbits = (R - 1).bit_length()
    want = max(n // 512, 1)
    gshift = 0
    while (R + (1 << gshift) - 1) >> gshift > want:
        gshift += 1
    assert gshift + cbits == lb
    say(f"key width R={R} n={n} lb={lb}")
    rng = np.random.default_rng(R % 1009 + n)
    base = random_coo(rng, R, R, 400_000, 0.0, 0.0, "sym")
    batches = [random_coo(rng, R, R, n, 0.2, 0.01, "sym") for _ in range(2)]
    br, bc, bv = batches[1]
    br[:5000], bc[:5000] = base[0][:5000], base[1][:5000]
    run_chain(R, R, base, batches, 7)


def test_oversized_groups_sort_in_global_memory(torch):
    """Hot rows: one row takes 30,000 triples (with repeats) and a second
    group 9,000, far past the 4,096 a group sorts in LDS; the other groups
    stay small."""
    R, C = 20_000, 50_000
    say("oversized groups")
    rng = np.random.default_rng(31)
    r, c, v = random_coo(rng, R, C, 100_000, 0.0, 0.0, "sym")
    batches = []
    for k in range(2):
        hr = np.full(30_000, 777 + k, np.int32)
        hc = rng.integers(0, C, 30_000).astype(np.int32)
        wr = rng.integers(5000, 5008, 9_000).astype(np.int32)
        wc = rng.integers(0, C, 9_000).astype(np.int32)
        sr, sc, sv = random_coo(rng, R, C, 20_000, 0.1, 0.01, "sym")
        br = np.concatenate([hr, sr, wr])
        bc = np.concatenate([hc, sc, wc])
        bv = rng.uniform(-1, 1, br.size).astype(np.float32)
        perm = rng.permutation(br.size)
        batches.append((br[perm], bc[perm], bv[perm]))
    run_chain(R, C, (r, c, v), batches, 32)


def test_wide_local_keys(torch):
    """4M x 4M with a 6,000-triple batch: few groups of 2^19 rows, so a
    group's local key (19 row bits + 22 column bits) needs 64 bits."""
    R = C = 1 << 22
    say("64-bit local keys")
    rng = np.random.default_rng(41)
    r, c, v = random_coo(rng, R, C, 50_000, 0.0, 0.0, "sym")
    batches = [random_coo(rng, R, C, 6_000, 0.2, 0.0, "sym") for _ in range(2)]
    # also hit stored coordinates
    br, bc, bv = batches[1]
    br[:1000], bc[:1000] = r[:1000], c[:1000]
    run_chain(R, C, (r, c, v), batches, 42)


@pytest.mark.parametrize("R,C,n", [(1, 1, 5), (1, 300, 5000), (7, 9, 1), (3, 1_000_000, 12_000)])
def test_tiny_shapes(torch, R, C, n):
    """One group (a single row, or a handful of rows), batches from 1 to
    12,000 triples (past one LDS group: the global path on one group)."""
    say(f"tiny R={R} C={C} n={n}")
    rng = np.random.default_rng(R * 7 + n)
    base = random_coo(rng, R, C, min(n, 200), 0.0, 0.0, "sym")
    batches = [random_coo(rng, R, C, n, 0.3, 0.05, "small_int") for _ in range(2)]
    run_chain(R, C, base, batches, 5)


@pytest.mark.parametrize("n", [1 << 21, (1 << 21) + 1])
def test_sort_path_boundary(torch, n):
    """2^21 triples (row groups) and one more (LSD radix sort) give the same
    results as the oracle."""
    R = 50_000
    say(f"boundary n={n}")
    rng = np.random.default_rng(n % 97)
    base = random_coo(rng, R, R, 300_000, 0.0, 0.0, "sym")
    batches = [random_coo(rng, R, R, n, 0.1, 0.01, "sym")]
    run_chain(R, R, base, batches, 6)


def test_bad_index_rejects_grouped_batch(torch):
    """An out-of-range index anywhere in a grouped batch rejects it whole."""
    from dcsr_amd import DcsrError, DcsrMatrix
    R, C = 10_000, 10_000
    rng = np.random.default_rng(51)
    r, c, v = random_coo(rng, R, C, 50_000, 0.0, 0.0, "sym")
    m = DcsrMatrix(R, C)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    br, bc, bv = random_coo(rng, R, C, 40_000, 0.0, 0.0, "sym")
    br[12_345] = R  # one row past the end
    with pytest.raises(DcsrError):
        m.UpdateMatrix(br, bc, bv)
    assert_same_structure(to_host_csr(m), A)
    m.close()


def test_degree_drawn_chain_with_hot_row_and_cancellations(torch):
    """A chain of degree-drawn batches with exact cancellations and one hot
    row per batch (5,000 triples: its group sorts out of global memory),
    synchronous and NOSYNC calls alternating: the oracle after every batch."""
    from dcsr_amd import DcsrMatrix, UPDATE_NOSYNC
    R, C = 30_000, 30_000
    say("degree-drawn chain with a hot row")
    r, c, v = W.powerlaw(R, 300_000, max_degree=5000, seed=61, signed=True)
    rng = np.random.default_rng(62)
    m = DcsrMatrix(R, C, pool_factor=6.0)
    m.LoadMatrix(r, c, v)
    A = O.coo_to_csr(R, C, r, c, v)
    for k in range(4):
        br, bc, bv = W.batch_mixed(rng, r, c, R, C, 80_000, signed=True)
        rows_of = np.repeat(np.arange(R), np.diff(A.off))
        pick = rng.integers(0, A.nnz, 3000)
        hot = np.full(5000, 17 + k, np.int32)
        br = np.concatenate([br, rows_of[pick], hot]).astype(np.int32)
        bc = np.concatenate([bc, A.col[pick], rng.integers(0, C, 5000)]).astype(np.int32)
        bv = np.concatenate([bv, -A.val[pick], rng.uniform(-1, 1, 5000)]).astype(np.float32)
        perm = rng.permutation(br.size)
        br, bc, bv = br[perm], bc[perm], bv[perm]
        m.UpdateMatrix(br, bc, bv, flags=UPDATE_NOSYNC if k % 2 else 0)
        A = O.update(A, br, bc, bv)
        assert_same_structure(to_host_csr(m), A)
        info = check_layout(m)
        assert info["nnz"] == A.nnz
        x = rng.uniform(-1, 1, C).astype(np.float32)
        fp64_spmv_check(A, x, spmv_host(m, x))
    check_healthy(m)
    m.close()
