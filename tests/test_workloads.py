"""Synthetic workload generators: determinism and the stated shapes."""
import numpy as np

from dcsr_amd import workloads as W


def _unique(r, c, ncols):
    k = r.astype(np.int64) * ncols + c
    return np.unique(k).size == k.size


def test_banded_shape_and_determinism():
    r, c, v = W.banded(256, 40, seed=0)
    r2, c2, v2 = W.banded(256, 40, seed=0)
    assert np.array_equal(r, r2) and np.array_equal(c, c2) and np.array_equal(v, v2)
    assert 2500 < r.size < 3500
    assert (np.abs(r.astype(int) - c) <= 40).all()
    assert _unique(r, c, 256)
    s = set(zip(r.tolist(), c.tolist()))
    assert all((b, a) in s for a, b in s)  # symmetric pattern
    assert not (v == 0).any()


def test_zipf_degrees_hit_target_exactly():
    rng = np.random.default_rng(1)
    d, s = W.zipf_degrees(rng, 20000, 200000, 2.1, 2700)
    assert d.sum() == 200000 and d.min() >= 1 and d.max() <= 2700 and s > 0


def test_powerlaw_small():
    r, c, v = W.powerlaw(5000, 50000, max_degree=400, seed=1)
    assert r.size == 50000 and _unique(r, c, 5000)
    assert (np.diff(r.astype(np.int64) * 5000 + c) > 0).all()  # sorted
    assert (v > 0).all() and (v <= 1).all()


def test_batch_mixed_existing_fraction():
    r, c, v = W.powerlaw(5000, 50000, max_degree=400, seed=1)
    rng = np.random.default_rng(2)
    br, bc, bv = W.batch_mixed(rng, r, c, 5000, 5000, 10000)
    have = set((r.astype(np.int64) * 5000 + c).tolist())
    hits = np.mean([k in have for k in (br.astype(np.int64) * 5000 + bc).tolist()])
    assert 0.48 < hits < 0.56


def test_many_small():
    mats = W.many_small(count=16, seed=4)
    assert len(mats) == 16
    for n, r, c, v in mats:
        assert 64 <= n <= 4096 and r.size == c.size == v.size and _unique(r, c, n)
