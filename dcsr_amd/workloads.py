"""Deterministic synthetic workloads for the Dynamic-CSR engine.

All generators use numpy's PCG64 (``numpy.random.default_rng(seed)``), so the
same seed yields the same bits on every machine running the same numpy
version.  Matrices are returned as row-major sorted COO triples
``(rows, cols, vals)`` (int32, int32, float32) with unique coordinates; update
batches are returned in random (unsorted) order, as the API accepts them.

Workloads (sizes as in the reference's driver configurations):

* ``banded``      -- 256 x 256 symmetric banded pattern, half-bandwidth 40,
  ~3,000 nnz, the shape of the reference's bundled ``matrices/can_256.mtx``
  (src/Fill_Test.inl:75-99 loads it, then updates and multiplies).
* ``powerlaw``    -- Zipf(2.1) out-degrees scaled to a target nnz, uniform
  columns without replacement (web-graph shape of cnr-2000,
  build/run_tests.py:4-7); 325,557 x 325,557 with 3.2M nnz by default.
* ``poisson``     -- Poisson(lam) nnz per row, uniform columns (SpGEMM input).
* ``batch_mixed`` -- insert-or-add batches: a fraction of the triples hits
  existing coordinates (sampled uniformly among stored entries, hence rows in
  proportion to degree), the rest are new columns on degree-drawn rows.
* ``many_small``  -- many independent small matrices (batched launch).
* ``serving_matrix`` / ``serving_stream`` -- the resident 10M-row power-law
  matrix of the serving loop (coordinates made unique on the GPU) and its
  Poisson request mix of update batches and SpMV / SpMM queries.

Values are drawn from (0, 1] or (-1, 1] so that no generated value is an exact
zero (an exact zero is a structural event in the insert-or-add semantics).
"""
from __future__ import annotations

import numpy as np


def _u01(rng, n):
    # (0, 1]: 1 - U[0,1) in float32 never yields 0
    return (np.float32(1.0) - rng.random(n, dtype=np.float32)).astype(np.float32)


def _usym(rng, n):
    # (-1, 1] excluding 0
    v = (rng.random(n, dtype=np.float32) * np.float32(2.0) - np.float32(1.0)).astype(np.float32)
    v[v == 0] = np.float32(0.5)
    return v


def _unique_cols(rng, rows, ncols, degrees):
    """Uniform columns without replacement inside each row (resample clashes)."""
    n = rows.size
    cols = rng.integers(0, ncols, n, dtype=np.int64)
    for _ in range(64):
        key = rows.astype(np.int64) * ncols + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(n, bool)
        dup[order[1:]] = ks[1:] == ks[:-1]
        nd = int(dup.sum())
        if nd == 0:
            break
        cols[dup] = rng.integers(0, ncols, nd, dtype=np.int64)
    else:  # pragma: no cover - only for rows with degree close to ncols
        raise RuntimeError("could not draw unique columns")
    return cols


def _sorted_coo(rows, cols, vals, ncols):
    key = rows.astype(np.int64) * ncols + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    return (rows[order].astype(np.int32), cols[order].astype(np.int32), vals[order].astype(np.float32))


def zipf_degrees(rng, nrows, target_nnz, exponent=2.1, cap=None):
    """Zipf(exponent) degrees d_i = clip(ceil(s * z_i), 1, cap) with the scale
    s found by bisection, then +-1 corrections on random rows so that
    sum(d) == target_nnz exactly.  Returns (degrees int64, s)."""
    cap = nrows if cap is None else cap
    z = rng.zipf(exponent, nrows).astype(np.float64)
    lo, hi = 1e-3, 1e3
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if np.clip(np.ceil(mid * z), 1, cap).sum() < target_nnz:
            lo = mid
        else:
            hi = mid
    d = np.clip(np.ceil(lo * z), 1, cap).astype(np.int64)
    diff = int(target_nnz - d.sum())
    while diff != 0:
        k = min(abs(diff), nrows)
        pick = rng.choice(nrows, k, replace=False)
        if diff > 0:
            pick = pick[d[pick] < cap]
            d[pick] += 1
            diff -= pick.size
        else:
            pick = pick[d[pick] > 1]
            d[pick] -= 1
            diff += pick.size
    return d, lo


def banded(n=256, half_bw=40, p=0.1457, seed=0):
    """Full diagonal plus upper pairs |i-j| <= half_bw kept with probability p,
    mirrored (pattern symmetric), values U(-1,1] drawn independently."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n, 1)
    band = (ju - iu) <= half_bw
    iu, ju = iu[band], ju[band]
    keep = rng.random(iu.size) < p
    iu, ju = iu[keep], ju[keep]
    diag = np.arange(n)
    rows = np.concatenate([diag, iu, ju])
    cols = np.concatenate([diag, ju, iu])
    vals = _usym(rng, rows.size)
    return _sorted_coo(rows, cols, vals, n)


def powerlaw(nrows=325_557, nnz=3_200_000, exponent=2.1, max_degree=2700, seed=1, signed=False):
    rng = np.random.default_rng(seed)
    d, _ = zipf_degrees(rng, nrows, nnz, exponent, max_degree)
    rows = np.repeat(np.arange(nrows, dtype=np.int64), d)
    cols = _unique_cols(rng, rows, nrows, d)
    vals = _usym(rng, rows.size) if signed else _u01(rng, rows.size)
    return _sorted_coo(rows, cols, vals, nrows)


def poisson(nrows, ncols, lam=8.0, seed=3, normal=True):
    rng = np.random.default_rng(seed)
    d = np.minimum(rng.poisson(lam, nrows), ncols).astype(np.int64)
    rows = np.repeat(np.arange(nrows, dtype=np.int64), d)
    cols = _unique_cols(rng, rows, ncols, d)
    vals = rng.standard_normal(rows.size, dtype=np.float32) if normal else _usym(rng, rows.size)
    vals[vals == 0] = np.float32(1.0)
    return _sorted_coo(rows, cols, vals, ncols)


def batch_mixed(rng, rows, cols, nrows, ncols, size, frac_existing=0.5, signed=False,
                delete_frac=0.0, current_vals=None):
    """One insert-or-add batch.  ``frac_existing`` of the triples reuse stored
    coordinates (uniform over stored entries => rows in proportion to degree),
    the rest put a uniform column on a degree-drawn row.  With
    ``delete_frac`` > 0, that share of the existing-coordinate triples carry
    value = -current (deletion by zeroing; needs ``current_vals``)."""
    n_ex = int(round(size * frac_existing))
    n_new = size - n_ex
    nnz = rows.size
    if nnz == 0:
        n_new, n_ex = size, 0
    pick = rng.integers(0, max(nnz, 1), n_ex) if n_ex else np.zeros(0, np.int64)
    r_ex = rows[pick].astype(np.int64)
    c_ex = cols[pick].astype(np.int64)
    if nnz:
        r_new = rows[rng.integers(0, nnz, n_new)].astype(np.int64)
    else:
        r_new = rng.integers(0, nrows, n_new)
    c_new = rng.integers(0, ncols, n_new)
    v = _usym(rng, size) if signed else _u01(rng, size)
    if delete_frac > 0 and n_ex and current_vals is not None:
        ndel = int(round(n_ex * delete_frac))
        v[:ndel] = -current_vals[pick[:ndel]]
    r = np.concatenate([r_ex, r_new])
    c = np.concatenate([c_ex, c_new])
    perm = rng.permutation(size)
    return r[perm].astype(np.int32), c[perm].astype(np.int32), v[perm].astype(np.float32)


def banded_batches(rows, cols, n=256, nbatches=4, size=1000, seed=0):
    """Small-case batches: half existing coordinates, half uniform new ones
    outside the pattern, values U(-1,1]."""
    rng = np.random.default_rng(seed + 1000)
    have = set((rows.astype(np.int64) * n + cols).tolist())
    out = []
    for _ in range(nbatches):
        half = size // 2
        pick = rng.integers(0, rows.size, half)
        r1, c1 = rows[pick].astype(np.int64), cols[pick].astype(np.int64)
        r2 = np.empty(size - half, np.int64)
        c2 = np.empty(size - half, np.int64)
        k = 0
        while k < r2.size:
            rr, cc = int(rng.integers(0, n)), int(rng.integers(0, n))
            if rr * n + cc not in have:
                r2[k], c2[k] = rr, cc
                k += 1
        r = np.concatenate([r1, r2])
        c = np.concatenate([c1, c2])
        v = _usym(rng, size)
        perm = rng.permutation(size)
        out.append((r[perm].astype(np.int32), c[perm].astype(np.int32), v[perm]))
    return out


def many_small(count=4096, min_rows=64, max_rows=4096, lam=16.0, seed=4):
    """Independent square matrices with rows log-uniform in [min_rows,
    max_rows] and Poisson(lam) nnz per row.  Returns a list of
    (nrows, rows, cols, vals)."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(min_rows), np.log(max_rows), count))).astype(np.int64)
    mats = []
    for n in sizes.tolist():
        d = np.minimum(rng.poisson(lam, n), n).astype(np.int64)
        r = np.repeat(np.arange(n, dtype=np.int64), d)
        c = _unique_cols(rng, r, n, d)
        v = _usym(rng, r.size)
        rr, cc, vv = _sorted_coo(r, c, v, n)
        mats.append((n, rr, cc, vv))
    return mats


def serving_matrix(R, N, seed=5):
    """Zipf(2.1) out-degrees scaled to N (capped at R/10), uniform columns;
    coordinates made unique on the GPU.  Returns host (rows, cols, vals)."""
    import torch
    rng = np.random.default_rng(seed)
    d, _ = zipf_degrees(rng, R, N, 2.1, max(R // 10, 1))
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    rows = torch.repeat_interleave(torch.arange(R, device="cuda", dtype=torch.int64),
                                   torch.from_numpy(d).to("cuda"))
    cols = torch.randint(0, R, (rows.numel(),), device="cuda", generator=g, dtype=torch.int64)
    key = torch.unique(rows * R + cols, sorted=True)
    del rows, cols
    r = (key // R).to(torch.int32).cpu().numpy()
    c = (key % R).to(torch.int32).cpu().numpy()
    del key
    v = (1.0 - rng.random(r.size, dtype=np.float32)).astype(np.float32)
    return r, c, v


def serving_stream(rng, r, c, v, R, nreq, upd_frac=0.1):
    """Request list: (kind, k or n, arrays).  Updates: log-uniform 1e3..1e6
    triples, 30% deletions (-current of a coordinate no earlier request
    touched), 35% adds to stored coordinates, 35% new coordinates on
    degree-drawn rows.  Queries: k = 1, 4, 16, 32 with weights .55/.2/.15/.1."""
    dirty = np.zeros(r.size, bool)
    reqs = []
    ks = np.array([1, 4, 16, 32])
    for _ in range(nreq):
        if rng.random() < upd_frac:
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
            nd = int(0.3 * n)
            na = int(0.35 * n)
            nn = n - nd - na
            pick = rng.integers(0, r.size, 2 * nd + 16)
            pick = np.unique(pick[~dirty[pick]])[:nd]
            nd = pick.size
            add = rng.integers(0, r.size, na)
            dirty[pick] = True
            dirty[add] = True
            newrow = r[rng.integers(0, r.size, nn)]
            rows = np.concatenate([r[pick], r[add], newrow]).astype(np.int32)
            cols = np.concatenate([c[pick], c[add], rng.integers(0, R, nn)]).astype(np.int32)
            vals = np.concatenate([-v[pick], (1.0 - rng.random(na + nn, dtype=np.float32))]).astype(np.float32)
            perm = rng.permutation(rows.size)
            reqs.append(("update", rows.size, (rows[perm], cols[perm], vals[perm])))
        else:
            reqs.append(("query", int(rng.choice(ks, p=[0.55, 0.2, 0.15, 0.1])), None))
    return reqs
