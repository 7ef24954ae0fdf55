"""Real-matrix suite runner: the reference's FillTests flow over its list of
SuiteSparse matrices, on the Dynamic-CSR engine and its comparators.

Mirrors build/run_tests.py:4-7 (the 16 matrices) driving src/Fill_Test.inl
per matrix:

  1. read the matrix (MatrixMarket, include/dcsr_mtx.hpp: the file's values,
     symmetric storage expanded) and load it as DCSR, HYB and CSR;
  2. compact the DCSR matrix, then 100 SpMVs each for DCSR, HYB and CSR
     (Fill_Test.inl:120-178);
  3. 50 update slices of NNZ/500 uniform random triples, values in (0, 1]
     (Fill_Test.inl:182-210): DCSR insert-or-add with a compaction every 5
     slices, HYB append, CSR rebuilt whole (cusp's generic add), each
     followed by 5 SpMVs (Fill_Test.inl:212-333).

The SuiteSparse files are not shipped; a matrix whose file is not found in
--dir is replaced by a synthetic look-alike of the same shape and entry count
(Zipf(2.1) row degrees, uniform columns; e.g. cnr-2000 = 325,557 rows with
3.2M entries, the engine's power-law benchmark matrix).  With --check the
three formats' results are compared with each other after the slices (DCSR
and the rebuilt CSR bit-exactly, the SpMVs of HYB and CSR against DCSR's
within 2e-5 of (|A| |x|)_i).

  python -m dcsr_amd.suite [--dir PATH] [--only NAME ...] [--scale S] [--check]

prints one JSON line per matrix (seconds per phase, as the reference
prints its timings).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

# (name, rows, cols, entries) of build/run_tests.py:4-7, from the SuiteSparse
# collection's listings; used for the synthetic look-alikes
SUITE = [
    ("kron_g500-logn18", 262_144, 262_144, 21_165_908),
    ("internet", 124_651, 124_651, 207_214),
    ("webbase-1M", 1_000_005, 1_000_005, 3_105_536),
    ("rail4284", 4_284, 1_092_610, 11_279_748),
    ("eu-2005", 862_664, 862_664, 19_235_140),
    ("dblp-2010", 326_186, 326_186, 1_615_400),
    ("enron", 69_244, 69_244, 276_143),
    ("cnr-2000", 325_557, 325_557, 3_216_152),
    ("flickr", 820_878, 820_878, 9_837_214),
    ("in-2004", 1_382_908, 1_382_908, 16_917_053),
    ("amazon-2008", 735_323, 735_323, 5_158_388),
    ("wikipedia-20051105", 1_634_989, 1_634_989, 19_753_078),
    ("ljournal-2008", 5_363_260, 5_363_260, 79_023_142),
    ("hollywood-2009", 1_139_905, 1_139_905, 113_891_327),
    ("soc-LiveJournal1", 4_847_571, 4_847_571, 68_993_773),
    ("indochina-2004", 7_414_866, 7_414_866, 194_109_311),
]

SPMV_REPS = 100
SLICES = 50
SLICE_SPMV = 5
COMPACT_EVERY = 5


def lookalike(R, C, N, seed):
    """Zipf(2.1) row degrees scaled to N (each at most C), uniform columns,
    coordinates made unique on the GPU; values in (0, 1]."""
    import torch
    from dcsr_amd import workloads as W
    rng = np.random.default_rng(seed)
    cap = max(1, min(C, max(N // max(R, 1) * 50, 1000)))
    N = min(N, R * cap)
    d, _ = W.zipf_degrees(rng, R, N, 2.1, cap)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    rows = torch.repeat_interleave(torch.arange(R, device="cuda", dtype=torch.int64), torch.from_numpy(d).to("cuda"))
    cols = torch.randint(0, C, (rows.numel(),), device="cuda", generator=g, dtype=torch.int64)
    key = torch.unique(rows * C + cols, sorted=True)
    del rows, cols
    r = (key // C).to(torch.int32).cpu().numpy()
    c = (key % C).to(torch.int32).cpu().numpy()
    v = (1.0 - rng.random(r.size, dtype=np.float32)).astype(np.float32)
    return r, c, v


def timed(torch, fn, reps=1):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0, out


def run_matrix(name, R, C, r, c, v, check, seed=0, trace=None):
    """One matrix through the flow; returns the timings (seconds).  With a
    `trace` dict, the slices (host arrays) and the final canonical CSR are
    stored in it."""
    import torch
    import dcsr_amd as D

    N = r.size
    res = {"matrix": name, "rows": R, "cols": C, "nnz": int(N)}
    rng = np.random.default_rng(seed)
    x = torch.from_numpy((1.0 - rng.random(C, dtype=np.float32)).astype(np.float32)).cuda()
    rd, cd, vd = (torch.from_numpy(a).cuda() for a in (r, c, v))
    # loads (Fill_Test.inl:95-112)
    m = D.DcsrMatrix(R, C)
    res["dcsr_load_s"], _ = timed(torch, lambda: m.LoadMatrix(rd, cd, vd))
    avg = max(1, N // max(R, 1))
    hyb = D.DcsrHyb(R, C, width=avg, coo_cap=max(1024, N))
    res["hyb_load_s"], _ = timed(torch, lambda: hyb.LoadMatrix(rd, cd, vd))
    # the CSR baseline starts from the matrix as DCSR holds it after its
    # first update: columns sorted, duplicates merged, zeros dropped
    csr = m.to_csr(canonical=True)
    # compaction and SpMV (Fill_Test.inl:120-178)
    res["dcsr_sort_s"], _ = timed(torch, m.SortMatrixRow)
    y = torch.empty(R, dtype=torch.float32, device="cuda")
    res["dcsr_spmv_s"], _ = timed(torch, lambda: m.spmv(x, y), SPMV_REPS)
    res["hyb_spmv_s"], _ = timed(torch, lambda: hyb.spmv(x, y), SPMV_REPS)
    res["csr_spmv_s"], _ = timed(torch, lambda: D.csr_spmv(*csr, x, y), SPMV_REPS)
    # update slices (Fill_Test.inl:182-333)
    slice_n = max(1, N // 500)
    t_add = {"dcsr": 0.0, "hyb": 0.0, "csr": 0.0}
    t_spmv = {"dcsr": 0.0, "hyb": 0.0, "csr": 0.0}
    for i in range(SLICES):
        br = torch.from_numpy(rng.integers(0, R, slice_n).astype(np.int32)).cuda()
        bc = torch.from_numpy(rng.integers(0, C, slice_n).astype(np.int32)).cuda()
        bv = torch.from_numpy((1.0 - rng.random(slice_n, dtype=np.float32)).astype(np.float32)).cuda()
        if trace is not None:
            trace.setdefault("slices", []).append((br.cpu().numpy(), bc.cpu().numpy(), bv.cpu().numpy()))
        dt, _ = timed(torch, lambda: m.UpdateMatrix(br, bc, bv))
        if i > 0 and i % COMPACT_EVERY == 0:
            dc, _ = timed(torch, m.SortMatrixRow)
            dt += dc
        t_add["dcsr"] += dt
        t_spmv["dcsr"] += timed(torch, lambda: m.spmv(x, y), SLICE_SPMV)[0]
        t_add["hyb"] += timed(torch, lambda: hyb.UpdateMatrix(br, bc, bv))[0]
        t_spmv["hyb"] += timed(torch, lambda: hyb.spmv(x, y), SLICE_SPMV)[0]
        dt, csr = timed(torch, lambda: D.csr_add(R, C, *csr, br, bc, bv))
        t_add["csr"] += dt
        t_spmv["csr"] += timed(torch, lambda: D.csr_spmv(*csr, x, y), SLICE_SPMV)[0]
    for k in t_add:
        res[f"{k}_add_s"] = t_add[k]
        res[f"{k}_slice_spmv_s"] = t_spmv[k]
        res[f"{k}_iter_total_s"] = t_add[k] + t_spmv[k]
    res["slices"] = SLICES
    res["slice_triples"] = slice_n
    if check:
        off_d, col_d, val_d = m.to_csr(canonical=True)
        same = (torch.equal(off_d, csr[0]) and torch.equal(col_d, csr[1])
                and torch.equal(val_d.view(torch.int32), csr[2].view(torch.int32)))
        yd = m.spmv(x).to(torch.float64)
        yh = hyb.spmv(x).to(torch.float64)
        yc = D.csr_spmv(*csr, x).to(torch.float64)
        scale = D.csr_spmv(csr[0], csr[1], csr[2].abs(), x.abs()).to(torch.float64)
        tol = 2e-5 * scale + 1e-30
        close = bool(((yd - yh).abs() <= tol).all()) and bool(((yd - yc).abs() <= tol).all())
        res["check"] = "ok" if same and close else "MISMATCH"
    if trace is not None:
        trace["final"] = tuple(t.cpu().numpy() for t in m.to_csr(canonical=True))
    m.close()
    hyb.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", default=None, help="directory holding <name>.mtx files")
    ap.add_argument("--only", nargs="*", default=None, help="matrix names (default: all 16)")
    ap.add_argument("--files", nargs="*", default=[], help="extra .mtx files to run")
    ap.add_argument("--scale", type=float, default=1.0, help="scale look-alike rows and entries")
    ap.add_argument("--check", action="store_true", help="compare the DCSR, rebuilt CSR and HYB results")
    args = ap.parse_args(argv)
    import dcsr_amd as D
    D.lib()
    jobs = []
    for path in args.files:
        jobs.append((os.path.splitext(os.path.basename(path))[0], path, None))
    for k, (name, R, C, N) in enumerate(SUITE):
        if args.only is not None and name not in args.only:
            continue
        path = os.path.join(args.dir, name + ".mtx") if args.dir else None
        jobs.append((name, path if path and os.path.exists(path) else None, (R, C, N, k)))
    for name, path, shape in jobs:
        if path:
            R, C, r, c, v = D.read_mtx(path)
            src = "file"
        else:
            R0, C0, N0, k = shape
            R = max(1, int(R0 * args.scale))
            C = max(1, int(C0 * args.scale)) if R0 == C0 else C0
            if R0 == C0:
                C = R
            r, c, v = lookalike(R, C, max(1, int(N0 * args.scale)), 100 + k)
            src = "synthetic look-alike"
        res = run_matrix(name, R, C, r, c, v, args.check)
        res["source"] = src
        print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
