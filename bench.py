#!/usr/bin/env python3
"""Benchmark: insert-or-add update batches + SpMV on a power-law DCSR matrix.

Workload (one step = one 1,000,000-triple insert-or-add batch followed by one
SpMV y = A x, on the same resident matrix):
  * matrix: 325,557 x 325,557, 3,200,000 nnz, Zipf(2.1) out-degrees capped at
    2,700, uniform columns without replacement, FP32 values in (0, 1], seed 1;
  * batches: 50% existing coordinates (uniform over stored entries, i.e. rows
    in proportion to degree), 50% new columns on degree-drawn rows, values in
    (0, 1]; all batches are generated up front and resident in HBM.

The default run applies 8 batches (the matrix grows to ~7.2M nnz) after W
warmup steps on a separate copy.  Prints ONE JSON line.

value      = triples applied per second over whole steps (update + SpMV)
roofline   = the dominant kernel of the step, algorithmic bytes / measured
             per-launch time (HIP events on the launch stream)
cpu_baseline = the reference host path (cusp sequential, oracle/_ref) or its
             restatement (oracle/) timed on a bounded sample on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
R_CFG = 325_557
NNZ_CFG = 3_200_000
MAXDEG = 2700
BATCH = 1_000_000


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample-batches", type=int, default=8)
    ap.add_argument("--cpu-min-seconds", type=float, default=10.0)
    return ap.parse_args()


def spmv_bytes(R, C, N):
    return 4 * (R + 1) + 8 * N + 4 * C + 4 * R


def update_bytes(b, n_ins, n_add, touched):
    return 12 * b + 8 * n_ins + 8 * n_add + 8 * touched


def cpu_baseline(r, c, v, batches, x, nsample, min_seconds):
    """Reference host path (or its restatement) on the first nsample batches,
    the sequence repeated from the initial matrix until min_seconds of work."""
    import oracle as O
    use_ref = O.ref_available()
    t_upd = 0.0
    t_spmv = 0.0
    trip = 0
    reps = 0
    while reps == 0 or t_upd + t_spmv < min_seconds:
        A = O.coo_to_csr(R_CFG, R_CFG, r, c, v, use_ref=use_ref)
        for k in range(nsample):
            br, bc, bv = batches[k]
            t0 = time.perf_counter()
            A = O.update(A, br, bc, bv, use_ref=use_ref)
            t1 = time.perf_counter()
            O.spmv(A, x, use_ref=use_ref)
            t2 = time.perf_counter()
            t_upd += t1 - t0
            t_spmv += t2 - t1
            trip += br.size
        reps += 1
    return {
        "value": trip / (t_upd + t_spmv),
        "unit": "triples/s",
        "cores": 1,
        "kind": "reference" if use_ref else "port",
        "sample": f"{reps} x {nsample} steps (1M-triple batch: sort_by_row + COO->CSR + cusp::add, then CSR SpMV) on the "
                  f"325,557-row power-law matrix, single thread; update {t_upd:.2f}s, SpMV {t_spmv:.2f}s",
        "update_triples_per_s": trip / t_upd,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this engine runs on a single MI355X.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W

    dcsr_amd.lib()
    K, Wm = args.steps, args.warmup
    r, c, v = W.powerlaw(R_CFG, NNZ_CFG, 2.1, MAXDEG, seed=1)
    rng = np.random.default_rng(101)
    batches = [W.batch_mixed(rng, r, c, R_CFG, R_CFG, args.batch) for _ in range(K + Wm)]
    x = (1.0 - np.random.default_rng(7).random(R_CFG, dtype=np.float32)).astype(np.float32)

    dev = torch.device("cuda:0")
    d_r, d_c, d_v = (torch.from_numpy(a).to(dev) for a in (r, c, v))
    d_b = [tuple(torch.from_numpy(a).to(dev) for a in b) for b in batches]
    d_x = torch.from_numpy(x).to(dev)
    d_y = torch.empty(R_CFG, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    # warmup on the same handle (first-touch workspace allocation happens
    # here), then reload the initial matrix: the timed steps start from it
    m = dcsr_amd.DcsrMatrix(R_CFG, R_CFG, pool_factor=8.0)
    m.LoadMatrix(d_r, d_c, d_v)
    for k in range(Wm):
        m.UpdateMatrix(*d_b[K + k])
        m.spmv(d_x, d_y)
    torch.cuda.synchronize()
    m.LoadMatrix(d_r, d_c, d_v)
    m.BinRows()
    torch.cuda.synchronize()

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        e0, e1, e2 = ev[k]
        e0.record(stream)
        m.UpdateMatrix(*d_b[k])
        e1.record(stream)
        m.spmv(d_x, d_y)
        e2.record(stream)
        inf = m.info()
        stats.append(inf)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0

    upd_ms = np.array([e0.elapsed_time(e1) for e0, e1, _ in ev])
    spmv_ms = np.array([e1.elapsed_time(e2) for _, e1, e2 in ev])
    step_ms = np.array([e0.elapsed_time(e2) for e0, _, e2 in ev])
    trip = K * args.batch
    n_ins = sum(s["last_inserted"] for s in stats)
    n_add = sum(s["last_added"] for s in stats)
    touched = sum(s["last_touched_rows"] for s in stats)
    nnz_after = [s["nnz"] for s in stats]
    spmv_gb = [spmv_bytes(R_CFG, R_CFG, n) / 1e9 for n in nnz_after]
    spmv_gbs = float(np.mean([b / (t / 1e3) for b, t in zip(spmv_gb, spmv_ms)]))
    upd_algo = [update_bytes(args.batch, s["last_inserted"], s["last_added"], s["last_touched_rows"]) for s in stats]
    upd_gbs = float(np.mean([b / 1e9 / (t / 1e3) for b, t in zip(upd_algo, upd_ms)]))

    # dominant piece of the step
    if upd_ms.sum() >= spmv_ms.sum():
        dom = {"kernel": "update pipeline (dcsr_update)", "bound": "hbm", "achieved": upd_gbs}
        traffic = None
    else:
        dom = {"kernel": "k_spmv_rows (dcsr_spmv)", "bound": "hbm", "achieved": spmv_gbs}
        traffic = None
    roof = {"bound": dom["bound"], "achieved": round(dom["achieved"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(dom["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom["kernel"]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(r, c, v, batches, x, min(args.cpu_sample_batches, len(batches)), args.cpu_min_seconds)

    out = {
        "metric": "inserted nnz/s, SpMV GB/s (% HBM peak), p99 latency on mixed update/SpMV stream",
        "value": trip / elapsed,
        "unit": "triples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "powerlaw-update-spmv", "rows": R_CFG, "nnz0": NNZ_CFG, "zipf": 2.1,
                   "max_degree": MAXDEG, "batch_triples": args.batch, "nnz_final": nnz_after[-1]},
        "inserted_nnz_per_s": n_ins / elapsed,
        "update_triples_per_s_gpu": trip / (upd_ms.sum() / 1e3),
        "spmv_gbs": spmv_gbs,
        "spmv_hbm_frac": spmv_gbs / HBM_PEAK_GBS,
        "update_algo_gbs": upd_gbs,
        "p50_step_ms": float(np.percentile(step_ms, 50)),
        "p99_step_ms": float(np.percentile(step_ms, 99)),
        "mean_update_ms": float(upd_ms.mean()),
        "mean_spmv_ms": float(spmv_ms.mean()),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
