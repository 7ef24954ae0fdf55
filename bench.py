#!/usr/bin/env python3
"""Benchmarks of the Dynamic-CSR engine on one MI355X.  Prints ONE JSON line.

--workload powerlaw (default; the headline):
  one step = one 1,000,000-triple insert-or-add batch followed by one SpMV
  y = A x, on the same resident matrix:
  * matrix: 325,557 x 325,557, 3,200,000 nnz, Zipf(2.1) out-degrees capped at
    2,700, uniform columns without replacement, FP32 values in (0, 1], seed 1;
  * batches: 50% existing coordinates (uniform over stored entries, i.e. rows
    in proportion to degree), 50% new columns on degree-drawn rows, values in
    (0, 1]; all batches are generated up front and resident in HBM.
  The default run applies 8 batches (the matrix grows to ~7.2M nnz) after W
  warmup steps on the same handle, reloaded before the timed steps.

--workload batched: 4,096 square matrices, rows log-uniform in [64, 4096],
  Poisson(16) nnz per row; one step = every matrix applies its own
  2,000-triple batch and computes its SpMV, in one persistent launch replayed
  from a hipGraph.

--workload spgemm: 2,000,000 x 2,000,000, Poisson(8) nnz per row, N(0,1)
  values; one 4M-triple update, then one step = C = A * A.

--workload spmv: SpMV and SpMM (k = 4, 16, 32) on a resident 10M x 10M
  power-law matrix with 160M nnz (HBM-resident: A alone is 1.3 GB, far
  beyond the 256 MB last-level cache); one step = one SpMV.

--workload serving: resident power-law matrix; a Poisson stream of update
  batches (log-uniform 1k..1M triples, 30% deletions by -current) and
  SpMV/SpMM queries (k in {1, 4, 16, 32}); the server (continuous batching)
  against the plain per-call path on the same stream.

value        = whole-job throughput of the workload's step (unit in "unit")
roofline     = the dominant kernel (or kernel pipeline) of the step:
               algorithmic bytes / measured time (HIP events on the launch
               stream), against 8 TB/s of HBM
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
METRIC = "inserted nnz/s, SpMV GB/s (% HBM peak), p99 latency on mixed update/SpMV stream"
R_CFG = 325_557
NNZ_CFG = 3_200_000
MAXDEG = 2700
BATCH = 1_000_000
WORKLOADS = ("powerlaw", "batched", "spgemm", "serving", "spmv")
DEFAULT_STEPS = {"powerlaw": 8, "batched": 10, "spgemm": 3, "serving": 400, "spmv": 20}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default per workload)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=WORKLOADS, default="powerlaw")
    ap.add_argument("--batch", type=int, default=BATCH)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-hbm-spmv", action="store_true",
                    help="skip the 10M-row / 160M-entry SpMV reported beside the power-law headline")
    ap.add_argument("--cpu-sample-batches", type=int, default=8)
    ap.add_argument("--cpu-min-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (calibrates HBM traffic counters)")
    ap.add_argument("--serving-rows", type=int, default=10_000_000)
    ap.add_argument("--serving-nnz", type=int, default=160_000_000)
    ap.add_argument("--serving-max-cols", type=int, default=64, help="server group width (packed columns)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = DEFAULT_STEPS[args.workload]
    return args


def spmv_bytes(R, C, N):
    return 4 * (R + 1) + 8 * N + 4 * C + 4 * R


def spmm_bytes(R, C, N, k):
    return 4 * (R + 1) + 8 * N + 4 * k * C + 4 * k * R


def update_bytes(b, n_ins, n_add, touched):
    return 12 * b + 8 * n_ins + 8 * n_add + 8 * touched


def roofline(achieved_gbs, kernel, traffic=None):
    return {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel}


def measured_traffic(name, key):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (copy-calibrated) and committed under profiles/; None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith(key + ":"):
                    return int(ln.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def line(args, value, unit, ms_per_step, config, higher=True, **extra):
    out = {
        "metric": METRIC,
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": higher,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


# --------------------------------------------------------------------------
# powerlaw: update + SpMV steps (headline)
# --------------------------------------------------------------------------

def cpu_baseline_powerlaw(r, c, v, batches, x, nsample, min_seconds):
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
    out = {
        "value": trip / (t_upd + t_spmv),
        "unit": "triples/s",
        "cores": 1,
        "kind": "reference" if use_ref else "port",
        "sample": f"{reps} x {nsample} steps (1M-triple batch: sort_by_row + COO->CSR + cusp::add, then CSR SpMV) on the "
                  f"325,557-row power-law matrix, single thread; update {t_upd:.2f}s, SpMV {t_spmv:.2f}s",
        "update_triples_per_s": trip / t_upd,
        "host_cores": host_cores(),
    }
    # the same steps on every host core: the C restatement with its per-row
    # work split over OpenMP threads (bit-identical results)
    th = host_cores()
    tm_upd = tm_spmv = 0.0
    trip_m = 0
    reps_m = 0
    while reps_m == 0 or tm_upd + tm_spmv < min_seconds / 2:
        A = O.coo_to_csr(R_CFG, R_CFG, r, c, v)
        for k in range(nsample):
            br, bc, bv = batches[k]
            t0 = time.perf_counter()
            A = O.update_mt(A, br, bc, bv, th)
            t1 = time.perf_counter()
            O.spmv_mt(A, x, th)
            t2 = time.perf_counter()
            tm_upd += t1 - t0
            tm_spmv += t2 - t1
            trip_m += br.size
        reps_m += 1
    out["multicore"] = {
        "value": trip_m / (tm_upd + tm_spmv), "unit": "triples/s", "cores": th, "kind": "port",
        "sample": f"{reps_m} x {nsample} steps, the same flow on {th} OpenMP threads (rows split per thread, same "
                  f"results bit for bit); update {tm_upd:.2f}s, SpMV {tm_spmv:.2f}s",
        "update_triples_per_s": trip_m / tm_upd,
    }
    return out


def host_cores():
    """Host cores this process may use: the OpenMP thread count the box
    grants (OMP_NUM_THREADS), else the CPU affinity set."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def bench_powerlaw(args):
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W

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
    # pool: 64 x the initial 3.2M entries (1.6 GB of the 288 GB of HBM).
    # Every batch rewrites its touched rows (~85% of the live entries, hub
    # rows included) into fresh chunks, so the old chunks pile up as garbage
    # until an in-order compaction (about 40 us at 13M entries) reclaims
    # them.  At 64x the 20-step run (13.2M live entries at its end) compacts
    # about once; at 16x it compacted every two to three batches past ~10M
    # entries (0.333 vs 0.315 ms per step).
    m = dcsr_amd.DcsrMatrix(R_CFG, R_CFG, pool_factor=64.0)
    m.LoadMatrix(d_r, d_c, d_v)
    for k in range(Wm):
        m.UpdateMatrix(*d_b[K + k], flags=dcsr_amd.UPDATE_NOSYNC)
        m.spmv(d_x, d_y)
        m.UpdateMatrix(*d_b[K + k])
        m.spmv(d_x, d_y)
    torch.cuda.synchronize()

    def reload():
        m.LoadMatrix(d_r, d_c, d_v)
        m.BinRows()
        torch.cuda.synchronize()

    def timed(flags, split, with_info):
        # split: events before the update, between update and SpMV and after
        # the SpMV of every step; otherwise one event per step boundary (each
        # event costs the stream a few microseconds: 3 per step measured
        # 0.333 ms per step where 1 per step measured 0.325)
        nev = 3 * K if split else K + 1
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(nev)]
        stats = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(K):
            ev[3 * k if split else k].record(stream)
            m.UpdateMatrix(*d_b[k], flags=flags)
            if split:
                ev[3 * k + 1].record(stream)
            m.spmv(d_x, d_y)
            if split:
                ev[3 * k + 2].record(stream)
            if with_info:
                stats.append(m.info())
        if not split:
            ev[K].record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ev, stats

    # timed: the asynchronous update (as the reference's UpdateMatrix is a
    # bare launch): no host read-back per step, pool room from host bounds;
    # the batch status is read once after the timed steps.  The throughput
    # and step-latency pass records one event per step boundary; a second
    # pass of the same steps records the update / SpMV split.
    reload()
    compactions0 = m.info()["compactions"]
    elapsed, ev_b, _ = timed(dcsr_amd.UPDATE_NOSYNC, False, False)
    if m.status() != 0:
        raise RuntimeError(f"update status {m.status()} after the timed steps")
    compactions = m.info()["compactions"] - compactions0
    y_async = d_y.clone()
    reload()
    elapsed_split, ev, _ = timed(dcsr_amd.UPDATE_NOSYNC, True, False)
    if m.status() != 0 or not torch.equal(d_y, y_async):
        raise RuntimeError("the split-timed pass differs from the throughput pass")
    # the same steps through the synchronous call (status per call), also
    # timed; its per-call counters give the algorithmic bytes, and its last
    # SpMV must equal the asynchronous run's bit for bit
    reload()
    elapsed_sync, _, stats = timed(0, False, True)
    if not torch.equal(d_y, y_async):
        raise RuntimeError("asynchronous and synchronous runs differ")

    upd_ms = np.array([ev[3 * k].elapsed_time(ev[3 * k + 1]) for k in range(K)])
    spmv_ms = np.array([ev[3 * k + 1].elapsed_time(ev[3 * k + 2]) for k in range(K)])
    step_ms = np.array([ev_b[k].elapsed_time(ev_b[k + 1]) for k in range(K)])
    trip = K * args.batch
    n_ins = sum(s["last_inserted"] for s in stats)
    nnz_after = [s["nnz"] for s in stats]
    spmv_gb = [spmv_bytes(R_CFG, R_CFG, n) / 1e9 for n in nnz_after]
    spmv_gbs = float(np.mean([b / (t / 1e3) for b, t in zip(spmv_gb, spmv_ms)]))
    upd_algo = [update_bytes(args.batch, s["last_inserted"], s["last_added"], s["last_touched_rows"]) for s in stats]
    upd_gbs = float(np.mean([b / 1e9 / (t / 1e3) for b, t in zip(upd_algo, upd_ms)]))

    if upd_ms.sum() >= spmv_ms.sum():
        roof = roofline(upd_gbs, "update pipeline (dcsr_update, all kernels of one call)",
                        measured_traffic("update_powerlaw_325k_traffic.txt", "traffic_bytes_per_update"))
        if roof["traffic"] is not None and args.batch == BATCH:
            roof["traffic_source"] = "profiles/update_powerlaw_325k_traffic.txt"
        else:
            roof["traffic"] = None
    else:
        roof = roofline(spmv_gbs, "flat SpMV (dcsr_spmv)")
    roof["algorithmic_bytes_per_launch"] = float(np.mean(upd_algo))
    roof["avg_launch_ms"] = float(upd_ms.mean())

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_powerlaw(r, c, v, batches, x, min(args.cpu_sample_batches, len(batches)),
                                    args.cpu_min_seconds)
    hbm = None
    if not args.no_hbm_spmv:
        m.close()
        del d_b
        torch.cuda.empty_cache()
        hbm = spmv_hbm_leg(args)
    line(args, trip / elapsed, "triples/s", elapsed * 1e3 / K,
         {"workload": "powerlaw-update-spmv", "rows": R_CFG, "nnz0": NNZ_CFG, "zipf": 2.1, "max_degree": MAXDEG,
          "batch_triples": args.batch, "nnz_final": nnz_after[-1]},
         inserted_nnz_per_s=n_ins / elapsed,
         update_triples_per_s_gpu=trip / (upd_ms.sum() / 1e3),
         spmv_gbs=spmv_gbs,
         spmv_hbm_frac=spmv_gbs / HBM_PEAK_GBS,
         update_algo_gbs=upd_gbs,
         p50_step_ms=float(np.percentile(step_ms, 50)),
         p99_step_ms=float(np.percentile(step_ms, 99)),
         mean_update_ms=float(upd_ms.mean()),
         mean_spmv_ms=float(spmv_ms.mean()),
         update_call="DCSR_UPDATE_NOSYNC, status read once after the timed steps",
         split_pass_ms_per_step=elapsed_split * 1e3 / K,
         compactions_in_timed_steps=compactions,
         sync_call_triples_per_s=trip / elapsed_sync,
         sync_call_ms_per_step=elapsed_sync * 1e3 / K,
         roofline=roof,
         cpu_baseline=cpu,
         spmv_hbm=hbm)


def spmv_hbm_leg(args, reps=10):
    """SpMV at HBM scale beside the headline: the 10M x 10M power-law matrix
    with 160M entries (A is 1.3 GB, x 40 MB: far past the 256 MB last-level
    cache), median of `reps` SpMVs after two warm-up calls (the second builds
    the column-paneled copy), HIP events on the launch stream."""
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W
    R, N = args.serving_rows, args.serving_nnz
    r, c, v = W.serving_matrix(R, N)
    dev = torch.device("cuda:0")
    m = dcsr_amd.DcsrMatrix(R, R, pool_factor=1.0)
    m.LoadMatrix(*(torch.from_numpy(a).to(dev) for a in (r, c, v)))
    nnz = int(r.size)
    del r, c, v
    m.BinRows()
    g = torch.Generator(device="cuda").manual_seed(9)
    x = torch.rand(R, generator=g, device=dev)
    y = torch.empty(R, device=dev)
    for _ in range(2):
        m.spmv(x, y)
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for e0, e1 in ev:
        e0.record(stream)
        m.spmv(x, y)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
    gbs = spmv_bytes(R, R, nnz) / 1e9 / (ms / 1e3)
    m.close()
    return {"rows": R, "nnz": nnz, "ms": ms, "gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
            "traffic_bytes": measured_traffic("spmv_powerlaw_10m_160m_traffic.txt", "traffic_bytes_per_spmv"),
            "traffic_source": "profiles/spmv_powerlaw_10m_160m_traffic.txt"}


# --------------------------------------------------------------------------
# batched: 4,096 small matrices, one graph-replayed persistent launch per step
# --------------------------------------------------------------------------

def bench_batched(args):
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W

    K, Wm = args.steps, args.warmup
    count, per = 4096, 2000
    mats = W.many_small(count=count, min_rows=64, max_rows=4096, lam=16.0, seed=4)
    sizes = np.array([mm[0] for mm in mats], np.int32)
    rng = np.random.default_rng(202)

    def concat(parts):
        off = np.zeros(len(parts) + 1, np.int64)
        off[1:] = np.cumsum([p[0].size for p in parts])
        return off, np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]), \
            np.concatenate([p[2] for p in parts])

    steps = [concat([W.batch_mixed(rng, mm[1], mm[2], mm[0], mm[0], per, signed=True) for mm in mats])
             for _ in range(K + Wm)]
    dev = torch.device("cuda:0")
    b = dcsr_amd.DcsrBatch(sizes, sizes, pool_factor=8.0)
    off0, r0, c0, v0 = concat([(mm[1], mm[2], mm[3]) for mm in mats])
    b.load_coo(off0, r0, c0, v0)
    nnz0 = int(r0.size)
    d_steps = [tuple(torch.from_numpy(a).to(dev) for a in st) for st in steps]
    x = torch.from_numpy(rng.uniform(-1, 1, b.total_cols).astype(np.float32)).to(dev)
    y = torch.empty(b.total_rows, dtype=torch.float32, device=dev)
    team, thresh = b.autotune(x, y)
    for k in range(Wm):
        b.step(*d_steps[K + k], x=x, y=y, graph=True)
    torch.cuda.synchronize()
    code, _ = b.status()
    if code != 0:
        raise RuntimeError(f"batched warmup step failed with status {code}")
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    infos = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        ev[k][0].record(stream)
        b.step(*d_steps[k], x=x, y=y, graph=True)
        ev[k][1].record(stream)
        code, _ = b.status()
        if code != 0:
            raise RuntimeError(f"batched step {k} failed with status {code}")
        infos.append(b.info())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_ms = np.array([e0.elapsed_time(e1) for e0, e1 in ev])
    trip = count * per  # triples per step
    R = b.total_rows
    algo = []
    for inf in infos:
        # SpMV bytes summed over matrices = formula on the block-diagonal whole
        algo.append(update_bytes(trip, inf["last_inserted"], inf["last_added"], inf["last_touched_rows"])
                    + spmv_bytes(R, b.total_cols, inf["nnz"]) + 4 * count)
    gbs = float(np.mean([a / 1e9 / (t / 1e3) for a, t in zip(algo, step_ms)]))
    b.close()
    line(args, K * trip / elapsed, "triples/s", elapsed * 1e3 / K,
         {"workload": "batched-4096-update-spmv", "matrices": count, "rows_total": int(R), "nnz0": nnz0,
          "triples_per_matrix": per, "team": team, "thresh": thresh},
         step_wall_ms=elapsed * 1e3 / K,
         kernel_step_ms=float(step_ms.mean()),
         aggregate_gbs=gbs,
         roofline=roofline(gbs, "k_batch_step (persistent, graph replay)"),
         cpu_baseline=None)


# --------------------------------------------------------------------------
# spgemm: 2M x 2M Poisson(8), one 4M-triple update, then C = A * A
# --------------------------------------------------------------------------

def bench_spgemm(args):
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W

    K, Wm = args.steps, args.warmup
    R = 2_000_000
    r, c, v = W.poisson(R, R, 8.0, seed=3, normal=True)
    rng = np.random.default_rng(303)
    br, bc, bv = W.batch_mixed(rng, r, c, R, R, 4_000_000, signed=True)
    dev = torch.device("cuda:0")
    A = dcsr_amd.DcsrMatrix(R, R, pool_factor=2.0)
    A.LoadMatrix(*(torch.from_numpy(a).to(dev) for a in (r, c, v)))
    A.UpdateMatrix(*(torch.from_numpy(a).to(dev) for a in (br, bc, bv)))
    off, col, val = (t.cpu().numpy() for t in A.to_csr(canonical=True))
    nA = int(col.size)
    rowlen = np.diff(off.astype(np.int64))
    P = int(np.dot(np.bincount(col, minlength=R).astype(np.int64), rowlen))
    C = dcsr_amd.DcsrMatrix(R, R)
    for _ in range(Wm):
        dcsr_amd.spgemm(A, A, C)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    t0 = time.perf_counter()
    for k in range(K):
        ev[k][0].record(stream)
        dcsr_amd.spgemm(A, A, C)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = np.array([e0.elapsed_time(e1) for e0, e1 in ev])
    nC = C.info()["nnz"]
    b_comp = 8 * nA + 8 * nA + 8 * nC + 12 * R
    b_gather = 8 * nA + 8 * P + 8 * nC + 12 * R
    t = float(ms.mean()) / 1e3
    cpu = None
    if not args.no_cpu:
        import oracle as O
        use_ref = O.ref_available()
        Acsr = O.Csr(R, R, off.astype(np.int32), col.astype(np.int32), val.astype(np.float32))
        rows_s = 50_000
        tt, pp, reps = 0.0, 0, 0
        while reps == 0 or tt < args.cpu_min_seconds:
            r0 = (reps * rows_s) % (R - rows_s)
            o = off[r0:r0 + rows_s + 1].astype(np.int64)
            sub = O.Csr(rows_s, R, (o - o[0]).astype(np.int32), col[o[0]:o[-1]], val[o[0]:o[-1]])
            pp += int(rowlen[col[o[0]:o[-1]]].sum())
            t1 = time.perf_counter()
            O.spgemm(sub, Acsr, use_ref=use_ref)
            tt += time.perf_counter() - t1
            reps += 1
        cpu = {"value": 2 * pp / tt / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": "reference" if use_ref else "port",
               "sample": f"{reps} blocks of {rows_s} rows of A times A (host Gustavson SpGEMM, single thread, with its "
                         f"symbolic count pass); {pp} partial products in {tt:.2f}s"}
    line(args, 2 * P / t / 1e9, "GFLOP/s", elapsed * 1e3 / K,
         {"workload": "spgemm-2M-poisson8", "rows": R, "nnz_A": nA, "partials": P, "nnz_C": int(nC),
          "update_triples": 4_000_000},
         spgemm_ms=float(ms.mean()),
         b_comp_gbs=b_comp / t / 1e9,
         b_gather_gbs=b_gather / t / 1e9,
         roofline=roofline(b_comp / t / 1e9, "SpGEMM pipeline (dcsr_spgemm, all kernels of one call)"),
         cpu_baseline=cpu)


# --------------------------------------------------------------------------
# serving: resident power-law matrix, Poisson request stream, server vs
# the plain per-call path
# --------------------------------------------------------------------------

def bench_serving(args):
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W

    R, N = args.serving_rows, args.serving_nnz
    nreq = args.steps
    t_gen = time.perf_counter()
    r, c, v = W.serving_matrix(R, N)
    rng = np.random.default_rng(55)
    reqs = W.serving_stream(rng, r, c, v, R, nreq + args.warmup * 4)
    warm, reqs = reqs[:args.warmup * 4], reqs[args.warmup * 4:]
    dev = torch.device("cuda:0")
    d_r, d_c, d_v = (torch.from_numpy(a).to(dev) for a in (r, c, v))
    nnz0 = int(r.size)
    del r, c, v
    upd = {i: tuple(torch.from_numpy(a).to(dev) for a in q[2]) for i, q in enumerate(reqs) if q[0] == "update"}
    wupd = {i: tuple(torch.from_numpy(a).to(dev) for a in q[2]) for i, q in enumerate(warm) if q[0] == "update"}
    # query inputs are shared per k (read-only); outputs rotate over a ring
    X = {k: torch.rand((R, k), generator=torch.Generator(device="cuda").manual_seed(k), device=dev)
         for k in (1, 4, 16, 32)}
    ring = 4
    Y = {k: [torch.empty((R, k), device=dev) for _ in range(ring)] for k in (1, 4, 16, 32)}
    t_gen = time.perf_counter() - t_gen

    zr = torch.from_numpy(np.random.default_rng(3).integers(0, R, 1_000_000).astype(np.int32)).to(dev)
    zc = torch.from_numpy(np.random.default_rng(4).integers(0, R, 1_000_000).astype(np.int32)).to(dev)
    zv = torch.zeros(1_000_000, dtype=torch.float32, device=dev)

    def fresh():
        """The resident matrix as loaded, with workspaces sized for the
        largest request: a 1M-triple all-zero batch changes nothing (a
        missing coordinate with sum 0.0 is not inserted, an existing one
        keeps old + 0.0) but grows every update buffer, and one packed
        128-column product sizes the query buffers."""
        m = dcsr_amd.DcsrMatrix(R, R, pool_factor=8.0)
        m.LoadMatrix(d_r, d_c, d_v)
        m.UpdateMatrix(zr, zc, zv)
        for k in (1, 4, 16, 32):
            m.spmm(X[k], Y[k][0]) if k > 1 else m.spmv(X[1][:, 0], Y[1][0][:, 0])
        torch.cuda.synchronize()
        return m

    def service_times():
        """Per-request service time of the per-call path, run back to back."""
        m = fresh()
        t = []
        for i, q in enumerate(reqs):
            t0 = time.perf_counter()
            if q[0] == "update":
                m.UpdateMatrix(*upd[i])
            elif q[1] == 1:
                m.spmv(X[1][:, 0], Y[1][i % ring][:, 0])
            else:
                m.spmm(X[q[1]], Y[q[1]][i % ring])
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        m.close()
        return np.array(t)

    def arrivals(rate):
        gaps = np.random.default_rng(77).exponential(1.0 / rate, len(reqs))
        return np.cumsum(gaps)

    def run_per_call(rate):
        m = fresh()
        arr = arrivals(rate)
        lat = np.zeros(len(reqs))
        t_start = time.perf_counter()
        for i, q in enumerate(reqs):
            now = time.perf_counter() - t_start
            if now < arr[i]:
                time.sleep(arr[i] - now)
            if q[0] == "update":
                m.UpdateMatrix(*upd[i])
            elif q[1] == 1:
                m.spmv(X[1][:, 0], Y[1][i % ring][:, 0])
            else:
                m.spmm(X[q[1]], Y[q[1]][i % ring])
            torch.cuda.synchronize()
            lat[i] = time.perf_counter() - t_start - arr[i]
        span = time.perf_counter() - t_start
        m.close()
        return len(reqs) / span, lat * 1e3

    def run_server(rate, warmup=False):
        m = fresh()
        srv = dcsr_amd.DcsrServer(m, max_cols=args.serving_max_cols)
        if warmup:
            for i, q in enumerate(warm):
                if q[0] == "update":
                    srv.wait(srv.update(*wupd[i]))
                else:
                    srv.wait(srv.spmm(X[q[1]], Y[q[1]][0])[0])
        arr = arrivals(rate)
        tickets = []
        t_start = time.perf_counter()
        for i, q in enumerate(reqs):
            now = time.perf_counter() - t_start
            if now < arr[i]:
                time.sleep(arr[i] - now)
            if q[0] == "update":
                tickets.append(srv.update(*upd[i]))
            elif q[1] == 1:
                tickets.append(srv.spmv(X[1][:, 0], Y[1][i % ring][:, 0])[0])
            else:
                tickets.append(srv.spmm(X[q[1]], Y[q[1]][i % ring])[0])
        lat = np.array([srv.wait(t) for t in tickets])
        span = time.perf_counter() - t_start
        st = srv.stats()
        srv.close()
        m.close()
        return len(reqs) / span, lat, st

    is_q = np.array([q[0] == "query" for q in reqs])

    run_server(1e9, warmup=True)  # first-touch workspaces, server threads
    svc = service_times()
    cap_pc = 1.0 / svc.mean()  # per-call capacity (requests/s)
    kinds = ["update" if q[0] == "update" else f"k{q[1]}" for q in reqs]
    service_ms = {k: float(np.mean([t for t, kk in zip(svc, kinds) if kk == k]) * 1e3) for k in sorted(set(kinds))}
    # per-call path at 90% of its capacity: its p99 is the latency bar
    rate_pc = 0.9 * cap_pc
    thr_pc, lat_pc = run_per_call(rate_pc)
    p99_pc = float(np.percentile(lat_pc, 99))
    # server at rising offered rates: the best throughput whose p99 stays
    # within the per-call path's p99 (equal-latency comparison)
    best = None
    sweep = []
    for mult in (1, 1.15, 1.3, 1.45, 1.6, 1.8, 2, 3, 4, 6, 8, 12):
        thr, lat, st = run_server(rate_pc * mult)
        p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
        sweep.append({"offered_x": mult, "throughput": thr, "p50_ms": p50, "p99_ms": p99,
                      "query_p99_ms": float(np.percentile(lat[is_q], 99)),
                      "update_p99_ms": float(np.percentile(lat[~is_q], 99)),
                      "groups": st["query_groups"], "queries": st["queries"], "compactions": st["compactions"]})
        if p99 <= p99_pc and (best is None or thr > best["throughput"]):
            best = sweep[-1]
        if p99 > 4 * p99_pc:
            break
    ratio = best["throughput"] / thr_pc if best else 0.0
    if best is None:
        best = min(sweep, key=lambda e: e["p99_ms"])
    line(args, best["throughput"], "requests/s", 1e3 / best["throughput"],
         {"workload": "serving-powerlaw", "rows": R, "nnz0": nnz0, "requests": len(reqs), "update_share": 0.1,
          "max_cols": args.serving_max_cols,
          "update_triples": "log-uniform 1e3..1e6, 30% deletes", "query_k": [1, 4, 16, 32]},
         p50_ms=best["p50_ms"], p99_ms=best["p99_ms"],
         per_call={"throughput": thr_pc, "p50_ms": float(np.percentile(lat_pc, 50)), "p99_ms": p99_pc,
                   "query_p99_ms": float(np.percentile(lat_pc[is_q], 99)),
                   "update_p99_ms": float(np.percentile(lat_pc[~is_q], 99)), "capacity": cap_pc},
         throughput_vs_per_call_at_equal_p99=ratio,
         per_call_service_ms=service_ms,
         sweep=sweep,
         setup_s=t_gen,
         roofline=None,
         cpu_baseline=None)


def bench_spmv(args):
    import torch
    import dcsr_amd
    from dcsr_amd import workloads as W

    R, N = args.serving_rows, args.serving_nnz
    r, c, v = W.serving_matrix(R, N)
    dev = torch.device("cuda:0")
    m = dcsr_amd.DcsrMatrix(R, R, pool_factor=1.0)
    m.LoadMatrix(*(torch.from_numpy(a).to(dev) for a in (r, c, v)))
    nnz = int(r.size)
    del r, c, v
    m.BinRows()
    stream = torch.cuda.current_stream()
    g = torch.Generator(device="cuda").manual_seed(9)
    res = {}
    for k in (1, 4, 16, 32):
        X = torch.rand((R, k), generator=g, device=dev)
        Y = torch.empty((R, k), device=dev)
        for _ in range(args.warmup):
            m.spmv(X[:, 0], Y[:, 0]) if k == 1 else m.spmm(X, Y)
        reps = args.steps if k == 1 else max(args.steps // 2, 3)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        for e0, e1 in ev:
            e0.record(stream)
            m.spmv(X[:, 0], Y[:, 0]) if k == 1 else m.spmm(X, Y)
            e1.record(stream)
        torch.cuda.synchronize()
        ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in ev]))
        b = spmv_bytes(R, R, nnz) if k == 1 else spmm_bytes(R, R, nnz, k)
        res[k] = {"ms": ms, "gbs": b / 1e9 / (ms / 1e3), "hbm_frac": b / 1e9 / (ms / 1e3) / HBM_PEAK_GBS}
        del X, Y
    gbs = res[1]["gbs"]
    line(args, gbs, "GB/s", res[1]["ms"],
         {"workload": "spmv-powerlaw-10M", "rows": R, "nnz": nnz, "zipf": 2.1},
         spmv_hbm_frac=res[1]["hbm_frac"],
         spmm={str(k): res[k] for k in (4, 16, 32)},
         roofline=roofline(gbs, "SpMV over the paneled copy (dcsr_spmv: flat kernel, carry, panel reduce)",
                           measured_traffic("spmv_powerlaw_10m_160m_traffic.txt", "traffic_bytes_per_spmv")),
         cpu_baseline=None)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); this engine runs on a single MI355X.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        x = torch.rand(64 * 1024 * 1024, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "launches": 3}))
        return
    import dcsr_amd
    dcsr_amd.lib()
    if args.workload == "powerlaw":
        bench_powerlaw(args)
    elif args.workload == "batched":
        bench_batched(args)
    elif args.workload == "spgemm":
        bench_spgemm(args)
    elif args.workload == "spmv":
        bench_spmv(args)
    else:
        bench_serving(args)


if __name__ == "__main__":
    main()
