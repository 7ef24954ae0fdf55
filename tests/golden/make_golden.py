"""Writes tests/golden/*.npz: input/output vectors produced by the reference's
own host path (cusp 0.5.1 sequential backend, compiled from the reference
checkout into oracle/_ref/libdcsr_ref.so by oracle/Makefile).

Run here (where the reference checkout exists):
    make -C oracle && python tests/golden/make_golden.py

Every fixture is data: inputs plus the reference outputs, stored order
included.  Nothing from the reference's sources is stored.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from dcsr_amd import workloads as W  # noqa: E402


def csr_dict(prefix, A):
    return {f"{prefix}_off": A.off, f"{prefix}_col": A.col, f"{prefix}_val": A.val}


def banded_case():
    """Small plain-DCSR case: 256 x 256 banded, 4 insert-or-add batches,
    SpMV, SpGEMM A*A, HYB conversion (reference driver flow,
    src/Fill_Test.inl:75-160, 194-333)."""
    n = 256
    r, c, v = W.banded(n=n, seed=0)
    out = {"R": np.int64(n), "C": np.int64(n), "in_row": r, "in_col": c, "in_val": v}
    A = O.coo_to_csr(n, n, r, c, v, use_ref=True)
    out.update(csr_dict("load", A))
    batches = W.banded_batches(r, c, n=n, nbatches=4, size=1000, seed=0)
    for i, (br, bc, bv) in enumerate(batches):
        out[f"b{i}_row"], out[f"b{i}_col"], out[f"b{i}_val"] = br, bc, bv
        A = O.update(A, br, bc, bv, use_ref=True)
        out.update(csr_dict(f"upd{i}", A))
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    out["x"] = x
    out["y"] = O.spmv(A, x, use_ref=True)
    G = O.spgemm(A, A, use_ref=True)
    out.update(csr_dict("spgemm", G))
    width = max(1, A.nnz // n)
    ec, ev, cr, cc, cv, pitch = O.csr_to_hyb(A, width, use_ref=True)
    out.update({"hyb_width": np.int64(width), "hyb_pitch": np.int64(pitch), "hyb_ell_col": ec, "hyb_ell_val": ev,
                "hyb_coo_row": cr, "hyb_coo_col": cc, "hyb_coo_val": cv})
    return out


def edge_case():
    """Duplicates, explicit zeros, unsorted input, empty rows, cancellation
    to exact zero, deletes by -current, maximum indices."""
    rng = np.random.default_rng(7)
    R, C = 113, 113
    n = 900
    r = rng.integers(0, R, n).astype(np.int32)
    c = rng.integers(0, C // 3, n).astype(np.int32)  # many duplicates
    v = rng.integers(-2, 3, n).astype(np.float32)   # includes zeros, cancellations
    r[:5] = R - 1
    c[:5] = C - 1
    r[r == 13] = 12                                  # row 13 stays empty
    out = {"R": np.int64(R), "C": np.int64(C), "in_row": r, "in_col": c, "in_val": v}
    A = O.coo_to_csr(R, C, r, c, v, use_ref=True)
    out.update(csr_dict("load", A))
    # empty batch: canonicalizes every row (merge duplicates, drop zeros)
    E = Csr0(R, C)
    A1 = O.csr_add(A, E, use_ref=True)
    out.update(csr_dict("canon", A1))
    # batch: deletes by -current for 40 stored entries, sums that cancel,
    # plus fresh inserts with duplicates
    pick = rng.integers(0, A1.nnz, 40)
    rows_of = np.repeat(np.arange(R), np.diff(A1.off))
    dr, dc, dv = rows_of[pick], A1.col[pick], -A1.val[pick]
    fr = rng.integers(0, R, 200)
    fc = rng.integers(0, C, 200)
    fv = rng.integers(-2, 3, 200).astype(np.float32)
    cr_ = np.array([5, 5, 5, 6, 6], np.int32)
    cc_ = np.array([112, 112, 112, 0, 0], np.int32)
    cv_ = np.array([1.5, -1.0, -0.5, 2.0, -2.0], np.float32)
    br = np.concatenate([dr, fr, cr_]).astype(np.int32)
    bc = np.concatenate([dc, fc, cc_]).astype(np.int32)
    bv = np.concatenate([dv, fv, cv_]).astype(np.float32)
    perm = rng.permutation(br.size)
    br, bc, bv = br[perm], bc[perm], bv[perm]
    out.update({"b0_row": br, "b0_col": bc, "b0_val": bv})
    A2 = O.update(A1, br, bc, bv, use_ref=True)
    out.update(csr_dict("upd0", A2))
    # integer-valued SpGEMM with exact cancellations
    G = O.spgemm(A2, A2, use_ref=True)
    out.update(csr_dict("spgemm", G))
    x = rng.integers(-3, 4, C).astype(np.float32)
    out["x"] = x
    out["y"] = O.spmv(A2, x, use_ref=True)
    return out


def Csr0(R, C):
    return O.Csr(R, C, np.zeros(R + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32))


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libdcsr_ref.so is missing: run make -C oracle first")
    np.savez_compressed(os.path.join(HERE, "banded256.npz"), **banded_case())
    np.savez_compressed(os.path.join(HERE, "edge.npz"), **edge_case())
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
