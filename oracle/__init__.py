"""CPU oracle for the Dynamic-CSR path (test infrastructure only).

Wraps two shared libraries through ctypes:

* ``oracle/liboracle.so``  -- ``dcsr_oracle.c``, the restatement of the
  reference's host path (cusp 0.5.1 sequential backend); always built.
* ``oracle/_ref/libdcsr_ref.so`` -- ``ref_harness.cpp`` compiled against the
  reference's own cusp/thrust headers; present only where the reference
  checkout was available at build time (it travels with the tree).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this package, and only as the checker or the timed CPU baseline.
The product library never calls it.

Matrices cross this boundary as host CSR triples ``(off, col, val)`` with
int32 indices and float32 values, exactly the reference's ``<int, float>``
instantiation (src/Fill_Test.inl:11-16, PRECISION 32).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_I64 = ctypes.c_int64
_P = ctypes.c_void_p

_orc = None
_ref = None


def _load(path):
    return ctypes.CDLL(path)


def lib():
    """The restatement library (builds it on first use if missing)."""
    global _orc
    if _orc is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-C", _HERE, "liboracle.so"])
        _orc = _load(path)
        _orc.orc_sort_by_row.restype = ctypes.c_int
        _orc.orc_sort_by_row.argtypes = [_I64, _P, _P, _P]
        _orc.orc_indices_to_offsets.restype = None
        _orc.orc_indices_to_offsets.argtypes = [_I64, _P, _I64, _P]
        _orc.orc_csr_add.restype = _I64
        _orc.orc_csr_add.argtypes = [_I64, _I64] + [_P] * 9
        _orc.orc_csr_spmv.restype = None
        _orc.orc_csr_spmv.argtypes = [_I64, _P, _P, _P, _P, _P]
        _orc.orc_spmm_dense.restype = None
        _orc.orc_spmm_dense.argtypes = [_I64, _P, _P, _P, _I64, _P, _I64, _P, _I64]
        _orc.orc_spgemm_count.restype = _I64
        _orc.orc_spgemm_count.argtypes = [_I64, _I64, _P, _P, _P, _P]
        _orc.orc_spgemm.restype = _I64
        _orc.orc_spgemm.argtypes = [_I64, _I64] + [_P] * 9
        _orc.orc_csr_to_hyb.restype = _I64
        _orc.orc_csr_to_hyb.argtypes = [_I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P, _P]
        _orc.orc_coo_to_csr.restype = ctypes.c_int
        _orc.orc_coo_to_csr.argtypes = [_I64, _I64, _P, _P, _P, _P]
        _orc.orc_update.restype = _I64
        _orc.orc_update.argtypes = [_I64, _I64, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P]
        _orc.orc_csr_spmv_mt.restype = None
        _orc.orc_csr_spmv_mt.argtypes = [_I64, _P, _P, _P, _P, _P, ctypes.c_int]
        _orc.orc_update_mt.restype = _I64
        _orc.orc_update_mt.argtypes = [_I64, _I64, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, ctypes.c_int]
        _orc.orc_sort_rows_by_col.restype = ctypes.c_int
        _orc.orc_sort_rows_by_col.argtypes = [_I64, _P, _P, _P]
        _orc.orc_hyb_spmv.restype = None
        _orc.orc_hyb_spmv.argtypes = [_I64, _I64, _I64, _P, _P, _I64, _P, _P, _P, _P, _P]
        _orc.orc_hyb_update.restype = _I64
        _orc.orc_hyb_update.argtypes = [_I64, _I64, _I64, _P, _P, _P, _I64, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P]
        _orc.orc_spgemm_check_f64.restype = _I64
        _orc.orc_spgemm_check_f64.argtypes = [_I64, _I64] + [_P] * 9 + [ctypes.c_double, ctypes.POINTER(_I64)]
    return _orc


def ref_path():
    return os.path.join(_HERE, "_ref", "libdcsr_ref.so")


def ref_available() -> bool:
    return os.path.exists(ref_path())


def ref():
    """The cusp-backed reference harness, or None when it was not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        _ref = _load(ref_path())
        _ref.ref_sort_by_row.restype = ctypes.c_int
        _ref.ref_sort_by_row.argtypes = [_I64, _P, _P, _P]
        _ref.ref_coo_to_csr.restype = ctypes.c_int
        _ref.ref_coo_to_csr.argtypes = [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P]
        _ref.ref_csr_add.restype = _I64
        _ref.ref_csr_add.argtypes = [_I64, _I64] + [_P] * 9
        _ref.ref_csr_spmv.restype = None
        _ref.ref_csr_spmv.argtypes = [_I64, _I64, _P, _P, _P, _P, _P]
        _ref.ref_spgemm.restype = _I64
        _ref.ref_spgemm.argtypes = [_I64, _I64, _I64] + [_P] * 9
        _ref.ref_hyb_spmv.restype = None
        _ref.ref_hyb_spmv.argtypes = [_I64, _I64, _I64, _P, _P, _I64, _P, _P, _P, _P, _P]
        _ref.ref_csr_to_hyb.restype = _I64
        _ref.ref_csr_to_hyb.argtypes = [_I64, _I64, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P]
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Csr:
    """Host CSR triple (int32 offsets/columns, float32 values)."""

    __slots__ = ("rows", "cols", "off", "col", "val")

    def __init__(self, rows, cols, off, col, val):
        self.rows = int(rows)
        self.cols = int(cols)
        self.off = _i32(off)
        self.col = _i32(col)
        self.val = _f32(val)

    @property
    def nnz(self):
        return int(self.off[-1]) if self.off.size else 0

    def row(self, r):
        s, e = self.off[r], self.off[r + 1]
        return self.col[s:e], self.val[s:e]

    def canonical(self):
        """Rows stably sorted by column (the comparison form; the stored order
        inside a row is not part of the structure).  Sorted per row in C
        (orc_sort_rows_by_col); canonical_np is the numpy statement of the
        same order, which pins it."""
        col = self.col.copy()
        val = self.val.copy()
        if col.size and lib().orc_sort_rows_by_col(self.rows, _p(self.off), _p(col), _p(val)) != 0:
            raise MemoryError("orc_sort_rows_by_col failed")
        return Csr(self.rows, self.cols, self.off, col, val)

    def canonical_np(self):
        off = self.off
        rowid = np.repeat(np.arange(self.rows, dtype=np.int64), np.diff(off))
        order = np.lexsort((np.arange(self.col.size), self.col.astype(np.int64), rowid))
        return Csr(self.rows, self.cols, off, self.col[order], self.val[order])

    def to_scipy(self, dtype=np.float64):
        import scipy.sparse as sp
        # copies: scipy may sort/merge indices in place
        return sp.csr_matrix((self.val.astype(dtype), self.col.copy(), self.off.copy()),
                             shape=(self.rows, self.cols))


# ----------------------------------------------------------------------------
# restatement entry points
# ----------------------------------------------------------------------------

def sort_by_row(rows, cols, vals, use_ref=False):
    r, c, v = _i32(rows).copy(), _i32(cols).copy(), _f32(vals).copy()
    L = ref() if use_ref else lib()
    fn = L.ref_sort_by_row if use_ref else L.orc_sort_by_row
    rc = fn(r.size, _p(r), _p(c), _p(v))
    if rc != 0:
        raise ValueError("sort_by_row failed (%d)" % rc)
    return r, c, v


def coo_to_csr(R, C, rows, cols, vals, use_ref=False):
    """cusp::sort_by_row + COO->CSR (duplicates and zeros kept)."""
    r, c, v = _i32(rows).copy(), _i32(cols).copy(), _f32(vals).copy()
    off = np.zeros(R + 1, np.int32)
    if use_ref:
        r, c, v = sort_by_row(r, c, v, use_ref=True)
        oc = np.zeros(max(r.size, 1), np.int32)
        ov = np.zeros(max(r.size, 1), np.float32)
        ref().ref_coo_to_csr(R, C, r.size, _p(r), _p(c), _p(v), _p(off), _p(oc), _p(ov))
        return Csr(R, C, off, oc[: r.size], ov[: r.size])
    rc = lib().orc_coo_to_csr(R, r.size, _p(r), _p(c), _p(v), _p(off))
    if rc != 0:
        raise ValueError("coo_to_csr failed (%d)" % rc)
    return Csr(R, C, off, c, v)


def csr_add(A: Csr, B: Csr, use_ref=False) -> Csr:
    cap = max(A.nnz + B.nnz, 1)
    off = np.zeros(A.rows + 1, np.int32)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float32)
    fn = ref().ref_csr_add if use_ref else lib().orc_csr_add
    nnz = fn(A.rows, A.cols, _p(A.off), _p(A.col), _p(A.val), _p(B.off), _p(B.col), _p(B.val),
             _p(off), _p(col), _p(val))
    if nnz < 0:
        raise MemoryError("csr_add failed")
    return Csr(A.rows, A.cols, off, col[:nnz], val[:nnz])


def update(A: Csr, rows, cols, vals, use_ref=False) -> Csr:
    """One insert-or-add batch on the CPU path: sort_by_row, COO->CSR, add."""
    if use_ref:
        B = coo_to_csr(A.rows, A.cols, rows, cols, vals, use_ref=True) if len(rows) else \
            Csr(A.rows, A.cols, np.zeros(A.rows + 1, np.int32), [], [])
        return csr_add(A, B, use_ref=True)
    r, c, v = _i32(rows), _i32(cols), _f32(vals)
    cap = max(A.nnz + r.size, 1)
    off = np.zeros(A.rows + 1, np.int32)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float32)
    nnz = lib().orc_update(A.rows, A.cols, _p(A.off), _p(A.col), _p(A.val), r.size, _p(r), _p(c), _p(v),
                           _p(off), _p(col), _p(val))
    if nnz < 0:
        raise ValueError("update failed (%d)" % nnz)
    return Csr(A.rows, A.cols, off, col[:nnz], val[:nnz])


def update_mt(A: Csr, rows, cols, vals, threads) -> Csr:
    """update() on `threads` OpenMP threads (orc_update_mt): the same result,
    bit for bit; the multi-core CPU baseline."""
    r, c, v = _i32(rows), _i32(cols), _f32(vals)
    cap = max(A.nnz + r.size, 1)
    off = np.zeros(A.rows + 1, np.int32)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float32)
    nnz = lib().orc_update_mt(A.rows, A.cols, _p(A.off), _p(A.col), _p(A.val), r.size, _p(r), _p(c), _p(v),
                              _p(off), _p(col), _p(val), int(threads))
    if nnz < 0:
        raise ValueError("update_mt failed (%d)" % nnz)
    return Csr(A.rows, A.cols, off, col[:nnz], val[:nnz])


def spmv(A: Csr, x, use_ref=False):
    x = _f32(x)
    y = np.zeros(A.rows, np.float32)
    if use_ref:
        ref().ref_csr_spmv(A.rows, A.cols, _p(A.off), _p(A.col), _p(A.val), _p(x), _p(y))
    else:
        lib().orc_csr_spmv(A.rows, _p(A.off), _p(A.col), _p(A.val), _p(x), _p(y))
    return y


def spmv_mt(A: Csr, x, threads):
    """spmv() with rows split over `threads` OpenMP threads (same result)."""
    x = _f32(x)
    y = np.zeros(A.rows, np.float32)
    lib().orc_csr_spmv_mt(A.rows, _p(A.off), _p(A.col), _p(A.val), _p(x), _p(y), int(threads))
    return y


def spmm_dense(A: Csr, X):
    X = _f32(X)
    k = X.shape[1]
    Y = np.zeros((A.rows, k), np.float32)
    lib().orc_spmm_dense(A.rows, _p(A.off), _p(A.col), _p(A.val), k, _p(X), X.strides[0] // 4, _p(Y), k)
    return Y


def spgemm(A: Csr, B: Csr, use_ref=False) -> Csr:
    cnt = lib().orc_spgemm_count(A.rows, B.cols, _p(A.off), _p(A.col), _p(B.off), _p(B.col))
    cap = max(int(cnt), 1)
    off = np.zeros(A.rows + 1, np.int32)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float32)
    if use_ref:
        nnz = ref().ref_spgemm(A.rows, A.cols, B.cols, _p(A.off), _p(A.col), _p(A.val),
                               _p(B.off), _p(B.col), _p(B.val), _p(off), _p(col), _p(val))
    else:
        nnz = lib().orc_spgemm(A.rows, B.cols, _p(A.off), _p(A.col), _p(A.val),
                               _p(B.off), _p(B.col), _p(B.val), _p(off), _p(col), _p(val))
    return Csr(A.rows, B.cols, off, col[:nnz], val[:nnz])


def spgemm_check_f64(A: Csr, B: Csr, C: Csr, rtol=1e-5):
    """Number of stored entries of C outside rtol * (|A||B|) of the FP64
    product A*B (entries without any partial product count as failures),
    and the flat index of the first one (-1 if none)."""
    first = ctypes.c_int64(-1)
    bad = lib().orc_spgemm_check_f64(A.rows, B.cols, _p(A.off), _p(A.col), _p(A.val), _p(B.off), _p(B.col),
                                     _p(B.val), _p(C.off), _p(C.col), _p(C.val), float(rtol), ctypes.byref(first))
    if bad < 0:
        raise MemoryError("orc_spgemm_check_f64 failed")
    return int(bad), int(first.value)


def spgemm_count(A: Csr, B: Csr) -> int:
    return int(lib().orc_spgemm_count(A.rows, B.cols, _p(A.off), _p(A.col), _p(B.off), _p(B.col)))


def hyb_pitch(R, alignment=32):
    return ((R + alignment - 1) // alignment) * alignment


def csr_to_hyb(A: Csr, width, use_ref=False):
    """Returns (ell_col[width*pitch], ell_val, coo_row, coo_col, coo_val, pitch)."""
    pitch = hyb_pitch(A.rows)
    ell_c = np.zeros(max(width * pitch, 1), np.int32)
    ell_v = np.zeros(max(width * pitch, 1), np.float32)
    cap = max(A.nnz, 1)
    cr = np.zeros(cap, np.int32)
    cc = np.zeros(cap, np.int32)
    cv = np.zeros(cap, np.float32)
    if use_ref:
        po = ctypes.c_int64(0)
        nc = ref().ref_csr_to_hyb(A.rows, A.cols, _p(A.off), _p(A.col), _p(A.val), width,
                                  ctypes.byref(po), _p(ell_c), _p(ell_v), _p(cr), _p(cc), _p(cv))
        pitch = po.value
    else:
        nc = lib().orc_csr_to_hyb(A.rows, _p(A.off), _p(A.col), _p(A.val), width, pitch,
                                  _p(ell_c), _p(ell_v), _p(cr), _p(cc), _p(cv))
    n = width * pitch
    return ell_c[:n], ell_v[:n], cr[:nc], cc[:nc], cv[:nc], pitch


class Hyb:
    """Host HYB matrix in the reference's dynamic-HYB layout
    (include/dcsr_matrix.h:148-172): ELL block of `width` slots per row,
    column-major with pitch (padding column -1, value 0), COO overflow sorted
    by row, and row_sizes[r] = entries of row r (ELL + COO), row_sizes[R] =
    COO count."""

    __slots__ = ("rows", "cols", "width", "pitch", "row_sizes", "ell_col", "ell_val", "coo_row", "coo_col",
                 "coo_val")

    def __init__(self, rows, cols, width, pitch, row_sizes, ell_col, ell_val, coo_row, coo_col, coo_val):
        self.rows, self.cols, self.width, self.pitch = int(rows), int(cols), int(width), int(pitch)
        self.row_sizes = _i32(row_sizes)
        self.ell_col, self.ell_val = _i32(ell_col), _f32(ell_val)
        self.coo_row, self.coo_col, self.coo_val = _i32(coo_row), _i32(coo_col), _f32(coo_val)

    @property
    def ncoo(self):
        return int(self.coo_row.size)

    def to_coo(self):
        """Every stored entry (ELL slots in slot order per row, then COO)."""
        R, w, p = self.rows, self.width, self.pitch
        ec = self.ell_col[: w * p].reshape(w, p)[:, :R]
        ev = self.ell_val[: w * p].reshape(w, p)[:, :R]
        k, rr = np.nonzero(ec != -1)
        order = np.lexsort((k, rr))
        r = np.concatenate([rr[order], self.coo_row])
        c = np.concatenate([ec[k, rr][order], self.coo_col])
        v = np.concatenate([ev[k, rr][order], self.coo_val])
        return r.astype(np.int32), c.astype(np.int32), v.astype(np.float32)


def csr_to_hyb_full(A: Csr, width, use_ref=False) -> Hyb:
    """CSR -> HYB (cusp's generic conversion) with the row sizes."""
    ec, ev, cr, cc, cv, pitch = csr_to_hyb(A, width, use_ref=use_ref)
    rs = np.zeros(A.rows + 1, np.int32)
    rs[: A.rows] = np.diff(A.off)
    rs[A.rows] = cr.size
    return Hyb(A.rows, A.cols, width, pitch, rs, ec, ev, cr, cc, cv)


def coo_to_hyb(R, C, rows, cols, vals, width, use_ref=False) -> Hyb:
    """COO -> HYB: sort_by_row + COO->CSR, then the CSR -> HYB conversion
    (the first `width` entries of each row, in input order, fill its ELL
    slots; the rest go to COO in row order)."""
    return csr_to_hyb_full(coo_to_csr(R, C, rows, cols, vals, use_ref=use_ref), width, use_ref=use_ref)


def hyb_update(H: Hyb, rows, cols, vals) -> Hyb:
    """The reference's append-only HYB update (orc_hyb_update)."""
    r, c, v = _i32(rows), _i32(cols), _f32(vals)
    rs = H.row_sizes.copy()
    ec, ev = H.ell_col.copy(), H.ell_val.copy()
    cap = max(H.ncoo + r.size, 1)
    orow, ocol, oval = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.float32)
    k = lib().orc_hyb_update(H.rows, H.width, H.pitch, _p(rs), _p(ec), _p(ev), H.ncoo, _p(H.coo_row), _p(H.coo_col),
                             _p(H.coo_val), r.size, _p(r), _p(c), _p(v), _p(orow), _p(ocol), _p(oval))
    if k < 0:
        raise MemoryError("orc_hyb_update failed")
    return Hyb(H.rows, H.cols, H.width, H.pitch, rs, ec, ev, orow[:k], ocol[:k], oval[:k])


def hyb_spmv(H: Hyb, x, use_ref=False):
    x = _f32(x)
    y = np.zeros(H.rows, np.float32)
    if use_ref:
        ref().ref_hyb_spmv(H.rows, H.cols, H.width, _p(H.ell_col), _p(H.ell_val), H.ncoo, _p(H.coo_row),
                           _p(H.coo_col), _p(H.coo_val), _p(x), _p(y))
    else:
        lib().orc_hyb_spmv(H.rows, H.width, H.pitch, _p(H.ell_col), _p(H.ell_val), H.ncoo, _p(H.coo_row),
                           _p(H.coo_col), _p(H.coo_val), _p(x), _p(y))
    return y
