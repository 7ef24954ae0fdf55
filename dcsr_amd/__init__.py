"""dcsr_amd -- MI355X-native Dynamic-CSR engine (Python side).

The compute path is the in-tree HIP library ``dcsr_amd/libdcsr.so`` (gfx950),
reached through its C ABI (``include/dcsr.h``) with ctypes.  PyTorch only
provides device memory and the current HIP stream.  There is no CPU fallback:
if the library is missing, every call raises.

``DcsrMatrix`` mirrors the reference's ``dcsr_matrix`` container and the
``device::`` entry points that operate on it (reference include/load.h,
include/spmv.h): ``Initialize_Matrix``, ``LoadMatrix`` (COO / CSR -> DCSR),
``to_csr`` (DCSR -> CSR), ``UpdateMatrix``, ``SortMatrixRow`` (defrag),
``BinRows`` and ``spmv``.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdcsr.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "dcsr.h")

DCSR_OK = 0
DCSR_ERR_TOO_LARGE = 9
STATUS_NAMES = {
    0: "DCSR_OK", 1: "DCSR_ERR_INVALID", 2: "DCSR_ERR_BAD_INDEX", 3: "DCSR_ERR_POOL_EXHAUSTED",
    4: "DCSR_ERR_SEGMENT_OVERFLOW", 5: "DCSR_ERR_HIP", 6: "DCSR_ERR_OOM", 7: "DCSR_ERR_CAPACITY",
    8: "DCSR_ERR_UNSUPPORTED", 9: "DCSR_ERR_TOO_LARGE",
}
UPDATE_NOSYNC = 0x1
UPDATE_NOGROW = 0x2
EXPORT_CANONICAL = 0x1


class DcsrError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = int(code)
        super().__init__(f"{what}: {STATUS_NAMES.get(self.code, self.code)}")


class DcsrConfig(ctypes.Structure):
    _fields_ = [("pool_factor", ctypes.c_double), ("min_pool", ctypes.c_int64),
                ("allow_grow", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class DcsrRequest(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("k", ctypes.c_int32), ("n", ctypes.c_int64),
                ("rows", ctypes.c_void_p), ("cols", ctypes.c_void_p), ("vals", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("ldx", ctypes.c_int64), ("ldy", ctypes.c_int64)]


class DcsrServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("updates", "update_triples", "queries", "query_groups", "compactions")]


class DcsrHybView(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("width", ctypes.c_int64),
                ("pitch", ctypes.c_int64), ("ell_col", ctypes.c_void_p), ("ell_val", ctypes.c_void_p),
                ("row_sizes", ctypes.c_void_p), ("coo_row", ctypes.c_void_p), ("coo_col", ctypes.c_void_p),
                ("coo_val", ctypes.c_void_p), ("num_overflow", ctypes.c_int64), ("coo_cap", ctypes.c_int64)]


class DcsrInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "rows", "cols", "nnz", "pool_cap", "pool_used", "garbage", "hub_rows",
        "last_inserted", "last_added", "last_deleted", "last_touched_rows")] + \
        [("canonical", ctypes.c_int32), ("max_segments", ctypes.c_int32),
         ("scan_fallback", ctypes.c_int32), ("spmv_paneled", ctypes.c_int32),
         ("compactions", ctypes.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32
_F = ctypes.c_float
_H = ctypes.c_void_p


def header_functions():
    """Names of the functions declared in include/dcsr.h."""
    with open(HEADER_PATH) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(dcsr_[a-z0-9_]+)\s*\(", txt, re.M)))


def lib():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run python -m dcsr_amd.build (no CPU fallback exists)")
    # torch ships its own HIP runtime under the same SONAME (libamdhip64.so.7):
    # whichever is loaded first serves the whole process, so load torch's
    # before this library pulls in the system one
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "dcsr_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "dcsr_get_status": (ctypes.c_int, [_H]),
        "dcsr_create": (ctypes.c_int, [ctypes.POINTER(_H), _I64, _I64, ctypes.POINTER(DcsrConfig)]),
        "dcsr_destroy": (ctypes.c_int, [_H]),
        "dcsr_initialize": (ctypes.c_int, [_H, ctypes.c_int32, _P]),
        "dcsr_load_coo": (ctypes.c_int, [_H, _P, _P, _P, _I64, _P]),
        "dcsr_load_csr": (ctypes.c_int, [_H, _P, _P, _P, _P]),
        "dcsr_export_csr": (ctypes.c_int, [_H, _P, _P, _P, _I64, ctypes.POINTER(_I64), _U32, _P]),
        "dcsr_update": (ctypes.c_int, [_H, _P, _P, _P, _I64, _U32, _P]),
        "dcsr_compact": (ctypes.c_int, [_H, _P]),
        "dcsr_rebin": (ctypes.c_int, [_H, _P]),
        "dcsr_spmv": (ctypes.c_int, [_H, _P, _P, _F, _F, _P]),
        "dcsr_spmm_dense": (ctypes.c_int, [_H, _P, _I64, _I64, _P, _I64, _F, _F, _P]),
        "dcsr_spgemm": (ctypes.c_int, [_H, _H, _H, _P]),
        "dcsr_export_hyb": (ctypes.c_int, [_H, ctypes.c_int32, _I64, _P, _P, _P, _P, _P, _I64,
                                           ctypes.POINTER(_I64), _P]),
        "dcsr_load_hyb": (ctypes.c_int, [_H, ctypes.c_int32, _I64, _P, _P, _P, _P, _P, _I64, _P]),
        "dcsr_batch_create": (ctypes.c_int, [ctypes.POINTER(_H), ctypes.c_int32, _P, _P, ctypes.POINTER(DcsrConfig)]),
        "dcsr_batch_destroy": (ctypes.c_int, [_H]),
        "dcsr_batch_handle": (_H, [_H]),
        "dcsr_batch_info": (ctypes.c_int, [_H, _P, _P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "dcsr_batch_load_coo": (ctypes.c_int, [_H, _P, _P, _P, _P, _I64, _P]),
        "dcsr_batch_step": (ctypes.c_int, [_H, _P, _P, _P, _P, _I64, _P, _P, _U32, _P]),
        "dcsr_batch_status": (ctypes.c_int, [_H, _P]),
        "dcsr_batch_autotune": (ctypes.c_int, [_H, _P, _P, ctypes.POINTER(ctypes.c_int32),
                                               ctypes.POINTER(ctypes.c_int32), _P]),
        "dcsr_batch_set_shape": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int32]),
        "dcsr_server_create": (ctypes.c_int, [ctypes.POINTER(_H), _H, ctypes.c_int32]),
        "dcsr_server_submit": (ctypes.c_int, [_H, ctypes.POINTER(DcsrRequest), ctypes.POINTER(ctypes.c_uint64)]),
        "dcsr_server_wait": (ctypes.c_int, [_H, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
        "dcsr_server_drain": (ctypes.c_int, [_H]),
        "dcsr_server_get_stats": (ctypes.c_int, [_H, ctypes.POINTER(DcsrServerStats)]),
        "dcsr_server_destroy": (ctypes.c_int, [_H]),
        "dcsr_query": (ctypes.c_int, [_H, ctypes.POINTER(DcsrInfo)]),
        "dcsr_csr_spmv": (ctypes.c_int, [_I64, _P, _P, _P, _P, _P, _P]),
        "dcsr_export_layout": (ctypes.c_int, [_H, _P, _P, _P, _P]),
        "dcsr_hyb_create": (ctypes.c_int, [ctypes.POINTER(_H), _I64, _I64, ctypes.c_int32, _I64]),
        "dcsr_hyb_destroy": (ctypes.c_int, [_H]),
        "dcsr_hyb_view_get": (ctypes.c_int, [_H, ctypes.POINTER(DcsrHybView)]),
        "dcsr_hyb_load_coo": (ctypes.c_int, [_H, _P, _P, _P, _I64, _P]),
        "dcsr_hyb_load_csr": (ctypes.c_int, [_H, _P, _P, _P, _P]),
        "dcsr_hyb_update": (ctypes.c_int, [_H, _P, _P, _P, _I64, _P]),
        "dcsr_hyb_spmv": (ctypes.c_int, [_H, _P, _P, _F, _F, _P]),
        "dcsr_hyb_export": (ctypes.c_int, [_H, _P, _P, _P, _P, _P, _P, _I64, ctypes.POINTER(_I64), _P]),
        "dcsr_hyb_to_dcsr": (ctypes.c_int, [_H, _H, _P]),
        "dcsr_hyb_from_dcsr": (ctypes.c_int, [_H, _H, _P]),
        "dcsr_set_scan_wait_limit": (None, [ctypes.c_int32]),
        "dcsr_update_noalloc": (ctypes.c_int, [_H, _P, _P, _P, _I64, _U32, _P]),
        "dcsr_update_noalloc_max_triples": (_I64, []),
        "dcsr_csr_add": (ctypes.c_int, [_I64, _I64, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _I64,
                                        ctypes.POINTER(_I64), _P]),
        "dcsr_csr_spgemm": (ctypes.c_int, [_I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64,
                                           ctypes.POINTER(_I64), _P]),
        "dcsr_mtx_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_H), ctypes.POINTER(_I64),
                                         ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
        "dcsr_mtx_copy": (ctypes.c_int, [_H, _P, _P, _P]),
        "dcsr_mtx_free": (ctypes.c_int, [_H]),
        "dcsr_mtx_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code, what=""):
    if code != DCSR_OK:
        raise DcsrError(code, what)


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else ctypes.c_void_p(0)


def _dev(a, dtype):
    """Host array / tensor -> contiguous CUDA tensor of dtype."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a.to(device="cuda", dtype=dtype)
    else:
        import numpy as np
        t = torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)
    return t.contiguous()


class DcsrMatrix:
    """Owning handle of one device-resident Dynamic-CSR matrix.

    Mirrors ``dcsr_matrix<int, float, device_memory, BINS>`` (reference
    include/dcsr_matrix.h:8-109) with its ``device::`` operations."""

    def __init__(self, num_rows, num_cols, pool_factor=2.0, min_pool=1024, allow_grow=True):
        torch = _torch()
        torch.cuda.init()
        self.num_rows = int(num_rows)
        self.num_cols = int(num_cols)
        cfg = DcsrConfig(pool_factor, min_pool, 1 if allow_grow else 0, 0)
        h = _H()
        check(lib().dcsr_create(ctypes.byref(h), self.num_rows, self.num_cols, ctypes.byref(cfg)), "dcsr_create")
        self._h = h

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().dcsr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # -- reference API mirror ---------------------------------------------
    def Initialize_Matrix(self, bin_length):
        """device::Initialize_Matrix (reference include/load.h:8-22)."""
        check(lib().dcsr_initialize(self._h, int(bin_length), _stream()), "dcsr_initialize")

    def LoadMatrix(self, rows, cols, vals):
        """device::LoadMatrix(dcsr&, rows, cols, vals, N) -- COO -> DCSR
        (reference include/load.h:27-59).  Rows need not be sorted."""
        torch = _torch()
        r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        if not (r.numel() == c.numel() == v.numel()):
            raise ValueError("rows/cols/vals length mismatch")
        check(lib().dcsr_load_coo(self._h, _ptr(r), _ptr(c), _ptr(v), r.numel(), _stream()), "dcsr_load_coo")

    def load_csr(self, row_offsets, cols, vals):
        """device::LoadMatrix(dcsr&, const csr_matrix&) (reference include/load.h:207-227)."""
        torch = _torch()
        o, c, v = _dev(row_offsets, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        if o.numel() != self.num_rows + 1:
            raise ValueError("row_offsets must have rows+1 entries")
        check(lib().dcsr_load_csr(self._h, _ptr(o), _ptr(c), _ptr(v), _stream()), "dcsr_load_csr")

    def to_csr(self, canonical=False):
        """device::LoadMatrix(csr_matrix&, const dcsr&) -- DCSR -> CSR in
        natural row order with columns ascending (reference include/load.h:232-253)."""
        torch = _torch()
        info = self.info()
        cap = max(info["nnz"], 1)
        off = torch.empty(self.num_rows + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(cap, dtype=torch.int32, device="cuda")
        val = torch.empty(cap, dtype=torch.float32, device="cuda")
        nnz = _I64(0)
        check(lib().dcsr_export_csr(self._h, _ptr(off), _ptr(col), _ptr(val), cap, ctypes.byref(nnz),
                                    EXPORT_CANONICAL if canonical else 0, _stream()), "dcsr_export_csr")
        return off, col[: nnz.value], val[: nnz.value]

    def UpdateMatrix(self, rows, cols, vals, flags=0):
        """device::UpdateMatrix (reference include/load.h:258-325), with
        insert-or-add semantics; rows need not be sorted."""
        torch = _torch()
        r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        if not (r.numel() == c.numel() == v.numel()):
            raise ValueError("rows/cols/vals length mismatch")
        check(lib().dcsr_update(self._h, _ptr(r), _ptr(c), _ptr(v), r.numel(), int(flags), _stream()),
              "dcsr_update")

    def SortMatrixRow(self):
        """device::SortMatrixRow -- defragmentation (reference include/load.h:404-581)."""
        check(lib().dcsr_compact(self._h, _stream()), "dcsr_compact")

    def BinRows(self):
        """device::BinRows (reference include/load.h:583-610)."""
        check(lib().dcsr_rebin(self._h, _stream()), "dcsr_rebin")

    def spmv(self, x, y=None, alpha=1.0, beta=0.0):
        """device::spmv(const dcsr&, x, y) (reference include/spmv.h:121-248):
        y = alpha * A x + beta * y."""
        torch = _torch()
        x = _dev(x, torch.float32)
        if x.numel() != self.num_cols:
            raise ValueError("x must have num_cols entries")
        if y is None:
            y = torch.empty(self.num_rows, dtype=torch.float32, device="cuda")
            beta = 0.0
        elif y.numel() != self.num_rows or y.dtype != torch.float32 or not y.is_cuda or not y.is_contiguous():
            raise ValueError("y must be a contiguous float32 CUDA tensor with num_rows entries")
        check(lib().dcsr_spmv(self._h, _ptr(x), _ptr(y), float(alpha), float(beta), _stream()), "dcsr_spmv")
        return y

    def spmm(self, X, Y=None, alpha=1.0, beta=0.0):
        """Y = alpha * A X + beta * Y for a dense row-major X (num_cols x k)."""
        torch = _torch()
        X = X if isinstance(X, torch.Tensor) else torch.from_numpy(X)
        X = X.to(device="cuda", dtype=torch.float32)
        if X.dim() != 2 or X.shape[0] != self.num_cols or X.stride(1) != 1:
            raise ValueError("X must be a (num_cols, k) row-major float32 tensor")
        k = X.shape[1]
        if Y is None:
            Y = torch.empty((self.num_rows, k), dtype=torch.float32, device="cuda")
            beta = 0.0
        elif Y.shape != (self.num_rows, k) or Y.dtype != torch.float32 or Y.stride(1) != 1 or not Y.is_cuda:
            raise ValueError("Y must be a (num_rows, k) row-major float32 CUDA tensor")
        check(lib().dcsr_spmm_dense(self._h, _ptr(X), k, X.stride(0), _ptr(Y), Y.stride(0), float(alpha),
                                    float(beta), _stream()), "dcsr_spmm_dense")
        return Y

    def to_hyb(self, width, pitch=None):
        """DCSR -> HYB: (ell_col, ell_val) of shape (width, pitch) column-major
        slots and COO (row, col, val) overflow; pitch defaults to rows rounded
        up to 32, as cusp's hyb_matrix."""
        torch = _torch()
        pitch = ((self.num_rows + 31) // 32) * 32 if pitch is None else int(pitch)
        nnz = self.info()["nnz"]
        ec = torch.empty(max(width * pitch, 1), dtype=torch.int32, device="cuda")
        ev = torch.empty(max(width * pitch, 1), dtype=torch.float32, device="cuda")
        cap = max(nnz, 1)
        cr = torch.empty(cap, dtype=torch.int32, device="cuda")
        cc = torch.empty(cap, dtype=torch.int32, device="cuda")
        cv = torch.empty(cap, dtype=torch.float32, device="cuda")
        n = _I64(0)
        check(lib().dcsr_export_hyb(self._h, int(width), pitch, _ptr(ec), _ptr(ev), _ptr(cr), _ptr(cc), _ptr(cv),
                                    cap, ctypes.byref(n), _stream()), "dcsr_export_hyb")
        k = n.value
        return ec[: width * pitch], ev[: width * pitch], cr[:k], cc[:k], cv[:k], pitch

    def load_hyb(self, width, pitch, ell_col, ell_val, coo_row, coo_col, coo_val):
        """HYB -> DCSR (ELL slots with column != -1, then the COO part)."""
        torch = _torch()
        ec, ev = _dev(ell_col, torch.int32), _dev(ell_val, torch.float32)
        cr, cc, cv = _dev(coo_row, torch.int32), _dev(coo_col, torch.int32), _dev(coo_val, torch.float32)
        check(lib().dcsr_load_hyb(self._h, int(width), int(pitch), _ptr(ec), _ptr(ev), _ptr(cr), _ptr(cc), _ptr(cv),
                                  cr.numel(), _stream()), "dcsr_load_hyb")

    def info(self):
        inf = DcsrInfo()
        check(lib().dcsr_query(self._h, ctypes.byref(inf)), "dcsr_query")
        return inf.as_dict()

    def layout(self):
        """The allocator's row table (start, len, cap) as CUDA tensors."""
        torch = _torch()
        R = max(self.num_rows, 1)
        st = torch.empty(R, dtype=torch.int32, device="cuda")
        ln = torch.empty(R, dtype=torch.int32, device="cuda")
        cp = torch.empty(R, dtype=torch.int32, device="cuda")
        check(lib().dcsr_export_layout(self._h, _ptr(st), _ptr(ln), _ptr(cp), _stream()), "dcsr_export_layout")
        n = self.num_rows
        return st[:n], ln[:n], cp[:n]

    def status(self):
        return lib().dcsr_get_status(self._h)

    @property
    def nnz(self):
        return self.info()["nnz"]


def spgemm(A: DcsrMatrix, B: DcsrMatrix, C: DcsrMatrix = None) -> DcsrMatrix:
    """C = A * B (device::spmm, reference include/spmm.h:30-391)."""
    if A.num_cols != B.num_rows:
        raise ValueError("inner dimensions differ")
    if C is None:
        C = DcsrMatrix(A.num_rows, B.num_cols)
    check(lib().dcsr_spgemm(A.handle, B.handle, C.handle, _stream()), "dcsr_spgemm")
    return C


BATCH_GRAPH = 0x1


class DcsrBatch:
    """Many small DCSR matrices stepped (update + SpMV) in one persistent
    launch.  Matrix m is rows[m] x cols[m] (each <= 4096), local indices."""

    def __init__(self, rows, cols, pool_factor=3.0, allow_grow=True, min_pool=1024):
        import numpy as np
        torch = _torch()
        torch.cuda.init()
        r = np.ascontiguousarray(rows, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        assert r.size == c.size
        self.count = r.size
        cfg = DcsrConfig(pool_factor, min_pool, 1 if allow_grow else 0, 0)
        h = _H()
        check(lib().dcsr_batch_create(ctypes.byref(h), r.size, r.ctypes.data_as(_P), c.ctypes.data_as(_P),
                                      ctypes.byref(cfg)), "dcsr_batch_create")
        self._b = h
        rb = np.zeros(r.size + 1, np.int64)
        cb = np.zeros(r.size + 1, np.int64)
        check(lib().dcsr_batch_info(h, rb.ctypes.data_as(_P), cb.ctypes.data_as(_P), None, None), "dcsr_batch_info")
        self.row_base, self.col_base = rb, cb
        self.total_rows, self.total_cols = int(rb[-1]), int(cb[-1])

    def close(self):
        if getattr(self, "_b", None):
            lib().dcsr_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_lengths(self, o, r, c, v):
        if o.numel() != self.count + 1:
            raise DcsrError(1, f"tri_off needs {self.count + 1} entries, got {o.numel()}")
        if not (r.numel() == c.numel() == v.numel()):
            raise DcsrError(1, "rows, cols and vals differ in length")

    def load_coo(self, tri_off, rows, cols, vals):
        torch = _torch()
        o = _dev(tri_off, torch.int64)
        r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        self._check_lengths(o, r, c, v)
        check(lib().dcsr_batch_load_coo(self._b, _ptr(o), _ptr(r), _ptr(c), _ptr(v), r.numel(), _stream()),
              "dcsr_batch_load_coo")

    def step(self, tri_off=None, rows=None, cols=None, vals=None, x=None, y=None, graph=False):
        """Update every matrix with its triples, then y = A x per matrix."""
        torch = _torch()
        if tri_off is not None:
            o = _dev(tri_off, torch.int64)
            r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
            self._check_lengths(o, r, c, v)
            n = r.numel()
        else:
            o = r = c = v = None
            n = 0
        if x is not None:
            x = _dev(x, torch.float32)
            if y is None:
                y = torch.empty(self.total_rows, dtype=torch.float32, device="cuda")
        check(lib().dcsr_batch_step(self._b, _ptr(o), _ptr(r), _ptr(c), _ptr(v), n, _ptr(x), _ptr(y),
                                    BATCH_GRAPH if graph else 0, _stream()), "dcsr_batch_step")
        self._keep = (o, r, c, v, x, y)  # inputs must outlive a captured graph
        return y

    def status(self):
        import numpy as np
        per = np.zeros(self.count, np.uint32)
        code = lib().dcsr_batch_status(self._b, per.ctypes.data_as(_P))
        return code, per

    def autotune(self, x, y):
        t, th = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib().dcsr_batch_autotune(self._b, _ptr(x), _ptr(y), ctypes.byref(t), ctypes.byref(th), _stream()),
              "dcsr_batch_autotune")
        return t.value, th.value

    def set_shape(self, team, thresh):
        check(lib().dcsr_batch_set_shape(self._b, int(team), int(thresh)), "dcsr_batch_set_shape")

    def info(self):
        """Counters of the whole batch (dcsr_query on its storage); the
        last_* fields are those of the last step once status() has run."""
        inf = DcsrInfo()
        check(lib().dcsr_query(lib().dcsr_batch_handle(self._b), ctypes.byref(inf)), "dcsr_query")
        return inf.as_dict()

    def layout(self):
        """The allocator's row table of the block-diagonal storage."""
        torch = _torch()
        n = max(self.total_rows, 1)
        st, ln, cp = (torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3))
        check(lib().dcsr_export_layout(lib().dcsr_batch_handle(self._b), _ptr(st), _ptr(ln), _ptr(cp), _stream()),
              "dcsr_export_layout")
        return st[: self.total_rows], ln[: self.total_rows], cp[: self.total_rows]

    def to_csr(self):
        """The whole block-diagonal batch as CSR (local column indices)."""
        torch = _torch()
        h = lib().dcsr_batch_handle(self._b)
        inf = DcsrInfo()
        check(lib().dcsr_query(h, ctypes.byref(inf)), "dcsr_query")
        cap = max(inf.nnz, 1)
        off = torch.empty(self.total_rows + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(cap, dtype=torch.int32, device="cuda")
        val = torch.empty(cap, dtype=torch.float32, device="cuda")
        nnz = _I64(0)
        check(lib().dcsr_export_csr(h, _ptr(off), _ptr(col), _ptr(val), cap, ctypes.byref(nnz), 0, _stream()),
              "dcsr_export_csr")
        return off, col[: nnz.value], val[: nnz.value]


class DcsrHyb:
    """Dynamic HYB matrix (ELL + COO overflow): the reference's hyb_matrix
    (include/dcsr_matrix.h:148-172) with LoadMatrix (COO / CSR -> HYB),
    UpdateMatrix (append) and SpMV (include/load.h:61-106, 174-205,
    327-373)."""

    def __init__(self, num_rows, num_cols, width, coo_cap=1024):
        torch = _torch()
        torch.cuda.init()
        self.num_rows, self.num_cols, self.width = int(num_rows), int(num_cols), int(width)
        h = _H()
        check(lib().dcsr_hyb_create(ctypes.byref(h), self.num_rows, self.num_cols, self.width, int(coo_cap)),
              "dcsr_hyb_create")
        self._h = h
        self.pitch = self.view()["pitch"]

    def close(self):
        if getattr(self, "_h", None):
            lib().dcsr_hyb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def view(self):
        v = DcsrHybView()
        check(lib().dcsr_hyb_view_get(self._h, ctypes.byref(v)), "dcsr_hyb_view_get")
        return {n: getattr(v, n) for n, _ in v._fields_}

    def LoadMatrix(self, rows, cols, vals):
        """LoadMatrix(hyb&, rows, cols, vals, N) (reference include/load.h:61-106)."""
        torch = _torch()
        r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        check(lib().dcsr_hyb_load_coo(self._h, _ptr(r), _ptr(c), _ptr(v), r.numel(), _stream()), "dcsr_hyb_load_coo")

    def load_csr(self, row_offsets, cols, vals):
        """LoadMatrix(hyb&, const csr&) (reference include/load.h:174-205)."""
        torch = _torch()
        o, c, v = _dev(row_offsets, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        check(lib().dcsr_hyb_load_csr(self._h, _ptr(o), _ptr(c), _ptr(v), _stream()), "dcsr_hyb_load_csr")

    def UpdateMatrix(self, rows, cols, vals):
        """UpdateMatrix(hyb&, rows, cols, vals, N) (reference include/load.h:327-373): append."""
        torch = _torch()
        r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        check(lib().dcsr_hyb_update(self._h, _ptr(r), _ptr(c), _ptr(v), r.numel(), _stream()), "dcsr_hyb_update")

    def spmv(self, x, y=None, alpha=1.0, beta=0.0):
        torch = _torch()
        x = _dev(x, torch.float32)
        if y is None:
            y = torch.empty(self.num_rows, dtype=torch.float32, device="cuda")
            beta = 0.0
        check(lib().dcsr_hyb_spmv(self._h, _ptr(x), _ptr(y), float(alpha), float(beta), _stream()), "dcsr_hyb_spmv")
        return y

    def export(self):
        """(ell_col, ell_val, row_sizes, coo_row, coo_col, coo_val) as CUDA tensors."""
        torch = _torch()
        n = self.view()["num_overflow"]
        ncell = max(self.width * self.pitch, 1)
        ec = torch.empty(ncell, dtype=torch.int32, device="cuda")
        ev = torch.empty(ncell, dtype=torch.float32, device="cuda")
        rs = torch.empty(self.num_rows + 1, dtype=torch.int32, device="cuda")
        cr, cc = (torch.empty(max(n, 1), dtype=torch.int32, device="cuda") for _ in range(2))
        cv = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
        got = _I64(0)
        check(lib().dcsr_hyb_export(self._h, _ptr(ec), _ptr(ev), _ptr(rs), _ptr(cr), _ptr(cc), _ptr(cv), max(n, 1),
                                    ctypes.byref(got), _stream()), "dcsr_hyb_export")
        k = got.value
        m = self.width * self.pitch
        return ec[:m], ev[:m], rs, cr[:k], cc[:k], cv[:k]

    def to_dcsr(self, m):
        check(lib().dcsr_hyb_to_dcsr(self._h, m.handle, _stream()), "dcsr_hyb_to_dcsr")

    def from_dcsr(self, m):
        check(lib().dcsr_hyb_from_dcsr(self._h, m.handle, _stream()), "dcsr_hyb_from_dcsr")


REQ_UPDATE, REQ_SPMV, REQ_SPMM = 0, 1, 2


class DcsrServer:
    """Resident-matrix request loop with continuous batching (see
    include/dcsr.h).  Request tensors are kept alive until completion."""

    def __init__(self, matrix: DcsrMatrix, max_cols=32):
        h = _H()
        check(lib().dcsr_server_create(ctypes.byref(h), matrix.handle, int(max_cols)), "dcsr_server_create")
        self._s = h
        self.matrix = matrix
        self._live = {}

    def _submit(self, req, keep):
        # the server reads on its own streams: inputs must be complete now
        _torch().cuda.current_stream().synchronize()
        t = ctypes.c_uint64(0)
        check(lib().dcsr_server_submit(self._s, ctypes.byref(req), ctypes.byref(t)), "dcsr_server_submit")
        self._live[t.value] = keep
        return t.value

    def update(self, rows, cols, vals):
        torch = _torch()
        r, c, v = _dev(rows, torch.int32), _dev(cols, torch.int32), _dev(vals, torch.float32)
        req = DcsrRequest(REQ_UPDATE, 0, r.numel(), r.data_ptr() if r.numel() else None,
                          c.data_ptr() if c.numel() else None, v.data_ptr() if v.numel() else None, None, None, 0, 0)
        return self._submit(req, (r, c, v))

    def spmv(self, x, y=None):
        torch = _torch()
        x = _dev(x, torch.float32)
        if y is None:
            y = torch.empty(self.matrix.num_rows, dtype=torch.float32, device="cuda")
        req = DcsrRequest(REQ_SPMV, 1, 0, None, None, None, x.data_ptr(), y.data_ptr(), 1, 1)
        return self._submit(req, (x, y)), y

    def spmm(self, X, Y=None):
        torch = _torch()
        X = _dev(X, torch.float32)
        k = X.shape[1]
        if Y is None:
            Y = torch.empty((self.matrix.num_rows, k), dtype=torch.float32, device="cuda")
        req = DcsrRequest(REQ_SPMM, k, 0, None, None, None, X.data_ptr(), Y.data_ptr(), X.stride(0), Y.stride(0))
        return self._submit(req, (X, Y)), Y

    def wait(self, ticket):
        lat = ctypes.c_double(0.0)
        code = lib().dcsr_server_wait(self._s, ticket, ctypes.byref(lat))
        self._live.pop(ticket, None)
        check(code, "server request")
        return lat.value

    def drain(self):
        check(lib().dcsr_server_drain(self._s), "dcsr_server_drain")

    def stats(self):
        st = DcsrServerStats()
        check(lib().dcsr_server_get_stats(self._s, ctypes.byref(st)), "dcsr_server_get_stats")
        return {n: getattr(st, n) for n, _ in st._fields_}

    def close(self):
        if getattr(self, "_s", None):
            lib().dcsr_server_destroy(self._s)
            self._s = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_scan_wait_limit(polls):
    """Poll limit of the scans' bounded look-back waits (negative: the
    default); 0 sends every tile down its exact fallback without waiting."""
    lib().dcsr_set_scan_wait_limit(int(polls))


def csr_spmv(row_offsets, cols, vals, x, y=None):
    """Plain CSR SpMV comparator (reference include/spmv.h:310-332)."""
    torch = _torch()
    o = _dev(row_offsets, torch.int32)
    c = _dev(cols, torch.int32)
    v = _dev(vals, torch.float32)
    x = _dev(x, torch.float32)
    R = o.numel() - 1
    if y is None:
        y = torch.empty(R, dtype=torch.float32, device="cuda")
    check(lib().dcsr_csr_spmv(R, _ptr(o), _ptr(c), _ptr(v), _ptr(x), _ptr(y), _stream()), "dcsr_csr_spmv")
    return y


def read_mtx(path):
    """MatrixMarket file -> (rows, cols, row, col, val): 0-based int32 / float32
    numpy arrays sorted by (row, column), symmetric storage expanded, the
    file's values kept (include/dcsr_mtx.hpp; host only, no GPU needed)."""
    import numpy as np
    L = lib()
    h = _H()
    R, C, N = _I64(), _I64(), _I64()
    code = L.dcsr_mtx_read(os.fsencode(path), ctypes.byref(h), ctypes.byref(R), ctypes.byref(C), ctypes.byref(N))
    if code != DCSR_OK:
        raise DcsrError(code, f"dcsr_mtx_read({path}): {L.dcsr_mtx_error().decode()}")
    try:
        n = N.value
        r = np.empty(n, np.int32)
        c = np.empty(n, np.int32)
        v = np.empty(n, np.float32)
        check(L.dcsr_mtx_copy(h, r.ctypes.data, c.ctypes.data, v.ctypes.data), "dcsr_mtx_copy")
    finally:
        L.dcsr_mtx_free(h)
    return R.value, C.value, r, c, v


def csr_add(rows, cols, row_offsets, col_idx, vals, b_rows, b_cols, b_vals):
    """CSR-rebuild comparator C = A + B (cusp's generic add; include/dcsr.h
    dcsr_csr_add).  Returns device tensors (row_offsets, cols, vals)."""
    torch = _torch()
    o, c, v = _dev(row_offsets, torch.int32), _dev(col_idx, torch.int32), _dev(vals, torch.float32)
    br, bc, bv = _dev(b_rows, torch.int32), _dev(b_cols, torch.int32), _dev(b_vals, torch.float32)
    cap = c.numel() + br.numel()
    co = torch.empty(rows + 1, dtype=torch.int32, device="cuda")
    cc = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
    cv = torch.empty(max(cap, 1), dtype=torch.float32, device="cuda")
    nnz = _I64()
    check(lib().dcsr_csr_add(rows, cols, _ptr(o), _ptr(c), _ptr(v), _ptr(br), _ptr(bc), _ptr(bv), br.numel(),
                             _ptr(co), _ptr(cc), _ptr(cv), cap, ctypes.byref(nnz), _stream()), "dcsr_csr_add")
    return co, cc[:nnz.value], cv[:nnz.value]


def csr_spgemm(a_rows, a_cols, b_cols, a_off, a_col, a_val, b_off, b_col, b_val):
    """CSR global-ESC SpGEMM comparator C = A * B (include/dcsr.h
    dcsr_csr_spgemm).  Returns device tensors (row_offsets, cols, vals)."""
    torch = _torch()
    ao, ac, av = _dev(a_off, torch.int32), _dev(a_col, torch.int32), _dev(a_val, torch.float32)
    bo, bc, bv = _dev(b_off, torch.int32), _dev(b_col, torch.int32), _dev(b_val, torch.float32)
    blen = (bo[1:] - bo[:-1]).to(torch.int64)
    cap = int(blen[ac.to(torch.int64)].sum().item()) if ac.numel() else 0
    co = torch.empty(a_rows + 1, dtype=torch.int32, device="cuda")
    cc = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
    cv = torch.empty(max(cap, 1), dtype=torch.float32, device="cuda")
    nnz = _I64()
    check(lib().dcsr_csr_spgemm(a_rows, a_cols, b_cols, _ptr(ao), _ptr(ac), _ptr(av), _ptr(bo), _ptr(bc), _ptr(bv),
                                _ptr(co), _ptr(cc), _ptr(cv), cap, ctypes.byref(nnz), _stream()), "dcsr_csr_spgemm")
    return co, cc[:nnz.value], cv[:nnz.value]
