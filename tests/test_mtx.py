"""MatrixMarket ingest (include/dcsr_mtx.hpp through the C ABI): the
reference's bundled matrices (matrices/can_256.mtx, a symmetric pattern
file, and matrices/tols90.mtx, real general) read against scipy's
independent reader, plus crafted files for each banner case and error.

The reference's drivers read these files with ReadMatrixFile
(include/load_matrix.h:54-116), which sets every value to 1.0 and keeps
symmetric storage unexpanded; the reader here follows the MatrixMarket
format as cusp::io::read_matrix_market_file does (matrix_market.inl:117-269)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io

import dcsr_amd

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def scipy_sorted(path):
    S = scipy.io.mmread(path).tocoo()
    r, c = S.row.astype(np.int64), S.col.astype(np.int64)
    v = S.data.astype(np.float32)
    order = np.lexsort((c, r))
    return S.shape, r[order], c[order], v[order]


@pytest.mark.parametrize("name,sym", [("can_256.mtx", True), ("tols90.mtx", False)])
def test_bundled_matrices_match_scipy(name, sym):
    path = os.path.join(GOLD, name)
    R, C, r, c, v = dcsr_amd.read_mtx(path)
    (sr, sc), er, ec, ev = scipy_sorted(path)
    assert (R, C) == (sr, sc)
    np.testing.assert_array_equal(r, er)
    np.testing.assert_array_equal(c, ec)
    np.testing.assert_array_equal(v.view(np.uint32), ev.view(np.uint32))
    if sym:
        # pattern symmetric: both triangles present, every value 1
        A = np.zeros((R, C), bool)
        A[r, c] = True
        assert (A == A.T).all() and (v == 1.0).all()
    else:
        assert not (v == 1.0).all()  # real values kept (no all-ones)


def write(tmp_path, text, name="m.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_symmetric_real_expansion(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real symmetric\n% c\n\n3 3 4\n1 1 2.5\n3 1 -1\n"
                        "2 2 4\n3 2 0.5\n")
    R, C, r, c, v = dcsr_amd.read_mtx(p)
    assert (R, C) == (3, 3)
    got = list(zip(r.tolist(), c.tolist(), v.tolist()))
    assert got == [(0, 0, 2.5), (0, 2, -1.0), (1, 1, 4.0), (1, 2, 0.5), (2, 0, -1.0), (2, 1, 0.5)]


def test_skew_symmetric_and_integer(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate integer skew-symmetric\n2 2 1\n2 1 7\n")
    R, C, r, c, v = dcsr_amd.read_mtx(p)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [(0, 1, -7.0), (1, 0, 7.0)]


def test_duplicates_keep_file_order(tmp_path):
    p = write(tmp_path, "%%MatrixMarket matrix coordinate real general\n2 3 4\n2 3 1\n1 2 5\n2 3 2\n1 1 3\n")
    R, C, r, c, v = dcsr_amd.read_mtx(p)
    assert list(zip(r.tolist(), c.tolist(), v.tolist())) == [(0, 0, 3.0), (0, 1, 5.0), (1, 2, 1.0), (1, 2, 2.0)]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "storage"),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "data type"),
    ("%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n", "symmetry"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", "end of"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "row index"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n", "column index"),
    ("%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n", "banner"),
])
def test_errors(tmp_path, text, msg):
    p = write(tmp_path, text)
    with pytest.raises(dcsr_amd.DcsrError, match=msg):
        dcsr_amd.read_mtx(p)


def test_missing_file():
    with pytest.raises(dcsr_amd.DcsrError, match="cannot open"):
        dcsr_amd.read_mtx("/nonexistent/x.mtx")


def test_header_compiles_standalone(tmp_path):
    """include/dcsr_mtx.hpp in a plain g++ program reads tols90 like the
    library does."""
    src = tmp_path / "rd.cpp"
    src.write_text('#include "dcsr_mtx.hpp"\n#include <cstdio>\n'
                   'int main(int, char** a){dcsr::MtxCoo m; std::string e;\n'
                   ' if(!dcsr::read_mtx(a[1], m, &e)){std::printf("ERR %s\\n", e.c_str()); return 1;}\n'
                   ' double s=0; for(size_t k=0;k<m.row.size();++k) s += (m.row[k]+1.0)*(m.col[k]+2.0)*m.val[k];\n'
                   ' std::printf("%lld %lld %zu %.17g\\n", (long long)m.rows, (long long)m.cols, m.row.size(), s);}\n')
    exe = tmp_path / "rd"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), os.path.join(GOLD, "tols90.mtx")], capture_output=True, text=True, check=True)
    R, C, r, c, v = dcsr_amd.read_mtx(os.path.join(GOLD, "tols90.mtx"))
    s = float(np.sum((r + 1.0) * (c + 2.0) * v.astype(np.float64)))
    rr, cc, nn, ss = out.stdout.split()
    assert (int(rr), int(cc), int(nn)) == (R, C, r.size)
    assert abs(float(ss) - s) <= 1e-9 * abs(s)
