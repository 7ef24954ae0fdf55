// mtx.hip -- C ABI over the MatrixMarket reader of include/dcsr_mtx.hpp
// (host code only).  Replaces cusp::io::read_matrix_market_file
// (cusp/io/detail/matrix_market.inl:117-269) and the drivers' ReadMatrixFile
// (include/load_matrix.h:54-116) for callers of the C ABI.
#include <string>

#include "../../include/dcsr.h"
#include "../../include/dcsr_mtx.hpp"

struct dcsr_mtx_s {
    dcsr::MtxCoo m;
};

static thread_local std::string g_mtx_error;

extern "C" dcsr_status dcsr_mtx_read(const char *path, dcsr_mtx *out, int64_t *rows, int64_t *cols, int64_t *nnz)
{
    if (!path || !out) return DCSR_ERR_INVALID;
    *out = nullptr;
    dcsr_mtx_s *h = new (std::nothrow) dcsr_mtx_s;
    if (!h) return DCSR_ERR_OOM;
    std::string err;
    if (!dcsr::read_mtx(std::string(path), h->m, &err)) {
        g_mtx_error = err;
        delete h;
        return DCSR_ERR_INVALID;
    }
    g_mtx_error.clear();
    if (rows) *rows = h->m.rows;
    if (cols) *cols = h->m.cols;
    if (nnz) *nnz = (int64_t)h->m.row.size();
    *out = h;
    return DCSR_OK;
}

extern "C" dcsr_status dcsr_mtx_copy(dcsr_mtx m, int32_t *row, int32_t *col, float *val)
{
    if (!m) return DCSR_ERR_INVALID;
    const size_t n = m->m.row.size();
    if (n && (!row || !col || !val)) return DCSR_ERR_INVALID;
    for (size_t k = 0; k < n; ++k) {
        row[k] = m->m.row[k];
        col[k] = m->m.col[k];
        val[k] = m->m.val[k];
    }
    return DCSR_OK;
}

extern "C" dcsr_status dcsr_mtx_free(dcsr_mtx m)
{
    delete m;
    return DCSR_OK;
}

extern "C" const char *dcsr_mtx_error(void) { return g_mtx_error.c_str(); }
