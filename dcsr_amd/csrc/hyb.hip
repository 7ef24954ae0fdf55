// hyb.hip -- HYB (ELL + COO) <-> DCSR conversions.
//
// The reference converts COO/CSR into its hyb_matrix wrapper
// (include/load.h:61-106, 174-205 -> LoadMatrix_hyb_coo / LoadMatrix_hyb_csr,
// include/sparse_update.inl:569-698) and has no direct HYB <-> DCSR path.
// Layout follows cusp's hyb_matrix, the reference's container
// (include/dcsr_matrix.h:148-172): an ELL block of `width` entries per row in
// column-major order (slot k of row r at k * pitch + r, padding column -1 and
// value 0) and a COO overflow in row order.
//
// DCSR -> HYB puts the first `width` entries of every row (columns ascending)
// into the ELL block and the rest into COO, as the host CSR -> HYB conversion
// does with the same width (cusp/system/detail/generic/conversions/
// csr_to_other.h:226-303).  HYB -> DCSR loads every ELL entry with a valid
// column (slot order) followed by the COO part, like COO -> DCSR.
#include "dcsr_internal.h"
#include "prims.h"
#include "segments.h"

namespace {

constexpr int TB = 256;

__global__ void k_hyb_pad(int32_t *__restrict__ ell_col, float *__restrict__ ell_val, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        ell_col[i] = -1;
        ell_val[i] = 0.0f;
    }
}

__global__ void k_hyb_overflow(uint32_t R, const int32_t *__restrict__ len, uint32_t width, uint32_t *__restrict__ ov)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) {
        const uint32_t L = (uint32_t)len[r];
        ov[r] = L > width ? L - width : 0u;
    }
}

__global__ __launch_bounds__(TB) void k_hyb_scatter(uint32_t R, const uint32_t *__restrict__ foff, uint32_t total,
                                                    const uint32_t *__restrict__ row_start,
                                                    const int32_t *__restrict__ pcol, const float *__restrict__ pval,
                                                    uint32_t width, uint64_t pitch, int32_t *__restrict__ ell_col,
                                                    float *__restrict__ ell_val, const uint32_t *__restrict__ coff,
                                                    int32_t *__restrict__ coo_row, int32_t *__restrict__ coo_col,
                                                    float *__restrict__ coo_val)
{
    __shared__ uint32_t s_off[TB / 64][65];
    const int w = threadIdx.x >> 6;
    const uint64_t wave = (uint64_t)blockIdx.x * (TB / 64) + w;
    const uint32_t per = 256;
    const uint32_t dev_total = foff[R];
    if (total > dev_total) total = dev_total;
    const uint64_t e0 = wave * per;
    if (e0 >= total) return;
    const uint64_t e1 = e0 + per < total ? e0 + per : total;
    wave_visit(foff, R, (uint32_t)e0, (uint32_t)e1, s_off[w], [&](bool valid, uint32_t r, uint32_t k, uint32_t) {
        if (!valid) return;
        const uint64_t sp = (uint64_t)row_start[r] + k;
        const int32_t c = pcol[sp];
        const float v = pval[sp];
        if (k < width) {
            const uint64_t i = (uint64_t)k * pitch + r;
            ell_col[i] = c;
            ell_val[i] = v;
        } else {
            const uint64_t o = (uint64_t)coff[r] + (k - width);
            coo_row[o] = (int32_t)r;
            coo_col[o] = c;
            coo_val[o] = v;
        }
    });
}

// ELL slots that hold an entry (row < R and column != -1).
__global__ void k_hyb_ell_flags(const int32_t *__restrict__ ell_col, uint64_t n, uint64_t pitch, uint32_t R,
                                uint32_t *__restrict__ flag)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    flag[i] = (i < n && (i % pitch) < R && ell_col[i] != -1) ? 1u : 0u;
}

__global__ void k_hyb_ell_gather(const int32_t *__restrict__ ell_col, const float *__restrict__ ell_val, uint64_t n,
                                 uint64_t pitch, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos,
                                 int32_t *__restrict__ rows, int32_t *__restrict__ cols, float *__restrict__ vals)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t p = pos[i];
    rows[p] = (int32_t)(i % pitch);
    cols[p] = ell_col[i];
    vals[p] = ell_val[i];
}

}  // namespace

dcsr_status load_coo_internal(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                              hipStream_t s);

dcsr_status dcsr_export_hyb(dcsr_handle h, int32_t width, int64_t pitch, int32_t *ell_col, float *ell_val,
                            int32_t *coo_row, int32_t *coo_col, float *coo_val, int64_t coo_cap, int64_t *coo_n,
                            dcsr_stream_t stream)
{
    if (!h || width < 0 || pitch < h->rows || !coo_n) return DCSR_ERR_INVALID;
    if (width > 0 && h->rows > 0 && (!ell_col || !ell_val)) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    DCSR_TRY(ensure_flat(h, s));
    const uint32_t R = (uint32_t)h->rows;
    const uint64_t ncell = (uint64_t)width * (uint64_t)pitch;
    uint32_t *ov, *coff, *part;
    DCSR_TRY(ws_get(h, WS_TMP1, (size_t)(R + 1) * 4, (void **)&ov));
    DCSR_TRY(ws_get(h, WS_TMP0, (size_t)(R + 1) * 4, (void **)&coff));
    DCSR_TRY(ws_get_zeroed(h, WS_PART, scan_part_words(R + 1, 1) * 4, (void **)&part, s));
    if (R > 0) k_hyb_overflow<<<grid_for(R, TB), TB, 0, s>>>(R, h->row_len, (uint32_t)width, ov);
    DCSR_HIP(scan_exclusive_u32(ov, coff, R, 1, part, coff + R, s, fault_word(h)));
    uint32_t ncoo = 0;
    DCSR_HIP(hipMemcpyAsync(&ncoo, coff + R, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    *coo_n = ncoo;
    if ((int64_t)ncoo > coo_cap) return DCSR_ERR_CAPACITY;
    if (ncoo > 0 && (!coo_row || !coo_col || !coo_val)) return DCSR_ERR_INVALID;
    if (ncell > 0) {
        const uint64_t g = (ncell + TB - 1) / TB;
        k_hyb_pad<<<(unsigned)(g < 65536 ? g : 65536), TB, 0, s>>>(ell_col, ell_val, ncell);
    }
    const int64_t nnz = h->nnz_h;
    if (R > 0 && nnz > 0) {
        const uint32_t *foff = (const uint32_t *)h->ws.ptr[WS_FOFF];
        k_hyb_scatter<<<grid_for(nnz, 4 * 256), TB, 0, s>>>(R, foff, (uint32_t)nnz, h->row_start, h->pool_col,
                                                            h->pool_val, (uint32_t)width, (uint64_t)pitch, ell_col,
                                                            ell_val, coff, coo_row, coo_col, coo_val);
    }
    DCSR_HIP(hipGetLastError());
    DCSR_HIP(hipStreamSynchronize(s));
    return DCSR_OK;
}

dcsr_status dcsr_load_hyb(dcsr_handle h, int32_t width, int64_t pitch, const int32_t *ell_col, const float *ell_val,
                          const int32_t *coo_row, const int32_t *coo_col, const float *coo_val, int64_t coo_n,
                          dcsr_stream_t stream)
{
    if (!h || width < 0 || pitch < h->rows || coo_n < 0) return DCSR_ERR_INVALID;
    if (width > 0 && h->rows > 0 && (!ell_col || !ell_val)) return DCSR_ERR_INVALID;
    if (coo_n > 0 && (!coo_row || !coo_col || !coo_val)) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    const uint64_t ncell = (uint64_t)width * (uint64_t)pitch;
    if (ncell >= 0xfffffff0ull) return DCSR_ERR_UNSUPPORTED;
    uint32_t *flag, *pos, *part;
    DCSR_TRY(ws_get(h, WS_FLAGS, (size_t)(ncell + 2) * 4, (void **)&flag));
    DCSR_TRY(ws_get(h, WS_TMP0, (size_t)(ncell + 2) * 4, (void **)&pos));
    DCSR_TRY(ws_get_zeroed(h, WS_PART, scan_part_words(ncell + 2, 1) * 4, (void **)&part, s));
    k_hyb_ell_flags<<<grid_for(ncell + 1, TB), TB, 0, s>>>(ell_col, ncell, (uint64_t)pitch, (uint32_t)h->rows, flag);
    DCSR_HIP(scan_exclusive_u32(flag, pos, ncell + 1, 1, part, nullptr, s, fault_word(h)));
    uint32_t nell = 0;
    DCSR_HIP(hipMemcpyAsync(&nell, pos + ncell, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    const int64_t n = (int64_t)nell + coo_n;
    int32_t *rows, *cols;
    float *vals;
    DCSR_TRY(ws_get(h, WS_UCOL, (size_t)(n + 1) * 4, (void **)&rows));
    DCSR_TRY(ws_get(h, WS_UPOS, (size_t)(n + 1) * 4, (void **)&cols));
    DCSR_TRY(ws_get(h, WS_UVAL, (size_t)(n + 1) * 4, (void **)&vals));
    if (ncell > 0) k_hyb_ell_gather<<<grid_for(ncell, TB), TB, 0, s>>>(ell_col, ell_val, ncell, (uint64_t)pitch, flag,
                                                                       pos, rows, cols, vals);
    if (coo_n > 0) {
        DCSR_HIP(hipMemcpyAsync(rows + nell, coo_row, (size_t)coo_n * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(cols + nell, coo_col, (size_t)coo_n * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(vals + nell, coo_val, (size_t)coo_n * 4, hipMemcpyDeviceToDevice, s));
    }
    DCSR_HIP(hipGetLastError());
    return load_coo_internal(h, rows, cols, vals, n, s);
}
