// dcsr_matrix.hpp -- C++ drop-in shim over the C ABI (dcsr.h) that restates
// the reference's container and `device::` entry points with the same names
// and argument order:
//   dcsr_matrix<I,V,MEM,BINS>           include/dcsr_matrix.h:8-109
//   device::Initialize_Matrix           include/load.h:8-22
//   device::LoadMatrix (COO/CSR->DCSR, DCSR->CSR, COO/CSR->HYB, DCSR<->HYB)
//                                        include/load.h:27-59, 61-106, 174-253
//   device::UpdateMatrix (DCSR, HYB)    include/load.h:258-325, 327-373
//   device::SortMatrixRow               include/load.h:404-581
//   device::BinRows                     include/load.h:583-610
//   device::spmv (DCSR, CSR, HYB)      include/spmv.h:121-248, 310-332;
//                                        src/Fill_Test.inl:167-170
//   device::spmm (SpGEMM)               include/spmm.h:30-391
// cusp containers are replaced by the repo-owned dcsr::device_array<T> and the
// small csr_matrix / hyb_matrix holders below; errors throw std::runtime_error
// (the reference printed and continued).  Only <int, float> is provided, the
// reference's instantiation (src/Fill_Test.inl:11-16, PRECISION 32).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "dcsr.h"

namespace dcsr {

inline void throw_if(dcsr_status s, const char *what)
{
    if (s != DCSR_OK) throw std::runtime_error(std::string(what) + ": " + dcsr_status_string(s));
}

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device array (the role cusp::array1d<T, device_memory> played).
template <class T>
class device_array {
    T *p_ = nullptr;
    size_t n_ = 0;

public:
    device_array() = default;
    explicit device_array(size_t n) { resize(n); }
    explicit device_array(const std::vector<T> &h)
    {
        resize(h.size());
        copy_from(h.data(), h.size());
    }
    ~device_array()
    {
        if (p_) (void)hipFree(p_);
    }
    device_array(const device_array &) = delete;
    device_array &operator=(const device_array &) = delete;
    device_array(device_array &&o) noexcept : p_(o.p_), n_(o.n_)
    {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    device_array &operator=(device_array &&o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    void resize(size_t n)
    {
        if (n == n_) return;
        if (p_) hip_check(hipFree(p_), "hipFree");
        p_ = nullptr;
        n_ = n;
        if (n) hip_check(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)), "hipMalloc");
    }
    size_t size() const { return n_; }
    T *data() { return p_; }
    const T *data() const { return p_; }
    void copy_from(const T *host, size_t n)
    {
        if (n > n_) resize(n);
        if (n) hip_check(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
    }
    std::vector<T> to_host() const
    {
        std::vector<T> h(n_);
        if (n_) hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
        return h;
    }
};

}  // namespace dcsr

struct device_memory {};  // memory-space tag kept for source compatibility

// Device CSR holder (the cusp::csr_matrix<int, float, device_memory> role).
template <typename I = int, typename V = float, typename MEM = device_memory>
struct csr_matrix {
    size_t num_rows = 0, num_cols = 0, num_entries = 0;
    dcsr::device_array<I> row_offsets, column_indices;
    dcsr::device_array<V> values;
    void resize(size_t r, size_t c, size_t nnz)
    {
        num_rows = r;
        num_cols = c;
        num_entries = nnz;
        row_offsets.resize(r + 1);
        column_indices.resize(nnz);
        values.resize(nnz);
    }
};

// Dynamic HYB matrix (the reference's hyb_matrix wrapper,
// include/dcsr_matrix.h:148-172): an owning handle of the library's HYB store
// (ELL block of num_cols_per_row slots per row, column-major with pitch, COO
// overflow sorted by row, row sizes).  The device arrays are reached through
// view(); num_entries / num_overflow are refreshed by every operation.
template <typename I = int, typename V = float, typename MEM = device_memory>
struct hyb_matrix {
    static_assert(std::is_same<I, int>::value && std::is_same<V, float>::value, "the engine is built for <int, float>");
    dcsr_hyb h = nullptr;
    size_t num_rows = 0, num_cols = 0, num_cols_per_row = 0, pitch = 0, num_entries = 0, num_overflow = 0;

    hyb_matrix() = default;
    ~hyb_matrix()
    {
        if (h) (void)dcsr_hyb_destroy(h);
    }
    hyb_matrix(const hyb_matrix &) = delete;
    hyb_matrix &operator=(const hyb_matrix &) = delete;

    // resize(rows, cols, num_coo_entries, num_cols_per_row): an empty store;
    // the COO part starts with room for 2 * num_coo_entries, as the reference
    // reserved, and grows on demand
    void resize(size_t rows, size_t cols, size_t num_coo_entries, size_t cols_per_row)
    {
        if (h) (void)dcsr_hyb_destroy(h);
        h = nullptr;
        dcsr::throw_if(dcsr_hyb_create(&h, (int64_t)rows, (int64_t)cols, (int32_t)cols_per_row,
                                       (int64_t)(2 * num_coo_entries)),
                       "hyb_matrix::resize");
        num_rows = rows;
        num_cols = cols;
        num_cols_per_row = cols_per_row;
        refresh();
    }
    dcsr_hyb_view view() const
    {
        dcsr_hyb_view v;
        dcsr::throw_if(dcsr_hyb_view_get(h, &v), "dcsr_hyb_view_get");
        return v;
    }
    void refresh()
    {
        const dcsr_hyb_view v = view();
        pitch = (size_t)v.pitch;
        num_overflow = (size_t)v.num_overflow;
    }
};

template <typename INDEX_TYPE, typename VALUE_TYPE, typename MEM_TYPE = device_memory, size_t BINS = 4>
struct dcsr_matrix {
    static_assert(std::is_same<INDEX_TYPE, int>::value && std::is_same<VALUE_TYPE, float>::value,
                  "the engine is built for <int, float>");
    dcsr_handle h = nullptr;
    size_t num_rows = 0, num_cols = 0;
    size_t bin_length = 0;  // per-row preallocation used by Initialize_Matrix
    float alpha = 0.0f;     // kept for source compatibility (unused)

    dcsr_matrix() = default;
    ~dcsr_matrix()
    {
        if (h) (void)dcsr_destroy(h);
    }
    dcsr_matrix(const dcsr_matrix &) = delete;  // the reference's copy was shallow
    dcsr_matrix &operator=(const dcsr_matrix &) = delete;

    // resize(rows, cols, bin_size, factor): pool = bin_size * rows * factor
    void resize(size_t n_rows, size_t n_cols, size_t bin_size, float factor)
    {
        reset(n_rows, n_cols, (double)factor, (int64_t)(bin_size * n_rows * factor));
        bin_length = bin_size;
    }
    // resize(rows, cols, NNZ): pool of exactly NNZ entries (grown on demand)
    void resize(size_t n_rows, size_t n_cols, size_t NNZ)
    {
        reset(n_rows, n_cols, 1.0, (int64_t)NNZ);
        bin_length = 0;
    }
    size_t num_entries() const
    {
        dcsr_info i;
        dcsr::throw_if(dcsr_query(h, &i), "dcsr_query");
        return (size_t)i.nnz;
    }

private:
    void reset(size_t r, size_t c, double factor, int64_t pool)
    {
        if (h) (void)dcsr_destroy(h);
        h = nullptr;
        num_rows = r;
        num_cols = c;
        dcsr_config cfg{factor < 1.0 ? 1.0 : factor, pool > 1024 ? pool : 1024, 1, 0};
        dcsr::throw_if(dcsr_create(&h, (int64_t)r, (int64_t)c, &cfg), "dcsr_create");
    }
};

namespace device {

using dcsr::device_array;

template <typename I, typename V, size_t BINS>
void Initialize_Matrix(dcsr_matrix<I, V, device_memory, BINS> &mat, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_initialize(mat.h, (int32_t)mat.bin_length, s), "Initialize_Matrix");
}

// COO -> DCSR (rows need not be sorted)
template <typename I, typename V, size_t BINS>
void LoadMatrix(dcsr_matrix<I, V, device_memory, BINS> &dst, const device_array<I> &rows, const device_array<I> &cols,
                const device_array<V> &vals, const int N, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_load_coo(dst.h, rows.data(), cols.data(), vals.data(), N, s), "LoadMatrix(COO->DCSR)");
}

// CSR -> DCSR
template <typename I, typename V, size_t BINS>
void LoadMatrix(dcsr_matrix<I, V, device_memory, BINS> &dst, const csr_matrix<I, V> &src, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_load_csr(dst.h, src.row_offsets.data(), src.column_indices.data(), src.values.data(), s),
                   "LoadMatrix(CSR->DCSR)");
}

// DCSR -> CSR (natural row order, columns ascending)
template <typename I, typename V, size_t BINS>
void LoadMatrix(csr_matrix<I, V> &dst, const dcsr_matrix<I, V, device_memory, BINS> &src, hipStream_t s = nullptr)
{
    const size_t nnz = src.num_entries();
    dst.resize(src.num_rows, src.num_cols, nnz);
    int64_t got = 0;
    dcsr::throw_if(dcsr_export_csr(src.h, dst.row_offsets.data(), dst.column_indices.data(), dst.values.data(),
                                   (int64_t)nnz, &got, 0, s),
                   "LoadMatrix(DCSR->CSR)");
    dst.num_entries = (size_t)got;
}

// DCSR -> HYB with the holder's num_cols_per_row (rows with columns ascending)
template <typename I, typename V, size_t BINS>
void LoadMatrix(hyb_matrix<I, V> &dst, const dcsr_matrix<I, V, device_memory, BINS> &src, hipStream_t s = nullptr)
{
    if (!dst.h || dst.num_rows != src.num_rows || dst.num_cols != src.num_cols)
        dst.resize(src.num_rows, src.num_cols, 0, dst.num_cols_per_row);
    dcsr::throw_if(dcsr_hyb_from_dcsr(dst.h, src.h, s), "LoadMatrix(DCSR->HYB)");
    dst.refresh();
}

// HYB -> DCSR (every stored entry)
template <typename I, typename V, size_t BINS>
void LoadMatrix(dcsr_matrix<I, V, device_memory, BINS> &dst, const hyb_matrix<I, V> &src, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_hyb_to_dcsr(src.h, dst.h, s), "LoadMatrix(HYB->DCSR)");
}

// COO -> HYB: the first num_cols_per_row entries of each row fill its ELL
// slots, the rest go to the COO overflow (rows need not be sorted)
template <typename I, typename V>
void LoadMatrix(hyb_matrix<I, V> &mat, const device_array<I> &rows, const device_array<I> &cols,
                const device_array<V> &vals, const int N, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_hyb_load_coo(mat.h, rows.data(), cols.data(), vals.data(), N, s), "LoadMatrix(COO->HYB)");
    mat.num_entries = (size_t)N;
    mat.refresh();
}

// CSR -> HYB
template <typename I, typename V>
void LoadMatrix(hyb_matrix<I, V> &dst, const csr_matrix<I, V> &src, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_hyb_load_csr(dst.h, src.row_offsets.data(), src.column_indices.data(), src.values.data(), s),
                   "LoadMatrix(CSR->HYB)");
    dst.num_entries = src.num_entries;
    dst.refresh();
}

// Append update of the HYB matrix (entries past a row's ELL slots overflow
// into the COO part; duplicates are kept and add up in SpMV)
template <typename I, typename V>
void UpdateMatrix(hyb_matrix<I, V> &mat, const device_array<I> &rows, const device_array<I> &cols,
                  const device_array<V> &vals, const int N, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_hyb_update(mat.h, rows.data(), cols.data(), vals.data(), N, s), "UpdateMatrix(HYB)");
    mat.num_entries += (size_t)N;
    mat.refresh();
}

// insert-or-add; `permuted` selected a kernel variant in the reference and
// has no effect on the result here
template <typename I, typename V, size_t BINS>
void UpdateMatrix(dcsr_matrix<I, V, device_memory, BINS> &mat, const device_array<I> &rows,
                  const device_array<I> &cols, const device_array<V> &vals, const bool permuted = false,
                  hipStream_t s = nullptr)
{
    (void)permuted;
    dcsr::throw_if(dcsr_update(mat.h, rows.data(), cols.data(), vals.data(), (int64_t)rows.size(), 0, s),
                   "UpdateMatrix");
}

template <typename I, typename V, size_t BINS>
void SortMatrixRow(dcsr_matrix<I, V, device_memory, BINS> &mat, const hipStream_t *streams = nullptr)
{
    dcsr::throw_if(dcsr_compact(mat.h, streams ? streams[0] : nullptr), "SortMatrixRow");
}

template <typename I, typename V, size_t BINS>
void BinRows(dcsr_matrix<I, V, device_memory, BINS> &mat, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_rebin(mat.h, s), "BinRows");
}

template <typename I, typename V, size_t BINS>
void spmv(const dcsr_matrix<I, V, device_memory, BINS> &A, const device_array<V> &x, device_array<V> &y,
          const hipStream_t *streams = nullptr)
{
    dcsr::throw_if(dcsr_spmv(A.h, x.data(), y.data(), 1.0f, 0.0f, streams ? streams[0] : nullptr), "spmv");
}

// y = A x on the HYB matrix (the reference's driver called cusp::multiply)
template <typename I, typename V>
void spmv(const hyb_matrix<I, V> &A, const device_array<V> &x, device_array<V> &y, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_hyb_spmv(A.h, x.data(), y.data(), 1.0f, 0.0f, s), "spmv(HYB)");
}

template <typename I, typename V>
void spmv(const csr_matrix<I, V> &A, const device_array<V> &x, device_array<V> &y, hipStream_t s = nullptr)
{
    dcsr::throw_if(dcsr_csr_spmv((int64_t)A.num_rows, A.row_offsets.data(), A.column_indices.data(), A.values.data(),
                                 x.data(), y.data(), s),
                   "spmv(CSR)");
}

// SpGEMM C = A * B (the reference calls it spmm)
template <typename I, typename V, size_t BINS>
void spmm(const dcsr_matrix<I, V, device_memory, BINS> &A, const dcsr_matrix<I, V, device_memory, BINS> &B,
          dcsr_matrix<I, V, device_memory, BINS> &C, const hipStream_t *streams = nullptr)
{
    if (!C.h || C.num_rows != A.num_rows || C.num_cols != B.num_cols) C.resize(A.num_rows, B.num_cols, 1024);
    dcsr::throw_if(dcsr_spgemm(A.h, B.h, C.h, streams ? streams[0] : nullptr), "spmm");
}

}  // namespace device
