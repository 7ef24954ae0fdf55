// ref_harness.cpp -- C entry points over the reference's own host path
// (cusp 0.5.1 / thrust 1.8.1 headers, included from the reference checkout at
// build time; nothing from it is copied here).  Built by oracle/Makefile into
// oracle/_ref/libdcsr_ref.so and used ONLY as a checker of oracle/dcsr_oracle.c
// and as the optional "reference" CPU baseline.  Every call uses cusp host
// containers, which dispatch to cusp::system::detail::sequential.
#include <cusp/array1d.h>
#include <cusp/coo_matrix.h>
#include <cusp/csr_matrix.h>
#include <cusp/hyb_matrix.h>
#include <cusp/elementwise.h>
#include <cusp/multiply.h>
#include <cusp/sort.h>
#include <cusp/convert.h>
#include <cstdint>
#include <cstring>

typedef cusp::csr_matrix<int, float, cusp::host_memory> HostCsr;
typedef cusp::array1d<int, cusp::host_memory> HostI;
typedef cusp::array1d<float, cusp::host_memory> HostF;

static void fill_csr(HostCsr &A, int64_t R, int64_t C, const int *off, const int *col, const float *val)
{
    int64_t nnz = off[R];
    A.resize(R, C, nnz);
    std::memcpy(&A.row_offsets[0], off, (R + 1) * sizeof(int));
    if (nnz) {
        std::memcpy(&A.column_indices[0], col, nnz * sizeof(int));
        std::memcpy(&A.values[0], val, nnz * sizeof(float));
    }
}

static int64_t drain_csr(const HostCsr &A, int *off, int *col, float *val)
{
    int64_t R = A.num_rows, nnz = A.num_entries;
    std::memcpy(off, &A.row_offsets[0], (R + 1) * sizeof(int));
    if (nnz) {
        std::memcpy(col, &A.column_indices[0], nnz * sizeof(int));
        std::memcpy(val, &A.values[0], nnz * sizeof(float));
    }
    return nnz;
}

extern "C" {

int ref_sort_by_row(int64_t n, int *rows, int *cols, float *vals)
{
    if (n <= 0) return 0;
    HostI r(rows, rows + n), c(cols, cols + n);
    HostF v(vals, vals + n);
    cusp::sort_by_row(r, c, v);
    std::memcpy(rows, &r[0], n * sizeof(int));
    std::memcpy(cols, &c[0], n * sizeof(int));
    std::memcpy(vals, &v[0], n * sizeof(float));
    return 0;
}

// COO (already row-sorted) -> CSR through cusp's generic conversion.
int ref_coo_to_csr(int64_t R, int64_t C, int64_t n, const int *rows, const int *cols,
                   const float *vals, int *off, int *ocol, float *oval)
{
    cusp::coo_matrix<int, float, cusp::host_memory> A(R, C, n);
    if (n) {
        std::memcpy(&A.row_indices[0], rows, n * sizeof(int));
        std::memcpy(&A.column_indices[0], cols, n * sizeof(int));
        std::memcpy(&A.values[0], vals, n * sizeof(float));
    }
    HostCsr B;
    cusp::convert(A, B);
    if (n == 0) { for (int64_t r = 0; r <= R; ++r) off[r] = 0; return 0; }
    drain_csr(B, off, ocol, oval);
    return 0;
}

int64_t ref_csr_add(int64_t R, int64_t C,
                    const int *Aoff, const int *Acol, const float *Aval,
                    const int *Boff, const int *Bcol, const float *Bval,
                    int *Coff, int *Ccol, float *Cval)
{
    HostCsr A, B, Cm;
    fill_csr(A, R, C, Aoff, Acol, Aval);
    fill_csr(B, R, C, Boff, Bcol, Bval);
    cusp::add(A, B, Cm);
    return drain_csr(Cm, Coff, Ccol, Cval);
}

void ref_csr_spmv(int64_t R, int64_t C, const int *off, const int *col, const float *val,
                  const float *x, float *y)
{
    HostCsr A;
    fill_csr(A, R, C, off, col, val);
    HostF xv(x, x + C), yv(R, 0.0f);
    cusp::multiply(A, xv, yv);
    std::memcpy(y, &yv[0], R * sizeof(float));
}

// C = A * B; C arrays sized by the caller from ref_spgemm_count.
int64_t ref_spgemm(int64_t R, int64_t K, int64_t Cb,
                   const int *Aoff, const int *Acol, const float *Aval,
                   const int *Boff, const int *Bcol, const float *Bval,
                   int *Coff, int *Ccol, float *Cval)
{
    HostCsr A, B, Cm;
    fill_csr(A, R, K, Aoff, Acol, Aval);
    fill_csr(B, K, Cb, Boff, Bcol, Bval);
    cusp::multiply(A, B, Cm);
    return drain_csr(Cm, Coff, Ccol, Cval);
}

// Host CSR -> HYB with an explicit ELL width through the generic conversion.
int64_t ref_csr_to_hyb(int64_t R, int64_t C, const int *off, const int *col, const float *val,
                       int64_t width, int64_t *pitch_out,
                       int *ell_col, float *ell_val, int *coo_row, int *coo_col, float *coo_val)
{
    HostCsr A;
    fill_csr(A, R, C, off, col, val);
    cusp::hyb_matrix<int, float, cusp::host_memory> H;
    cusp::csr_format f1;
    cusp::hyb_format f2;
    thrust::cpp::tag exec;
    cusp::system::detail::generic::convert(exec, A, H, f1, f2, (size_t)width, (size_t)32);
    int64_t pitch = H.ell.column_indices.pitch;
    *pitch_out = pitch;
    int64_t ne = (int64_t)H.ell.column_indices.values.size();
    if (ell_col) {
        std::memcpy(ell_col, &H.ell.column_indices.values[0], ne * sizeof(int));
        std::memcpy(ell_val, &H.ell.values.values[0], ne * sizeof(float));
    }
    int64_t nc = H.coo.num_entries;
    if (coo_row && nc) {
        std::memcpy(coo_row, &H.coo.row_indices[0], nc * sizeof(int));
        std::memcpy(coo_col, &H.coo.column_indices[0], nc * sizeof(int));
        std::memcpy(coo_val, &H.coo.values[0], nc * sizeof(float));
    }
    return nc;
}

// y = A x for a host HYB matrix given as arrays (ELL column-major with
// pitch = rows rounded up to 32, padding column -1; COO sorted by row):
// cusp::multiply on cusp::hyb_matrix<int, float, host_memory>.
void ref_hyb_spmv(int64_t R, int64_t C, int64_t width, const int *ell_col, const float *ell_val, int64_t ncoo,
                  const int *coo_row, const int *coo_col, const float *coo_val, const float *x, float *y)
{
    cusp::hyb_matrix<int, float, cusp::host_memory> H;
    int64_t ne = 0;
    const int64_t pitch = (R + 31) / 32 * 32;
    for (int64_t t = 0; t < width * pitch; ++t) ne += ell_col[t] != -1;
    H.resize(R, C, ne, ncoo, width, 32);
    std::memcpy(&H.ell.column_indices.values[0], ell_col, width * pitch * sizeof(int));
    std::memcpy(&H.ell.values.values[0], ell_val, width * pitch * sizeof(float));
    if (ncoo) {
        std::memcpy(&H.coo.row_indices[0], coo_row, ncoo * sizeof(int));
        std::memcpy(&H.coo.column_indices[0], coo_col, ncoo * sizeof(int));
        std::memcpy(&H.coo.values[0], coo_val, ncoo * sizeof(float));
    }
    HostF xv(x, x + C), yv(R, 0.0f);
    cusp::multiply(H, xv, yv);
    std::memcpy(y, &yv[0], R * sizeof(float));
}

}  // extern "C"
