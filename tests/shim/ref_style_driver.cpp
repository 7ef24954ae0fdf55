// ref_style_driver.cpp -- a caller written the way the reference's drivers
// call its API (src/Fill_Test.inl:75-99, 141-170, 296-321;
// src/SPMM_Test.inl:63-88, 128-148), compiled against the C++ drop-in shim
// include/dcsr_matrix.hpp instead of the reference's headers.  Test
// infrastructure: tests/test_gpu_shim.py writes the inputs, runs it and
// compares what it writes with the golden fixtures.
//
// usage: ref_style_driver <dir>
//   reads  <dir>/meta.i32 = {R, nbatch, hyb_width},
//          in_{row,col}.i32, in_val.f32, b<k>_{row,col}.i32, b<k>_val.f32, x.f32
//   writes load_*, upd<k>_*, compact_* (CSR: _off/_col .i32, _val .f32),
//          y.f32 (DCSR SpMV), spgemm_* (C = A*A), ycsr.f32 (CSR SpMV),
//          hyb_{ell_col,rs}.i32 hyb_ell_val.f32 hyb_coo_{row,col}.i32
//          hyb_coo_val.f32 (HYB of the final matrix), yhyb.f32 (its SpMV),
//          hupd_* and yhupd.f32 (the HYB store after an append update)
#include <cstdio>
#include <string>
#include <vector>

#include "dcsr_matrix.hpp"

template <class T>
static std::vector<T> rd(const std::string &path)
{
    std::vector<T> v;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)n / sizeof(T));
    if (n && std::fread(v.data(), 1, (size_t)n, f) != (size_t)n) throw std::runtime_error("short read " + path);
    std::fclose(f);
    return v;
}

template <class T>
static void wr(const std::string &path, const std::vector<T> &v)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

typedef dcsr_matrix<int, float, device_memory, 4> DCSR;

static void dump_csr(const std::string &dir, const std::string &name, const DCSR &A)
{
    csr_matrix<int, float> C;
    device::LoadMatrix(C, A);
    std::vector<int> off = C.row_offsets.to_host(), col = C.column_indices.to_host();
    std::vector<float> val = C.values.to_host();
    col.resize(C.num_entries);
    val.resize(C.num_entries);
    wr(dir + "/" + name + "_off.i32", off);
    wr(dir + "/" + name + "_col.i32", col);
    wr(dir + "/" + name + "_val.f32", val);
}

static void dump_hyb(const std::string &dir, const std::string &name, const hyb_matrix<int, float> &H)
{
    const dcsr_hyb_view v = H.view();
    std::vector<int> ec((size_t)(v.width * v.pitch)), rs((size_t)v.rows + 1), cr((size_t)v.num_overflow),
        cc((size_t)v.num_overflow);
    std::vector<float> ev(ec.size()), cv(cr.size());
    dcsr::hip_check(hipMemcpy(ec.data(), v.ell_col, ec.size() * 4, hipMemcpyDeviceToHost), "copy");
    dcsr::hip_check(hipMemcpy(ev.data(), v.ell_val, ev.size() * 4, hipMemcpyDeviceToHost), "copy");
    dcsr::hip_check(hipMemcpy(rs.data(), v.row_sizes, rs.size() * 4, hipMemcpyDeviceToHost), "copy");
    if (!cr.empty()) {
        dcsr::hip_check(hipMemcpy(cr.data(), v.coo_row, cr.size() * 4, hipMemcpyDeviceToHost), "copy");
        dcsr::hip_check(hipMemcpy(cc.data(), v.coo_col, cc.size() * 4, hipMemcpyDeviceToHost), "copy");
        dcsr::hip_check(hipMemcpy(cv.data(), v.coo_val, cv.size() * 4, hipMemcpyDeviceToHost), "copy");
    }
    wr(dir + "/" + name + "_ell_col.i32", ec);
    wr(dir + "/" + name + "_ell_val.f32", ev);
    wr(dir + "/" + name + "_rs.i32", rs);
    wr(dir + "/" + name + "_coo_row.i32", cr);
    wr(dir + "/" + name + "_coo_col.i32", cc);
    wr(dir + "/" + name + "_coo_val.f32", cv);
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    try {
        const std::vector<int> meta = rd<int>(dir + "/meta.i32");
        const int R = meta.at(0), nbatch = meta.at(1), width = meta.at(2);
        dcsr::device_array<int> rows(rd<int>(dir + "/in_row.i32")), cols(rd<int>(dir + "/in_col.i32"));
        dcsr::device_array<float> vals(rd<float>(dir + "/in_val.f32"));

        // load (the drivers: resize, LoadMatrix, BinRows)
        DCSR A;
        A.resize(R, R, 16, 2.0f);
        device::LoadMatrix(A, rows, cols, vals, (int)rows.size());
        device::BinRows(A);
        dump_csr(dir, "load", A);

        // insert-or-add batches
        for (int k = 0; k < nbatch; ++k) {
            const std::string b = dir + "/b" + std::to_string(k);
            dcsr::device_array<int> br(rd<int>(b + "_row.i32")), bc(rd<int>(b + "_col.i32"));
            dcsr::device_array<float> bv(rd<float>(b + "_val.f32"));
            device::UpdateMatrix(A, br, bc, bv);
            dump_csr(dir, "upd" + std::to_string(k), A);
        }

        // SpMV on the DCSR matrix, and the CSR comparator on its export
        dcsr::device_array<float> x(rd<float>(dir + "/x.f32")), y((size_t)R);
        hipStream_t streams[8] = {};
        device::spmv(A, x, y, streams);
        wr(dir + "/y.f32", y.to_host());
        csr_matrix<int, float> Ac;
        device::LoadMatrix(Ac, A);
        dcsr::device_array<float> yc((size_t)R);
        device::spmv(Ac, x, yc);
        wr(dir + "/ycsr.f32", yc.to_host());

        // SpGEMM C = A * A (SPMM_Test's spmm), then defragmentation
        DCSR C;
        device::spmm(A, A, C, streams);
        dump_csr(dir, "spgemm", C);
        device::SortMatrixRow(A, streams);
        dump_csr(dir, "compact", A);

        // HYB: DCSR -> HYB, its SpMV, then an append update and SpMV again
        hyb_matrix<int, float> H;
        H.resize(R, R, 0, width);
        device::LoadMatrix(H, A);
        dump_hyb(dir, "hyb", H);
        device::spmv(H, x, y);
        wr(dir + "/yhyb.f32", y.to_host());
        {
            const std::string b = dir + "/b0";
            dcsr::device_array<int> br(rd<int>(b + "_row.i32")), bc(rd<int>(b + "_col.i32"));
            dcsr::device_array<float> bv(rd<float>(b + "_val.f32"));
            device::UpdateMatrix(H, br, bc, bv, (int)br.size());
        }
        dump_hyb(dir, "hupd", H);
        device::spmv(H, x, y);
        wr(dir + "/yhupd.f32", y.to_host());
        // and back: HYB -> DCSR
        DCSR B;
        B.resize(R, R, 16, 2.0f);
        device::LoadMatrix(B, H);
        dump_csr(dir, "hback", B);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "driver failed: %s\n", e.what());
        return 1;
    }
    std::printf("driver ok\n");
    return 0;
}
