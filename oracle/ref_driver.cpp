// Thin extern "C" driver over the reference's host-side C++ that compiles
// without FPGA tooling (see oracle/Makefile, target `ref`).  TEST
// INFRASTRUCTURE ONLY: it is built into oracle/_ref/libxjpcg_ref.so from the
// reference sources where they lie, and only tests/ load it, to pin the
// oracle's vector set-up, record codec and matrix-ingest restatements.
//
// Wrapped reference code:
//   GenCgVector::loadVec/init         pcg/sw/include/impl/cgVector.hpp:55-114
//   GenCgInstr::setInstr/updateInstr  pcg/sw/include/impl/cgVector.hpp:146-175
//   CGSolverInstr::store/load         pcg/hw/include/impl/cgInstr.hpp:82-156
//   SparseMatrix::loadCoo/loadCscSym  L2/sparse/include/sw/fp64/matrix_params.hpp:51-134
//   SpmPar::getCooDatFromCscSym       L2/sparse/include/sw/fp64/gen_signature.hpp:88-112
//   SpmPar::partitionCooMat / partitionCscSymMat / updateMat
//                                     L2/sparse/include/sw/fp64/gen_signature.hpp:41-69
//   Signature::store_rbParam/store_parParam/store_nnz/store_info
//                                     L2/sparse/include/sw/fp64/signature.hpp:557-573
#include <algorithm>
#include <array>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pcg.h"
// the standard headers above are complete before this point, so only the
// reference classes see it: getCooDatFromCscSym is private and the driver calls it directly
#define private public
#include "gen_signature.hpp"
#undef private
#include "impl/cgVector.hpp"

using namespace xilinx_apps::pcg;

extern "C" {

// r0/J/z0/p0 set-up; writes dot, rz, dimAligned and the J, z, p vectors (dimAligned entries each)
int ref_cg_init(uint32_t n, const double* b, const double* diag, double* dot, double* rz, uint32_t* dim_aligned,
                double* J, double* z, double* p) {
    try {
        GenCgVector<double, 4> v;
        v.loadVec(n, b, diag);
        v.init();
        *dot = v.getDot();
        *rz = v.getRz();
        *dim_aligned = v.getDimAligned();
        CgVector cv = v.getVec();
        size_t bytes = cv.vecBytes;
        if (J) std::memcpy(J, cv.h_jacobi, bytes);
        if (z) std::memcpy(z, cv.h_zk, bytes);
        if (p) std::memcpy(p, cv.h_pk, bytes);
        return 0;
    } catch (...) {
        return -1;
    }
}

// record 0 exactly as the host writes it before a run (maxIter is the API value; the host adds 1)
int ref_record0(uint32_t max_iter, uint32_t dim_aligned, double dot, double tol, double rz, uint8_t* out64) {
    GenCgInstr<double, 64> g;
    g.setInstr(max_iter + 1, dim_aligned, dot, tol, rz);
    g.updateInstr();
    CgInstr ci = g.getInstrPtr();
    std::memcpy(out64, ci.h_instr, 64);
    return (int)ci.h_instrBytes;
}

int ref_record_encode(uint32_t id, uint32_t iter, uint32_t vec_size, uint32_t max_iter, double tols, double res,
                      double rz, double alpha, double beta, uint64_t clock, uint8_t* out64) {
    xf::hpc::cg::CGSolverInstr<double> r;
    r.setID(id);
    r.setIter(iter);
    r.setVecSize(vec_size);
    r.setMaxIter(max_iter);
    r.setTols(tols);
    r.setRes(res);
    r.setRZ(rz);
    r.setAlpha(alpha);
    r.setBeta(beta);
    r.setClock(clock);
    xf::hpc::MemInstr<64> m;
    r.store(out64, m);
    return 0;
}

int ref_record_decode(const uint8_t* in64, uint32_t* u32x4, double* f64x5, uint64_t* clock) {
    xf::hpc::cg::CGSolverInstr<double> r;
    xf::hpc::MemInstr<64> m;
    r.load(in64, m);
    u32x4[0] = r.getID();
    u32x4[1] = r.getIter();
    u32x4[2] = r.getVecSize();
    u32x4[3] = r.getMaxIter();
    f64x5[0] = r.getTols();
    f64x5[1] = r.getRes();
    f64x5[2] = r.getRZ();
    f64x5[3] = r.getAlpha();
    f64x5[4] = r.getBeta();
    *clock = r.getClock();
    return 0;
}

// index lists after loadCoo (index base removed); returns 0, -1 on exception
int ref_load_coo(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base, uint32_t* row_out,
                 uint32_t* col_out) {
    try {
        xf::sparse::SparseMatrix s;
        s.loadCoo(n, n, nnz, row, col, base);
        std::memcpy(row_out, s.m_row_list.data(), sizeof(uint32_t) * nnz);
        std::memcpy(col_out, s.m_col_list.data(), sizeof(uint32_t) * nnz);
        return 0;
    } catch (...) {
        return -1;
    }
}

// full COO of a half-stored symmetric CSC (indices via loadCscSym, values via getCooDatFromCscSym)
int ref_cscsym_to_coo(uint32_t n, uint32_t nnz_full, const int64_t* row_idx, const int64_t* col_ptr,
                      const double* data, int base, uint32_t* row_out, uint32_t* col_out, double* val_out) {
    try {
        xf::sparse::SparseMatrix s;
        s.loadCscSym<int64_t>(n, nnz_full, row_idx, col_ptr, base);
        xf::sparse::SpmPar<double> par(4, 8, 16, 4096, 4096, 256);
        std::vector<double> v = par.getCooDatFromCscSym<int64_t>(n, nnz_full, row_idx, col_ptr, data, base);
        std::memcpy(row_out, s.m_row_list.data(), sizeof(uint32_t) * nnz_full);
        std::memcpy(col_out, s.m_col_list.data(), sizeof(uint32_t) * nnz_full);
        std::memcpy(val_out, v.data(), sizeof(double) * nnz_full);
        return 0;
    } catch (...) {
        return -1;
    }
}

// Signature files written by the reference partitioner with the PCG build's
// geometry (4 values per word, 8-word accumulation, 16 channels, 4096 x 4096
// partitions, 256-bit words) into dir.  val2, when not null, is applied with
// the value-only update (updateMat / updateCscSymMat) before the files are
// written.  Returns 0, -1 on exception (message in err, err_len bytes).
static void ref_store_sig(xf::sparse::SpmPar<double>& par, const std::string& dir) {
    xf::sparse::Signature& sg = par.m_sig;
    sg.store_rbParam(dir + "/rbParam.dat");
    sg.store_parParam(dir + "/parParam.dat");
    std::string names[16];
    for (int i = 0; i < 16; ++i) names[i] = dir + "/nnzVal_" + std::to_string(i) + ".dat";
    sg.store_nnz(names);
    sg.store_info(dir + "/info.dat");
}

static void ref_copy_err(const char* what, char* err, int err_len) {
    if (err && err_len > 0) {
        std::strncpy(err, what, (size_t)err_len - 1);
        err[err_len - 1] = 0;
    }
}

int ref_sig_coo(uint32_t m, uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, const double* val,
                const double* val2, int base, const char* dir, char* err, int err_len) {
    try {
        xf::sparse::SpmPar<double> par(4, 8, 16, 4096, 4096, 256);
        par.partitionCooMat(m, n, nnz, row, col, val, base);
        if (val2) {
            par.checkUpdateDim(m, n, nnz);
            par.updateMat(val2);
        }
        ref_store_sig(par, dir);
        return 0;
    } catch (const std::exception& e) {
        ref_copy_err(e.what(), err, err_len);
        return -1;
    }
}

int ref_sig_cscsym(uint32_t n, uint32_t nnz_full, const int64_t* row_idx, const int64_t* col_ptr, const double* val,
                   const double* val2, int base, const char* dir, char* err, int err_len) {
    try {
        xf::sparse::SpmPar<double> par(4, 8, 16, 4096, 4096, 256);
        par.partitionCscSymMat<int64_t>(n, nnz_full, row_idx, col_ptr, val, base);
        if (val2) par.updateCscSymMat<int64_t>(n, nnz_full, row_idx, col_ptr, val2, base);
        ref_store_sig(par, dir);
        return 0;
    } catch (const std::exception& e) {
        ref_copy_err(e.what(), err, err_len);
        return -1;
    }
}
}
