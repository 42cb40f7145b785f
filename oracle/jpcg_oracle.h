/*
 * CPU oracle for the xJPCG solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's Jacobi-preconditioned CG semantics
 * (Xilinx/HPC pcg/ and the L1/L2 templates it instantiates).  Only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it; the
 * product library never links or calls it.
 *
 * Pins, per function (what checks each restatement; see tests/test_oracle.py):
 *   orc_coo_to_csr       ref SparseMatrix::loadCoo row/col lists (oracle/_ref), dense rebuild
 *   orc_cscsym_to_coo    ref SparseMatrix::loadCscSym + SpmPar::getCooDatFromCscSym (oracle/_ref), bit-exact
 *   orc_spmv             dense numpy products on tiny matrices (duplicates, empty rows),
 *                        closed-form stencil row sums A*1
 *   orc_dot_fixed        brute-force pairwise/tree re-association on small vectors (exact integers),
 *                        and agreement with a naive sum to rounding
 *   orc_cg_init          ref GenCgVector::loadVec+init (oracle/_ref), bit-exact
 *   orc_record_encode    ref CGSolverInstr::store/MemInstr (oracle/_ref), byte-exact
 *   orc_jpcg             dense Cholesky solutions (tests/golden fixtures), diagonal A => 1 iteration,
 *                        k distinct eigenvalues of D^-1 A => at most k iterations, manufactured
 *                        solutions, the maxIter+1 count convention on 2x2 systems worked by hand
 */
#ifndef JPCG_ORACLE_H
#define JPCG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 64-byte per-iteration record (pcg/hw/include/impl/cgInstr.hpp:117-132) */
typedef struct {
    uint32_t id, iter, vec_size, max_iter;
    double tols, res, rz, alpha, beta;
    uint64_t clock;
} orc_record_t;

int orc_coo_to_csr(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, const double* val, int base,
                   int32_t* rowptr, int32_t* col_out, double* val_out, int64_t* perm);
int64_t orc_cscsym_to_coo(int64_t n, int64_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                          const double* data, int base, uint32_t* row, uint32_t* col, double* val);
void orc_spmv(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* x, double* y);
void orc_spmv_omp(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* x,
                  double* y, int threads);
double orc_dot_fixed(int64_t n, const double* a, const double* b);
void orc_cg_init(uint32_t n, const double* b, const double* diag, double* dot, double* rz, uint32_t* dim_aligned);
void orc_record_encode(const orc_record_t* r, uint8_t out[64]);
void orc_record_decode(const uint8_t in[64], orc_record_t* r);
int orc_jpcg(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* diag,
             const double* b, uint32_t max_iter, double tol, double* x, uint32_t* iter, double* res,
             uint8_t* records, double* rr_hist, int threads);

#ifdef __cplusplus
}
#endif
#endif
