/*
 * Extensions of the pcg.h handle API (libxjpcg.so).
 */
#ifndef XJPCG_EXT_H
#define XJPCG_EXT_H

#include <stdint.h>

#include "pcg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-iteration history of the last solve on `handle`: 64-byte records laid
 * out as the reference's CGSolverInstr (pcg/hw/include/impl/cgInstr.hpp:117-132,
 * little-endian u32 ID, Iter, VecSize, MaxIter; f64 Tols, Res, RZ, Alpha,
 * Beta; u64 Clock).  Record 0 is the set-up record the reference host writes
 * (cgVector.hpp:149-163); record k is iteration k as krnl_control stores it
 * (control.hpp:75-90), with Clock from the GPU's 100 MHz wall clock.
 * The reference reads these records back but exposes only the final residual
 * and count (pcgImp.hpp:171-180).  Copies min(count, capacity) records to
 * `records` (may be NULL to query) and sets *count = executed iterations + 1;
 * *count = 0 when the last solve kept no history (maxIter + 2 > 2^20). */
XJPCG_Status_t xJPCG_getHistory(const XJPCG_Handle_t* handle, void* records, uint64_t capacity, uint64_t* count);

/* The steps of one solve as separate calls (the reference's PCGImpl
 * methods, pcgImp.hpp:49-204, which xJPCG_cooSolver / xJPCG_cscSymSolver
 * chain): load the matrix, load the vectors, solve, query the iteration
 * count and residual, retrieve x.  Same checks, statuses, messages and
 * metrics as the solver calls; x stays on the GPU until xJPCG_getX.
 *
 * Matrix load: `mode & 0x0f` = XJPCG_MODE_DEFAULT loads a new matrix
 * (PCGImpl::setCooMat / setCscSymMat, pcgImp.hpp:49-94);
 * XJPCG_MODE_KEEP_NZ_LAYOUT refreshes only the values of the loaded pattern
 * (row/col arrays may be NULL; PCGImpl::updateMat / updateCscSymMat,
 * pcgImp.hpp:95-145); XJPCG_MODE_KEEP_MATRIX keeps everything.
 * `mode & 0xf0` = XJPCG_MODE_FORTRAN_INDEX for 1-based indices. */
XJPCG_Status_t xJPCG_setCooMat(XJPCG_Handle_t* handle, uint32_t n, uint32_t nnz, const uint32_t* rowIdx,
                               const uint32_t* colIdx, const double* data, XJPCG_Mode_t mode);
XJPCG_Status_t xJPCG_setCscSymMat(XJPCG_Handle_t* handle, int64_t n, int64_t nnz, const int64_t* rowIdx,
                                  const int64_t* colPtr, const double* data, XJPCG_Mode_t mode);
/* Vector load (PCGImpl::setVec, pcgImp.hpp:147-155): b and the Jacobi
 * diagonal diagA, n entries each, n equal to the loaded matrix's. */
XJPCG_Status_t xJPCG_setVec(XJPCG_Handle_t* handle, uint32_t n, const double* b, const double* diagA);
/* Solve from x0 = 0 (PCGImpl::run, pcgImp.hpp:157-183).  *iter / *res (may
 * be NULL) get the reported count and the recursive relative residual;
 * XJPCG_STATUS_EXECUTION_FAILED when res > tol, results kept (pcg.cpp:159-166). */
XJPCG_Status_t xJPCG_run(XJPCG_Handle_t* handle, uint32_t maxIter, double tol, uint32_t* iter, double* res);
/* Iteration count and residual of the last xJPCG_run or solver call
 * (Results::m_nIters / m_residual, pcgImp.hpp:30-35). */
XJPCG_Status_t xJPCG_getResults(const XJPCG_Handle_t* handle, uint32_t* iter, double* res);
/* x of the last solve into x[n] (PCGImpl::getRes, pcgImp.hpp:200-213). */
XJPCG_Status_t xJPCG_getX(const XJPCG_Handle_t* handle, double* x, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
