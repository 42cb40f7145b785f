/*
 * xJPCG C API -- drop-in replacement for the reference's pcg/sw/include/pcg.h,
 * implemented for one MI355X GPU by libxjpcg.so (HIP kernels for gfx950).
 *
 * Every type, enumerator value and function signature below matches the
 * reference header so that existing callers (pcgtest.cpp, examples/c/example.c,
 * the dlopen loader) build unchanged; each declaration cites the reference
 * line it replaces.  Behaviour follows pcg/sw/src/pcg.cpp:
 *   - x0 = 0 always; p_x is an output (written although declared const);
 *   - Jacobi uses the caller's p_diagA, not the matrix diagonal;
 *   - iteration stops when ||r||^2 < tol^2 ||b||^2 on the recursive residual,
 *     or after maxIter+1 iterations; *p_iter is capped at maxIter and
 *     *p_res = ||r|| / ||b|| of the last iteration;
 *   - *p_res > tol returns XJPCG_STATUS_EXECUTION_FAILED with outputs written.
 * Deliberate differences are listed with each function.
 */
#ifndef XILINX_PCG_H
#define XILINX_PCG_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

/* reference pcg.h:28-32: inline vs extern linkage for the loader */
#ifdef XILINX_PCG_INLINE
#define XILINX_PCG_LINKAGE_DECL inline
#else
#define XILINX_PCG_LINKAGE_DECL extern
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* reference pcg.h:41-49 */
typedef enum {
    XJPCG_STATUS_SUCCESS,              /* success */
    XJPCG_STATUS_NOT_INITIALIZED,      /* null handle */
    XJPCG_STATUS_ALLOC_FAILED,         /* device (HBM) allocation failure */
    XJPCG_STATUS_INVALID_VALUE,        /* invalid arguments */
    XJPCG_STATUS_EXECUTION_FAILED,     /* solver did not reach the tolerance */
    XJPCG_STATUS_INTERNAL_ERROR,       /* internal error, HIP runtime error, no GPU */
    XJPCG_STATUS_DYNAMIC_LOADING_ERROR /* the dlopen loader could not reach the library */
} XJPCG_Status_t;

/* reference pcg.h:56-57 -- name of a status code, e.g. "XJPCG_STATUS_SUCCESS" */
XILINX_PCG_LINKAGE_DECL
const char* xJPCG_getErrorString(const XJPCG_Status_t code);

/* reference pcg.h:62-67 -- wall-clock seconds of the last createHandle / solve phases */
typedef struct {
    double m_objInit; /* handle creation */
    double m_matProc; /* matrix ingest: COO / CSC-sym -> CSR, upload, tile plan */
    double m_vecProc; /* b and diagonal upload */
    double m_solver;  /* GPU solve + copy of x to the caller */
} XJPCG_Metric_t;

/* reference pcg.h:69-74 */
struct XJPCG_ObjectStruct;
typedef struct XJPCG_ObjectStruct XJPCG_Handle_t;

/* reference pcg.h:79-85 -- low nibble: matrix reuse; high nibble: index base */
typedef enum XJPCG_Mode_t {
    XJPCG_MODE_DEFAULT = 0x00,        /* new matrix */
    XJPCG_MODE_KEEP_NZ_LAYOUT = 0x01, /* same sparsity pattern, new values */
    XJPCG_MODE_KEEP_MATRIX = 0x02,    /* reuse the matrix of the previous call */
    XJPCG_MODE_C_INDEX = 0x00,        /* indices start at 0 */
    XJPCG_MODE_FORTRAN_INDEX = 0x10   /* indices start at 1 */
} XJPCG_Mode_t;

/* reference pcg.h:94-95.  xclbinPath is accepted for source compatibility:
 * a string holding only a decimal integer selects that HIP device, anything
 * else (including NULL) selects device 0.  The handle is returned even when
 * initialisation fails (reference pcg.cpp:38-40), so its message can be read;
 * no GPU visible => XJPCG_STATUS_INTERNAL_ERROR, and later solves on that
 * handle fail the same way after argument checks. */
XILINX_PCG_LINKAGE_DECL
XJPCG_Status_t xJPCG_createHandle(XJPCG_Handle_t** handle, const char* xclbinPath);

/* reference pcg.h:103-104 */
XILINX_PCG_LINKAGE_DECL
XJPCG_Status_t xJPCG_destroyHandle(XJPCG_Handle_t* handle);

/* reference pcg.h:163-177 -- A given as the lower or upper triangle plus the
 * full diagonal in CSC; p_nnz = (total_nnz + p_n)/2.  The first call on a
 * handle must use XJPCG_MODE_DEFAULT.  KEEP_NZ_LAYOUT refreshes the values
 * through the stored pattern and requires identical n and nnz (else
 * XJPCG_STATUS_INTERNAL_ERROR, as the reference's partitioner raises).
 * Limits: n and the expanded nnz below 2^31. */
XILINX_PCG_LINKAGE_DECL
XJPCG_Status_t xJPCG_cscSymSolver(XJPCG_Handle_t* handle,
                                  const int64_t p_n,
                                  const int64_t p_nnz,
                                  const int64_t* p_rowIdx,
                                  const int64_t* p_colptr,
                                  const double* p_data,
                                  const double* p_diagA,
                                  const double* p_b,
                                  const double* p_x,
                                  const int64_t p_maxIter,
                                  const double p_tol,
                                  int64_t* p_iter,
                                  double* p_res,
                                  const XJPCG_Mode_t mode);

/* reference pcg.h:201-215 -- full matrix in COO; duplicates are summed.  The
 * index base follows (mode & 0xf0) >> 4 as in the reference implementation
 * (pcg.cpp:139).  KEEP_NZ_LAYOUT may pass NULL index arrays.
 * Differences: b == 0 returns x = 0, iter = 0, res = 0 (the reference
 * divides 0/0); an out-of-range index returns XJPCG_STATUS_INVALID_VALUE. */
XILINX_PCG_LINKAGE_DECL
XJPCG_Status_t xJPCG_cooSolver(XJPCG_Handle_t* handle,
                               const uint32_t p_n,
                               const uint32_t p_nnz,
                               const uint32_t* p_rowIdx,
                               const uint32_t* p_colIdx,
                               const double* p_data,
                               const double* p_diagA,
                               const double* p_b,
                               const double* p_x,
                               const uint32_t p_maxIter,
                               const double p_tol,
                               uint32_t* p_iter,
                               double* p_res,
                               const XJPCG_Mode_t mode);

/* reference pcg.h:223-224 */
XILINX_PCG_LINKAGE_DECL
XJPCG_Status_t xJPCG_peekAtLastStatus(const XJPCG_Handle_t* handle);

/* reference pcg.h:240-241 -- pointer valid until the next call on the handle */
XILINX_PCG_LINKAGE_DECL
const char* xJPCG_getLastMessage(const XJPCG_Handle_t* handle);

/* reference pcg.h:253-254 */
XILINX_PCG_LINKAGE_DECL
XJPCG_Status_t xJPCG_getMetrics(const XJPCG_Handle_t* handle, XJPCG_Metric_t* metric);

#ifdef __cplusplus
}
#endif

#endif /* XILINX_PCG_H */
