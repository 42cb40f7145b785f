/*
 * Device-level C ABI of libxjpcg.so: the thin layer between host code (the
 * pcg.h engine, the Python package, the PyTorch autograd function) and the
 * HIP kernels.  Arguments are plain pointers into HBM (d_*), host pointers
 * (h_*) and sizes; `stream` is a hipStream_t passed as void* (NULL = the null
 * stream).  Calls return an XJPCG_Status_t value; xjpcg_dev_last_error()
 * describes the last failure on the calling thread.
 *
 * These entry points split the reference's PCGImpl host operations
 * (pcg/sw/include/impl/pcgImp.hpp:49-204) -- matrix load, vector load, run,
 * iteration / residual query, result retrieval -- over device buffers, and
 * add three forms with no reference counterpart: the mixed-precision refined
 * solve, the batched solve and the SDDMM gradient.
 *
 * Operators reference the caller's device memory (nothing is copied); the
 * caller keeps it alive until xjpcg_op_destroy.  Not thread-safe per operator.
 */
#ifndef XJPCG_DEV_H
#define XJPCG_DEV_H

#include <stdint.h>

#include "pcg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xjpcg_op xjpcg_op_t;

/* message of the last failing call on this thread ("" if none) */
const char* xjpcg_dev_last_error(void);

/* CSR operator A (n x n): d_rowptr[n+1], d_col[nnz] int32; d_val[nnz] fp64
 * (val_f32 = 0) or fp32 (val_f32 = 1).  With d_dsplit != NULL the CSR holds
 * only off-diagonal entries and d_dsplit[n] (fp64) is the diagonal.
 * h_rowptr (n+1 host entries) may be NULL; the row pointers are then read
 * back from the device to plan the row tiles.
 * Replaces the matrix-load step PCGImpl::setCooMat (pcgImp.hpp:49-70). */
XJPCG_Status_t xjpcg_op_create(xjpcg_op_t** op, int32_t n, int64_t nnz, const int32_t* d_rowptr,
                               const int32_t* d_col, const void* d_val, int val_f32, const double* d_dsplit,
                               const int32_t* h_rowptr, void* stream);
XJPCG_Status_t xjpcg_op_destroy(xjpcg_op_t* op);
/* Compact column stream: when every row tile's columns lie within 32767 of
 * the tile's first row, the operator keeps an int16 copy of the column
 * indices (col - first row of the tile; 2 bytes per entry instead of 4, in
 * the spirit of the reference's 16-bit local indices, SPARSE_indexBits=16 in
 * pcg/hw/Makefile:103) and its kernels read that stream.  on = 1 uses it when
 * it exists (the default), 0 uses the int32 stream, -1 only queries;
 * *h_active (may be NULL) tells which stream is in use. */
XJPCG_Status_t xjpcg_op_set_compact(xjpcg_op_t* op, int on, int* h_active);
/* Rebind the operator to another device value array of the same type and
 * length; the pattern and row tiles are kept (the device-side counterpart of
 * the reference's KEEP_NZ_LAYOUT value refresh, signature.hpp:427-469).
 * Call it also, with the same pointer, after changing values in place: an
 * operator on the panel layout keeps its own copy of the values, refreshed
 * on the stream of its next use. */
XJPCG_Status_t xjpcg_op_set_values(xjpcg_op_t* op, const void* d_val);
/* SpMV layout (no counterpart in the reference, whose SpMV streams its own
 * partitioned format from HBM): 1 = row tiles over the CSR arrays (int16 or
 * int32 column stream, see xjpcg_op_set_compact); rows of up to 64 entries
 * are summed left to right by one lane, longer rows by a wavefront in a fixed
 * order (strided lane sums, xor butterfly), rows above a tile's 2044 entries
 * by the workgroup;
 * 2 = panel-swept copy for matrices whose columns scatter over x (row blocks
 * of up to 4096 rows keep their sums in LDS and sweep 65536-column panels of
 * x in order, so the gathered window stays in L2; 12 bytes per entry plus an
 * 8-byte value map; a row's run inside a staged chunk is added left to right
 * when shorter than 15 entries, as one wavefront total otherwise);
 * 3 = row tiles on column-offset patterns (stencil operators: each tile's at
 * most 32 distinct offsets col - row plus a 32-bit mask per row replace the
 * column stream, 4 bytes per row instead of 2-4 per entry; same sums as 1);
 * 4 = diagonals (one value array per global offset col - row with zeros where
 * a row has no entry, a 32-bit presence mask per row; each row's entries are
 * summed in column order, two rows per lane for fp64 values from 327,680 rows,
 * one row per lane otherwise: same sums as 1).
 * At creation the operator times the layouts that apply and keeps the fastest
 * (panels are built for operators of at least 2^18 rows whose columns do not
 * fit the int16 stream); layout = 0 restores
 * that choice, layout < 0 only queries.  Layout 2 needs each row's columns in
 * ascending order; layout 3 needs strictly ascending columns, rows of at most
 * 32 entries and at most 32 offsets per tile; layout 4 strictly ascending
 * columns, at most 32 offsets over the whole matrix and diagonals holding at
 * most twice the stored entries.  *h_active (may be NULL) receives the layout
 * in use.  Building a layout synchronises `stream`. */
XJPCG_Status_t xjpcg_op_set_layout(xjpcg_op_t* op, int layout, void* stream, int* h_active);
/* The layout the solves run on.  It follows from the matrix, not from
 * timing, so that operators built from the same matrix give bit-identical
 * iterates: diagonals when they apply with at most 25% padding slots, else
 * the panels when they were built (the rule above), else the tiles or the pattern tiles
 * (the faster SpMV of the two; both give the same iterates).  It can differ
 * from the standalone SpMV's layout.  xjpcg_op_set_layout with layout > 0
 * sets both; layout = 0 restores both creation-time choices. */
XJPCG_Status_t xjpcg_op_solve_layout(const xjpcg_op_t* op, int* h_layout);
/* Work vectors of the operator's solves: bits = 64 (default) or 32, which
 * stores r, p, A p and the Jacobi factors 1/diag in fp32 while x, the
 * arithmetic and the dot products stay fp64 (meant for the inner operator of
 * xjpcg_solve_mixed; halves the vector traffic of an iteration).  32 needs the
 * diagonal or the panel solve layout (INVALID_VALUE otherwise) and runs the
 * multi-kernel loop; while another solve layout is set later, the solves use
 * fp64 vectors.
 * bits = 0 queries.  *h_bits (may be NULL) receives the bits in effect.
 * Extension of the device API: the reference's solver keeps all vectors in
 * fp64 (pcg/sw/include/pcgImp.hpp:86-120). */
XJPCG_Status_t xjpcg_op_set_vectors(xjpcg_op_t* op, int bits, int* h_bits);
/* Host-only view of the panel layout an operator would build (no device
 * calls): xjpcg_panel_dims gives the rows per block, the row blocks and the
 * column panels for n rows; xjpcg_panel_plan fills seg[nb*P+1], off[nnz],
 * lrow[nnz] and perm[nnz] (layout entry q holds CSR entry perm[q]) from a host
 * CSR.  lrow[q] & 0xFFF is the entry's row offset in its block; lrow[q] >> 12
 * is, on the first entry of a row's run inside each 2048-entry chunk of a
 * segment, min(run length, 15), else 0.  *h_ordered = 0 when some row's
 * columns are not in ascending panel order.  No counterpart in the
 * reference; for inspection and tests. */
XJPCG_Status_t xjpcg_panel_dims(int32_t n, int32_t* block_rows, int32_t* nb, int32_t* npanels);
XJPCG_Status_t xjpcg_panel_plan(int32_t n, const int32_t* h_rowptr, const int32_t* h_col, int32_t* seg,
                                uint16_t* off, uint16_t* lrow, int64_t* perm, int* h_ordered);
/* Host-only view of the row tiles and their column-offset patterns:
 * xjpcg_tile_plan writes the first row of each tile (tiles[ntiles+1], at
 * most n+1 entries) and *h_ntiles; xjpcg_pattern_plan fills mask[n] and
 * offs[ntiles*32] for that plan and sets *h_ok = 0 when layout 3 does not
 * apply.  No counterpart in the reference; for inspection and tests. */
XJPCG_Status_t xjpcg_tile_plan(int32_t n, const int32_t* h_rowptr, int32_t* tiles, int32_t* h_ntiles);
XJPCG_Status_t xjpcg_pattern_plan(int32_t n, const int32_t* h_rowptr, const int32_t* h_col, uint32_t* mask,
                                  int32_t* offs, int* h_ok);
/* n, nnz, number of row tiles, grid of the persistent kernel */
XJPCG_Status_t xjpcg_op_info(const xjpcg_op_t* op, int32_t* n, int64_t* nnz, int32_t* ntiles, int32_t* grid);
/* Bytes one SpMV launch on `layout` (0 = the layout in use, 1-4 as for
 * xjpcg_op_set_layout) must move: the layout's matrix streams including its
 * zero padding (diagonals), its per-row tables (row pointers, masks,
 * segment table), x read once, y written once and the split diagonal read
 * once.  *h_bytes = -1 when that layout has not been built.  Divided by a
 * launch's duration it gives the SpMV's HBM rate on the bytes its layout
 * needs (no counterpart in the reference). */
XJPCG_Status_t xjpcg_op_spmv_bytes(const xjpcg_op_t* op, int layout, int64_t* h_bytes);

/* Loop used by xjpcg_solve on this operator: 0 = automatic (the faster
 * measured loop), 1 = one persistent cooperative launch per solve with grid
 * barriers, 2 = three launches per iteration (SpMV with the p.Ap partials,
 * residual update, x / p update) with the loop state on the device.  A
 * runtime that refuses the cooperative launch gets loop 2.  *h_last (may be
 * NULL) receives the loop the last solve used; path < 0 only queries. */
XJPCG_Status_t xjpcg_op_set_path(xjpcg_op_t* op, int path, int* h_last);

/* Matrix ingest on the GPU (the reference does this on the host:
 * matrix_params.hpp:51-134 loadCoo / loadCscSym, gen_signature.hpp:72-112).
 * From the caller's host COO (row, col with index base 0 or 1) build a device
 * CSR whose rows are sorted, whose columns ascend within a row and whose
 * duplicate entries keep their input order (the SpMV sums them, pcg.h:189),
 * plus d_perm[k] = index of entry k's value in the caller's value array, so
 * values can be placed on the device.  d_rowptr[n+1], d_col[nnz],
 * d_perm[nnz].  The indices are checked on the host first with the checks
 * and messages of xJPCG_cooSolver; synchronises `stream`. */
XJPCG_Status_t xjpcg_csr_from_coo(uint32_t n, uint32_t nnz, const uint32_t* h_row, const uint32_t* h_col, int base,
                                  int32_t* d_rowptr, int32_t* d_col, int64_t* d_perm, void* stream);
/* The same from a half-stored symmetric CSC (xJPCG_cscSymSolver's input,
 * pcg.h:129-133), expanded to 2*nnz_half - n entries (*h_nnz), which d_col
 * and d_perm must hold; off-diagonal entries appear twice with the same value
 * index. */
XJPCG_Status_t xjpcg_csr_from_cscsym(uint32_t n, uint32_t nnz_half, const int64_t* h_row_idx,
                                     const int64_t* h_col_ptr, int base, int32_t* d_rowptr, int32_t* d_col,
                                     int64_t* d_perm, int64_t* h_nnz, void* stream);

/* d_dst = d_src for n doubles (16-byte aligned), with 16-byte non-temporal
 * loads and stores: the stream kernel that gives the achievable HBM rate the
 * SpMV is compared with.  No counterpart in the reference. */
XJPCG_Status_t xjpcg_stream_copy(const double* d_src, double* d_dst, int64_t n, void* stream);

/* d_y = A d_x  (the reference's SpMV compute units, L2/sparse/src/hw/fp64) */
XJPCG_Status_t xjpcg_spmv(xjpcg_op_t* op, const double* d_x, double* d_y, void* stream);

/* d_r = d_b - A d_x and *h_rr = r.r (synchronises the stream) */
XJPCG_Status_t xjpcg_residual(xjpcg_op_t* op, const double* d_b, const double* d_x, double* d_r, double* h_rr,
                              void* stream);

/* Jacobi-preconditioned CG from x0 = 0 in one persistent launch, semantics of
 * the reference device loop and result scan (control.hpp:65-91,
 * pcgImp.hpp:157-183, pcg.cpp:159-160): at most maxIter+1 iterations,
 * *h_iter = min(executed, maxIter), *h_res = ||r||/||b||; *h_executed (may be
 * NULL) = iterations actually run.  d_records, if not NULL, receives
 * (maxIter+2) 64-byte records laid out as the reference's CGSolverInstr
 * (cgInstr.hpp:117-132): record 0 holds the set-up, record k iteration k.
 * Returns SUCCESS whether or not the tolerance was met (compare *h_res). */
XJPCG_Status_t xjpcg_solve(xjpcg_op_t* op, const double* d_diag, const double* d_b, double* d_x, uint32_t max_iter,
                           double tol, uint32_t* h_iter, double* h_res, int64_t* h_executed, void* d_records,
                           void* stream);

/* d_g[k] = -d_lam[row(k)] * d_x[col(k)] for every stored entry k: the
 * gradient of a loss through x = A^-1 b with respect to A's stored values,
 * given the adjoint solution lam = A^-T dL/dx. */
XJPCG_Status_t xjpcg_sddmm(xjpcg_op_t* op, const double* d_lam, const double* d_x, double* d_g, void* stream);

/* Mixed-precision iterative refinement.
 *   outer: fp64-accurate operator for the true residual r = b - A x (fp64
 *          values, or fp32 off-diagonal values with an fp64 split diagonal
 *          when those values are exact in fp32);
 *   inner: fp32-valued operator solved by the persistent JPCG (fp64 vectors,
 *          fp64 accumulation, fp64 alpha/beta dot products).
 * Loop: r = b - A x (fp64); stop when ||r||/||b|| < tol; solve inner d = r to
 * the relative tolerance max(inner_tol, 0.5 tol ||b|| / ||r||), capped at
 * 0.5 (at most inner_max_iter iterations); x += d.  inner_tol = 0 lets one
 * step finish when the inner operator equals A; a larger inner_tol bounds
 * the work per step when it only approximates A.
 * d_x receives the solution.  Counts and the final true residual are returned
 * through the h_ pointers (any may be NULL).  Returns EXECUTION_FAILED when
 * tol is not reached within max_outer refinement steps. */
XJPCG_Status_t xjpcg_solve_mixed(xjpcg_op_t* outer, xjpcg_op_t* inner, const double* d_diag, const double* d_b,
                                 double* d_x, uint32_t max_outer, double tol, double inner_tol,
                                 uint32_t inner_max_iter, uint32_t* h_outer, int64_t* h_inner_total,
                                 double* h_res, void* stream);

/* Rows of the static workspace of xjpcg_csr_solve_static. */
#define XJPCG_STATIC_MAX_ROWS 262144

/* Allocation-free Jacobi PCG on the caller's device CSR (fp64 values, int32
 * indices, the diagonal stored in the rows), with the semantics and outputs
 * of xjpcg_solve: x0 = 0, at most maxIter+1 iterations, *h_iter =
 * min(executed, maxIter), *h_res = ||r||/||b||, SUCCESS whether or not tol
 * was met.  No counterpart in the reference: for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc).
 * It creates no operator, event, graph or device buffer: the work vectors
 * (r, p, Ap, 1/diag), the partial sums and the loop state live in one
 * static device array per GPU of 8.4 MB, sized for n <= XJPCG_STATIC_MAX_ROWS;
 * larger n returns INVALID_VALUE.  The SpMV sums each row left to right on
 * one lane (no row-tile plan), so iterates agree with xjpcg_solve on the
 * same matrix within rounding of the p.Ap sum.  Synchronises `stream` every
 * 16 iterations to read the stop flag.  Calls that overlap on one GPU share
 * the array and race; the last call's contents stay in it. */
XJPCG_Status_t xjpcg_csr_solve_static(int32_t n, int64_t nnz, const int32_t* d_rowptr, const int32_t* d_col,
                                      const double* d_val, const double* d_diag, const double* d_b, double* d_x,
                                      uint32_t max_iter, double tol, uint32_t* h_iter, double* h_res,
                                      int64_t* h_executed, void* stream);

#ifdef __cplusplus
}
#endif
#endif
