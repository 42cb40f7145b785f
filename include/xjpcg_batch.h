/*
 * Batched JPCG (libxjpcg.so): many independent SPD systems of uneven size,
 * each with R right-hand sides, solved together on one GPU.
 *
 * No reference counterpart: the reference's control unit can carry several
 * tasks with per-task IDs and stop tests (pcg/hw/include/control.hpp:54-63,
 * :83-88) but the product builds it with CG_numTasks=1 (pcg/hw/Makefile:103).
 *
 * Layout in HBM (all device pointers):
 *   the S systems form one block-diagonal CSR: d_rowptr[N+1], d_col[nnz]
 *   (global column indices), d_val[nnz] (fp64 or fp32), optional split fp64
 *   diagonal d_dsplit[N] (the CSR then holds off-diagonal entries only);
 *   h_sys_row[S+1] (host) gives each system's row range;
 *   vectors are row-major [N][R]: the R right-hand sides of a row are contiguous.
 * Each (system, rhs) pair runs the reference's JPCG semantics (x0 = 0, Jacobi
 * with the caller's diagonal d_diag[N], stop when ||r||^2 < tol^2 ||b||^2 or
 * after maxIter+1 iterations) with its own alpha, beta and convergence mask;
 * finished pairs stop updating and finished systems drop out of the launches.
 */
#ifndef XJPCG_BATCH_H
#define XJPCG_BATCH_H

#include <stdint.h>

#include "pcg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xjpcg_batch xjpcg_batch_t;

/* message of the last failing batch call on this thread */
const char* xjpcg_batch_last_error(void);

/* nrhs must be 1, 2, 4, 8, 16 or 32.  h_rowptr (host copy of d_rowptr) is required. */
XJPCG_Status_t xjpcg_batch_create(xjpcg_batch_t** batch, int32_t nsys, int32_t nrhs, const int32_t* h_sys_row,
                                  const int32_t* h_rowptr, const int32_t* d_rowptr, const int32_t* d_col,
                                  const void* d_val, int val_f32, const double* d_dsplit, void* stream);
XJPCG_Status_t xjpcg_batch_destroy(xjpcg_batch_t* batch);

/* rows N, stored entries, tiles, rows per tile, threads per workgroup */
XJPCG_Status_t xjpcg_batch_info(const xjpcg_batch_t* batch, int32_t* nrows, int32_t* nnz, int32_t* ntiles,
                                int32_t* tile_rows, int32_t* block);

/* Launch configuration: tile_rows (rows per tile, 16..1024) and block (threads
 * per workgroup: 64, 128, 256 or 512); tune != 0 times the candidate
 * configurations on this batch's SpMM and keeps the fastest (the autotuner),
 * ignoring the other two arguments. */
XJPCG_Status_t xjpcg_batch_set_config(xjpcg_batch_t* batch, int32_t tile_rows, int32_t block, int32_t tune);
/* Storage of the SpMM: 1 = the block-diagonal CSR; 2 = per-system diagonals
 * (each system's at most 8 distinct offsets col - row, one value array per
 * offset with zeros where a row lacks it, a presence byte per row: every
 * load of a row goes out at once; the sums are those of the CSR, in column
 * order).  At creation the batch builds the diagonals when every system
 * allows them and keeps the faster SpMM; the autotuner tunes the layout with
 * the tile configuration.  layout = 0 restores that choice, < 0 only queries;
 * *h_active (may be NULL) receives the layout in use. */
XJPCG_Status_t xjpcg_batch_set_layout(xjpcg_batch_t* batch, int32_t layout, int32_t* h_active, void* stream);

/* Solve every (system, rhs) pair from x0 = 0.  d_b, d_x: [N][R].
 * d_iters (int32 [S][R], may be NULL) receives the iterations each pair ran,
 * d_res (fp64 [S][R], may be NULL) its final ||r||/||b|| (recursive residual).
 * *h_launched (may be NULL) = batched iterations launched. */
XJPCG_Status_t xjpcg_batch_solve(xjpcg_batch_t* batch, const double* d_diag, const double* d_b, double* d_x,
                                 uint32_t max_iter, double tol, int32_t* d_iters, double* d_res,
                                 int64_t* h_launched, void* stream);

/* Iteration loop of xjpcg_batch_solve: path 0 (auto) or 2 = graph-replayed
 * launches, five per batched iteration (SpMM + p.Ap partials, alpha per
 * pair, r update + r.r / r.z partials, beta / stop rule / masks per pair,
 * x / p update), kernel boundaries carrying the data between them; path 1 =
 * one persistent cooperative launch for all iterations, the same five steps
 * separated by grid barriers, with `grid` co-resident workgroups (0: the
 * largest grid that stays placeable, at most one workgroup per tile).  Both
 * produce the same bits; converged systems and pairs drop out through the
 * same per-system and per-pair masks.  A runtime that refuses the
 * cooperative launch gets path 2; a grid barrier that waits longer than 5 s
 * aborts the solve with XJPCG_STATUS_INTERNAL_ERROR.  path < 0 only queries:
 * *h_last (may be NULL) = the loop the last solve used (0 before any solve),
 * *h_grid (may be NULL) = its persistent grid.  The reference's control unit
 * runs its whole loop from one launch (pcg/sw/src/cgHost.cpp:447-471). */
XJPCG_Status_t xjpcg_batch_set_path(xjpcg_batch_t* batch, int32_t path, int32_t grid, int32_t* h_last,
                                    int32_t* h_grid);

/* Storage of the work vectors r, p and Ap of xjpcg_batch_solve: bits = 64
 * (the default) or 32.  With 32 they are stored in fp32 -- every value is
 * widened to fp64 before it is used; x, all arithmetic, the per-pair p.Ap,
 * r.r and r.z sums, alpha and beta stay fp64 -- so an iteration moves about
 * 40% fewer vector bytes, while the recursive residual keeps about seven
 * significant digits: meant for the inner batch of xjpcg_batch_solve_mixed,
 * whose fp64 outer loop (true residual, x += d) restores full accuracy.
 * bits <= 0 only queries; *h_bits (may be NULL) receives the setting.  No
 * counterpart in the reference. */
XJPCG_Status_t xjpcg_batch_set_vectors(xjpcg_batch_t* batch, int32_t bits, int32_t* h_bits);

/* Iterative refinement: `outer` (fp64-accurate values, same pattern and
 * sizes) computes the true residual r = b - A x in fp64; `inner` (typically
 * fp32 values) solves A d = r to max(inner_tol, 0.5 tol ||b||/||r||) for the
 * unfinished pair furthest from tol (capped at 0.5; at least 1e-7 when the
 * inner batch has fp32 work vectors), finished pairs masked
 * to a zero right-hand side; x += d; repeat until every pair
 * has ||r||/||b|| < tol or max_outer steps.  d_res (fp64 [S][R], may be NULL)
 * receives the final true relative residuals. */
XJPCG_Status_t xjpcg_batch_solve_mixed(xjpcg_batch_t* outer, xjpcg_batch_t* inner, const double* d_diag,
                                       const double* d_b, double* d_x, uint32_t max_outer, double tol,
                                       double inner_tol, uint32_t inner_max_iter, int32_t* d_iters, double* d_res,
                                       uint32_t* h_outer, int64_t* h_inner_launched, void* stream);

#ifdef __cplusplus
}
#endif
#endif
