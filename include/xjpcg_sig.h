/*
 * Partitioned-matrix ("signature") files of the reference's SpMV engine
 * (libxjpcg.so): read matrices the reference's partitioner wrote, and write
 * the same files, byte for byte, from a COO or half-stored symmetric CSC.
 *
 * Files in one directory (reference L2/sparse/src/sw/c++/gen_signature.cpp:58-81):
 *   info.dat         6 x uint32: m, n, nnz, m, n rounded up to 4, padded entries
 *   rbParam.dat      header word + 8 words per row block (<= 4096 rows)
 *   parParam.dat     header word + 8 words per column partition (<= 4096 columns)
 *   nnzVal_<c>.dat   c = 0..15: header word, then per partition row-index,
 *                    column and value words (4 doubles per word)
 * (32-byte words; layout in matrix_params.hpp:329-688, signature.hpp:92-421.)
 *
 * Calls return an XJPCG_Status_t; xjpcg_sig_last_error() gives the message
 * of the last failing call on the thread.  Partitioner errors the reference
 * raises as SpmException return XJPCG_STATUS_INTERNAL_ERROR (as pcg.cpp:112-113
 * maps them); argument and file errors return XJPCG_STATUS_INVALID_VALUE.
 */
#ifndef XJPCG_SIG_H
#define XJPCG_SIG_H

#include <stdint.h>

#include "pcg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xjpcg_sig xjpcg_sig_t;

const char* xjpcg_sig_last_error(void);

/* Partition an m x n COO matrix (base 0 or 1; duplicates stay separate
 * entries, as the reference stores them).  Replaces
 * SpmPar<double>::partitionCooMat (L2/sparse/include/sw/fp64/gen_signature.hpp:41-47)
 * with the PCG build's geometry (4, 8, 16, 4096, 4096, 256). */
XJPCG_Status_t xjpcg_sig_from_coo(xjpcg_sig_t** sig, uint32_t m, uint32_t n, uint32_t nnz, const uint32_t* row,
                                  const uint32_t* col, const double* val, int base);
/* Partition a half-stored symmetric CSC with the diagonal (xJPCG_cscSymSolver's
 * input, pcg.h:129-133), expanded to 2*nnz_half - n entries in the reference's
 * order.  Replaces SpmPar<double>::partitionCscSymMat (gen_signature.hpp:49-56). */
XJPCG_Status_t xjpcg_sig_from_cscsym(xjpcg_sig_t** sig, uint32_t n, uint32_t nnz_half, const int64_t* row_idx,
                                     const int64_t* col_ptr, const double* val, int base);
/* New values, same pattern: rewrites only the value words.  len is the
 * length of the value array the partition was made from (nnz, or nnz_half
 * for CSC-sym).  Replaces SpmPar::updateMat / updateCscSymMat
 * (gen_signature.hpp:61-69, signature.hpp:427-469, :534-555).  Not
 * available for a signature read from files. */
XJPCG_Status_t xjpcg_sig_update(xjpcg_sig_t* sig, const double* val, uint64_t len);
/* Write the files into an existing directory (storeMatPar,
 * gen_signature.cpp:58-81); read them back, checking every size, count and
 * header against the parameter files. */
XJPCG_Status_t xjpcg_sig_store(const xjpcg_sig_t* sig, const char* dir);
XJPCG_Status_t xjpcg_sig_load(xjpcg_sig_t** sig, const char* dir);
XJPCG_Status_t xjpcg_sig_destroy(xjpcg_sig_t* sig);

/* info[6] = info.dat; *entries = entries xjpcg_sig_to_coo returns (non-zero
 * values); row blocks and partitions.  Any output may be NULL. */
XJPCG_Status_t xjpcg_sig_info(const xjpcg_sig_t* sig, uint32_t* info, uint64_t* entries, uint32_t* row_blocks,
                              uint32_t* partitions);
/* The raw buffers (the reference's MatPartition, matrix_params.hpp:703-712):
 * which = 0 rbParam, 1 parParam, 2 + c nnzVal_c.  The pointer stays valid
 * until the next update or destroy. */
XJPCG_Status_t xjpcg_sig_buffer(const xjpcg_sig_t* sig, int which, const void** ptr, uint64_t* bytes);

/* Decode to COO with 0-based indices, in stream order (channel, partition,
 * slot); cap = capacity of the arrays (>= entries).  Zero values are not
 * returned: the format does not tell them from padding (the reference's
 * decoder, L2/sparse/tests/fp64/spmv/python/signature.py:373-409, drops them
 * the same way). */
XJPCG_Status_t xjpcg_sig_to_coo(const xjpcg_sig_t* sig, uint32_t* row, uint32_t* col, double* val, uint64_t cap);
/* Build a device CSR of a square signature (m == n) on the GPU: the streams
 * are uploaded as they are and decoded there, then sorted into
 * d_rowptr[n+1], d_col[*h_entries], d_val[*h_entries] (fp64); rows sorted,
 * columns ascending in a row, duplicates adjacent in stream order, so the
 * result is the CSR xjpcg_csr_from_coo builds from the original COO (without
 * its zero values).  cap = capacity of d_col and d_val; info.dat's nnz
 * (info[2]) always suffices for files the reference wrote.  Synchronises
 * `stream`. */
XJPCG_Status_t xjpcg_sig_to_csr(const xjpcg_sig_t* sig, int32_t* d_rowptr, int32_t* d_col, double* d_val,
                                uint64_t cap, uint64_t* h_entries, void* stream);

#ifdef __cplusplus
}
#endif

#endif
