// Matrix ingest: COO or half-stored symmetric CSC from the caller's host
// arrays to a CSR with a value map, on the host (small inputs, unusual
// column-pointer layouts) or on the GPU (ingest.hip).  Shared by the pcg.h
// drop-in (pcg_api.cpp) and the device-level C ABI (dev_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "operator.hpp"

namespace xj {

// Errors the reference's partitioner raises as SpmException (not a
// CgException), which pcg.cpp:112-113 maps to XJPCG_STATUS_INTERNAL_ERROR.
inline Error spm_invalid(const std::string& m) {
    return Error(XJPCG_STATUS_INTERNAL_ERROR, "INVALID VALUE ERROR: " + m);
}
inline Error spm_internal(const std::string& m) { return Error(XJPCG_STATUS_INTERNAL_ERROR, "INTERNAL ERROR: " + m); }
inline Error spm_not_supported(const std::string& m) {
    return Error(XJPCG_STATUS_INTERNAL_ERROR, "Matrix Operation ERROR: " + m);
}

// Host CSR pattern with a map from CSR slot to the caller's value array.
struct Pattern {
    uint32_t n = 0;
    int64_t nnz = 0;       // stored entries after expansion
    int64_t src_len = 0;   // length of the caller's value array
    std::vector<int32_t> rowptr, col;
    std::vector<int64_t> perm;
};

// Host builds (every check and message of the reference's loaders, in their order).
void pattern_from_coo(Pattern& P, uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base);
void pattern_from_cscsym(Pattern& P, uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                         int base);

// The same checks without building anything.  validate_cscsym returns the
// expanded count and sets *standard when col_ptr runs from 0 to nnz_half
// without decreasing (each stored entry then belongs to exactly one column,
// the layout the GPU expansion handles).
void validate_coo(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base);
uint32_t validate_cscsym(uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr, int base,
                         bool* standard);

// Device builds into caller device arrays d_rowptr[n+1], d_col[m], d_perm[m]
// (m = nnz, or the expanded count for CSC-sym); validate first (unless the
// caller already ran validate_coo / validate_cscsym on the same arrays:
// validated = true), then build on the GPU (CSC-sym with a non-standard
// col_ptr: on the host, then upload).  The CSR equals the host build's.
// Synchronises `s`.
void csr_from_coo_device(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base,
                         int32_t* d_rowptr, int32_t* d_col, int64_t* d_perm, hipStream_t s, bool validated = false);
uint32_t csr_from_cscsym_device(uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                                int base, int32_t* d_rowptr, int32_t* d_col, int64_t* d_perm, hipStream_t s,
                                bool validated = false);

// kernels and hipCUB passes (ingest.hip)
hipError_t ingest_coo_keys(int64_t m, uint32_t n, uint32_t base, const uint32_t* row, const uint32_t* col,
                           uint64_t* key, int32_t* idx, hipStream_t s);
hipError_t ingest_csc_keys(int64_t m_half, uint32_t n, int64_t base, const int64_t* row_idx, const int64_t* col_ptr,
                           uint32_t* colof, uint32_t* cnt, uint32_t* pos, void* temp, size_t temp_bytes,
                           uint64_t* key, int32_t* idx, hipStream_t s);
hipError_t ingest_sort_build(int64_t m, uint32_t n, uint64_t* key, int32_t* idx, uint64_t* key2, int32_t* idx2,
                             void* temp, size_t temp_bytes, int32_t* rowptr, int32_t* col, int64_t* perm,
                             hipStream_t s);
size_t ingest_sort_temp_bytes(int64_t m, uint32_t n);
size_t ingest_scan_temp_bytes(int64_t m);
// out = exclusive prefix sum of in[0..m) (temp sized by ingest_scan_temp_bytes)
hipError_t ingest_exclusive_sum(void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out, int64_t m,
                                hipStream_t s);

}  // namespace xj
