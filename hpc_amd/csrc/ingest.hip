// Matrix ingest on the GPU: host-validated COO, or half-stored symmetric CSC
// expanded to full COO, becomes a device CSR whose rows are sorted, whose
// columns ascend within a row and whose duplicate entries keep their input
// order (the SpMV then sums them, as the reference accumulates duplicates,
// pcg.h:189), plus the value map perm[k] = index of entry k's value in the
// caller's value array.  The order matches the host build of the pcg.h
// drop-in (stable bucket by row, stable sort by column): entries are keyed
// row*n + col and sorted with a stable LSD radix sort.
//
//   coo_keys     key = (row-base)*n + (col-base), idx = k
//   csc_cols     column of every half-stored entry (one thread per column)
//   csc_counts   1 for a diagonal entry, 2 for an off-diagonal one
//   csc_emit     (i,j) then (j,i) at the scanned output position, as the
//                reference's getCooDatFromCscSym emits them (gen_signature.hpp:88-112)
//   csr_build    rowptr / col / perm from the sorted keys
#include <hipcub/hipcub.hpp>

#include "ingest.hpp"
#include "kernels.hpp"

namespace xj {

namespace {

__global__ __launch_bounds__(kBlock) void coo_keys(int64_t m, uint32_t n, uint32_t off,
                                                   const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                                                   uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t i = row[k] - off, j = col[k] - off;
        key[k] = (uint64_t)i * n + j;
        idx[k] = (int32_t)k;
    }
}

__global__ __launch_bounds__(kBlock) void csc_cols(uint32_t n, int64_t off, const int64_t* __restrict__ cp,
                                                   uint32_t* __restrict__ colof) {
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n; j += (int64_t)gridDim.x * kBlock)
        for (int64_t k = cp[j] - off; k < cp[j + 1] - off; ++k) colof[k] = (uint32_t)j;
}

__global__ __launch_bounds__(kBlock) void csc_counts(int64_t m, int64_t off, const int64_t* __restrict__ ri,
                                                     const uint32_t* __restrict__ colof, uint32_t* __restrict__ cnt) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock)
        cnt[k] = (uint32_t)(ri[k] - off) != colof[k] ? 2u : 1u;
}

__global__ __launch_bounds__(kBlock) void csc_emit(int64_t m, uint32_t n, int64_t off, const int64_t* __restrict__ ri,
                                                   const uint32_t* __restrict__ colof,
                                                   const uint32_t* __restrict__ pos, uint64_t* __restrict__ key,
                                                   int32_t* __restrict__ idx) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock) {
        const uint32_t i = (uint32_t)(ri[k] - off), j = colof[k];
        const uint32_t p = pos[k];
        key[p] = (uint64_t)i * n + j;
        idx[p] = (int32_t)k;
        if (i != j) {
            key[p + 1] = (uint64_t)j * n + i;
            idx[p + 1] = (int32_t)k;
        }
    }
}

__global__ __launch_bounds__(kBlock) void csr_build(int64_t m, uint32_t n, const uint64_t* __restrict__ key,
                                                    const int32_t* __restrict__ idx, int32_t* __restrict__ rowptr,
                                                    int32_t* __restrict__ col, int64_t* __restrict__ perm) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock) {
        const uint64_t kk = key[k];
        const int64_t r = (int64_t)(kk / n);
        col[k] = (int32_t)(kk - (uint64_t)r * n);
        perm[k] = idx[k];
        // rows (previous entry's row, r] start at k; rows after the last entry start at m
        const int64_t rp = k == 0 ? -1 : (int64_t)(key[k - 1] / n);
        for (int64_t q = rp + 1; q <= r; ++q) rowptr[q] = (int32_t)k;
        if (k == m - 1)
            for (int64_t q = r + 1; q <= (int64_t)n; ++q) rowptr[q] = (int32_t)m;
    }
}

__global__ __launch_bounds__(kBlock) void fill_zero(int64_t m, int32_t* __restrict__ a) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock) a[k] = 0;
}

int grid_of(int64_t m) {
    int64_t g = (m + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

int key_bits(uint32_t n) {
    const uint64_t maxkey = (uint64_t)n * n - 1;
    int b = 1;
    while (b < 64 && (maxkey >> b)) ++b;
    return b;
}

}  // namespace

hipError_t ingest_sort_build(int64_t m, uint32_t n, uint64_t* key, int32_t* idx, uint64_t* key2, int32_t* idx2,
                             void* temp, size_t temp_bytes, int32_t* rowptr, int32_t* col, int64_t* perm,
                             hipStream_t s) {
    if (m == 0) {
        fill_zero<<<grid_of((int64_t)n + 1), kBlock, 0, s>>>((int64_t)n + 1, rowptr);
        return hipGetLastError();
    }
    size_t bytes = temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, bytes, key, key2, idx, idx2, (int)m, 0, key_bits(n), s);
    if (e != hipSuccess) return e;
    csr_build<<<grid_of(m), kBlock, 0, s>>>(m, n, key2, idx2, rowptr, col, perm);
    return hipGetLastError();
}

size_t ingest_sort_temp_bytes(int64_t m, uint32_t n) {
    size_t bytes = 0;
    if (m <= 0) return 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const int32_t*)nullptr, (int32_t*)nullptr, (int)m, 0, key_bits(n),
                                             (hipStream_t)0);
    return bytes;
}

size_t ingest_scan_temp_bytes(int64_t m) {
    size_t bytes = 0;
    if (m <= 0) return 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)m,
                                           (hipStream_t)0);
    return bytes;
}

hipError_t ingest_exclusive_sum(void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out, int64_t m,
                                hipStream_t s) {
    if (m <= 0) return hipSuccess;
    size_t bytes = temp_bytes;
    return hipcub::DeviceScan::ExclusiveSum(temp, bytes, in, out, (int)m, s);
}

hipError_t ingest_coo_keys(int64_t m, uint32_t n, uint32_t base, const uint32_t* row, const uint32_t* col,
                           uint64_t* key, int32_t* idx, hipStream_t s) {
    if (m == 0) return hipSuccess;
    coo_keys<<<grid_of(m), kBlock, 0, s>>>(m, n, base, row, col, key, idx);
    return hipGetLastError();
}

hipError_t ingest_csc_keys(int64_t m_half, uint32_t n, int64_t base, const int64_t* row_idx, const int64_t* col_ptr,
                           uint32_t* colof, uint32_t* cnt, uint32_t* pos, void* temp, size_t temp_bytes,
                           uint64_t* key, int32_t* idx, hipStream_t s) {
    if (m_half == 0) return hipSuccess;
    csc_cols<<<grid_of(n), kBlock, 0, s>>>(n, base, col_ptr, colof);
    csc_counts<<<grid_of(m_half), kBlock, 0, s>>>(m_half, base, row_idx, colof, cnt);
    size_t bytes = temp_bytes;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, bytes, cnt, pos, (int)m_half, s);
    if (e != hipSuccess) return e;
    csc_emit<<<grid_of(m_half), kBlock, 0, s>>>(m_half, n, base, row_idx, colof, pos, key, idx);
    return hipGetLastError();
}

}  // namespace xj
