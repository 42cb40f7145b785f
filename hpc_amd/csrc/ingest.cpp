// Matrix ingest on the host and on the GPU (see ingest.hpp).
#include "ingest.hpp"
#include "staging.hpp"

#include <algorithm>
#include <thread>

namespace xj {

namespace {

// COO entries in input order (reference matrix_params.hpp:51-79 loadCoo):
// emit(i, j, k) for every entry after its index check.
template <typename Emit>
void coo_walk(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base, Emit emit) {
    if (base != 0 && base != 1)
        throw spm_not_supported("from getCooDatFromCscSym in gen_signature.hpp, unsupported storage mode.");
    if ((int64_t)nnz >= (int64_t)INT32_MAX) throw invalid("number of non-zeros must be below 2^31.");
    const uint32_t off = (uint32_t)base;
    for (uint32_t k = 0; k < nnz; ++k) {
        const uint32_t i = row[k] - off, j = col[k] - off;
        if (i >= n || j >= n) throw invalid("matrix index out of range at entry " + std::to_string(k) + ".");
        emit(i, j, (int64_t)k);
    }
}

// Half-stored symmetric CSC expanded in column order (reference
// matrix_params.hpp:81-134 loadCscSym, gen_signature.hpp:88-112
// getCooDatFromCscSym): each off-diagonal entry is emitted twice with the same
// value index; the expansion must produce exactly 2*nnz_half - n entries.
template <typename Emit>
uint32_t cscsym_walk(uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr, int base,
                     Emit emit) {
    if (base != 0 && base != 1)
        throw spm_not_supported("from getCooDatFromCscSym in gen_signature.hpp, unsupported storage mode.");
    const int64_t off = base;
    if (base == 1 && col_ptr[0] < 1)
        throw spm_invalid(
            "from getCooDatFromCscSym in gen_signature.hpp, colPtr[0] start from 1 in FORTRAN storage mode.");
    const uint32_t full = nnz_half * 2u - n;  // uint32 arithmetic as pcgImp.hpp:87
    if ((int64_t)full >= (int64_t)INT32_MAX) throw invalid("number of non-zeros must be below 2^31.");
    uint32_t count = 0;
    for (uint32_t j = 0; j < n; ++j) {
        for (int64_t k = col_ptr[j] - off; k < col_ptr[j + 1] - off; ++k) {
            if (k < 0 || k >= (int64_t)nnz_half) throw spm_internal("from loadCscSym in matrix_params.hpp, index >= nnz.");
            const int64_t i = row_idx[k] - off;
            if (i < 0 || i >= (int64_t)n) throw invalid("matrix index out of range at entry " + std::to_string(k) + ".");
            if (count >= full) throw spm_internal("from loadCscSym in matrix_params.hpp, index >= nnz.");
            emit((uint32_t)i, j, k);
            ++count;
            if (i != (int64_t)j) {
                if (count >= full) throw spm_internal("from loadCscSym in matrix_params.hpp, index >= nnz.");
                emit(j, (uint32_t)i, k);
                ++count;
            }
        }
    }
    if (count != full)
        throw spm_internal(
            "from loadCscSym in matrix_params.hpp, failed to convert indices from CscSym to COO format.");
    return full;
}

// Bucket (row, col, src) triples by row, stably, then order each row by
// column (ties keep input order, so duplicates stay adjacent and are summed
// by the SpMV like the reference accumulates them).
void finish_pattern(Pattern& P, const std::vector<uint32_t>& r, const std::vector<uint32_t>& c,
                    const std::vector<int64_t>& src) {
    const uint32_t n = P.n;
    const int64_t m = (int64_t)r.size();
    P.nnz = m;
    P.rowptr.assign((size_t)n + 1, 0);
    for (int64_t k = 0; k < m; ++k) P.rowptr[r[k] + 1]++;
    for (uint32_t i = 0; i < n; ++i) P.rowptr[i + 1] += P.rowptr[i];
    P.col.resize(m);
    P.perm.resize(m);
    {
        std::vector<int64_t> pos(P.rowptr.begin(), P.rowptr.end() - 1);
        for (int64_t k = 0; k < m; ++k) {
            int64_t p = pos[r[k]]++;
            P.col[p] = (int32_t)c[k];
            P.perm[p] = src[k];
        }
    }
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (m < (1 << 20)) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            std::vector<std::pair<int32_t, int64_t>> tmp;
            for (uint32_t i = (uint32_t)((uint64_t)n * t / nt); i < (uint32_t)((uint64_t)n * (t + 1) / nt); ++i) {
                const int32_t a = P.rowptr[i], b = P.rowptr[i + 1];
                bool sorted = true;
                for (int32_t k = a + 1; k < b; ++k)
                    if (P.col[k] < P.col[k - 1]) {
                        sorted = false;
                        break;
                    }
                if (sorted) continue;
                tmp.clear();
                for (int32_t k = a; k < b; ++k) tmp.emplace_back(P.col[k], P.perm[k]);
                std::stable_sort(tmp.begin(), tmp.end(),
                                 [](const std::pair<int32_t, int64_t>& x, const std::pair<int32_t, int64_t>& y) {
                                     return x.first < y.first;
                                 });
                for (int32_t k = a; k < b; ++k) {
                    P.col[k] = tmp[k - a].first;
                    P.perm[k] = tmp[k - a].second;
                }
            }
        });
    }
    for (auto& x : th) x.join();
}

}  // namespace

void pattern_from_coo(Pattern& P, uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base) {
    std::vector<uint32_t> r, c;
    std::vector<int64_t> src;
    r.reserve(nnz);
    c.reserve(nnz);
    src.reserve(nnz);
    coo_walk(n, nnz, row, col, base, [&](uint32_t i, uint32_t j, int64_t k) {
        r.push_back(i);
        c.push_back(j);
        src.push_back(k);
    });
    P.n = n;
    P.src_len = nnz;
    finish_pattern(P, r, c, src);
}

void pattern_from_cscsym(Pattern& P, uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                         int base) {
    std::vector<uint32_t> r, c;
    std::vector<int64_t> src;
    const uint32_t expect = nnz_half * 2u - n;
    if ((int64_t)expect < (int64_t)INT32_MAX) {
        r.reserve(expect);
        c.reserve(expect);
        src.reserve(expect);
    }
    cscsym_walk(n, nnz_half, row_idx, col_ptr, base, [&](uint32_t i, uint32_t j, int64_t k) {
        r.push_back(i);
        c.push_back(j);
        src.push_back(k);
    });
    P.n = n;
    P.src_len = nnz_half;
    finish_pattern(P, r, c, src);
}

// The checks of coo_walk, with the entries split over host threads: each
// thread finds the first out-of-range entry of its slice, and the earliest
// slice's hit is the entry the serial walk stops at (the same error and
// message).  One walk over 109M entries took 93 ms on one thread.
void validate_coo(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base) {
    constexpr uint32_t kSlice = 1u << 22;  // entries below which one thread walks them all
    if (nnz < 2 * kSlice) {
        coo_walk(n, nnz, row, col, base, [](uint32_t, uint32_t, int64_t) {});
        return;
    }
    coo_walk(n, 0, row, col, base, [](uint32_t, uint32_t, int64_t) {});  // the base and size checks
    if ((int64_t)nnz >= (int64_t)INT32_MAX) throw invalid("number of non-zeros must be below 2^31.");
    const uint32_t off = (uint32_t)base;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned T = (unsigned)std::min<uint64_t>(std::min(16u, hw), nnz / kSlice);
    std::vector<int64_t> first(T, -1);
    std::vector<std::thread> pool;
    pool.reserve(T);
    for (unsigned t = 0; t < T; ++t) {
        pool.emplace_back([&, t] {
            const uint32_t lo = (uint32_t)((uint64_t)nnz * t / T), hi = (uint32_t)((uint64_t)nnz * (t + 1) / T);
            for (uint32_t k = lo; k < hi; ++k)
                if (row[k] - off >= n || col[k] - off >= n) {
                    first[t] = k;
                    return;
                }
        });
    }
    for (auto& th : pool) th.join();
    for (unsigned t = 0; t < T; ++t)
        if (first[t] >= 0) throw invalid("matrix index out of range at entry " + std::to_string(first[t]) + ".");
}

namespace {
// col_ptr runs from 0 to nnz_half (after the index base) without decreasing
bool cscsym_standard(uint32_t n, uint32_t nnz_half, const int64_t* col_ptr, int base) {
    const int64_t off = base;
    bool ok = col_ptr[0] - off == 0 && col_ptr[n] - off == (int64_t)nnz_half;
    for (uint32_t j = 0; ok && j < n; ++j) ok = col_ptr[j + 1] >= col_ptr[j];
    return ok;
}
}  // namespace

uint32_t validate_cscsym(uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr, int base,
                         bool* standard) {
    const uint32_t full = cscsym_walk(n, nnz_half, row_idx, col_ptr, base, [](uint32_t, uint32_t, int64_t) {});
    if (standard) *standard = cscsym_standard(n, nnz_half, col_ptr, base);
    return full;
}

void csr_from_coo_device(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base,
                         int32_t* d_rowptr, int32_t* d_col, int64_t* d_perm, hipStream_t s, bool validated) {
    if (!validated) validate_coo(n, nnz, row, col, base);
    const int64_t m = nnz;
    DevBuf<uint32_t> drow, dcol;
    DevBuf<uint64_t> key, key2;
    DevBuf<int32_t> idx, idx2;
    DevBuf<uint8_t> temp;
    drow.ensure((size_t)m);
    dcol.ensure((size_t)m);
    key.ensure((size_t)m);
    key2.ensure((size_t)m);
    idx.ensure((size_t)m);
    idx2.ensure((size_t)m);
    const size_t tb = ingest_sort_temp_bytes(m, n);
    temp.ensure(std::max<size_t>(tb, 1));
    if (m > 0) {
        const HostPiece up[2] = {{drow.ptr, row, sizeof(uint32_t) * (size_t)m},
                                 {dcol.ptr, col, sizeof(uint32_t) * (size_t)m}};
        staged_upload(up, 2, s);
    }
    hip_check(ingest_coo_keys(m, n, (uint32_t)base, drow.ptr, dcol.ptr, key.ptr, idx.ptr, s), "COO keys");
    hip_check(ingest_sort_build(m, n, key.ptr, idx.ptr, key2.ptr, idx2.ptr, temp.ptr, tb, d_rowptr, d_col, d_perm, s),
              "sort and CSR build");
    hip_check(hipStreamSynchronize(s), "COO ingest");
}

uint32_t csr_from_cscsym_device(uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                                int base, int32_t* d_rowptr, int32_t* d_col, int64_t* d_perm, hipStream_t s,
                                bool validated) {
    bool standard = false;
    uint32_t full = 0;
    if (validated) {
        full = nnz_half * 2u - n;  // the walk's count, as it checked
        standard = cscsym_standard(n, nnz_half, col_ptr, base);
    } else {
        full = validate_cscsym(n, nnz_half, row_idx, col_ptr, base, &standard);
    }
    if (!standard) {
        // overlapping or unordered column ranges: expand in the loader's own order on the host
        Pattern P;
        pattern_from_cscsym(P, n, nnz_half, row_idx, col_ptr, base);
        const HostPiece up[3] = {{d_rowptr, P.rowptr.data(), sizeof(int32_t) * P.rowptr.size()},
                                 {d_col, P.col.data(), sizeof(int32_t) * P.col.size()},
                                 {d_perm, P.perm.data(), sizeof(int64_t) * P.perm.size()}};
        staged_upload(up, P.nnz > 0 ? 3 : 1, s);
        hip_check(hipStreamSynchronize(s), "CSC-sym ingest");
        return full;
    }
    const int64_t mh = nnz_half, m = full;
    DevBuf<int64_t> dri, dcp;
    DevBuf<uint32_t> colof, cnt, pos;
    DevBuf<uint64_t> key, key2;
    DevBuf<int32_t> idx, idx2;
    DevBuf<uint8_t> temp;
    dri.ensure((size_t)mh);
    dcp.ensure((size_t)n + 1);
    colof.ensure((size_t)mh);
    cnt.ensure((size_t)mh);
    pos.ensure((size_t)mh);
    key.ensure((size_t)m);
    key2.ensure((size_t)m);
    idx.ensure((size_t)m);
    idx2.ensure((size_t)m);
    const size_t tscan = ingest_scan_temp_bytes(mh), tsort = ingest_sort_temp_bytes(m, n);
    temp.ensure(std::max<size_t>(std::max(tscan, tsort), 1));
    const HostPiece up[2] = {{dri.ptr, row_idx, sizeof(int64_t) * (size_t)mh},
                             {dcp.ptr, col_ptr, sizeof(int64_t) * ((size_t)n + 1)}};
    staged_upload(up, 2, s);
    hip_check(ingest_csc_keys(mh, n, base, dri.ptr, dcp.ptr, colof.ptr, cnt.ptr, pos.ptr, temp.ptr, tscan, key.ptr,
                              idx.ptr, s),
              "CSC-sym expansion");
    hip_check(ingest_sort_build(m, n, key.ptr, idx.ptr, key2.ptr, idx2.ptr, temp.ptr, tsort, d_rowptr, d_col, d_perm, s),
              "sort and CSR build");
    hip_check(hipStreamSynchronize(s), "CSC-sym ingest");
    return full;
}

}  // namespace xj
