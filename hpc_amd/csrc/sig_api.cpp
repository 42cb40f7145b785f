// extern "C" ABI of the signature-file codec (include/xjpcg_sig.h).
#include "xjpcg_sig.h"

#include <algorithm>
#include <cstring>
#include <string>

#include "ingest.hpp"
#include "launch.hpp"
#include "operator.hpp"
#include "signature.hpp"
#include "staging.hpp"

struct xjpcg_sig {
    xj::sig::Image im;
};

namespace {

thread_local std::string g_sig_error;

template <typename F>
XJPCG_Status_t guarded(F f) {
    try {
        f();
        g_sig_error.clear();
        return XJPCG_STATUS_SUCCESS;
    } catch (const xj::Error& e) {
        g_sig_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_sig_error = "Alloc Failed ERROR: host allocation failed";
        return XJPCG_STATUS_ALLOC_FAILED;
    } catch (const std::exception& e) {
        g_sig_error = std::string("INTERNAL ERROR: ") + e.what();
        return XJPCG_STATUS_INTERNAL_ERROR;
    }
}

const xj::sig::Image& image_of(const xjpcg_sig_t* s) {
    if (s == nullptr) throw xj::Error(XJPCG_STATUS_NOT_INITIALIZED, "null signature");
    return s->im;
}

// Builds into a fresh object and hands it out only on success.
template <typename F>
XJPCG_Status_t create(xjpcg_sig_t** out, F build) {
    return guarded([&] {
        if (out == nullptr) throw xj::invalid("null output pointer.");
        *out = nullptr;
        xjpcg_sig_t* s = new xjpcg_sig_t;
        try {
            build(s->im);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

}  // namespace

extern "C" {

const char* xjpcg_sig_last_error(void) { return g_sig_error.c_str(); }

XJPCG_Status_t xjpcg_sig_from_coo(xjpcg_sig_t** sig, uint32_t m, uint32_t n, uint32_t nnz, const uint32_t* row,
                                  const uint32_t* col, const double* val, int base) {
    return create(sig, [&](xj::sig::Image& im) {
        if (m == 0 || n == 0) throw xj::invalid("Wrong dimension size.");
        if (nnz > 0 && (!row || !col || !val)) throw xj::invalid("Matrix is nullptr.");
        xj::sig::encode_coo(im, m, n, nnz, row, col, val, base);
    });
}

XJPCG_Status_t xjpcg_sig_from_cscsym(xjpcg_sig_t** sig, uint32_t n, uint32_t nnz_half, const int64_t* row_idx,
                                     const int64_t* col_ptr, const double* val, int base) {
    return create(sig, [&](xj::sig::Image& im) {
        if (n == 0) throw xj::invalid("Wrong dimension size.");
        if (!col_ptr || (nnz_half > 0 && (!row_idx || !val))) throw xj::invalid("Matrix is nullptr.");
        xj::sig::encode_cscsym(im, n, nnz_half, row_idx, col_ptr, val, base);
    });
}

XJPCG_Status_t xjpcg_sig_update(xjpcg_sig_t* sig, const double* val, uint64_t len) {
    return guarded([&] {
        if (sig == nullptr) throw xj::Error(XJPCG_STATUS_NOT_INITIALIZED, "null signature");
        if (!val && len > 0) throw xj::invalid("Matrix is nullptr.");
        xj::sig::update_values(sig->im, val, len);
    });
}

XJPCG_Status_t xjpcg_sig_store(const xjpcg_sig_t* sig, const char* dir) {
    return guarded([&] {
        const xj::sig::Image& im = image_of(sig);
        if (dir == nullptr) throw xj::invalid("null directory.");
        xj::sig::store(im, dir);
    });
}

XJPCG_Status_t xjpcg_sig_load(xjpcg_sig_t** sig, const char* dir) {
    return create(sig, [&](xj::sig::Image& im) {
        if (dir == nullptr) throw xj::invalid("null directory.");
        xj::sig::load(im, dir);
    });
}

XJPCG_Status_t xjpcg_sig_destroy(xjpcg_sig_t* sig) {
    delete sig;
    return XJPCG_STATUS_SUCCESS;
}

XJPCG_Status_t xjpcg_sig_info(const xjpcg_sig_t* sig, uint32_t* info, uint64_t* entries, uint32_t* row_blocks,
                              uint32_t* partitions) {
    return guarded([&] {
        const xj::sig::Image& im = image_of(sig);
        if (info) {
            const uint32_t v[6] = {im.m, im.n, im.nnz, im.m_pad, im.n_pad, im.nnz_pad};
            std::memcpy(info, v, sizeof(v));
        }
        if (entries) *entries = xj::sig::count_entries(im);
        if (row_blocks || partitions) {
            const xj::sig::Summary s = xj::sig::summarize(im);
            if (row_blocks) *row_blocks = s.row_blocks;
            if (partitions) *partitions = s.partitions;
        }
    });
}

XJPCG_Status_t xjpcg_sig_buffer(const xjpcg_sig_t* sig, int which, const void** ptr, uint64_t* bytes) {
    return guarded([&] {
        const xj::sig::Image& im = image_of(sig);
        if (which < 0 || which >= 2 + (int)xj::sig::kChannels) throw xj::invalid("buffer index out of range.");
        const std::vector<uint8_t>& b = which == 0 ? im.rb_param : which == 1 ? im.par_param : im.nnz_val[which - 2];
        if (ptr) *ptr = b.data();
        if (bytes) *bytes = b.size();
    });
}

XJPCG_Status_t xjpcg_sig_to_coo(const xjpcg_sig_t* sig, uint32_t* row, uint32_t* col, double* val, uint64_t cap) {
    return guarded([&] {
        const xj::sig::Image& im = image_of(sig);
        xj::sig::Coo c;
        xj::sig::decode(im, c);
        if (c.row.size() > cap)
            throw xj::invalid("output capacity " + std::to_string(cap) + " below the " + std::to_string(c.row.size()) +
                              " decoded entries.");
        if (!c.row.empty() && (!row || !col || !val)) throw xj::invalid("null output array.");
        if (c.row.empty()) return;
        std::memcpy(row, c.row.data(), c.row.size() * sizeof(uint32_t));
        std::memcpy(col, c.col.data(), c.col.size() * sizeof(uint32_t));
        std::memcpy(val, c.val.data(), c.val.size() * sizeof(double));
    });
}

XJPCG_Status_t xjpcg_sig_to_csr(const xjpcg_sig_t* sig, int32_t* d_rowptr, int32_t* d_col, double* d_val,
                                uint64_t cap, uint64_t* h_entries, void* stream) {
    return guarded([&] {
        const xj::sig::Image& im = image_of(sig);
        if (im.m != im.n) throw xj::invalid("the signature's matrix is not square.");
        if (!d_rowptr) throw xj::invalid("Matrix is nullptr.");
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint64_t vw = 0;
        const std::vector<xj::sig::Segment> seg = xj::sig::segments(im, &vw);
        if (vw >= (uint64_t)INT32_MAX) throw xj::invalid("more than 2^31 value words.");
        const uint32_t nseg = (uint32_t)seg.size();
        size_t words_bytes = 0;
        for (const auto& st : im.nnz_val) words_bytes += st.size();

        xj::DevBuf<uint8_t> words;
        xj::DevBuf<xj::sig::Segment> dseg;
        xj::DevBuf<uint32_t> cnt, pos;
        xj::DevBuf<int32_t> bad;
        words.ensure(words_bytes);
        dseg.ensure(nseg);
        cnt.ensure((size_t)vw);
        pos.ensure((size_t)vw);
        bad.ensure(1);
        std::vector<xj::HostPiece> up;
        size_t off = 0;
        for (const auto& st : im.nnz_val) {
            up.push_back({words.ptr + off, st.data(), st.size()});
            off += st.size();
        }
        up.push_back({dseg.ptr, seg.data(), sizeof(xj::sig::Segment) * nseg});
        xj::staged_upload(up.data(), up.size(), s);
        xj::hip_check(xj::sig::decode_count(vw, dseg.ptr, nseg, words.ptr, cnt.ptr, s), "signature count");
        uint64_t total = 0;
        if (vw > 0) {
            xj::DevBuf<uint8_t> scan_tmp;
            const size_t tb = xj::ingest_scan_temp_bytes((int64_t)vw);
            scan_tmp.ensure(std::max<size_t>(tb, 1));
            xj::hip_check(xj::ingest_exclusive_sum(scan_tmp.ptr, tb, cnt.ptr, pos.ptr, (int64_t)vw, s),
                          "signature scan");
            uint32_t last[2] = {0, 0};
            xj::hip_check(hipMemcpyAsync(&last[0], pos.ptr + vw - 1, 4, hipMemcpyDeviceToHost, s), "count");
            xj::hip_check(hipMemcpyAsync(&last[1], cnt.ptr + vw - 1, 4, hipMemcpyDeviceToHost, s), "count");
            xj::hip_check(hipStreamSynchronize(s), "signature count");
            total = (uint64_t)last[0] + last[1];
        }
        if (h_entries) *h_entries = total;
        if (total > cap)
            throw xj::invalid("output capacity " + std::to_string(cap) + " below the " + std::to_string(total) +
                              " decoded entries.");
        if (total >= (uint64_t)INT32_MAX) throw xj::invalid("number of non-zeros must be below 2^31.");
        if (total > 0 && (!d_col || !d_val)) throw xj::invalid("Matrix is nullptr.");
        const int64_t m = (int64_t)total;
        xj::DevBuf<uint64_t> key, key2;
        xj::DevBuf<int32_t> idx, idx2;
        xj::DevBuf<double> dv;
        xj::DevBuf<int64_t> perm;
        xj::DevBuf<uint8_t> temp;
        key.ensure((size_t)m);
        key2.ensure((size_t)m);
        idx.ensure((size_t)m);
        idx2.ensure((size_t)m);
        dv.ensure((size_t)m);
        perm.ensure((size_t)m);
        const size_t tsort = xj::ingest_sort_temp_bytes(m, im.n);
        temp.ensure(std::max<size_t>(tsort, 1));
        int32_t zero = 0, flag = 0;
        xj::hip_check(hipMemcpyAsync(bad.ptr, &zero, sizeof(zero), hipMemcpyHostToDevice, s), "flag");
        xj::hip_check(xj::sig::decode_emit(vw, dseg.ptr, nseg, words.ptr, pos.ptr, im.m, im.n, key.ptr, idx.ptr,
                                           dv.ptr, bad.ptr, s),
                      "signature decode");
        xj::hip_check(hipMemcpyAsync(&flag, bad.ptr, sizeof(flag), hipMemcpyDeviceToHost, s), "flag");
        xj::hip_check(hipStreamSynchronize(s), "signature decode");
        if (flag)
            throw xj::invalid("malformed signature files: an entry lies outside the " + std::to_string(im.m) + " x " +
                              std::to_string(im.n) + " matrix.");
        xj::hip_check(xj::ingest_sort_build(m, im.n, key.ptr, idx.ptr, key2.ptr, idx2.ptr, temp.ptr, tsort, d_rowptr,
                                            d_col, perm.ptr, s),
                      "sort and CSR build");
        xj::hip_check(xj::launch_gather_vals(m, perm.ptr, dv.ptr, d_val, false, s), "value gather");
        xj::hip_check(hipStreamSynchronize(s), "signature CSR build");
    });
}

}  // extern "C"
