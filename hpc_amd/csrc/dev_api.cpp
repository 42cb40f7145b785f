// extern "C" device-level ABI (include/xjpcg_dev.h) over the C++ engine core.
#include "xjpcg_dev.h"

#include <algorithm>
#include <cmath>
#include <string>

#include "ingest.hpp"
#include "operator.hpp"

struct xjpcg_op {
    xj::Operator* impl;
};

namespace {

thread_local std::string g_last_error;

// smallest relative tolerance a refinement step asks of an inner operator with fp32 work vectors
constexpr double kW32StepTol = 1e-7;

template <typename F>
XJPCG_Status_t guarded(F f) {
    try {
        f();
        g_last_error.clear();
        return XJPCG_STATUS_SUCCESS;
    } catch (const xj::Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_last_error = std::string("INTERNAL ERROR: ") + e.what();
        return XJPCG_STATUS_INTERNAL_ERROR;
    }
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

xj::Operator& impl_of(xjpcg_op_t* op) {
    if (op == nullptr || op->impl == nullptr) throw xj::Error(XJPCG_STATUS_NOT_INITIALIZED, "null operator");
    return *op->impl;
}

}  // namespace

namespace xj {
void set_dev_error(const std::string& m) { g_last_error = m; }
}  // namespace xj

extern "C" {

const char* xjpcg_dev_last_error(void) { return g_last_error.c_str(); }

XJPCG_Status_t xjpcg_op_create(xjpcg_op_t** op, int32_t n, int64_t nnz, const int32_t* d_rowptr,
                               const int32_t* d_col, const void* d_val, int val_f32, const double* d_dsplit,
                               const int32_t* h_rowptr, void* stream) {
    return guarded([&] {
        if (op == nullptr) throw xj::invalid("null output pointer.");
        *op = nullptr;
        auto* impl = new xj::Operator(n, nnz, d_rowptr, d_col, d_val, val_f32 != 0, d_dsplit, h_rowptr,
                                      as_stream(stream));
        *op = new xjpcg_op{impl};
    });
}

XJPCG_Status_t xjpcg_op_destroy(xjpcg_op_t* op) {
    if (op == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    delete op->impl;
    delete op;
    return XJPCG_STATUS_SUCCESS;
}

XJPCG_Status_t xjpcg_op_set_compact(xjpcg_op_t* op, int on, int* h_active) {
    return guarded([&] {
        xj::Operator& o = impl_of(op);
        if (on >= 0) o.set_compact(on != 0);
        if (h_active) *h_active = o.compact() ? 1 : 0;
    });
}

XJPCG_Status_t xjpcg_op_set_values(xjpcg_op_t* op, const void* d_val) {
    return guarded([&] { impl_of(op).set_values(d_val); });
}

XJPCG_Status_t xjpcg_panel_dims(int32_t n, int32_t* block_rows, int32_t* nb, int32_t* npanels) {
    return guarded([&] {
        if (n <= 0) throw xj::invalid("Wrong dimension size.");
        const int br = xj::panel_block_rows(n);
        if (block_rows) *block_rows = br;
        if (nb) *nb = (int32_t)(((int64_t)n + br - 1) / br);
        if (npanels) *npanels = (int32_t)(((int64_t)n + 65535) / 65536);
    });
}

XJPCG_Status_t xjpcg_panel_plan(int32_t n, const int32_t* h_rowptr, const int32_t* h_col, int32_t* seg,
                                uint16_t* off, uint16_t* lrow, int64_t* perm, int* h_ordered) {
    return guarded([&] {
        if (n <= 0) throw xj::invalid("Wrong dimension size.");
        if (!h_rowptr || !seg || (h_rowptr[n] > 0 && (!h_col || !off || !lrow || !perm)))
            throw xj::invalid("Matrix is nullptr.");
        const xj::PanelPlan pl = xj::plan_panels(n, h_rowptr, h_col, xj::panel_block_rows(n));
        std::copy(pl.seg.begin(), pl.seg.end(), seg);
        std::copy(pl.off.begin(), pl.off.end(), off);
        std::copy(pl.lrow.begin(), pl.lrow.end(), lrow);
        std::copy(pl.perm.begin(), pl.perm.end(), perm);
        if (h_ordered) *h_ordered = pl.ordered ? 1 : 0;
    });
}

XJPCG_Status_t xjpcg_tile_plan(int32_t n, const int32_t* h_rowptr, int32_t* tiles, int32_t* h_ntiles) {
    return guarded([&] {
        if (n <= 0) throw xj::invalid("Wrong dimension size.");
        if (!h_rowptr || !tiles) throw xj::invalid("Matrix is nullptr.");
        const std::vector<int32_t> t = xj::plan_tiles(n, h_rowptr);
        std::copy(t.begin(), t.end(), tiles);
        if (h_ntiles) *h_ntiles = (int32_t)t.size() - 1;
    });
}

XJPCG_Status_t xjpcg_pattern_plan(int32_t n, const int32_t* h_rowptr, const int32_t* h_col, uint32_t* mask,
                                  int32_t* offs, int* h_ok) {
    return guarded([&] {
        if (n <= 0) throw xj::invalid("Wrong dimension size.");
        if (!h_rowptr || !mask || !offs || (h_rowptr[n] > 0 && !h_col)) throw xj::invalid("Matrix is nullptr.");
        const std::vector<int32_t> t = xj::plan_tiles(n, h_rowptr);
        std::vector<uint32_t> m;
        std::vector<int32_t> o;
        const bool ok = xj::plan_patterns(n, h_rowptr, h_col, t, m, o);
        if (ok) {
            std::copy(m.begin(), m.end(), mask);
            std::copy(o.begin(), o.end(), offs);
        }
        if (h_ok) *h_ok = ok ? 1 : 0;
    });
}

XJPCG_Status_t xjpcg_op_set_layout(xjpcg_op_t* op, int layout, void* stream, int* h_active) {
    return guarded([&] {
        xj::Operator& o = impl_of(op);
        if (layout >= 0) o.set_layout(layout, as_stream(stream));
        if (h_active) *h_active = o.layout();
    });
}

XJPCG_Status_t xjpcg_op_solve_layout(const xjpcg_op_t* op, int* h_layout) {
    return guarded([&] {
        const xj::Operator& o = impl_of(const_cast<xjpcg_op_t*>(op));
        if (h_layout) *h_layout = o.solve_layout();
    });
}

XJPCG_Status_t xjpcg_op_set_vectors(xjpcg_op_t* op, int bits, int* h_bits) {
    return guarded([&] {
        xj::Operator& o = impl_of(op);
        if (bits != 0) o.set_vector_bits(bits);
        if (h_bits) *h_bits = o.vector_bits();
    });
}

XJPCG_Status_t xjpcg_op_spmv_bytes(const xjpcg_op_t* op, int layout, int64_t* h_bytes) {
    return guarded([&] {
        const xj::Operator& o = impl_of(const_cast<xjpcg_op_t*>(op));
        if (layout < 0 || layout > 4) throw xj::invalid("layout must be 0 (in use), 1, 2, 3 or 4.");
        if (h_bytes) *h_bytes = o.spmv_bytes(layout);
    });
}

XJPCG_Status_t xjpcg_op_info(const xjpcg_op_t* op, int32_t* n, int64_t* nnz, int32_t* ntiles, int32_t* grid) {
    return guarded([&] {
        const xj::Operator& o = impl_of(const_cast<xjpcg_op_t*>(op));
        if (n) *n = o.csr().n;
        if (nnz) *nnz = o.csr().nnz;
        if (ntiles) *ntiles = o.csr().ntiles;
        if (grid) *grid = o.grid();
    });
}

XJPCG_Status_t xjpcg_op_set_path(xjpcg_op_t* op, int path, int* h_last) {
    return guarded([&] {
        xj::Operator& o = impl_of(op);
        if (h_last) *h_last = o.last_path();
        if (path < 0) return;  // query only
        if (path > 2) throw xj::invalid("path must be 0, 1 or 2.");
        o.set_path(path);
    });
}

XJPCG_Status_t xjpcg_csr_from_coo(uint32_t n, uint32_t nnz, const uint32_t* h_row, const uint32_t* h_col, int base,
                                  int32_t* d_rowptr, int32_t* d_col, int64_t* d_perm, void* stream) {
    return guarded([&] {
        if (n == 0) throw xj::invalid("Wrong dimension size.");
        if (!d_rowptr || (nnz > 0 && (!h_row || !h_col || !d_col || !d_perm))) throw xj::invalid("Matrix is nullptr.");
        xj::csr_from_coo_device(n, nnz, h_row, h_col, base, d_rowptr, d_col, d_perm, as_stream(stream));
    });
}

XJPCG_Status_t xjpcg_csr_from_cscsym(uint32_t n, uint32_t nnz_half, const int64_t* h_row_idx,
                                     const int64_t* h_col_ptr, int base, int32_t* d_rowptr, int32_t* d_col,
                                     int64_t* d_perm, int64_t* h_nnz, void* stream) {
    return guarded([&] {
        if (n == 0) throw xj::invalid("Wrong dimension size.");
        if (!h_col_ptr || !d_rowptr || (nnz_half > 0 && (!h_row_idx || !d_col || !d_perm)))
            throw xj::invalid("Matrix is nullptr.");
        const uint32_t full = xj::csr_from_cscsym_device(n, nnz_half, h_row_idx, h_col_ptr, base, d_rowptr, d_col,
                                                         d_perm, as_stream(stream));
        if (h_nnz) *h_nnz = full;
    });
}

XJPCG_Status_t xjpcg_stream_copy(const double* d_src, double* d_dst, int64_t n, void* stream) {
    return guarded([&] {
        if (n < 0) throw xj::invalid("length must not be negative.");
        if (n > 0 && (!d_src || !d_dst)) throw xj::invalid("Vector is nullptr.");
        if ((reinterpret_cast<uintptr_t>(d_src) | reinterpret_cast<uintptr_t>(d_dst)) & 15u)
            throw xj::invalid("arrays must be 16-byte aligned.");
        xj::hip_check(xj::launch_stream_copy(n, d_src, d_dst, as_stream(stream)), "stream copy launch");
    });
}

XJPCG_Status_t xjpcg_spmv(xjpcg_op_t* op, const double* d_x, double* d_y, void* stream) {
    return guarded([&] {
        if (d_x == nullptr || d_y == nullptr) throw xj::invalid("Vector is nullptr.");
        impl_of(op).spmv(d_x, d_y, as_stream(stream));
    });
}

XJPCG_Status_t xjpcg_residual(xjpcg_op_t* op, const double* d_b, const double* d_x, double* d_r, double* h_rr,
                              void* stream) {
    return guarded([&] {
        if (d_b == nullptr || d_x == nullptr || d_r == nullptr) throw xj::invalid("Vector is nullptr.");
        double rr = impl_of(op).residual(d_b, d_x, d_r, as_stream(stream));
        if (h_rr) *h_rr = rr;
    });
}

XJPCG_Status_t xjpcg_solve(xjpcg_op_t* op, const double* d_diag, const double* d_b, double* d_x, uint32_t max_iter,
                           double tol, uint32_t* h_iter, double* h_res, int64_t* h_executed, void* d_records,
                           void* stream) {
    return guarded([&] {
        xj::SolveResult R = impl_of(op).solve(d_diag, d_b, d_x, max_iter, tol, d_records, as_stream(stream));
        if (h_iter) *h_iter = R.iter;
        if (h_res) *h_res = R.res;
        if (h_executed) *h_executed = R.executed;
    });
}

XJPCG_Status_t xjpcg_sddmm(xjpcg_op_t* op, const double* d_lam, const double* d_x, double* d_g, void* stream) {
    return guarded([&] {
        if (d_lam == nullptr || d_x == nullptr || d_g == nullptr) throw xj::invalid("Vector is nullptr.");
        impl_of(op).sddmm(d_lam, d_x, d_g, as_stream(stream));
    });
}

XJPCG_Status_t xjpcg_solve_mixed(xjpcg_op_t* outer, xjpcg_op_t* inner, const double* d_diag, const double* d_b,
                                 double* d_x, uint32_t max_outer, double tol, double inner_tol,
                                 uint32_t inner_max_iter, uint32_t* h_outer, int64_t* h_inner_total,
                                 double* h_res, void* stream) {
    uint32_t outer_done = 0;
    int64_t inner_total = 0;
    double res = 0.0;
    XJPCG_Status_t st = guarded([&] {
        xj::Operator& O = impl_of(outer);
        xj::Operator& I = impl_of(inner);
        const int n = O.csr().n;
        if (I.csr().n != n) throw xj::invalid("inner and outer operators differ in size.");
        if (d_diag == nullptr || d_b == nullptr || d_x == nullptr) throw xj::invalid("Vector is nullptr.");
        hipStream_t s = as_stream(stream);
        double* r = O.aux(0);
        double* d = O.aux(1);
        xj::hip_check(hipMemsetAsync(d_x, 0, sizeof(double) * (size_t)n, s), "zero x");
        // x0 = 0 => r0 = b exactly
        xj::hip_check(hipMemcpyAsync(r, d_b, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s), "r0 = b");
        const double bb = O.dot(d_b, d_b, n, s);
        if (!(bb > 0.0)) {
            res = 0.0;
            return;
        }
        double rr = bb;
        res = 1.0;
        while (true) {
            res = std::sqrt(rr / bb);
            if (res < tol || outer_done >= max_outer) break;
            // The correction only has to bring ||r|| below tol ||b||: relative
            // target 0.5 tol ||b|| / ||r||, but never tighter than requested by
            // inner_tol (which bounds the work per step when the inner operator
            // is only an approximation of A) and never looser than 0.5.
            double tin = std::max(inner_tol, 0.5 * tol / res);
            tin = std::min(tin, 0.5);
            // fp32 work vectors: the recursive residual of the correction
            // drifts from its true residual at the fp32 level, so a step is
            // asked for no more than kW32StepTol
            if (I.vector_bits() == 32) tin = std::max(tin, kW32StepTol);
            xj::SolveResult R = I.solve(d_diag, r, d, inner_max_iter, tin, nullptr, s);
            inner_total += R.executed;
            xj::hip_check(xj::launch_axpy(n, 1.0, d, d_x, s), "x += d");
            rr = O.residual(d_b, d_x, r, s);
            ++outer_done;
        }
        if (!(res < tol))
            throw xj::exec_failed("refinement stopped at relative residual " + std::to_string(res) + " after " +
                                  std::to_string(outer_done) + " steps.");
    });
    if (h_outer) *h_outer = outer_done;
    if (h_inner_total) *h_inner_total = inner_total;
    if (h_res) *h_res = res;
    return st;
}

}  // extern "C"
