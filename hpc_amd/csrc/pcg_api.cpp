// pcg.h drop-in: the xJPCG C API over the C++ engine core.
//
// Mirrors pcg/sw/src/pcg.cpp (entry points, mode dispatch, metrics, the
// exception -> status mapping) and pcg/sw/include/impl/pcgImp.hpp (PCGImpl:
// argument checks, matrix / vector set-up, run, result scan), with the FPGA
// partitioner and OpenCL buffer plumbing replaced by a CSR build, HBM
// buffers and the persistent HIP solver.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ingest.hpp"
#include "operator.hpp"
#include "pcg.h"
#include "staging.hpp"
#include "xjpcg_ext.h"

namespace xj {
namespace {

using Clock = std::chrono::high_resolution_clock;

double elapsed(Clock::time_point& last) {
    auto now = Clock::now();
    double d = std::chrono::duration<double>(now - last).count();
    last = now;
    return d;
}

}  // namespace

// One solver session (reference PCGImpl, pcgImp.hpp:44-253).
class PcgHandle {
  public:
    XJPCG_Status_t set_status(XJPCG_Status_t s, const std::string& m) const {
        last_status_ = s;
        last_message_ = m;
        return s;
    }
    XJPCG_Status_t last_status() const { return last_status_; }
    const std::string& last_message() const { return last_message_; }
    XJPCG_Metric_t* metrics() { return &metrics_; }
    const XJPCG_Metric_t* metrics() const { return &metrics_; }
    bool first_call() const { return first_call_; }

    void init(const char* path) {
        device_ = 0;
        if (path != nullptr && *path != '\0') {
            char* end = nullptr;
            long v = std::strtol(path, &end, 10);
            if (end != path && *end == '\0' && v >= 0) device_ = (int)v;
        }
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) {
            init_error_ = "INTERNAL ERROR: no HIP device available.";
            throw Error(XJPCG_STATUS_INTERNAL_ERROR, init_error_);
        }
        if (device_ >= count) {
            init_error_ = "INVALID VALUE ERROR: HIP device " + std::to_string(device_) + " does not exist.";
            throw Error(XJPCG_STATUS_INVALID_VALUE, init_error_);
        }
        hip_check(hipSetDevice(device_), "select device");
        hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "create stream");
        ready_ = true;
    }

    ~PcgHandle() {
        if (ready_) {
            (void)hipSetDevice(device_);
            op_.reset();
            if (stream_) (void)hipStreamDestroy(stream_);
        }
    }

    // reference PCGImpl::setCooMat (pcgImp.hpp:49-70)
    void set_coo(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, const double* data, int base) {
        if (n == 0) throw invalid("Wrong dimension size.");
        if (nnz == 0) throw invalid("Wrong non-zero element size.");
        if (row == nullptr || col == nullptr || data == nullptr) throw invalid("Matrix is nullptr.");
        if (nnz < kDeviceIngestMin) {
            Pattern P;
            pattern_from_coo(P, n, nnz, row, col, base);
            require_device();
            upload_pattern(std::move(P), data);
            return;
        }
        validate_coo(n, nnz, row, col, base);  // every error before the device is touched
        require_device();
        build_on_device(n, nnz, nnz, data, [&](int32_t* rp, int32_t* c, int64_t* pm) {
            csr_from_coo_device(n, nnz, row, col, base, rp, c, pm, stream_, true);
        });
    }

    // reference PCGImpl::updateMat (pcgImp.hpp:95-116)
    void update_values(uint32_t n, uint32_t nnz, const double* data) {
        if (n == 0) throw invalid("Wrong dimension size.");
        if (nnz == 0) throw invalid("Wrong non-zero element size.");
        if (data == nullptr) throw invalid("Matrix is nullptr.");
        require_device();
        if (!op_ || n != pat_n_ || (int64_t)nnz != pat_nnz_)
            throw spm_invalid("Invalid dimension and/or number of nnzs used for updateing matrix.");
        refresh_values(data, src_len_);
    }

    // reference PCGImpl::setCscSymMat (pcgImp.hpp:71-94)
    void set_cscsym(uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr, const double* data,
                    int base) {
        if (n == 0) throw invalid("Wrong dimension size.");
        if (nnz_half == 0) throw invalid("Wrong non-zero element size.");
        if (row_idx == nullptr || col_ptr == nullptr || data == nullptr) throw invalid("Matrix is nullptr.");
        if (nnz_half < kDeviceIngestMin) {
            Pattern P;
            pattern_from_cscsym(P, n, nnz_half, row_idx, col_ptr, base);
            require_device();
            upload_pattern(std::move(P), data);
            return;
        }
        const uint32_t full = validate_cscsym(n, nnz_half, row_idx, col_ptr, base, nullptr);
        require_device();
        build_on_device(n, full, nnz_half, data, [&](int32_t* rp, int32_t* c, int64_t* pm) {
            csr_from_cscsym_device(n, nnz_half, row_idx, col_ptr, base, rp, c, pm, stream_, true);
        });
    }

    // reference PCGImpl::updateCscSymMat (pcgImp.hpp:117-145)
    void update_cscsym(uint32_t n, uint32_t nnz_half, const double* data) {
        if (n == 0) throw invalid("Wrong dimension size.");
        if (nnz_half == 0) throw invalid("Wrong non-zero element size.");
        if (data == nullptr) throw invalid("Matrix is nullptr.");
        require_device();
        const uint32_t full = nnz_half * 2u - n;
        if (!op_ || n != pat_n_ || (int64_t)full != pat_nnz_)
            throw spm_invalid("Invalid dimension and/or number of nnzs used for updateing matrix.");
        refresh_values(data, src_len_);
    }

    // reference PCGImpl::setVec (pcgImp.hpp:147-155)
    void set_vec(uint32_t n, const double* b, const double* diag) {
        if (b == nullptr || diag == nullptr) throw invalid("Vector is nullptr.");
        require_device();
        if (!op_ || n != pat_n_) throw invalid("vector size does not match the matrix.");
        d_b_.ensure(n);
        d_diag_.ensure(n);
        d_x_.ensure(n);
        const HostPiece up[2] = {{d_b_.ptr, b, sizeof(double) * n}, {d_diag_.ptr, diag, sizeof(double) * n}};
        staged_upload(up, 2, stream_);
        hip_check(hipStreamSynchronize(stream_), "upload vectors");
        vec_n_ = n;
    }

    // results of the last solve (Results, pcgImp.hpp:30-35)
    bool has_result() const { return has_result_; }
    const SolveResult& last_result() const { return last_result_; }
    void copy_x(double* out, uint32_t n) {
        if (!has_result_) throw invalid("no solve has run on this handle.");
        if (n != pat_n_) throw invalid("vector size does not match the matrix.");
        require_device();
        hip_check(hipMemcpyAsync(out, d_x_.ptr, sizeof(double) * pat_n_, hipMemcpyDeviceToHost, stream_), "copy x");
        hip_check(hipStreamSynchronize(stream_), "copy x");
    }
    bool has_vectors() const { return vec_n_ != 0 && vec_n_ == pat_n_ && op_ != nullptr; }

    // reference PCGImpl::run (pcgImp.hpp:157-183): solve, then copy x out (x_out may be null: x stays on
    // the GPU)
    SolveResult run(uint32_t max_iter, double tol, double* x_out) {
        require_device();
        void* rec = nullptr;
        const uint64_t nrec = (uint64_t)max_iter + 2;
        if (nrec <= kMaxRecords) {
            d_records_.ensure(nrec * 64);
            rec = d_records_.ptr;
        }
        SolveResult R = op_->solve(d_diag_.ptr, d_b_.ptr, d_x_.ptr, max_iter, tol, rec, stream_);
        if (x_out) {
            hip_check(hipMemcpyAsync(x_out, d_x_.ptr, sizeof(double) * pat_n_, hipMemcpyDeviceToHost, stream_),
                      "copy x");
            hip_check(hipStreamSynchronize(stream_), "copy x");
        }
        records_valid_ = rec != nullptr;
        last_executed_ = R.executed;
        first_call_ = false;
        last_result_ = R;
        has_result_ = true;
        return R;
    }

    // records 0..K of the last solve
    uint64_t history(void* out, uint64_t capacity) {
        if (!records_valid_) return 0;
        uint64_t count = (uint64_t)last_executed_ + 1;
        uint64_t m = std::min(count, capacity);
        if (out && m) {
            hip_check(hipMemcpyAsync(out, d_records_.ptr, m * 64, hipMemcpyDeviceToHost, stream_), "copy records");
            hip_check(hipStreamSynchronize(stream_), "copy records");
        }
        return count;
    }

  private:
    static constexpr uint64_t kMaxRecords = 1ull << 20;  // 64 MiB of records at most
    // inputs from this many stored entries are sorted into CSR on the GPU
    // (same CSR as the host build; below it the host build is quicker)
    static constexpr uint32_t kDeviceIngestMin = 1u << 16;

    void require_device() {
        if (!ready_) throw Error(XJPCG_STATUS_INTERNAL_ERROR, init_error_.empty() ? "INTERNAL ERROR: handle not initialised." : init_error_);
        hip_check(hipSetDevice(device_), "select device");
    }

    void upload_pattern(Pattern&& P, const double* data) {
        op_.reset();
        const uint32_t n = P.n;
        d_rowptr_.ensure((size_t)n + 1);
        d_col_.ensure((size_t)P.nnz);
        d_val_.ensure((size_t)P.nnz);
        d_perm_.ensure((size_t)P.nnz);
        const HostPiece up[3] = {{d_rowptr_.ptr, P.rowptr.data(), sizeof(int32_t) * ((size_t)n + 1)},
                                 {d_col_.ptr, P.col.data(), sizeof(int32_t) * (size_t)P.nnz},
                                 {d_perm_.ptr, P.perm.data(), sizeof(int64_t) * (size_t)P.nnz}};
        staged_upload(up, 3, stream_);
        pat_n_ = n;
        pat_nnz_ = P.nnz;
        src_len_ = P.src_len;
        refresh_values(data, P.src_len);
        op_.reset(new Operator((int32_t)n, P.nnz, d_rowptr_.ptr, d_col_.ptr, d_val_.ptr, false, nullptr,
                               P.rowptr.data(), stream_));
    }

    // CSR built on the GPU by `build` (into d_rowptr_, d_col_, d_perm_), then
    // values placed and the operator created as for a host-built pattern
    template <typename Build>
    void build_on_device(uint32_t n, int64_t nnz, int64_t src_len, const double* data, Build build) {
        op_.reset();
        d_rowptr_.ensure((size_t)n + 1);
        d_col_.ensure((size_t)nnz);
        d_val_.ensure((size_t)nnz);
        d_perm_.ensure((size_t)nnz);
        build(d_rowptr_.ptr, d_col_.ptr, d_perm_.ptr);
        std::vector<int32_t> rp((size_t)n + 1);
        hip_check(hipMemcpyAsync(rp.data(), d_rowptr_.ptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, stream_),
                  "copy row pointers");
        hip_check(hipStreamSynchronize(stream_), "copy row pointers");
        pat_n_ = n;
        pat_nnz_ = nnz;
        src_len_ = src_len;
        refresh_values(data, src_len);
        op_.reset(new Operator((int32_t)n, nnz, d_rowptr_.ptr, d_col_.ptr, d_val_.ptr, false, nullptr, rp.data(),
                               stream_));
    }

    // values reach their CSR slots on the GPU: val[k] = data[perm[k]]
    void refresh_values(const double* data, int64_t len) {
        d_raw_.ensure((size_t)len);
        staged_upload(d_raw_.ptr, data, sizeof(double) * (size_t)len, stream_);
        hip_check(launch_gather_f64(pat_nnz_, d_perm_.ptr, d_raw_.ptr, d_val_.ptr, stream_), "place values");
        hip_check(hipStreamSynchronize(stream_), "place values");
        if (op_) op_->set_values(d_val_.ptr);  // values changed in place: refresh the operator's own copies
    }

    mutable XJPCG_Status_t last_status_ = XJPCG_STATUS_SUCCESS;
    mutable std::string last_message_;
    XJPCG_Metric_t metrics_{0.0, 0.0, 0.0, 0.0};
    bool first_call_ = true;
    bool ready_ = false;
    std::string init_error_;
    int device_ = 0;
    hipStream_t stream_ = nullptr;

    uint32_t pat_n_ = 0;
    int64_t pat_nnz_ = 0, src_len_ = 0;
    DevBuf<int32_t> d_rowptr_, d_col_;
    DevBuf<double> d_val_, d_raw_;
    DevBuf<int64_t> d_perm_;
    DevBuf<double> d_b_, d_diag_, d_x_;
    DevBuf<uint8_t> d_records_;
    bool records_valid_ = false;
    int64_t last_executed_ = 0;
    uint32_t vec_n_ = 0;
    bool has_result_ = false;
    SolveResult last_result_;
    std::unique_ptr<Operator> op_;
};

}  // namespace xj

using xj::PcgHandle;

namespace {

PcgHandle* as_impl(XJPCG_Handle_t* h) { return reinterpret_cast<PcgHandle*>(h); }
const PcgHandle* as_impl(const XJPCG_Handle_t* h) { return reinterpret_cast<const PcgHandle*>(h); }

// catch blocks of pcg.cpp:46-50 / :110-114 / :167-171
template <typename F>
XJPCG_Status_t run_guarded(const PcgHandle* h, F f, const char* ok_message) {
    try {
        f();
    } catch (const xj::Error& e) {
        return h->set_status(e.status, e.what());
    } catch (const std::exception& e) {
        return h->set_status(XJPCG_STATUS_INTERNAL_ERROR, e.what());
    }
    return h->set_status(XJPCG_STATUS_SUCCESS, ok_message);
}

}  // namespace

extern "C" {

// reference pcg.cpp:34-52
XJPCG_Status_t xJPCG_createHandle(XJPCG_Handle_t** handle, const char* xclbinPath) {
    if (handle == nullptr) return XJPCG_STATUS_INVALID_VALUE;
    PcgHandle* h = new PcgHandle();
    *handle = reinterpret_cast<XJPCG_Handle_t*>(h);
    return run_guarded(h, [&] {
        auto last = xj::Clock::now();
        h->init(xclbinPath);
        h->metrics()->m_objInit = xj::elapsed(last);
    }, "Handle is created successfully");
}

// reference pcg.cpp:54-59
XJPCG_Status_t xJPCG_destroyHandle(XJPCG_Handle_t* handle) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    delete as_impl(handle);
    return XJPCG_STATUS_SUCCESS;
}

// reference pcg.cpp:61-116
XJPCG_Status_t xJPCG_cscSymSolver(XJPCG_Handle_t* handle, const int64_t p_n, const int64_t p_nnz,
                                  const int64_t* p_rowIdx, const int64_t* p_colPtr, const double* p_data,
                                  const double* p_diagA, const double* p_b, const double* p_x,
                                  const int64_t p_maxIter, const double p_tol, int64_t* p_iter, double* p_res,
                                  const XJPCG_Mode_t mode) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = as_impl(handle);
    return run_guarded(h, [&] {
        auto last = xj::Clock::now();
        const bool first = h->first_call();
        const uint32_t n = (uint32_t)p_n, nnz = (uint32_t)p_nnz;  // PCGImpl takes uint32 sizes
        switch (mode & 0x0f) {
            case XJPCG_MODE_DEFAULT:
                h->set_cscsym(n, nnz, p_rowIdx, p_colPtr, p_data, (mode & 0xf0) >> 4);
                break;
            case XJPCG_MODE_KEEP_NZ_LAYOUT:
                if (first) throw xj::invalid("wrong solver mode for the first call, please use XJPCG_MODEL_DEFAULT.");
                h->update_cscsym(n, nnz, p_data);
                break;
            default:
                if (first) throw xj::invalid("wrong solver mode for the first call, please use XJPCG_MODEL_DEFAULT.");
                break;
        }
        h->metrics()->m_matProc = xj::elapsed(last);
        h->set_vec(n, p_b, p_diagA);
        h->metrics()->m_vecProc = xj::elapsed(last);
        if (p_iter == nullptr || p_res == nullptr || p_x == nullptr) throw xj::invalid("output pointer is nullptr.");
        xj::SolveResult R = h->run((uint32_t)p_maxIter, p_tol, const_cast<double*>(p_x));
        *p_res = R.res;
        *p_iter = R.iter;
        h->metrics()->m_solver = xj::elapsed(last);
        if (*p_res > p_tol)
            throw xj::exec_failed("exit with divergent solution after " + std::to_string(*p_iter) + " iterations.");
    }, "Solver returns successfully");
}

// reference pcg.cpp:118-173
XJPCG_Status_t xJPCG_cooSolver(XJPCG_Handle_t* handle, const uint32_t p_n, const uint32_t p_nnz,
                               const uint32_t* p_rowIdx, const uint32_t* p_colIdx, const double* p_data,
                               const double* matJ, const double* b, const double* x, const uint32_t p_maxIter,
                               const double p_tol, uint32_t* p_iter, double* p_res, const XJPCG_Mode_t mode) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = as_impl(handle);
    return run_guarded(h, [&] {
        auto last = xj::Clock::now();
        const bool first = h->first_call();
        switch (mode & 0x0f) {
            case XJPCG_MODE_DEFAULT:
                h->set_coo(p_n, p_nnz, p_rowIdx, p_colIdx, p_data, (mode & 0xf0) >> 4);
                break;
            case XJPCG_MODE_KEEP_NZ_LAYOUT:
                if (first) throw xj::invalid("wrong solver mode for the first call, please use XJPCG_MODEL_DEFAULT.");
                h->update_values(p_n, p_nnz, p_data);
                break;
            default:
                if (first) throw xj::invalid("wrong solver mode for the first call, please use XJPCG_MODEL_DEFAULT.");
                break;
        }
        h->metrics()->m_matProc = xj::elapsed(last);
        h->set_vec(p_n, b, matJ);
        h->metrics()->m_vecProc = xj::elapsed(last);
        if (p_iter == nullptr || p_res == nullptr || x == nullptr) throw xj::invalid("output pointer is nullptr.");
        xj::SolveResult R = h->run(p_maxIter, p_tol, const_cast<double*>(x));
        *p_res = R.res;
        *p_iter = R.iter;
        h->metrics()->m_solver = xj::elapsed(last);
        if (*p_res > p_tol)
            throw xj::exec_failed("exit with divergent solution after " + std::to_string(*p_iter) + " iterations.");
    }, "Solver returns successfully");
}

// The solve steps as separate calls (PCGImpl, pcgImp.hpp:49-204).
XJPCG_Status_t xJPCG_setCooMat(XJPCG_Handle_t* handle, uint32_t n, uint32_t nnz, const uint32_t* rowIdx,
                               const uint32_t* colIdx, const double* data, XJPCG_Mode_t mode) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = as_impl(handle);
    return run_guarded(h, [&] {
        auto last = xj::Clock::now();
        switch (mode & 0x0f) {
            case XJPCG_MODE_DEFAULT: h->set_coo(n, nnz, rowIdx, colIdx, data, (mode & 0xf0) >> 4); break;
            case XJPCG_MODE_KEEP_NZ_LAYOUT: h->update_values(n, nnz, data); break;
            case XJPCG_MODE_KEEP_MATRIX: break;
            default: throw xj::invalid("unknown solver mode.");
        }
        h->metrics()->m_matProc = xj::elapsed(last);
    }, "Matrix is loaded successfully");
}

XJPCG_Status_t xJPCG_setCscSymMat(XJPCG_Handle_t* handle, int64_t n, int64_t nnz, const int64_t* rowIdx,
                                  const int64_t* colPtr, const double* data, XJPCG_Mode_t mode) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = as_impl(handle);
    return run_guarded(h, [&] {
        auto last = xj::Clock::now();
        const uint32_t n32 = (uint32_t)n, nnz32 = (uint32_t)nnz;  // PCGImpl takes uint32 sizes
        switch (mode & 0x0f) {
            case XJPCG_MODE_DEFAULT: h->set_cscsym(n32, nnz32, rowIdx, colPtr, data, (mode & 0xf0) >> 4); break;
            case XJPCG_MODE_KEEP_NZ_LAYOUT: h->update_cscsym(n32, nnz32, data); break;
            case XJPCG_MODE_KEEP_MATRIX: break;
            default: throw xj::invalid("unknown solver mode.");
        }
        h->metrics()->m_matProc = xj::elapsed(last);
    }, "Matrix is loaded successfully");
}

XJPCG_Status_t xJPCG_setVec(XJPCG_Handle_t* handle, uint32_t n, const double* b, const double* diagA) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = as_impl(handle);
    return run_guarded(h, [&] {
        auto last = xj::Clock::now();
        h->set_vec(n, b, diagA);
        h->metrics()->m_vecProc = xj::elapsed(last);
    }, "Vectors are loaded successfully");
}

XJPCG_Status_t xJPCG_run(XJPCG_Handle_t* handle, uint32_t maxIter, double tol, uint32_t* iter, double* res) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = as_impl(handle);
    return run_guarded(h, [&] {
        if (!h->has_vectors()) throw xj::invalid("load the matrix and the vectors before running the solver.");
        auto last = xj::Clock::now();
        const xj::SolveResult R = h->run(maxIter, tol, nullptr);
        if (iter) *iter = R.iter;
        if (res) *res = R.res;
        h->metrics()->m_solver = xj::elapsed(last);
        if (R.res > tol)
            throw xj::exec_failed("exit with divergent solution after " + std::to_string(R.iter) + " iterations.");
    }, "Solver returns successfully");
}

XJPCG_Status_t xJPCG_getResults(const XJPCG_Handle_t* handle, uint32_t* iter, double* res) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    const PcgHandle* h = as_impl(handle);
    if (!h->has_result()) return h->set_status(XJPCG_STATUS_INVALID_VALUE, "INVALID VALUE ERROR: no solve has run on this handle.");
    if (iter) *iter = h->last_result().iter;
    if (res) *res = h->last_result().res;
    return h->set_status(XJPCG_STATUS_SUCCESS, "Get results successfully");
}

XJPCG_Status_t xJPCG_getX(const XJPCG_Handle_t* handle, double* x, uint32_t n) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = const_cast<PcgHandle*>(as_impl(handle));
    return run_guarded(h, [&] {
        if (x == nullptr) throw xj::invalid("output pointer is nullptr.");
        h->copy_x(x, n);
    }, "Get x successfully");
}

// reference pcg.cpp:175-180
XJPCG_Status_t xJPCG_getMetrics(const XJPCG_Handle_t* handle, XJPCG_Metric_t* metric) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    const PcgHandle* h = as_impl(handle);
    if (metric == nullptr) return h->set_status(XJPCG_STATUS_INVALID_VALUE, "INVALID VALUE ERROR: metric is nullptr.");
    *metric = *h->metrics();
    return h->set_status(XJPCG_STATUS_SUCCESS, "Get metrics successfully");
}

// reference pcg.cpp:182-188
XJPCG_Status_t xJPCG_peekAtLastStatus(const XJPCG_Handle_t* handle) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    return as_impl(handle)->last_status();
}

// reference pcg.cpp:190-194
const char* xJPCG_getLastMessage(const XJPCG_Handle_t* handle) {
    if (handle == nullptr) return "Can't retrieve last error message from a null JPCG handle.";
    return as_impl(handle)->last_message().c_str();
}

// reference pcg.cpp:196-214
const char* xJPCG_getErrorString(const XJPCG_Status_t code) {
    switch (code) {
        case XJPCG_STATUS_SUCCESS: return "XJPCG_STATUS_SUCCESS";
        case XJPCG_STATUS_NOT_INITIALIZED: return "XJPCG_STATUS_NOT_INITIALIZED";
        case XJPCG_STATUS_ALLOC_FAILED: return "XJPCG_STATUS_ALLOC_FAILED";
        case XJPCG_STATUS_INVALID_VALUE: return "XJPCG_STATUS_INVALID_VALUE";
        case XJPCG_STATUS_EXECUTION_FAILED: return "XJPCG_STATUS_EXECUTION_FAILED";
        case XJPCG_STATUS_INTERNAL_ERROR: return "XJPCG_STATUS_INTERNAL_ERROR";
        case XJPCG_STATUS_DYNAMIC_LOADING_ERROR: return "XJPCG_STATUS_DYNAMIC_LOADING_ERROR";
    }
    return "Unknown Error Code!";
}

// per-iteration records of the last solve (extension, include/xjpcg_ext.h)
XJPCG_Status_t xJPCG_getHistory(const XJPCG_Handle_t* handle, void* records, uint64_t capacity, uint64_t* count) {
    if (handle == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    PcgHandle* h = const_cast<PcgHandle*>(as_impl(handle));
    return run_guarded(h, [&] {
        if (count == nullptr) throw xj::invalid("count is nullptr.");
        *count = h->history(records, capacity);
    }, "Get history successfully");
}

}  // extern "C"
