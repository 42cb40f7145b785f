// Allocation-free JPCG on a caller's device CSR (xjpcg_csr_solve_static in
// include/xjpcg_dev.h).
//
// Every work vector, partial sum and the loop state live in `scratch`, one
// file-scope device array per GPU, sized for kStaticMaxRows rows.  Its device
// address is looked up once per device with hipGetSymbolAddress and kept in
// g_scratch[device]; host code passes that address, never the array's name,
// to kernels.  The call launches kernels and copies 4-byte flags to the host
// stack: no hipMalloc, hipFree, event, graph or library allocation.  Calls
// that overlap on one GPU share the array and race.
//
// The loop is the multi-kernel JPCG of the operator solves (the same set-up,
// residual and x / p update kernels, the reference's update order and stop
// rule, pcg/hw/include/control.hpp:65-91) with a plan-free SpMV: one lane per
// row, products added left to right, one p.Ap partial per workgroup.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "kernels.hpp"
#include "launch.hpp"
#include "operator.hpp"
#include "xjpcg_dev.h"

namespace xj {

constexpr int kStaticMaxRows = XJPCG_STATIC_MAX_ROWS;
constexpr int kStaticSpmvBlocks = (kStaticMaxRows + kBlock - 1) / kBlock;  // p.Ap partials
constexpr int kStaticG2 = vector_groups(kStaticMaxRows);                   // vector-kernel workgroups
// r, p, Ap, J; the SpMV partials; the r.r and r.z partials; CgState (in doubles)
constexpr size_t kStaticDoubles = 4 * (size_t)kStaticMaxRows + kStaticSpmvBlocks + 2 * kStaticG2 +
                                  (sizeof(CgState) + 7) / 8;
static_assert(kStaticSpmvBlocks <= kMaxRedundantPartials, "partials are summed without the cg_alpha fold");

static __device__ __align__(256) double scratch[kStaticDoubles];

constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// address of this device's scratch array (looked up on the device's first call)
static double* scratch_ptr() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "current device");
    if (dev < 0 || dev >= kMaxDevices) throw internal("device index beyond the static workspace table.");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev] == nullptr) {
        double* ptr = nullptr;
        hip_check(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "static workspace address");
        if (ptr == nullptr) throw internal("static workspace has no device address.");
        g_scratch[dev] = ptr;
    }
    return g_scratch[dev];
}

// CG-loop SpMV without a plan: lane i sums row i left to right (the products
// rounded before each add, as in the tile kernels' short rows), stores Ap[i],
// and each workgroup leaves one p.Ap partial.
__global__ __launch_bounds__(kBlock) void cg_spmv_rows(CgParams P) {
#pragma clang fp contract(off)
    __shared__ double red[kWaves];
    CgState* st = P.st;
    const int stop = st->stop;
    if (stop) return;
    const int i = (int)blockIdx.x * kBlock + (int)threadIdx.x;
    const double* __restrict__ p = P.p;
    const double* __restrict__ val = static_cast<const double*>(P.val);
    const int* __restrict__ col = static_cast<const int*>(P.col);
    double pap = 0.0;
    if (i < P.n) {
        const int a = P.rowptr[i], b = P.rowptr[i + 1];
        double s = 0.0;
        for (int k = a; k < b; ++k) s += val[k] * p[col[k]];
        P.Ap[i] = s;
        pap = p[i] * s;
    }
    pap = block_sum(pap, red);
    if (threadIdx.x == 0) P.partA[blockIdx.x] = pap;
    // k advances after the last load; no other workgroup of this launch reads it
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = st->k + 1;
}

struct StaticSolve {
    int64_t executed;
    double rr, dot;
};

static StaticSolve csr_solve_static(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                                    const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                                    hipStream_t s) {
    double* w = scratch_ptr();
    CgParams P;
    P.n = n;
    P.ntiles = (n + kBlock - 1) / kBlock;
    P.tiles = P.ntiles;
    P.G2 = vector_groups(n);
    P.split_alpha = 0;
    P.rowptr = rowptr;
    P.col = col;
    P.val = val;
    P.b = b;
    P.diag = diag;
    P.x = x;
    P.r = w;
    P.p = w + (size_t)n;
    P.Ap = w + 2 * (size_t)n;
    P.J = w + 3 * (size_t)n;
    double* tail = w + 4 * (size_t)kStaticMaxRows;
    P.partA = tail;
    P.partB = tail + kStaticSpmvBlocks;
    P.partC = P.partB + kStaticG2;
    P.st = reinterpret_cast<CgState*>(P.partC + kStaticG2);
    P.records = nullptr;
    P.tol = tol;
    P.max_rounds = max_iter + 1u;  // pcgImp.hpp:158 programs maxIter+1 rounds
    // the set-up writes r, p, J, x and every CgState field; each iteration
    // writes every partial it reads, so nothing needs clearing
    hip_check(launch_cg_setup(P, s), "CG set-up launch");
    constexpr uint32_t kChunk = 16;  // iterations enqueued between stop-flag reads
    uint64_t enq = 0;
    while (enq < P.max_rounds) {
        const uint32_t c = (uint32_t)std::min<uint64_t>(P.max_rounds - enq, kChunk);
        for (uint32_t j = 0; j < c; ++j) {
            cg_spmv_rows<<<P.ntiles, kBlock, 0, s>>>(P);
            hip_check(hipGetLastError(), "CG SpMV launch");
            hip_check(launch_cg_vector_steps(P, s), "CG vector launches");
        }
        enq += c;
        int stop = 0;
        hip_check(hipMemcpyAsync(&stop, &P.st->stop, sizeof(int), hipMemcpyDeviceToHost, s), "read stop flag");
        hip_check(hipStreamSynchronize(s), "CG chunk");
        if (stop) break;
    }
    CgState hs;
    hip_check(hipMemcpyAsync(&hs, P.st, sizeof(CgState), hipMemcpyDeviceToHost, s), "read state");
    hip_check(hipStreamSynchronize(s), "JPCG solve");
    return StaticSolve{hs.K, hs.rr, hs.dot};
}

}  // namespace xj

extern "C" XJPCG_Status_t xjpcg_csr_solve_static(int32_t n, int64_t nnz, const int32_t* d_rowptr,
                                                 const int32_t* d_col, const double* d_val, const double* d_diag,
                                                 const double* d_b, double* d_x, uint32_t max_iter, double tol,
                                                 uint32_t* h_iter, double* h_res, int64_t* h_executed,
                                                 void* stream) {
    try {
        if (n <= 0) throw xj::invalid("Wrong dimension size.");
        if (n > xj::kStaticMaxRows)
            throw xj::invalid("n = " + std::to_string(n) + " exceeds the static workspace of " +
                              std::to_string(xj::kStaticMaxRows) + " rows; use xjpcg_op_create and xjpcg_solve.");
        if (nnz < 0 || nnz >= (int64_t)INT32_MAX) throw xj::invalid("number of stored entries must be in [0, 2^31-1).");
        if (d_rowptr == nullptr || (nnz > 0 && (d_col == nullptr || d_val == nullptr)))
            throw xj::invalid("Matrix is nullptr.");
        if (d_diag == nullptr || d_b == nullptr || d_x == nullptr) throw xj::invalid("Vector is nullptr.");
        if (max_iter == 0xFFFFFFFFu) throw xj::invalid("maxIter too large.");
        const xj::StaticSolve R = xj::csr_solve_static(n, d_rowptr, d_col, d_val, d_diag, d_b, d_x, max_iter, tol,
                                                       static_cast<hipStream_t>(stream));
        // result scan of pcgImp.hpp:171-180: count capped at maxIter; res = sqrt(rr / b.b)
        if (h_iter) *h_iter = R.executed < (int64_t)max_iter ? (uint32_t)R.executed : max_iter;
        if (h_res) *h_res = R.dot > 0.0 ? std::sqrt(R.rr / R.dot) : 0.0;
        if (h_executed) *h_executed = R.executed;
        xj::set_dev_error("");
        return XJPCG_STATUS_SUCCESS;
    } catch (const xj::Error& e) {
        xj::set_dev_error(e.what());
        return e.status;
    } catch (const std::exception& e) {
        xj::set_dev_error(std::string("INTERNAL ERROR: ") + e.what());
        return XJPCG_STATUS_INTERNAL_ERROR;
    }
}
