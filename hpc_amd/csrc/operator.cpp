// C++ host engine core (see operator.hpp).
#include "operator.hpp"

#include <cmath>
#include <cstring>

#include "kernels_consts.hpp"

namespace xj {

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) throw alloc_failed(m);
    throw internal(m);
}

std::vector<int32_t> plan_tiles(int32_t n, const int32_t* rp) {
    std::vector<int32_t> t;
    t.reserve((size_t)n / 32 + 2);
    t.push_back(0);
    int32_t r = 0;
    while (r < n) {
        const int32_t start = r;
        if (rp[r + 1] - rp[r] > kTileNnzHost) {
            r += 1;  // one long row
        } else {
            while (r < n && r - start < kTileRowsHost && rp[r + 1] - rp[start] <= kTileNnzHost) ++r;
        }
        t.push_back(r);
    }
    return t;
}

Operator::Operator(int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col, const void* val, bool val_f32,
                   const double* dsplit, const int32_t* rowptr_host, hipStream_t s) {
    if (n <= 0) throw invalid("Wrong dimension size.");
    if (nnz < 0 || nnz >= (int64_t)INT32_MAX) throw invalid("number of stored entries must be in [0, 2^31-1).");
    if (rowptr == nullptr || (nnz > 0 && (col == nullptr || val == nullptr)))
        throw invalid("Matrix is nullptr.");
    std::vector<int32_t> rp_local;
    const int32_t* rp = rowptr_host;
    if (rp == nullptr) {
        rp_local.resize((size_t)n + 1);
        hip_check(hipMemcpyAsync(rp_local.data(), rowptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s),
                  "copy row pointers");
        hip_check(hipStreamSynchronize(s), "copy row pointers");
        rp = rp_local.data();
    }
    if (rp[0] != 0 || (int64_t)rp[n] != nnz) throw invalid("row pointers do not span [0, nnz].");
    std::vector<int32_t> t = plan_tiles(n, rp);
    tiles_.ensure(t.size());
    hip_check(hipMemcpyAsync(tiles_.ptr, t.data(), sizeof(int32_t) * t.size(), hipMemcpyHostToDevice, s),
              "upload tile plan");
    A_.n = n;
    A_.nnz = nnz;
    A_.rowptr = rowptr;
    A_.col = col;
    A_.val = val;
    A_.val_f32 = val_f32;
    A_.dsplit = dsplit;
    A_.tile_row = tiles_.ptr;
    A_.ntiles = (int)t.size() - 1;
    hip_check(persistent_grid(val_f32, &G_), "persistent kernel occupancy");
    hip_check(hipHostMalloc((void**)&host_result_, 8 * sizeof(double), hipHostMallocDefault), "pinned result");
    // the vectors are only needed by solve(); allocate them now so a solve never allocates
    r_.ensure(n);
    p_.ensure(n);
    Ap_.ensure(n);
    J_.ensure(n);
    part_.ensure((size_t)5 * G_ + 4096);
    result_.ensure(8);
    bar_.ensure(1);
    abort_.ensure(1);
    hip_check(hipStreamSynchronize(s), "operator set-up");
}

Operator::~Operator() {
    if (host_result_) (void)hipHostFree(host_result_);
}

void Operator::spmv(const double* x, double* y, hipStream_t s) const {
    hip_check(launch_spmv(A_, x, y, nullptr, s), "spmv launch");
}

double Operator::dot(const double* a, const double* b, int n, hipStream_t s) {
    hip_check(launch_dot(n, a, b, part_.ptr, result_.ptr, s), "dot launch");
    hip_check(hipMemcpyAsync(host_result_, result_.ptr, sizeof(double), hipMemcpyDeviceToHost, s), "dot result");
    hip_check(hipStreamSynchronize(s), "dot");
    return host_result_[0];
}

double Operator::residual(const double* b, const double* x, double* r, hipStream_t s) {
    hip_check(launch_spmv(A_, x, r, b, s), "residual launch");
    return dot(r, r, A_.n, s);
}

void Operator::sddmm(const double* lam, const double* x, double* g, hipStream_t s) const {
    hip_check(launch_sddmm(A_, lam, x, g, s), "sddmm launch");
}

SolveResult Operator::solve(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                            void* records, hipStream_t s) {
    if (diag == nullptr || b == nullptr || x == nullptr) throw invalid("Vector is nullptr.");
    JpcgParams P;
    P.n = A_.n;
    P.ntiles = A_.ntiles;
    P.rowptr = A_.rowptr;
    P.col = A_.col;
    P.val = A_.val;
    P.dsplit = A_.dsplit;
    P.tile_row = A_.tile_row;
    P.b = b;
    P.diag = diag;
    P.x = x;
    P.r = r_.ptr;
    P.p = p_.ptr;
    P.Ap = Ap_.ptr;
    P.J = J_.ptr;
    P.part = part_.ptr;
    P.bar = bar_.ptr;
    P.abortf = abort_.ptr;
    P.records = static_cast<Record*>(records);
    P.result = result_.ptr;
    P.tol = tol;
    P.max_rounds = max_iter + 1u;  // pcgImp.hpp:158 programs maxIter+1 rounds
    if (P.max_rounds == 0) throw invalid("maxIter too large.");
    hip_check(hipMemsetAsync(bar_.ptr, 0, sizeof(unsigned long long), s), "reset barrier");
    hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset abort flag");
    if (records) hip_check(hipMemsetAsync(records, 0, (size_t)(max_iter + 2) * 64, s), "clear records");
    hip_check(launch_jpcg(P, A_.val_f32, G_, s), "persistent JPCG launch");
    hip_check(hipMemcpyAsync(host_result_, result_.ptr, 5 * sizeof(double), hipMemcpyDeviceToHost, s),
              "read solve result");
    hip_check(hipStreamSynchronize(s), "JPCG solve");
    if (host_result_[4] != 0.0) throw internal("grid barrier timed out; solve aborted.");
    SolveResult R;
    R.executed = (int64_t)host_result_[0];
    R.rr = host_result_[1];
    R.dot = host_result_[2];
    // result scan of pcgImp.hpp:171-180: count capped at maxIter
    R.iter = R.executed < (int64_t)max_iter ? (uint32_t)R.executed : max_iter;
    R.res = R.dot > 0.0 ? std::sqrt(R.rr / R.dot) : 0.0;
    return R;
}

}  // namespace xj
