// C++ host engine core: a CSR operator resident in HBM, its tile plan and
// the workspace of the persistent JPCG solver.  Shared by the pcg.h drop-in
// (pcg_api.cpp) and the device-level C ABI (dev_api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "launch.hpp"
#include "pcg.h"

namespace xj {

// Error carrying an XJPCG status; message prefixes follow the reference's
// exception classes (pcg/sw/include/impl/cgException.hpp:37-55).
struct Error : std::runtime_error {
    XJPCG_Status_t status;
    Error(XJPCG_Status_t s, const std::string& m) : std::runtime_error(m), status(s) {}
};
inline Error invalid(const std::string& m) { return Error(XJPCG_STATUS_INVALID_VALUE, "INVALID VALUE ERROR: " + m); }
inline Error internal(const std::string& m) { return Error(XJPCG_STATUS_INTERNAL_ERROR, "INTERNAL ERROR: " + m); }
inline Error alloc_failed(const std::string& m) { return Error(XJPCG_STATUS_ALLOC_FAILED, "Alloc Failed ERROR: " + m); }
inline Error exec_failed(const std::string& m) { return Error(XJPCG_STATUS_EXECUTION_FAILED, "Solver ERROR: " + m); }

void hip_check(hipError_t e, const char* what);

// message returned by xjpcg_dev_last_error() on the calling thread
void set_dev_error(const std::string& m);

// Owned device allocation.
template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void ensure(size_t n) {
        if (n <= count && ptr) return;
        release();
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void**)&ptr, n * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            // callers may catch this and carry on: clear the error, or the
            // next hipGetLastError (a launch check) would report it again
            (void)hipGetLastError();
            throw alloc_failed("device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
        }
        count = n;
    }
};

// Greedy row tiling (see tile_spmv in kernels.hpp): returns tile_row[ntiles+1].
std::vector<int32_t> plan_tiles(int32_t n, const int32_t* rowptr_host);

// Host plan of the panel-swept layout (PanelDev in launch.hpp): a stable
// counting sort of the CSR entries by (row block, column panel).  perm[q] is
// the CSR index of layout entry q.  Throws INVALID_VALUE for a column outside
// [0, n); `ordered` is false when some row's columns are not in ascending
// panel order (the layout would then change that row's summation order).
struct PanelPlan {
    int block_rows = 0, nb = 0, P = 0;
    bool ordered = true;
    std::vector<int32_t> seg;
    std::vector<uint16_t> off, lrow;
    std::vector<int64_t> perm;
};
// Host plan of the column-offset pattern tiles (tile_spmv_pattern in
// kernels.hpp) for a tile plan: false unless every tile is a short-row tile
// with at most 32 distinct column offsets and every row's columns are
// strictly ascending and inside [0, n).
bool plan_patterns(int32_t n, const int32_t* rowptr_host, const int32_t* col_host, const std::vector<int32_t>& tiles,
                   std::vector<uint32_t>& mask, std::vector<int32_t>& offs);

int panel_block_rows(int32_t n);
int64_t panel_segments(int32_t n, int block_rows);
PanelPlan plan_panels(int32_t n, const int32_t* rowptr_host, const int32_t* col_host, int block_rows);

struct SolveResult {
    int64_t executed = 0;  // iterations run (K)
    uint32_t iter = 0;     // reported count, min(K, maxIter)
    double rr = 0.0;       // last recursive r.r
    double dot = 0.0;      // b.b
    double res = 0.0;      // sqrt(rr/dot), 0 when b == 0
};

class Operator {
  public:
    // References caller device memory for rowptr/col/val/dsplit (not copied);
    // rowptr_host (n+1 entries) is used to plan tiles and may be null, in which
    // case rowptr is copied back from the device.  When every tile's columns
    // lie within 32767 of its first row, an int16 copy of the column stream
    // (col - tile's first row) is built and the SpMV kernels read it instead.
    Operator(int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col, const void* val, bool val_f32,
             const double* dsplit, const int32_t* rowptr_host, hipStream_t s);
    ~Operator();
    Operator(const Operator&) = delete;
    Operator& operator=(const Operator&) = delete;

    const CsrDev& csr() const { return A_; }
    int grid() const { return G_; }
    bool compact() const { return A_.col16 != nullptr; }
    // point the operator at another value array of the same type and length
    // (the pattern and tile plan stay); used when only A's values change, and
    // with the same pointer after values changed in place, so that the panel
    // layout's own copy of the values is refreshed before its next use
    void set_values(const void* val) {
        if (val == nullptr && A_.nnz > 0) throw invalid("Matrix is nullptr.");
        if (reinterpret_cast<uintptr_t>(val) & 15u) throw invalid("value array must be 16-byte aligned.");
        A_.val = val;
        pan_stale_ = true;
        dia_stale_ = true;
        drop_graph();
    }
    // SpMV layout: kTiles (row tiles over the CSR arrays) or kPanels (the
    // panel-swept copy, for columns scattered over x).  kLayoutAuto restores
    // the choice made at creation: the fastest layout measured (panels are
    // built when the columns do not fit the int16 stream).
    // kPatterns runs the row tiles on column-offset patterns (stencil
    // operators; bit-identical to kTiles), chosen at creation when faster.
    // kDiagonals runs on one value array per global column offset (stencils
    // with at most 32 distinct offsets over the whole matrix).
    enum Layout { kLayoutAuto = 0, kTiles = 1, kPanels = 2, kPatterns = 3, kDiagonals = 4 };
    void set_layout(int layout, hipStream_t s);
    // layout of the standalone SpMV / residual, and of the solves
    int layout() const {
        return A_.pan_on ? kPanels : A_.pat_on ? kPatterns : A_.dia_on ? kDiagonals : kTiles;
    }
    int solve_layout() const { return solve_layout_; }
    // Bytes one standalone SpMV launch on `layout` (0: the one in use) must
    // read and write: the layout's matrix streams (zero padding included),
    // its per-row tables, x once and y once; -1 when the layout is not built.
    int64_t spmv_bytes(int layout) const;
    // use (or stop using) the int16 column stream when it exists
    void set_compact(bool on);

    void spmv(const double* x, double* y, hipStream_t s) const;
    // r = b - A x; returns r.r (synchronises the stream)
    double residual(const double* b, const double* x, double* r, hipStream_t s);
    double dot(const double* a, const double* b, int n, hipStream_t s);
    void sddmm(const double* lam, const double* x, double* g, hipStream_t s) const;
    // extra n-vectors owned by the operator (refinement residual and correction)
    double* aux(int i) {
        aux_[i].ensure((size_t)A_.n);
        return aux_[i].ptr;
    }

    // Jacobi PCG from x0 = 0; records (device, (maxIter+2)*64 bytes) may be null.
    SolveResult solve(const double* diag, const double* b, double* x, uint32_t max_iter, double tol, void* records,
                      hipStream_t s);

    // Which loop solve() runs: kPersistent is one cooperative launch with grid
    // barriers, kMultiKernel is three or four launches per iteration replayed
    // from a captured graph; kAuto picks the faster measured one (the
    // multi-kernel loop at every size measured so far).
    enum Path { kAuto = 0, kPersistent = 1, kMultiKernel = 2 };
    void set_path(int p) { path_ = p; }
    int path() const { return path_; }
    int last_path() const { return last_path_; }

    // Work vectors of the solves: 64 (fp64) or 32 (r, p, Ap and the Jacobi
    // factors stored in fp32; x, the arithmetic and the dot products stay
    // fp64).  32 needs the diagonal or the panel solve layout and runs the
    // multi-kernel loop; while another layout is in use the solves run with
    // fp64 vectors.
    void set_vector_bits(int bits);
    int vector_bits() const {
        return w32_ && (solve_layout_ == kDiagonals || solve_layout_ == kPanels) ? 32 : 64;
    }

  private:
    SolveResult solve_persistent(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                                 void* records, hipStream_t s);
    SolveResult solve_multi(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                            void* records, hipStream_t s);
    CgParams cg_params(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                       void* records) const;
    void build_graph(const CgParams& P);
    void drop_graph();
    // panel layout: build (host plan + upload), release, refresh values, time
    bool build_panels(hipStream_t s, const int32_t* rp_host = nullptr, const int32_t* col_host = nullptr);
    bool build_patterns(hipStream_t s);
    bool build_diagonals(hipStream_t s);
    void free_diagonals();
    // the panel and diagonal layouts' own value copies follow A_.val before use
    void refresh_copies(hipStream_t s) const;
    void copy_pattern_to_host(hipStream_t s, std::vector<int32_t>& rp, std::vector<int32_t>& col) const;
    float time_layout(int layout, hipStream_t s);
    void free_panels();
    void free_patterns();
    void refresh_panels(hipStream_t s) const;

    const void* col_stream() const { return A_.col16 ? (const void*)A_.col16 : (const void*)A_.col; }

    // Iterations per captured graph: a chunk boundary costs a graph launch and
    // an event, and after convergence up to a chunk of launches return at
    // once.  Measured on one box: 32 iterations take the 256^2 Laplacian from
    // 9.89 to 9.27 ms and cost the 160^3 refined solve 43.9 -> 44.5 ms, so
    // operators up to 2^20 rows use 32, larger ones 8.
    static constexpr int chunk_for(int n) { return n <= (1 << 20) ? 32 : 8; }
    int chunk_ = 8;
    // panel layout considered from this many rows (x = 2 MiB, beyond what the
    // tile kernel's gathers keep in L2 once columns scatter)
    static constexpr int32_t kPanelAutoRows = 1 << 18;
    static constexpr int64_t kPanelMaxSegments = 1ll << 26;

    CsrDev A_;
    DevBuf<int32_t> tiles_, tilek_;
    DevBuf<int16_t> col16_;
    bool has_col16_ = false;
    int path_ = kAuto, last_path_ = 0;
    bool coop_ok_ = true;   // false once the runtime refused the cooperative launch
    bool graph_ok_ = true;  // false once graph capture failed; iterations are then launched directly
    bool w32_ = false;      // fp32 work vectors requested
    int G_ = 0, G2_ = 0;
    DevBuf<double> r_, p_, Ap_, J_, part_, result_;
    DevBuf<double> partA_, partA2_, partB_, partC_;
    DevBuf<CgState> state_;
    DevBuf<unsigned long long> bar_;
    DevBuf<int> abort_;
    DevBuf<double> aux_[2];
    DevBuf<int32_t> pseg_;
    DevBuf<uint16_t> poff_, plrow_;
    DevBuf<int64_t> pperm_;
    DevBuf<double> pval_;            // nnz values of the operator's type, in panel order
    bool has_pan_ = false;
    mutable bool pan_stale_ = false; // panel values differ from A_.val
    std::vector<int32_t> tile_plan_;  // host copy of the tile plan
    DevBuf<uint32_t> pmask_;
    DevBuf<int32_t> poffs_;
    bool has_pat_ = false;
    bool pat_tried_ = false;          // the pattern plan was attempted (and failed if !has_pat_)
    DevBuf<int32_t> doff_;
    DevBuf<uint32_t> dmask_;
    DevBuf<double> dval_;             // ndia*n values of the operator's type (fp32 ones packed two per double)
    bool has_dia_ = false;
    bool dia_tried_ = false;
    mutable bool dia_stale_ = false;  // diagonal values differ from A_.val
    int auto_layout_ = kTiles;        // creation-time choice for the standalone SpMV
    int auto_solve_ = kTiles;         // creation-time choice for the solves
    int solve_layout_ = kTiles;
    double* host_result_ = nullptr;  // pinned
    int* host_flags_ = nullptr;      // pinned
    hipStream_t work_ = nullptr;     // graphs are captured and replayed here
    hipEvent_t ev_[2] = {nullptr, nullptr};
    hipEvent_t ev_join_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    CgParams gkey_;                  // arguments the captured graph was built with
};

}  // namespace xj
