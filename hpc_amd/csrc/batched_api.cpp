// Batched JPCG host driver and its C ABI (include/xjpcg_batch.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "batched.hpp"
#include "operator.hpp"
#include "xjpcg_batch.h"

namespace xj {

// Two timing events, destroyed on every path out of the scope.
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    EventPair() {
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    float ms() const {
        float t = 0.f;
        hip_check(hipEventElapsedTime(&t, e0, e1), "event");
        return t;
    }
};

class Batch {
  public:
    Batch(int S, int R, const int32_t* h_sys_row, const int32_t* h_rowptr, const int32_t* d_rowptr,
          const int32_t* d_col, const void* d_val, bool f32, const double* d_dsplit, hipStream_t user) {
        if (S <= 0) throw invalid("number of systems must be positive.");
        if (!(R == 1 || R == 2 || R == 4 || R == 8 || R == 16 || R == 32))
            throw invalid("number of right-hand sides must be 1, 2, 4, 8, 16 or 32.");
        if (h_sys_row == nullptr || h_rowptr == nullptr || d_rowptr == nullptr) throw invalid("Matrix is nullptr.");
        if (h_sys_row[0] != 0) throw invalid("system row ranges must start at 0.");
        for (int s = 0; s < S; ++s)
            if (h_sys_row[s + 1] < h_sys_row[s]) throw invalid("system row ranges must not decrease.");
        N_ = h_sys_row[S];
        if (N_ <= 0) throw invalid("Wrong dimension size.");
        if (h_rowptr[0] != 0 || h_rowptr[N_] < 0) throw invalid("row pointers do not start at 0.");
        nnz_ = h_rowptr[N_];
        if (nnz_ > 0 && (d_col == nullptr || d_val == nullptr)) throw invalid("Matrix is nullptr.");
        S_ = S;
        R_ = R;
        f32_ = f32;
        sys_row_.assign(h_sys_row, h_sys_row + S + 1);
        P_.N = N_;
        P_.S = S;
        P_.rowptr = d_rowptr;
        P_.col = d_col;
        P_.val = d_val;
        P_.dsplit = d_dsplit;
        hip_check(hipStreamCreateWithFlags(&work_, hipStreamNonBlocking), "batch stream");
        hip_check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming), "event");
        hip_check(hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming), "event");
        hip_check(hipHostMalloc((void**)&host_flags_, 2 * sizeof(int), hipHostMallocDefault), "pinned flags");
        const size_t NR = (size_t)N_ * R;
        r_.ensure(NR);
        p_.ensure(NR);
        Ap_.ensure(NR);
        J_.ensure(N_);
        const size_t SR = (size_t)S * R;
        dot_.ensure(SR);
        tols_.ensure(SR);
        rz_.ensure(SR);
        rr_.ensure(SR);
        alpha_.ensure(SR);
        beta_.ensure(SR);
        active_.ensure(SR);
        upd_.ensure(SR);
        iters_.ensure(SR);
        it_.ensure(SR);
        sys_active_.ensure(S);
        sys_upd_.ensure(S);
        any_.ensure(1);
        sysrow_.ensure((size_t)S + 1);
        hip_check(hipMemcpyAsync(sysrow_.ptr, h_sys_row, sizeof(int32_t) * ((size_t)S + 1), hipMemcpyHostToDevice,
                                 work_),
                  "system rows");
        configure(default_tile_rows(), 256, user);
        // per-system diagonal layout when every system has at most kBDiaMax
        // offsets; kept when its SpMM is the faster one
        if (nnz_ > 0 && build_dia()) {
            dia_on_ = false;
            const float t_csr = time_spmm();
            dia_on_ = true;
            const float t_dia = time_spmm();
            dia_on_ = t_dia < t_csr;
            dia_auto_ = dia_on_;
        }
    }

    ~Batch() {
        drop_graph();
        if (host_flags_) (void)hipHostFree(host_flags_);
        for (hipEvent_t e : ev_)
            if (e) (void)hipEventDestroy(e);
        if (ev_join_) (void)hipEventDestroy(ev_join_);
        if (work_) (void)hipStreamDestroy(work_);
    }

    int default_tile_rows() const { return std::max(16, std::min(1024, 4096 / R_)); }

    void configure(int tile_rows, int block, hipStream_t user) {
        if (tile_rows < 16 || tile_rows > 1024) throw invalid("tile_rows must be in [16, 1024].");
        if (!(block == 64 || block == 128 || block == 256 || block == 512) || block < R_)
            throw invalid("block must be 64, 128, 256 or 512 threads and at least nrhs.");
        std::vector<int32_t> tr, ts, st;
        st.push_back(0);
        tr.push_back(0);
        for (int s = 0; s < S_; ++s) {
            for (int r = sys_row_[s]; r < sys_row_[s + 1]; r += tile_rows) {
                tr.push_back(std::min(r + tile_rows, sys_row_[s + 1]));
                ts.push_back(s);
            }
            st.push_back((int32_t)ts.size());
        }
        const int T = (int)ts.size();
        sys_tile_h_ = st;
        sync_user(user);
        tile_row_.ensure(tr.size());
        tile_sys_.ensure(std::max<size_t>(1, ts.size()));
        sys_tile_.ensure(st.size());
        hip_check(hipMemcpyAsync(tile_row_.ptr, tr.data(), tr.size() * 4, hipMemcpyHostToDevice, work_), "tiles");
        if (T) hip_check(hipMemcpyAsync(tile_sys_.ptr, ts.data(), ts.size() * 4, hipMemcpyHostToDevice, work_), "tiles");
        hip_check(hipMemcpyAsync(sys_tile_.ptr, st.data(), st.size() * 4, hipMemcpyHostToDevice, work_), "tiles");
        partA_.ensure((size_t)std::max(1, T) * R_);
        partB_.ensure((size_t)std::max(1, T) * R_);
        partC_.ensure((size_t)std::max(1, T) * R_);
        hip_check(hipStreamSynchronize(work_), "batch tile plan");
        T_ = T;
        tile_rows_ = tile_rows;
        block_ = block;
        drop_graph();
    }

    // Mean time of 5 batched SpMM launches (after 2 untimed ones) in the
    // current layout and tile configuration, on the work stream.
    float time_spmm() {
        double* xs = p_.ptr;
        double* ys = Ap_.ptr;
        hip_check(hipMemsetAsync(xs, 0, (size_t)N_ * R_ * sizeof(double), work_), "tune input");
        EventPair ev;
        BatchParams P = params(nullptr, nullptr, nullptr, 0.0, 1);
        for (int w = 0; w < 2; ++w) hip_check(batch_launch_spmm(P, R_, f32_, block_, xs, ys, nullptr, work_), "tune");
        hip_check(hipEventRecord(ev.e0, work_), "event");
        for (int w = 0; w < 5; ++w) hip_check(batch_launch_spmm(P, R_, f32_, block_, xs, ys, nullptr, work_), "tune");
        hip_check(hipEventRecord(ev.e1, work_), "event");
        hip_check(hipEventSynchronize(ev.e1), "event");
        return ev.ms() / 5.0f;
    }

    // Mean time of a whole batched iteration (SpMM, per-system alpha, r
    // update, beta, x/p update) in the current layout and tile configuration:
    // set-up on b = 1 (held in r) and a unit diagonal, 2 untimed iterations,
    // then 5 timed.  busy: tolerance 0, every pair runs; otherwise an
    // infinite tolerance, every pair stops in the first iteration and the
    // timed ones are the bare launches of a batch that has converged.  x and
    // the diagonal live in `v` (N*R + N doubles).
    float time_iteration(double* v, bool busy) {
        double* x = v;
        double* d = v + (size_t)N_ * R_;
        // b = 1 lives in r_ for fp64 work vectors (the set-up reads b[e] before
        // it writes r[e] = b[e]); fp32 work vectors are narrower than b, so b
        // then lives in Ap_, which the set-up does not touch
        double* bsrc = w32_ ? Ap_.ptr : r_.ptr;
        hip_check(launch_fill((int64_t)N_ * R_, 1.0, bsrc, work_), "tune input");
        BatchParams P = params(d, bsrc, x, busy ? 0.0 : INFINITY, 0xFFFFFFF0u);
        hip_check(hipMemsetAsync(any_.ptr, 0, sizeof(int), work_), "reset flag");
        hip_check(batch_launch_setup(P, R_, w32_, block_, work_), "tune set-up");
        EventPair ev;
        for (int w = 0; w < 2; ++w) hip_check(batch_launch_iteration(P, R_, f32_, w32_, block_, work_), "tune");
        hip_check(hipEventRecord(ev.e0, work_), "event");
        for (int w = 0; w < 5; ++w) hip_check(batch_launch_iteration(P, R_, f32_, w32_, block_, work_), "tune");
        hip_check(hipEventRecord(ev.e1, work_), "event");
        hip_check(hipEventSynchronize(ev.e1), "event");
        return ev.ms() / 5.0f;
    }

    // The autotuner: time whole batched iterations for each candidate
    // (layout, tile rows, block) and keep the one with the lowest
    // 0.25 * busy + 0.75 * idle iteration time.  The tile and block sizes
    // drive the vector kernels as well as the SpMM (together most of an
    // iteration), so timing the SpMM alone picked shapes that slowed the
    // iteration down; and a batch spends most of its iterations with few
    // pairs left (pair iterations 77..2184 in the 512-system workload), where
    // the cost is launching one workgroup per tile: timing busy iterations
    // alone once picked 32-row tiles (1.4M workgroups per launch) and lost 9%.
    void autotune(hipStream_t user) {
        static const int tiles[] = {32, 64, 128, 256, 512};
        static const int blocks[] = {128, 256, 512};
        sync_user(user);
        // scratch x and diagonal for the timing iterations; when the device
        // cannot hold them next to the batch, candidates are timed on the
        // SpMM alone
        DevBuf<double> tmp;
        bool whole = true;
        try {
            tmp.ensure((size_t)N_ * R_ + (size_t)N_);
        } catch (const Error& e) {
            if (e.status != XJPCG_STATUS_ALLOC_FAILED) throw;
            whole = false;
        }
        if (whole) hip_check(launch_fill((int64_t)N_ * R_ + N_, 1.0, tmp.ptr, work_), "tune input");
        float best = 1e30f;
        int bt = tile_rows_, bb = block_;
        bool bd = dia_on_;
        for (int lay = 0; lay < (has_dia_ ? 2 : 1); ++lay) {
            dia_on_ = lay == 1;
            for (int tr : tiles) {
                for (int bl : blocks) {
                    if (bl < R_ || bl / R_ > tr) continue;  // at most one row per lane group pass beyond the tile
                    configure(tr, bl, work_);
                    const float ms = whole ? 0.25f * time_iteration(tmp.ptr, true) + 0.75f * time_iteration(tmp.ptr, false)
                                           : time_spmm();
                    if (ms < best) {
                        best = ms;
                        bt = tr;
                        bb = bl;
                        bd = dia_on_;
                    }
                }
            }
        }
        dia_on_ = bd;
        dia_auto_ = bd;
        configure(bt, bb, work_);
    }

    // 0: keep the creation-time choice, 1: CSR, 2: per-system diagonals
    void set_layout(int layout, hipStream_t user) {
        if (layout < 0 || layout > 2) throw invalid("layout must be 0 (auto), 1 (CSR) or 2 (diagonals).");
        if (layout == 2 && !has_dia_)
            throw invalid("the diagonal layout needs at most 8 distinct offsets (col - row) per system, columns "
                          "inside the row's system and strictly ascending in every row.");
        sync_user(user);
        const bool want = layout == 0 ? dia_auto_ : layout == 2;
        if (want != dia_on_) {
            dia_on_ = want;
            drop_graph();
        }
    }
    int layout() const { return dia_on_ ? 2 : 1; }

    int64_t solve(const double* diag, const double* b, double* x, uint32_t max_iter, double tol, int32_t* d_iters,
                  double* d_res, hipStream_t user) {
        if (diag == nullptr || b == nullptr || x == nullptr) throw invalid("Vector is nullptr.");
        if (max_iter == 0xFFFFFFFFu) throw invalid("maxIter too large.");
        const uint32_t rounds = max_iter + 1u;
        sync_user(user);
        BatchParams P = params(diag, b, x, tol, rounds);
        hip_check(hipMemsetAsync(any_.ptr, 0, sizeof(int), work_), "reset flag");
        hip_check(batch_launch_setup(P, R_, w32_, block_, work_), "batch set-up");
        if (path_ == kPathPersistent) {
            int64_t launched = 0;
            if (solve_persistent(P, rounds, &launched)) {
                last_path_ = kPathPersistent;
                hip_check(batch_launch_finish(S_ * R_, dot_.ptr, rr_.ptr, iters_.ptr, d_iters, d_res, work_),
                          "finish");
                join_user(user);
                hip_check(hipStreamSynchronize(work_), "batch solve");
                return launched;
            }
        }
        last_path_ = kPathLaunches;
        // iterations run as replays of a captured graph of kChunk iterations
        if (!gexec_ || g_x_ != x || g_rounds_ != rounds) build_graph(P);
        int64_t launched = 0;
        int slot = 0;
        bool pending = false;
        while ((uint64_t)launched < rounds) {
            const uint32_t left = rounds - (uint32_t)launched;
            if (left >= (uint32_t)kChunk) {
                hip_check(hipGraphLaunch(gexec_, work_), "graph launch");
                launched += kChunk;
            } else {
                for (uint32_t c = 0; c < left; ++c)
                    hip_check(batch_launch_iteration(P, R_, f32_, w32_, block_, work_), "batch iteration");
                launched += left;
            }
            hip_check(hipMemcpyAsync(&host_flags_[slot], any_.ptr, sizeof(int), hipMemcpyDeviceToHost, work_),
                      "read flag");
            hip_check(hipEventRecord(ev_[slot], work_), "event");
            if (pending) {
                hip_check(hipEventSynchronize(ev_[slot ^ 1]), "event sync");
                if (!host_flags_[slot ^ 1]) break;
            }
            pending = true;
            slot ^= 1;
        }
        hip_check(batch_launch_finish(S_ * R_, dot_.ptr, rr_.ptr, iters_.ptr, d_iters, d_res, work_), "finish");
        join_user(user);
        hip_check(hipStreamSynchronize(work_), "batch solve");
        return launched;
    }

    // 0 / 2: iterations as graph-replayed launches (the default); 1: one
    // persistent cooperative launch of `grid` workgroups (0: the default size)
    void set_path(int path, int grid) {
        if (path < 0 || path > 2) throw invalid("path must be 0 (auto), 1 (persistent) or 2 (launches).");
        if (grid < 0) throw invalid("grid must be 0 (default) or a positive number of workgroups.");
        path_ = path == kPathPersistent ? kPathPersistent : kPathLaunches;
        grid_ = grid;
    }
    int last_path() const { return last_path_; }
    int last_grid() const { return last_grid_; }

    // 64: work vectors in fp64 (the default); 32: r, p and Ap stored in fp32
    // with x, the arithmetic and the dot products in fp64 (the inner batch of
    // a refinement, where they cut the vector traffic by about 40%)
    void set_vector_bits(int bits) {
        if (bits != 32 && bits != 64) throw invalid("vector bits must be 32 or 64.");
        if ((bits == 32) != w32_) {
            w32_ = bits == 32;
            drop_graph();
        }
    }
    int vector_bits() const { return w32_ ? 32 : 64; }

    // y = A x or y = b - A x on the caller's stream order
    void spmm(const double* x, double* y, const double* b, hipStream_t user) {
        sync_user(user);
        BatchParams P = params(nullptr, nullptr, nullptr, 0.0, 1);
        hip_check(batch_launch_spmm(P, R_, f32_, block_, x, y, b, work_), "batch spmm");
        join_user(user);
    }

    // out[s*R+j] = ||v_{s,j}||^2 (device)
    void norms(const double* v, double* out, hipStream_t user) {
        sync_user(user);
        BatchParams P = params(nullptr, nullptr, nullptr, 0.0, 1);
        hip_check(batch_launch_norms(P, R_, block_, v, partA_.ptr, out, work_), "batch norms");
        join_user(user);
    }

    void mask(const int* keep, double* v, hipStream_t user) {
        sync_user(user);
        BatchParams P = params(nullptr, nullptr, nullptr, 0.0, 1);
        hip_check(batch_launch_mask(P, R_, block_, keep, v, work_), "batch mask");
        join_user(user);
    }

    int N() const { return N_; }
    int S() const { return S_; }
    int R() const { return R_; }
    int T() const { return T_; }
    int nnz() const { return nnz_; }
    int tile_rows() const { return tile_rows_; }
    int block() const { return block_; }
    double* scratch(int i) {
        aux_[i].ensure((size_t)N_ * R_);
        return aux_[i].ptr;
    }

  private:
    static constexpr int kChunk = 16;
    static constexpr int kPathPersistent = 1, kPathLaunches = 2;

    // The iteration loop as one cooperative launch (set-up already queued on
    // work_).  Returns false, with nothing launched, when the runtime refuses
    // the cooperative launch (the caller then runs the launch-per-step loop);
    // throws when a grid barrier timed out.
    bool solve_persistent(const BatchParams& P, uint32_t rounds, int64_t* launched) {
        int gmax = 0;
        hip_check(batch_persistent_grid(R_, f32_, w32_, dia_on_, block_, &gmax), "persistent batch occupancy");
        int G = std::min(gmax, std::max(T_, S_));
        if (grid_ > 0) G = std::min(G, grid_);
        if (G < 1) return false;
        loop_bar_.ensure(1);
        loop_abort_.ensure(1);
        loop_res_.ensure(2);
        hip_check(hipMemsetAsync(loop_bar_.ptr, 0, sizeof(unsigned long long), work_), "reset barrier");
        hip_check(hipMemsetAsync(loop_abort_.ptr, 0, sizeof(int), work_), "reset abort flag");
        BatchLoop L;
        L.bar = loop_bar_.ptr;
        L.abortf = loop_abort_.ptr;
        L.result = loop_res_.ptr;
        L.rounds = rounds;
        const hipError_t e = batch_launch_persistent(P, L, R_, f32_, w32_, block_, G, work_);
        if (e == hipErrorCooperativeLaunchTooLarge || e == hipErrorNotSupported) {
            (void)hipGetLastError();
            return false;
        }
        hip_check(e, "persistent batch launch");
        long long res[2] = {0, 0};
        hip_check(hipMemcpyAsync(res, loop_res_.ptr, sizeof(res), hipMemcpyDeviceToHost, work_), "loop result");
        hip_check(hipStreamSynchronize(work_), "persistent batch");
        if (res[1]) throw Error(XJPCG_STATUS_INTERNAL_ERROR, "grid barrier timed out; batched solve aborted.");
        last_grid_ = G;
        *launched = res[0];
        return true;
    }

    BatchParams params(const double* diag, const double* b, double* x, double tol, uint32_t rounds) {
        BatchParams P = P_;
        P.T = T_;
        P.tile_row = tile_row_.ptr;
        P.tile_sys = tile_sys_.ptr;
        P.sys_tile = sys_tile_.ptr;
        P.b = b;
        P.diag = diag;
        P.x = x;
        P.r = r_.ptr;
        P.p = p_.ptr;
        P.Ap = Ap_.ptr;
        P.J = J_.ptr;
        P.partA = partA_.ptr;
        P.partB = partB_.ptr;
        P.partC = partC_.ptr;
        P.dot = dot_.ptr;
        P.tols = tols_.ptr;
        P.rz = rz_.ptr;
        P.rr = rr_.ptr;
        P.alpha = alpha_.ptr;
        P.beta = beta_.ptr;
        P.active = active_.ptr;
        P.upd = upd_.ptr;
        P.iters = iters_.ptr;
        P.it = it_.ptr;
        P.sys_active = sys_active_.ptr;
        P.sys_upd = sys_upd_.ptr;
        P.any_active = any_.ptr;
        P.sys_row = sysrow_.ptr;
        P.dia_on = dia_on_ ? 1 : 0;
        P.dsys = dsys_.ptr;
        P.dval = dval_.ptr;
        P.dmask = dmask_.ptr;
        P.tol = tol;
        P.max_rounds = rounds;
        return P;
    }

    // Per-system offsets from a per-tile plan on the GPU (union per system on
    // the host), then the value arrays and masks on the GPU.  Refused (false)
    // beyond kBDiaMax offsets in a system, for columns outside the row's system
    // or not strictly ascending, or value arrays beyond twice the stored entries.
    bool build_dia() {
        BatchParams P = params(nullptr, nullptr, nullptr, 0.0, 1);
        std::vector<int32_t> toffs((size_t)T_ * 16);
        {
            DevBuf<int32_t> dto;
            DevBuf<int> fail;
            dto.ensure(toffs.size());
            fail.ensure(1);
            hip_check(hipMemsetAsync(fail.ptr, 0, sizeof(int), work_), "reset flag");
            hip_check(batch_launch_tile_offsets(P, dto.ptr, fail.ptr, work_), "tile offsets");
            int f = 1;
            hip_check(hipMemcpyAsync(toffs.data(), dto.ptr, sizeof(int32_t) * toffs.size(), hipMemcpyDeviceToHost,
                                     work_),
                      "tile offsets");
            hip_check(hipMemcpyAsync(&f, fail.ptr, sizeof(int), hipMemcpyDeviceToHost, work_), "tile offsets");
            hip_check(hipStreamSynchronize(work_), "tile offsets");
            if (f) return false;
        }
        std::vector<BDiaSys> sys((size_t)S_);
        int64_t total = 0;
        for (int s = 0; s < S_; ++s) {
            std::vector<int32_t> u;
            for (int t = sys_tile_h_[s]; t < sys_tile_h_[s + 1]; ++t)
                for (int j = 0; j < 16; ++j)
                    if (toffs[(size_t)t * 16 + j] != INT32_MIN) u.push_back(toffs[(size_t)t * 16 + j]);
            std::sort(u.begin(), u.end());
            u.erase(std::unique(u.begin(), u.end()), u.end());
            if ((int)u.size() > kBDiaMax) return false;
            BDiaSys& Y = sys[s];
            std::memset(&Y, 0, sizeof(Y));
            for (size_t d = 0; d < u.size(); ++d) Y.off[d] = u[d];
            Y.r0 = sys_row_[s];
            Y.r1 = sys_row_[s + 1];
            Y.nd = (int)u.size();
            Y.base = total;
            total += (int64_t)Y.nd * (Y.r1 - Y.r0);
        }
        if (total > 2 * (int64_t)nnz_ + N_) return false;
        const size_t vbytes = (size_t)total * (f32_ ? 4 : 8);
        try {
            dsys_.ensure(sys.size());
            dval_.ensure(std::max<size_t>(1, (vbytes + 7) / 8));
            dmask_.ensure((size_t)N_);
        } catch (const Error& e) {
            if (e.status != XJPCG_STATUS_ALLOC_FAILED) throw;
            dsys_.release();
            dval_.release();
            dmask_.release();
            return false;
        }
        hip_check(hipMemcpyAsync(dsys_.ptr, sys.data(), sizeof(BDiaSys) * sys.size(), hipMemcpyHostToDevice, work_),
                  "system diagonals");
        hip_check(hipMemsetAsync(dval_.ptr, 0, vbytes, work_), "clear diagonals");
        DevBuf<int> fail;
        fail.ensure(1);
        hip_check(hipMemsetAsync(fail.ptr, 0, sizeof(int), work_), "reset flag");
        P = params(nullptr, nullptr, nullptr, 0.0, 1);
        hip_check(batch_launch_fill_dia(P, f32_, dval_.ptr, dmask_.ptr, fail.ptr, work_), "fill diagonals");
        int f = 1;
        hip_check(hipMemcpyAsync(&f, fail.ptr, sizeof(int), hipMemcpyDeviceToHost, work_), "fill diagonals");
        hip_check(hipStreamSynchronize(work_), "fill diagonals");  // host tables end with this scope
        if (f) {
            dsys_.release();
            dval_.release();
            dmask_.release();
            return false;
        }
        has_dia_ = true;
        return true;
    }

    void build_graph(const BatchParams& P) {
        drop_graph();
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(work_, hipStreamCaptureModeThreadLocal), "begin capture");
        for (int c = 0; c < kChunk; ++c) {
            hipError_t e = batch_launch_iteration(P, R_, f32_, w32_, block_, work_);
            if (e != hipSuccess) {
                (void)hipStreamEndCapture(work_, &g);
                if (g) (void)hipGraphDestroy(g);
                hip_check(e, "capture iteration");
            }
        }
        hip_check(hipStreamEndCapture(work_, &g), "end capture");
        hipError_t e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "instantiate graph");
        g_x_ = P.x;
        g_rounds_ = P.max_rounds;
    }

    void drop_graph() {
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
        g_x_ = nullptr;
        g_rounds_ = 0;
    }

    // work_ waits for everything already queued on the caller's stream
    void sync_user(hipStream_t user) {
        hip_check(hipEventRecord(ev_join_, user), "event");
        hip_check(hipStreamWaitEvent(work_, ev_join_, 0), "stream wait");
    }
    // the caller's stream waits for work_
    void join_user(hipStream_t user) {
        hip_check(hipEventRecord(ev_join_, work_), "event");
        hip_check(hipStreamWaitEvent(user, ev_join_, 0), "stream wait");
    }

    int N_ = 0, S_ = 0, R_ = 0, T_ = 0, nnz_ = 0, tile_rows_ = 0, block_ = 256;
    bool f32_ = false;
    bool w32_ = false;               // fp32 work vectors r, p, Ap
    std::vector<int32_t> sys_row_;
    BatchParams P_;
    hipStream_t work_ = nullptr;
    hipEvent_t ev_join_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    int* host_flags_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
    double* g_x_ = nullptr;
    uint32_t g_rounds_ = 0;
    DevBuf<int32_t> tile_row_, tile_sys_, sys_tile_, sysrow_;
    std::vector<int32_t> sys_tile_h_;
    DevBuf<BDiaSys> dsys_;
    DevBuf<double> dval_;            // per-system diagonals of the value type (fp32 ones packed two per double)
    DevBuf<uint8_t> dmask_;
    bool has_dia_ = false, dia_on_ = false, dia_auto_ = false;
    DevBuf<double> r_, p_, Ap_, J_, partA_, partB_, partC_;
    DevBuf<double> dot_, tols_, rz_, rr_, alpha_, beta_;
    DevBuf<int> active_, upd_, iters_, it_, sys_active_, sys_upd_, any_;
    DevBuf<double> aux_[3];
    int path_ = kPathLaunches, grid_ = 0, last_path_ = 0, last_grid_ = 0;
    DevBuf<unsigned long long> loop_bar_;
    DevBuf<int> loop_abort_;
    DevBuf<long long> loop_res_;
};

}  // namespace xj

struct xjpcg_batch {
    xj::Batch* impl;
};

namespace {
thread_local std::string g_batch_error;

// smallest relative tolerance a refinement step asks of an inner batch with fp32 work vectors
constexpr double kW32StepTol = 1e-7;

template <typename F>
XJPCG_Status_t bguard(F f) {
    try {
        f();
        return XJPCG_STATUS_SUCCESS;
    } catch (const xj::Error& e) {
        g_batch_error = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_batch_error = std::string("INTERNAL ERROR: ") + e.what();
        return XJPCG_STATUS_INTERNAL_ERROR;
    }
}

xj::Batch& impl_of(xjpcg_batch_t* b) {
    if (b == nullptr || b->impl == nullptr) throw xj::Error(XJPCG_STATUS_NOT_INITIALIZED, "null batch");
    return *b->impl;
}
}  // namespace

extern "C" const char* xjpcg_batch_last_error(void) { return g_batch_error.c_str(); }

extern "C" XJPCG_Status_t xjpcg_batch_create(xjpcg_batch_t** batch, int32_t nsys, int32_t nrhs,
                                             const int32_t* h_sys_row, const int32_t* h_rowptr,
                                             const int32_t* d_rowptr, const int32_t* d_col, const void* d_val,
                                             int val_f32, const double* d_dsplit, void* stream) {
    return bguard([&] {
        if (batch == nullptr) throw xj::invalid("null output pointer.");
        *batch = nullptr;
        auto* impl = new xj::Batch(nsys, nrhs, h_sys_row, h_rowptr, d_rowptr, d_col, d_val, val_f32 != 0, d_dsplit,
                                   static_cast<hipStream_t>(stream));
        *batch = new xjpcg_batch{impl};
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_destroy(xjpcg_batch_t* batch) {
    if (batch == nullptr) return XJPCG_STATUS_NOT_INITIALIZED;
    delete batch->impl;
    delete batch;
    return XJPCG_STATUS_SUCCESS;
}

extern "C" XJPCG_Status_t xjpcg_batch_info(const xjpcg_batch_t* batch, int32_t* nrows, int32_t* nnz,
                                           int32_t* ntiles, int32_t* tile_rows, int32_t* block) {
    return bguard([&] {
        xj::Batch& B = impl_of(const_cast<xjpcg_batch_t*>(batch));
        if (nrows) *nrows = B.N();
        if (nnz) *nnz = B.nnz();
        if (ntiles) *ntiles = B.T();
        if (tile_rows) *tile_rows = B.tile_rows();
        if (block) *block = B.block();
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_set_config(xjpcg_batch_t* batch, int32_t tile_rows, int32_t block,
                                                 int32_t tune) {
    return bguard([&] {
        xj::Batch& B = impl_of(batch);
        if (tune)
            B.autotune(nullptr);
        else
            B.configure(tile_rows, block, nullptr);
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_set_layout(xjpcg_batch_t* batch, int32_t layout, int32_t* h_active,
                                                 void* stream) {
    return bguard([&] {
        xj::Batch& B = impl_of(batch);
        if (layout >= 0) B.set_layout(layout, static_cast<hipStream_t>(stream));
        if (h_active) *h_active = B.layout();
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_solve(xjpcg_batch_t* batch, const double* d_diag, const double* d_b,
                                            double* d_x, uint32_t max_iter, double tol, int32_t* d_iters,
                                            double* d_res, int64_t* h_launched, void* stream) {
    return bguard([&] {
        int64_t n = impl_of(batch).solve(d_diag, d_b, d_x, max_iter, tol, d_iters, d_res,
                                         static_cast<hipStream_t>(stream));
        if (h_launched) *h_launched = n;
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_set_path(xjpcg_batch_t* batch, int32_t path, int32_t grid, int32_t* h_last,
                                               int32_t* h_grid) {
    return bguard([&] {
        xj::Batch& B = impl_of(batch);
        if (path >= 0) B.set_path(path, grid);
        if (h_last) *h_last = B.last_path();
        if (h_grid) *h_grid = B.last_grid();
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_set_vectors(xjpcg_batch_t* batch, int32_t bits, int32_t* h_bits) {
    return bguard([&] {
        xj::Batch& B = impl_of(batch);
        if (bits > 0) B.set_vector_bits(bits);
        if (h_bits) *h_bits = B.vector_bits();
    });
}

extern "C" XJPCG_Status_t xjpcg_batch_solve_mixed(xjpcg_batch_t* outer, xjpcg_batch_t* inner, const double* d_diag,
                                                  const double* d_b, double* d_x, uint32_t max_outer, double tol,
                                                  double inner_tol, uint32_t inner_max_iter, int32_t* d_iters,
                                                  double* d_res, uint32_t* h_outer, int64_t* h_inner_launched,
                                                  void* stream) {
    uint32_t steps = 0;
    int64_t launched = 0;
    XJPCG_Status_t st = bguard([&] {
        xj::Batch& O = impl_of(outer);
        xj::Batch& I = impl_of(inner);
        if (O.N() != I.N() || O.S() != I.S() || O.R() != I.R()) throw xj::invalid("inner and outer batches differ.");
        if (d_diag == nullptr || d_b == nullptr || d_x == nullptr) throw xj::invalid("Vector is nullptr.");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int SR = O.S() * O.R();
        const size_t NR = (size_t)O.N() * O.R();
        double* r = O.scratch(0);
        double* d = O.scratch(1);
        // small per-pair device arrays live in the third scratch vector
        double* pairs = O.scratch(2);
        double* d_bb = pairs;
        double* d_rr = pairs + SR;
        int* d_keep = reinterpret_cast<int*>(pairs + 2 * SR);
        int32_t* d_it = reinterpret_cast<int32_t*>(pairs + 3 * SR);
        if (NR < (size_t)4 * SR) throw xj::invalid("batch too small for the refinement workspace.");
        std::vector<double> bb(SR), rr(SR), res(SR);
        std::vector<int> keep(SR);
        std::vector<int32_t> it_total(SR, 0), it_step(SR);
        xj::hip_check(hipMemsetAsync(d_x, 0, NR * sizeof(double), s), "zero x");
        xj::hip_check(hipMemcpyAsync(r, d_b, NR * sizeof(double), hipMemcpyDeviceToDevice, s), "r0 = b");
        O.norms(d_b, d_bb, s);
        xj::hip_check(hipMemcpyAsync(bb.data(), d_bb, SR * sizeof(double), hipMemcpyDeviceToHost, s), "norms");
        xj::hip_check(hipStreamSynchronize(s), "norms");
        rr = bb;
        while (true) {
            int left = 0;
            for (int q = 0; q < SR; ++q) {
                res[q] = bb[q] > 0.0 ? std::sqrt(rr[q] / bb[q]) : 0.0;
                keep[q] = res[q] < tol ? 0 : 1;
                left += keep[q];
            }
            if (left == 0 || steps >= max_outer) break;
            // relative target for the correction: the unfinished pair furthest from tol sets it
            // (see xjpcg_solve_mixed), bounded below by inner_tol and above by 0.5
            double tin = 0.5;
            for (int q = 0; q < SR; ++q)
                if (keep[q]) tin = std::min(tin, 0.5 * tol / res[q]);
            // fp32 work vectors keep the true residual with the recursive one
            // down to about 1e-7 (x stays fp64): a step asks them for no more
            // than that, and the next step corrects the rest (on the
            // 512-system batch a floor of 1e-5 took 2,912 batched iterations
            // over two steps, 1e-7 took 2,496)
            if (I.vector_bits() == 32) tin = std::max(tin, kW32StepTol);
            tin = std::min(std::max(tin, inner_tol), 0.5);
            xj::hip_check(hipMemcpyAsync(d_keep, keep.data(), SR * sizeof(int), hipMemcpyHostToDevice, s), "mask");
            O.mask(d_keep, r, s);  // finished pairs get a zero right-hand side: they do not run
            launched += I.solve(d_diag, r, d, inner_max_iter, tin, d_it, nullptr, s);
            xj::hip_check(xj::batch_launch_axpy(NR, d, d_x, s), "x += d");
            O.spmm(d_x, r, d_b, s);
            O.norms(r, d_rr, s);
            xj::hip_check(hipMemcpyAsync(rr.data(), d_rr, SR * sizeof(double), hipMemcpyDeviceToHost, s), "norms");
            xj::hip_check(hipMemcpyAsync(it_step.data(), d_it, SR * sizeof(int32_t), hipMemcpyDeviceToHost, s),
                          "iterations");
            xj::hip_check(hipStreamSynchronize(s), "refinement step");
            for (int q = 0; q < SR; ++q) it_total[q] += it_step[q];
            ++steps;
        }
        if (d_res)
            xj::hip_check(hipMemcpyAsync(d_res, res.data(), SR * sizeof(double), hipMemcpyHostToDevice, s), "res");
        if (d_iters)
            xj::hip_check(hipMemcpyAsync(d_iters, it_total.data(), SR * sizeof(int32_t), hipMemcpyHostToDevice, s),
                          "iters");
        xj::hip_check(hipStreamSynchronize(s), "refinement");
        double worst = 0.0;
        for (double v : res) worst = std::max(worst, v);
        if (!(worst < tol))
            throw xj::exec_failed("batched refinement stopped at worst relative residual " + std::to_string(worst) +
                                  " after " + std::to_string(steps) + " steps.");
    });
    if (h_outer) *h_outer = steps;
    if (h_inner_launched) *h_inner_launched = launched;
    return st;
}
