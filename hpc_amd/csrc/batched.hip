// Batched JPCG kernels: S independent systems (one block-diagonal CSR), R
// right-hand sides per system stored row-major [N][R].
//
// Work unit: a tile of consecutive rows of one system.  Inside a workgroup,
// R lanes form a group that owns one row at a time -- lane j carries
// right-hand side j -- so a row's column index and value are read once and
// broadcast to the group while the R gathered x-values of a neighbour are one
// contiguous 8R-byte segment.  Per-(system, rhs) reductions go tile partial ->
// system total in fixed order (one workgroup per system), so every pair's
// alpha / beta are reproducible.
//
//   bsetup / bsys_setup      r = b, J = 1/diag, p = J b, x = 0; ||b||^2, rz per pair
//   bspmm_dot                Ap = A p, p.Ap tile partials          (one workgroup per tile)
//   bsys_alpha               alpha per pair                        (one workgroup per system)
//   bupdate_r                r -= alpha Ap; r.r, r.z tile partials
//   bsys_beta                beta, stop rule and masks per pair
//   bupdate_xp               x += alpha p; p = beta p + J r        (pairs that ran this iteration)
//   bfinish                  iterations and relative residual per pair
#include "batched.hpp"
#include "kernels.hpp"

namespace xj {

constexpr int kMaxBlock = 512;
constexpr int kRowsAhead = 4;  // rows a vector-kernel thread loads before it stores

// Sum over the row groups of a workgroup for each lane j < R; lanes with
// g == 0 return the total (fixed order g = 0, 1, ...).
template <int R>
__device__ __forceinline__ double group_sum(double v, double* red) {
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    red[g * R + j] = v;
    __syncthreads();
    double s = 0.0;
    if (g == 0)
        for (int q = 0; q < G; ++q) s += red[q * R + j];
    __syncthreads();
    return s;
}

// Sum of part[t*R + j] over the tiles t of system s in a fixed order that
// does not depend on the block size: kMaxBlock / R virtual groups, virtual
// group v adding the tiles t0 + v, t0 + v + kMaxBlock/R, ... in order, then
// the groups' sums added in order v = 0, 1, ...  A thread of a smaller
// workgroup carries several virtual groups, so the system kernels (512
// threads) and the persistent loop (the tile kernels' block size) produce
// the same bits.  Each virtual group issues four loads before adding them in
// order, so a large system's partials cost a quarter of the dependent round
// trips.  Lanes with g == 0 return the total.
template <int R>
__device__ __forceinline__ double system_sum(const double* part, int t0, int t1, double* red) {
    constexpr int GV = kMaxBlock / R;
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    for (int v = g; v < GV; v += G) {
        double s = 0.0;
        int t = t0 + v;
        for (; t + 3 * GV < t1; t += 4 * GV) {
            const double a0 = part[(size_t)t * R + j], a1 = part[(size_t)(t + GV) * R + j];
            const double a2 = part[(size_t)(t + 2 * GV) * R + j], a3 = part[(size_t)(t + 3 * GV) * R + j];
            s += a0;
            s += a1;
            s += a2;
            s += a3;
        }
        for (; t < t1; t += GV) s += part[(size_t)t * R + j];
        red[v * R + j] = s;
    }
    __syncthreads();
    double s = 0.0;
    if (g == 0)
        for (int v = 0; v < GV; ++v) s += red[v * R + j];
    __syncthreads();
    return s;
}

// The vector kernels stream their per-pair operands (r, Ap, x, p: several GB
// per vector for a large batch, far beyond L2 and the MALL) with
// non-temporal loads and stores; the per-row Jacobi factor, shared by the R
// lanes of a row, stays a cached load.  On 64 systems x 16 right-hand sides
// this took the fp64 batch from 1,788 to 1,620 ms and the fp32-vector
// refinement from 1,224 to 1,144 ms (same bits).  Single-system solves keep
// cached loads: their vectors fit the MALL between kernels.
//
// Work vectors r, p, Ap are stored as WT: double, or float for the inner
// batch of a refinement (WT = float cuts their traffic in half).  Every
// value is widened to double before any arithmetic and rounded to WT only
// when stored; dot products use the stored (rounded) values and accumulate
// in double.  x stays double: rounding x would leave a true residual of
// about eps32 * cond(A) whatever the recursive one (measured 3.5e-4 on the
// 512-system batch), while rounding r, p and Ap only perturbs the recursion
// at the size of the current residual.  With WT = double every expression
// below is the plain fp64 loop.
template <typename WT>
__device__ __forceinline__ WT* wvec(double* p) { return reinterpret_cast<WT*>(p); }

template <int R, typename WT>
__global__ __launch_bounds__(kMaxBlock) void bsetup(BatchParams P) {
    __shared__ double red[kMaxBlock];
    const int t = (int)blockIdx.x;
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    const int r0 = P.tile_row[t], r1 = P.tile_row[t + 1];
    WT* __restrict__ r = wvec<WT>(P.r);
    WT* __restrict__ p = wvec<WT>(P.p);
    double dot = 0.0, rz = 0.0;
    for (int i = r0 + g; i < r1; i += G) {
        const size_t e = (size_t)i * R + j;
        const double bi = P.b[e];
        const double ri = (double)(WT)bi;
        const double Ji = 1.0 / P.diag[i];
        const double zi = Ji * ri;
        if (j == 0) P.J[i] = Ji;
        r[e] = (WT)ri;
        p[e] = (WT)zi;
        P.x[e] = 0.0;
        dot += bi * bi;  // the stop rule's ||b|| from the exact right-hand side
        rz += ri * zi;
    }
    dot = group_sum<R>(dot, red);
    rz = group_sum<R>(rz, red);
    if (g == 0) {
        P.partB[(size_t)t * R + j] = dot;
        P.partC[(size_t)t * R + j] = rz;
    }
}

template <int R>
__global__ __launch_bounds__(kMaxBlock) void bsys_setup(BatchParams P) {
    __shared__ double red[kMaxBlock];
    __shared__ int any;
    const int s = blockIdx.x;
    const int j = threadIdx.x % R, g = threadIdx.x / R;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    const int t0 = P.sys_tile[s], t1 = P.sys_tile[s + 1];
    const double dot = system_sum<R>(P.partB, t0, t1, red);
    const double rz = system_sum<R>(P.partC, t0, t1, red);
    if (g == 0) {
        const int q = s * R + j;
        P.dot[q] = dot;
        P.tols[q] = dot * P.tol * P.tol;
        P.rz[q] = rz;
        P.rr[q] = dot;
        P.alpha[q] = 0.0;
        P.beta[q] = 0.0;
        P.iters[q] = 0;
        P.it[q] = 0;
        const int a = dot > 0.0 ? 1 : 0;  // b == 0: x = 0, the pair never runs
        P.active[q] = a;
        P.upd[q] = 0;
        if (a) atomicOr(&any, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        P.sys_active[s] = any;
        P.sys_upd[s] = 0;
        if (any) *P.any_active = 1;
    }
}

// SpMM lane mapping: a lane owns one row and Q = min(R, 4) consecutive
// right-hand sides (L = R / Q lanes per row), so a neighbour's values are one
// 8Q-byte vector load, the column index and value of an entry are loaded by L
// lanes instead of R, and the index arithmetic is done once per Q outputs.
// Entries are staged kRowStage at a time (indices and values first, then the
// gathers, then the adds), so a lane's gathers are in flight together.
// Each (row, rhs) is summed left to right from dsplit[i]*x (or 0).
constexpr int kRowStage = 8;

template <int R>
struct SpmmLanes {
    static constexpr int Q = R < 4 ? R : 4;
    static constexpr int L = R / Q;
};

template <int Q, typename T>
__device__ __forceinline__ void load_q(const T* __restrict__ p, double (&out)[Q]) {
#pragma unroll
    for (int q = 0; q < Q; ++q) out[q] = (double)p[q];
}

// the same load kept in T; converting at the use lets loads issued under a
// condition stay in flight together (a conversion under the condition waits
// for its load)
template <int Q, typename T>
__device__ __forceinline__ void load_raw(const T* __restrict__ p, T (&out)[Q]) {
#pragma unroll
    for (int q = 0; q < Q; ++q) out[q] = p[q];
}

// acc[q] = row i of A times column (q0 + q) of the [N][R] block xin
template <typename VT, int R, typename XT>
__device__ __forceinline__ void brow(const BatchParams& P, const VT* __restrict__ val, const XT* __restrict__ xin,
                                     int i, int q0, double (&acc)[SpmmLanes<R>::Q]) {
    constexpr int Q = SpmmLanes<R>::Q;
    const int k0 = P.rowptr[i], k1 = P.rowptr[i + 1];
    if (P.dsplit) {
        double xi[Q];
        load_q<Q>(xin + (size_t)i * R + q0, xi);
        const double d = P.dsplit[i];
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = d * xi[q];
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = 0.0;
    }
    for (int kb = k0; kb < k1; kb += kRowStage) {
        int c[kRowStage];
        double v[kRowStage];
#pragma unroll
        for (int u = 0; u < kRowStage; ++u) {
            const int k = kb + u < k1 ? kb + u : k0;
            c[u] = ld_stream(P.col + k);
            v[u] = (double)ld_stream(val + k);
        }
        double g[kRowStage][Q];
#pragma unroll
        for (int u = 0; u < kRowStage; ++u) load_q<Q>(xin + (size_t)c[u] * R + q0, g[u]);
#pragma unroll
        for (int u = 0; u < kRowStage; ++u)
            if (kb + u < k1) {
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] += v[u] * g[u][q];
            }
    }
}

// The same row on the per-system diagonal layout: the row's zero-padded
// values and the R/Q-wide x segments of its offsets (columns clamped into the
// system for offsets the row lacks) load with its mask, no index first; the
// entries are added in ascending offset order -- the CSR order of the row,
// so the sums are those of brow.
template <typename VT, int R, typename XT>
__device__ __forceinline__ void brow_dia(const BatchParams& P, const BDiaSys& Y, const VT* __restrict__ dval,
                                         const XT* __restrict__ xin, int i, int q0,
                                         double (&acc)[SpmmLanes<R>::Q]) {
    constexpr int Q = SpmmLanes<R>::Q;
    const uint32_t m = P.dmask[i];
    if (P.dsplit) {
        double xi[Q];
        load_q<Q>(xin + (size_t)i * R + q0, xi);
        const double d = P.dsplit[i];
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = d * xi[q];
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] = 0.0;
    }
    const int ns = Y.r1 - Y.r0, li = i - Y.r0;
    // four diagonals' loads at a time (eight in flight cost occupancy: measured
    // 2816 vs 3045 ms on 64 systems x 16 right-hand sides)
#pragma unroll
    for (int h = 0; h < kBDiaMax; h += 4) {
        VT v[4];
        XT g[4][Q];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int d = h + u;
            if (d < Y.nd) {
                v[u] = ld_stream(dval + Y.base + (size_t)d * ns + li);
                int c = i + Y.off[d];
                c = c < Y.r0 ? Y.r0 : (c >= Y.r1 ? Y.r1 - 1 : c);
                load_raw<Q>(xin + (size_t)c * R + q0, g[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int d = h + u;
            if (d < Y.nd && ((m >> d) & 1u)) {
                const double vu = (double)v[u];
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] += vu * (double)g[u][q];
            }
        }
    }
}

// Ap = A p on tile t and its p.Ap partials partA[t][R]; red: kMaxBlock * Q doubles of LDS
template <typename VT, int R, bool DIA, typename WT>
__device__ __forceinline__ void bspmm_dot_tile(const BatchParams& P, int t, double* red) {
    constexpr int Q = SpmmLanes<R>::Q, L = SpmmLanes<R>::L;
    const int s = P.tile_sys[t];
    BDiaSys Y;
    if (DIA) Y = P.dsys[s];
    if (!P.sys_active[s]) return;
    const int l = threadIdx.x % L, g = threadIdx.x / L, G = blockDim.x / L;
    const int q0 = l * Q;
    const int r0 = P.tile_row[t], r1 = P.tile_row[t + 1];
    const VT* __restrict__ val = static_cast<const VT*>(P.val);
    const WT* __restrict__ p = wvec<WT>(P.p);
    WT* __restrict__ Ap = wvec<WT>(P.Ap);
    double dl[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) dl[q] = 0.0;
    for (int i = r0 + g; i < r1; i += G) {
        double acc[Q];
        if (DIA)
            brow_dia<VT, R, WT>(P, Y, static_cast<const VT*>(P.dval), p, i, q0, acc);
        else
            brow<VT, R, WT>(P, val, p, i, q0, acc);
        double pi[Q];
        load_q<Q>(p + (size_t)i * R + q0, pi);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const WT a = (WT)acc[q];
            st_stream(Ap + (size_t)i * R + q0 + q, a);
            dl[q] += pi[q] * (double)a;
        }
    }
    // per right-hand side: sum over the row groups in fixed order g = 0, 1, ...
#pragma unroll
    for (int q = 0; q < Q; ++q) red[g * R + q0 + q] = dl[q];
    __syncthreads();
    if ((int)threadIdx.x < R) {
        const int j = threadIdx.x;
        double sum = 0.0;
        for (int gg = 0; gg < G; ++gg) sum += red[gg * R + j];
        P.partA[(size_t)t * R + j] = sum;
    }
    __syncthreads();  // red is reused by the workgroup's next tile
}

template <typename VT, int R, bool DIA, typename WT>
__global__ __launch_bounds__(kMaxBlock) void bspmm_dot(BatchParams P) {
    __shared__ double red[kMaxBlock * SpmmLanes<R>::Q];
    bspmm_dot_tile<VT, R, DIA, WT>(P, (int)blockIdx.x, red);
}

// plain batched SpMM y = A x (or y = b - A x) for the autotuner and the refinement residual
template <typename VT, int R, bool DIA>
__global__ __launch_bounds__(kMaxBlock) void bspmm(BatchParams P, const double* __restrict__ xin,
                                                   double* __restrict__ y, const double* __restrict__ b) {
    constexpr int Q = SpmmLanes<R>::Q, L = SpmmLanes<R>::L;
    const int t = (int)blockIdx.x;
    BDiaSys Y;
    if (DIA) Y = P.dsys[P.tile_sys[t]];
    const int l = threadIdx.x % L, g = threadIdx.x / L, G = blockDim.x / L;
    const int q0 = l * Q;
    const int r0 = P.tile_row[t], r1 = P.tile_row[t + 1];
    const VT* __restrict__ val = static_cast<const VT*>(P.val);
    for (int i = r0 + g; i < r1; i += G) {
        double acc[Q];
        if (DIA)
            brow_dia<VT, R>(P, Y, static_cast<const VT*>(P.dval), xin, i, q0, acc);
        else
            brow<VT, R>(P, val, xin, i, q0, acc);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const size_t e = (size_t)i * R + q0 + q;
            st_stream(y + e, b ? b[e] - acc[q] : acc[q]);
        }
    }
}

// alpha per pair of system s (red: kMaxBlock doubles of LDS)
template <int R>
__device__ __forceinline__ void bsys_alpha_sys(const BatchParams& P, int s, double* red) {
    if (P.sys_active[s]) {
        const int j = threadIdx.x % R, g = threadIdx.x / R;
        const double pap = system_sum<R>(P.partA, P.sys_tile[s], P.sys_tile[s + 1], red);
        if (g == 0) {
            const int q = s * R + j;
            P.alpha[q] = P.active[q] ? P.rz[q] / pap : 0.0;
        }
    }
    // cleared after this workgroup's loads (a store ahead of them would turn
    // their scalar loads into vector loads); bsys_beta sets it again for
    // systems still running
    if (s == 0 && threadIdx.x == 0) *P.any_active = 0;
}

template <int R>
__global__ __launch_bounds__(kMaxBlock) void bsys_alpha(BatchParams P) {
    __shared__ double red[kMaxBlock];
    bsys_alpha_sys<R>(P, (int)blockIdx.x, red);
}

// r -= alpha Ap on tile t and its r.r, r.z partials (red: kMaxBlock doubles of LDS)
template <int R, typename WT>
__device__ __forceinline__ void bupdate_r_tile(const BatchParams& P, int t, double* red) {
    const int s = P.tile_sys[t];
    if (!P.sys_active[s]) return;
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    const int r0 = P.tile_row[t], r1 = P.tile_row[t + 1];
    const double alpha = P.alpha[s * R + j];
    WT* __restrict__ r = wvec<WT>(P.r);
    const WT* __restrict__ Ap = wvec<WT>(P.Ap);
    double rr = 0.0, rz = 0.0;
    // four of this thread's rows at a time: their operands are all loaded
    // (clamped into the tile, no branch) before the first store to r, which
    // the compiler must otherwise keep ahead of the next row's loads; rows
    // are then updated and summed in order (contractions written out as fma)
    for (int i = r0 + g; i < r1; i += kRowsAhead * G) {
        double Jv[kRowsAhead], av[kRowsAhead], rv[kRowsAhead];
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const int iu = i + u * G < r1 ? i + u * G : i;
            const size_t e = (size_t)iu * R + j;
            Jv[u] = P.J[iu];
            av[u] = (double)ld_stream(Ap + e);
            rv[u] = (double)ld_stream(r + e);
        }
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            if (i + u * G < r1) {
                const size_t e = (size_t)(i + u * G) * R + j;
                const double ri = (double)(WT)__builtin_fma(-alpha, av[u], rv[u]);
                st_stream(r + e, (WT)ri);
                const double zi = Jv[u] * ri;
                rr = __builtin_fma(ri, ri, rr);
                rz = __builtin_fma(ri, zi, rz);
            }
        }
    }
    rr = group_sum<R>(rr, red);
    rz = group_sum<R>(rz, red);
    if (g == 0) {
        P.partB[(size_t)t * R + j] = rr;
        P.partC[(size_t)t * R + j] = rz;
    }
}

template <int R, typename WT>
__global__ __launch_bounds__(kMaxBlock) void bupdate_r(BatchParams P) {
    __shared__ double red[kMaxBlock];
    bupdate_r_tile<R, WT>(P, (int)blockIdx.x, red);
}

// beta, the stop rule and the masks per pair of system s (red: kMaxBlock
// doubles of LDS; flags: two ints of LDS)
template <int R>
__device__ __forceinline__ void bsys_beta_sys(const BatchParams& P, int s, double* red, int* flags) {
    int& any_active = flags[0];
    int& any_upd = flags[1];
    const int j = threadIdx.x % R, g = threadIdx.x / R;
    if (!P.sys_active[s]) {
        if (g == 0) P.upd[s * R + j] = 0;
        if (threadIdx.x == 0) P.sys_upd[s] = 0;
        return;
    }
    if (threadIdx.x == 0) {
        any_active = 0;
        any_upd = 0;
    }
    __syncthreads();
    const int t0 = P.sys_tile[s], t1 = P.sys_tile[s + 1];
    const double rr = system_sum<R>(P.partB, t0, t1, red);
    const double rzn = system_sum<R>(P.partC, t0, t1, red);
    if (g == 0) {
        const int q = s * R + j;
        const int a = P.active[q];
        P.upd[q] = a;
        double beta = 0.0;
        if (a) {
            beta = rzn / P.rz[q];
            P.rz[q] = rzn;
            P.rr[q] = rr;
            P.iters[q] = P.iters[q] + 1;
            // stop rule of the reference control loop (control.hpp:83-88, tasks.hpp:43-47)
            int it = P.it[q];
            bool more = false;
            if (!(rr < P.tols[q])) {
                ++it;
                more = (uint32_t)it < P.max_rounds;
            }
            P.it[q] = it;
            if (!more) P.active[q] = 0;
            else atomicOr(&any_active, 1);
            atomicOr(&any_upd, 1);
        }
        P.beta[q] = beta;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        P.sys_active[s] = any_active;
        P.sys_upd[s] = any_upd;
        if (any_active) *P.any_active = 1;
    }
    __syncthreads();  // flags are reused by the workgroup's next system
}

template <int R>
__global__ __launch_bounds__(kMaxBlock) void bsys_beta(BatchParams P) {
    __shared__ double red[kMaxBlock];
    __shared__ int flags[2];
    bsys_beta_sys<R>(P, (int)blockIdx.x, red, flags);
}

// x += alpha p; p = beta p + J r on tile t for the pairs that ran this iteration
template <int R, typename WT>
__device__ __forceinline__ void bupdate_xp_tile(const BatchParams& P, int t) {
    const int s = P.tile_sys[t];
    if (!P.sys_upd[s]) return;
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    const int q = s * R + j;
    if (!P.upd[q]) return;
    const int r0 = P.tile_row[t], r1 = P.tile_row[t + 1];
    const double alpha = P.alpha[q], beta = P.beta[q];
    WT* __restrict__ p = wvec<WT>(P.p);
    const WT* __restrict__ r = wvec<WT>(P.r);
    // four of this thread's rows at a time, every operand loaded (clamped into
    // the tile, no branch) before the first store: x could alias J or r for
    // all the compiler knows, which would order the next loads after it
    for (int i = r0 + g; i < r1; i += kRowsAhead * G) {
        double pv[kRowsAhead], xv[kRowsAhead], Jv[kRowsAhead], rv[kRowsAhead];
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const int iu = i + u * G < r1 ? i + u * G : i;
            const size_t e = (size_t)iu * R + j;
            pv[u] = (double)ld_stream(p + e);
            xv[u] = ld_stream(P.x + e);
            Jv[u] = P.J[iu];
            rv[u] = (double)ld_stream(r + e);
        }
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            if (i + u * G < r1) {
                const size_t e = (size_t)(i + u * G) * R + j;
                st_stream(P.x + e, __builtin_fma(alpha, pv[u], xv[u]));
                st_stream(p + e, (WT)__builtin_fma(beta, pv[u], Jv[u] * rv[u]));
            }
        }
    }
}

template <int R, typename WT>
__global__ __launch_bounds__(kMaxBlock) void bupdate_xp(BatchParams P) {
    bupdate_xp_tile<R, WT>(P, (int)blockIdx.x);
}

__global__ void bfinish(int nq, const double* __restrict__ dot, const double* __restrict__ rr,
                        const int* __restrict__ iters, int* __restrict__ out_iters, double* __restrict__ out_res) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
        if (out_iters) out_iters[q] = iters[q];
        if (out_res) out_res[q] = dot[q] > 0.0 ? sqrt(rr[q] / dot[q]) : 0.0;
    }
}

// per-pair ||v||^2 over each system (for the refinement's true residual)
template <int R>
__global__ __launch_bounds__(kMaxBlock) void btile_norm(BatchParams P, const double* __restrict__ v,
                                                        double* __restrict__ part) {
    __shared__ double red[kMaxBlock];
    const int t = (int)blockIdx.x;
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    double s = 0.0;
    for (int i = P.tile_row[t] + g; i < P.tile_row[t + 1]; i += G) {
        const double a = v[(size_t)i * R + j];
        s += a * a;
    }
    s = group_sum<R>(s, red);
    if (g == 0) part[(size_t)t * R + j] = s;
}

template <int R>
__global__ __launch_bounds__(kMaxBlock) void bsys_norm(BatchParams P, const double* __restrict__ part,
                                                       double* __restrict__ out) {
    __shared__ double red[kMaxBlock];
    const int s = blockIdx.x;
    const int j = threadIdx.x % R, g = threadIdx.x / R;
    const double v = system_sum<R>(part, P.sys_tile[s], P.sys_tile[s + 1], red);
    if (g == 0) out[s * R + j] = v;
}

// v[e] = 0 for the pairs whose flag is 0 (masks converged pairs out of a refinement step)
template <int R>
__global__ __launch_bounds__(kMaxBlock) void bmask(BatchParams P, const int* __restrict__ keep,
                                                   double* __restrict__ v) {
    const int t = (int)blockIdx.x;
    const int s = P.tile_sys[t];
    const int j = threadIdx.x % R, g = threadIdx.x / R, G = blockDim.x / R;
    if (keep[s * R + j]) return;
    for (int i = P.tile_row[t] + g; i < P.tile_row[t + 1]; i += G) v[(size_t)i * R + j] = 0.0;
}

__global__ void baxpy(size_t m, const double* __restrict__ d, double* __restrict__ x) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x)
        x[e] += d[e];
}

// --------------------------------------------------------------------------
// One persistent launch for the whole iteration loop of a batch.
//
// The grid (all workgroups co-resident, cooperative launch) runs the same
// five steps per iteration as the launch-per-step path, with a grid barrier
// after each: tiles w, w+G, ... of the SpMM, systems w, w+G, ... of the alpha
// step, tiles of the r update, systems of the beta / stop / mask step, tiles
// of the x / p update.  Every tile and system step is the launch-per-step
// path's own code with the same reduction orders (system sums do not depend
// on the block size), so both paths produce the same bits; converged
// systems drop out through the same per-system and per-pair masks.  After
// each iteration every workgroup reads the any-active flag the beta step
// left, so all of them leave the loop together.  A barrier wait longer than
// its timeout sets the abort flag and every workgroup leaves.
template <typename VT, int R, bool DIA, typename WT>
__global__ __launch_bounds__(kMaxBlock) void bpersistent(BatchParams P, BatchLoop Lp) {
    __shared__ double red[kMaxBlock * SpmmLanes<R>::Q];
    __shared__ int flags[2];
    __shared__ int sflag;
    const int G = (int)gridDim.x, w = (int)blockIdx.x;
    unsigned long long target = 0;
    long long k = 0;
    while (k < (long long)Lp.rounds) {
        for (int t = w; t < P.T; t += G) bspmm_dot_tile<VT, R, DIA, WT>(P, t, red);
        if (!grid_barrier(Lp.bar, target, G, Lp.abortf, &sflag)) break;
        for (int s = w; s < P.S; s += G) bsys_alpha_sys<R>(P, s, red);
        if (!grid_barrier(Lp.bar, target, G, Lp.abortf, &sflag)) break;
        for (int t = w; t < P.T; t += G) bupdate_r_tile<R, WT>(P, t, red);
        if (!grid_barrier(Lp.bar, target, G, Lp.abortf, &sflag)) break;
        for (int s = w; s < P.S; s += G) bsys_beta_sys<R>(P, s, red, flags);
        if (!grid_barrier(Lp.bar, target, G, Lp.abortf, &sflag)) break;
        for (int t = w; t < P.T; t += G) bupdate_xp_tile<R, WT>(P, t);
        ++k;
        if (!grid_barrier(Lp.bar, target, G, Lp.abortf, &sflag)) break;
        if (!__hip_atomic_load(P.any_active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    }
    if (w == 0 && threadIdx.x == 0) {
        Lp.result[0] = k;
        Lp.result[1] = __hip_atomic_load(Lp.abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// --------------------------------------------------------------------------
// launchers

template <int R, typename WT>
static void iteration_rw(const BatchParams& P, bool f32, int block, hipStream_t s) {
    if (P.dia_on) {
        if (f32)
            bspmm_dot<float, R, true, WT><<<P.T, block, 0, s>>>(P);
        else
            bspmm_dot<double, R, true, WT><<<P.T, block, 0, s>>>(P);
    } else {
        if (f32)
            bspmm_dot<float, R, false, WT><<<P.T, block, 0, s>>>(P);
        else
            bspmm_dot<double, R, false, WT><<<P.T, block, 0, s>>>(P);
    }
    bsys_alpha<R><<<P.S, kMaxBlock, 0, s>>>(P);
    bupdate_r<R, WT><<<P.T, block, 0, s>>>(P);
    bsys_beta<R><<<P.S, kMaxBlock, 0, s>>>(P);
    bupdate_xp<R, WT><<<P.T, block, 0, s>>>(P);
}

template <int R>
static void iteration_r(const BatchParams& P, bool f32, bool w32, int block, hipStream_t s) {
    if (w32)
        iteration_rw<R, float>(P, f32, block, s);
    else
        iteration_rw<R, double>(P, f32, block, s);
}

template <int R>
static void setup_r(const BatchParams& P, bool w32, int block, hipStream_t s) {
    if (w32)
        bsetup<R, float><<<P.T, block, 0, s>>>(P);
    else
        bsetup<R, double><<<P.T, block, 0, s>>>(P);
    bsys_setup<R><<<P.S, kMaxBlock, 0, s>>>(P);
}

template <int R>
static void spmm_r(const BatchParams& P, bool f32, int block, const double* x, double* y, const double* b,
                   hipStream_t s) {
    if (P.dia_on) {
        if (f32)
            bspmm<float, R, true><<<P.T, block, 0, s>>>(P, x, y, b);
        else
            bspmm<double, R, true><<<P.T, block, 0, s>>>(P, x, y, b);
    } else {
        if (f32)
            bspmm<float, R, false><<<P.T, block, 0, s>>>(P, x, y, b);
        else
            bspmm<double, R, false><<<P.T, block, 0, s>>>(P, x, y, b);
    }
}

template <int R>
static void norms_r(const BatchParams& P, int block, const double* v, double* part, double* out, hipStream_t s) {
    btile_norm<R><<<P.T, block, 0, s>>>(P, v, part);
    bsys_norm<R><<<P.S, kMaxBlock, 0, s>>>(P, part, out);
}

template <int R>
static void mask_r(const BatchParams& P, int block, const int* keep, double* v, hipStream_t s) {
    bmask<R><<<P.T, block, 0, s>>>(P, keep, v);
}

// the persistent kernel instantiated for (R, value type, layout, vector type)
template <int R>
static const void* persistent_fn(bool f32, bool w32, bool dia) {
    if (w32) {
        if (dia) return f32 ? (const void*)bpersistent<float, R, true, float> : (const void*)bpersistent<double, R, true, float>;
        return f32 ? (const void*)bpersistent<float, R, false, float> : (const void*)bpersistent<double, R, false, float>;
    }
    if (dia) return f32 ? (const void*)bpersistent<float, R, true, double> : (const void*)bpersistent<double, R, true, double>;
    return f32 ? (const void*)bpersistent<float, R, false, double> : (const void*)bpersistent<double, R, false, double>;
}

#define XJ_DISPATCH_R(R_, CALL)                       \
    switch (R_) {                                     \
        case 1: { constexpr int RR = 1; CALL; break; }  \
        case 2: { constexpr int RR = 2; CALL; break; }  \
        case 4: { constexpr int RR = 4; CALL; break; }  \
        case 8: { constexpr int RR = 8; CALL; break; }  \
        case 16: { constexpr int RR = 16; CALL; break; } \
        case 32: { constexpr int RR = 32; CALL; break; } \
        default: return hipErrorInvalidValue;         \
    }

hipError_t batch_launch_setup(const BatchParams& P, int R, bool w32, int block, hipStream_t s) {
    XJ_DISPATCH_R(R, setup_r<RR>(P, w32, block, s));
    return hipGetLastError();
}

hipError_t batch_launch_iteration(const BatchParams& P, int R, bool f32, bool w32, int block, hipStream_t s) {
    XJ_DISPATCH_R(R, iteration_r<RR>(P, f32, w32, block, s));
    return hipGetLastError();
}


// Largest co-resident grid of the persistent kernel for this instantiation
// and block size, one workgroup per CU below the occupancy limit (a grid
// that fills every slot cannot be placed while any other dispatch holds one).
hipError_t batch_persistent_grid(int R, bool f32, bool w32, bool dia, int block, int* G) {
    const void* fn = nullptr;
    XJ_DISPATCH_R(R, fn = persistent_fn<RR>(f32, w32, dia));
    int dev = 0, cus = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, 0);
    if (e != hipSuccess) return e;
    *G = cus * (per > 1 ? per - 1 : 1);
    return hipSuccess;
}

hipError_t batch_launch_persistent(const BatchParams& P0, const BatchLoop& L0, int R, bool f32, bool w32, int block,
                                   int G, hipStream_t s) {
    BatchParams P = P0;
    BatchLoop L = L0;
    void* args[] = {&P, &L};
    const void* fn = nullptr;
    XJ_DISPATCH_R(R, fn = persistent_fn<RR>(f32, w32, P.dia_on != 0));
    return hipLaunchCooperativeKernel(fn, dim3(G), dim3(block), args, 0, s);
}

hipError_t batch_launch_spmm(const BatchParams& P, int R, bool f32, int block, const double* x, double* y,
                             const double* b, hipStream_t s) {
    XJ_DISPATCH_R(R, spmm_r<RR>(P, f32, block, x, y, b, s));
    return hipGetLastError();
}

hipError_t batch_launch_norms(const BatchParams& P, int R, int block, const double* v, double* part, double* out,
                              hipStream_t s) {
    XJ_DISPATCH_R(R, norms_r<RR>(P, block, v, part, out, s));
    return hipGetLastError();
}

hipError_t batch_launch_mask(const BatchParams& P, int R, int block, const int* keep, double* v, hipStream_t s) {
    XJ_DISPATCH_R(R, mask_r<RR>(P, block, keep, v, s));
    return hipGetLastError();
}

hipError_t batch_launch_finish(int nq, const double* dot, const double* rr, const int* iters, int* out_iters,
                               double* out_res, hipStream_t s) {
    int g = (nq + 255) / 256;
    if (g < 1) g = 1;
    bfinish<<<g, 256, 0, s>>>(nq, dot, rr, iters, out_iters, out_res);
    return hipGetLastError();
}

// Distinct column offsets of one tile (one workgroup): an LDS hash of 16
// slots written out as is (the host forms each system's union).
__global__ __launch_bounds__(256) void btile_offsets(BatchParams P, int* __restrict__ toffs, int* __restrict__ fail) {
    constexpr int kSlots = 16;
    constexpr int kEmpty = INT32_MIN;
    __shared__ int tab[kSlots];
    __shared__ int cnt, bad;
    const int t = blockIdx.x;
    if (threadIdx.x < kSlots) tab[threadIdx.x] = kEmpty;
    if (threadIdx.x == 0) {
        cnt = 0;
        bad = 0;
    }
    __syncthreads();
    const int s = P.tile_sys[t];
    const int lo = P.sys_row[s], hi = P.sys_row[s + 1];
    for (int i = P.tile_row[t] + (int)threadIdx.x; i < P.tile_row[t + 1]; i += blockDim.x) {
        int prev = -1;
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            const int c = P.col[k];
            if (c < lo || c >= hi || c <= prev) {
                atomicOr(&bad, 1);
                break;
            }
            prev = c;
            const int d = c - i;
            int h = (int)(((uint32_t)d * 2654435761u) >> 28);
            for (int probe = 0; probe < kSlots; ++probe) {
                const int old = atomicCAS(&tab[h], kEmpty, d);
                if (old == kEmpty) {
                    if (atomicAdd(&cnt, 1) >= kBDiaMax) atomicOr(&bad, 1);
                    break;
                }
                if (old == d) break;
                h = (h + 1) & (kSlots - 1);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kSlots) toffs[t * kSlots + threadIdx.x] = tab[threadIdx.x];
    if (threadIdx.x == 0 && bad) atomicOr(fail, 1);
}

template <typename VT>
__global__ __launch_bounds__(256) void bfill_dia(BatchParams P, VT* __restrict__ dval, uint8_t* __restrict__ dmask,
                                                 int* __restrict__ fail) {
    const int t = blockIdx.x;
    const BDiaSys& Y = P.dsys[P.tile_sys[t]];  // read from memory: the offset scan indexes it at run time
    const VT* __restrict__ val = static_cast<const VT*>(P.val);
    const int ns = Y.r1 - Y.r0;
    for (int i = P.tile_row[t] + (int)threadIdx.x; i < P.tile_row[t + 1]; i += blockDim.x) {
        uint32_t m = 0;
        int prev = -1;
        bool bad = false;
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            const int off = P.col[k] - i;
            int d = 0;
            while (d < Y.nd && Y.off[d] != off) ++d;
            if (d >= Y.nd || d <= prev) {
                bad = true;
                break;
            }
            prev = d;
            m |= 1u << d;
            dval[Y.base + (size_t)d * ns + (i - Y.r0)] = val[k];
        }
        dmask[i] = (uint8_t)m;
        if (bad) atomicOr(fail, 1);
    }
}

hipError_t batch_launch_tile_offsets(const BatchParams& P, int* toffs, int* fail, hipStream_t s) {
    if (P.T <= 0) return hipSuccess;
    btile_offsets<<<P.T, 256, 0, s>>>(P, toffs, fail);
    return hipGetLastError();
}

hipError_t batch_launch_fill_dia(const BatchParams& P, bool f32, void* dval, uint8_t* dmask, int* fail,
                                 hipStream_t s) {
    if (P.T <= 0) return hipSuccess;
    if (f32)
        bfill_dia<float><<<P.T, 256, 0, s>>>(P, static_cast<float*>(dval), dmask, fail);
    else
        bfill_dia<double><<<P.T, 256, 0, s>>>(P, static_cast<double*>(dval), dmask, fail);
    return hipGetLastError();
}

hipError_t batch_launch_axpy(size_t m, const double* d, double* x, hipStream_t s) {
    size_t g = (m + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    baxpy<<<(unsigned)g, 256, 0, s>>>(m, d, x);
    return hipGetLastError();
}

}  // namespace xj
