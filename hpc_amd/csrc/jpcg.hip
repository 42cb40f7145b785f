// HIP kernels for the xJPCG path on gfx950 and their host launchers.
//
//  jpcg_persistent   one cooperative launch per solve: vector set-up, every
//                    CG iteration and the per-iteration records, separated by
//                    grid barriers (the reference likewise runs the whole loop
//                    on the card from one host launch, pcg/sw/src/cgHost.cpp:447-471)
//  cg_*              the same loop as three or four launches per iteration
//  spmv_tiles        standalone y = A x (or y = b - A x), one workgroup per tile
//  build_col16       compact int16 column stream (col - first row of the tile)
//  dot_partials / sum_partials   fixed-order two-pass dot product
//  gather_values     KEEP_NZ_LAYOUT value refresh through the stored permutation
//  sddmm_tiles       g_k = -lam[row_k] * x[col_k] on the sparsity pattern
//  spmv_pattern / cg_spmv_pattern SpMV on column-offset pattern tiles (stencils)
//  spmv_dia / cg_spmv_dia         SpMV on the diagonal layout (few global offsets)
//  spmv_panels / cg_spmv_panels   SpMV on the panel-swept layout (operators
//                    whose columns scatter over x), standalone and in the CG loop
//
// Kernels that read the matrix are instantiated for fp64 / fp32 values (VT)
// and int32 / int16-delta columns (CT).
#include "kernels.hpp"
#include "launch.hpp"

namespace xj {

// --------------------------------------------------------------------------
// persistent JPCG

template <typename VT, typename CT>
__global__ __launch_bounds__(kBlock) void jpcg_persistent(JpcgParams P) {
    __shared__ double lds[kTileNnz];
    __shared__ double red[kWaves];
    __shared__ int sflag;
    const int G = gridDim.x;
    const int L = xcd_remap(blockIdx.x, G);
    const int tid = threadIdx.x;
    const int gtid = L * kBlock + tid;
    const int gthreads = G * kBlock;
    const int n = P.n;
    const VT* __restrict__ val = static_cast<const VT*>(P.val);
    const CT* __restrict__ col = static_cast<const CT*>(P.col);
    double* __restrict__ x = P.x;
    double* __restrict__ r = P.r;
    double* __restrict__ p = P.p;
    double* __restrict__ Ap = P.Ap;
    double* __restrict__ J = P.J;
    double* part = P.part;
    unsigned long long target = 0;
    const TileSweep sw = tile_sweep(L, G, P.ntiles);

    // set-up (reference cgVector.hpp:101-114): r0 = b, J = 1/diag, z0 = J r0, p0 = z0, x0 = 0
    double dot = 0.0, rz = 0.0;
    for (int i = gtid; i < n; i += gthreads) {
        const double bi = P.b[i];
        const double Ji = 1.0 / P.diag[i];
        const double zi = Ji * bi;
        r[i] = bi;
        J[i] = Ji;
        p[i] = zi;
        x[i] = 0.0;
        dot += bi * bi;
        rz += bi * zi;
    }
    dot = block_sum(dot, red);
    rz = block_sum(rz, red);
    if (tid == 0) {
        part[3 * G + L] = dot;
        part[4 * G + L] = rz;
    }
    int K = 0;
    double rr = 0.0;
    bool ok = grid_barrier(P.bar, target, G, P.abortf, &sflag);
    if (ok) {
        dot = grid_total(part + 3 * G, G, red);
        rz = grid_total(part + 4 * G, G, red);
    }
    const double rz0 = rz;
    const double tols = dot * P.tol * P.tol;
    const uint32_t vec_size = (uint32_t)((n + 3) / 4 * 4);
    if (ok && L == 0 && tid == 0 && P.records) {
        Record rec;
        rec.id = 0; rec.iter = 0; rec.vec_size = vec_size; rec.max_iter = P.max_rounds;
        rec.tols = tols; rec.res = dot; rec.rz = rz; rec.alpha = 0.0; rec.beta = 0.0; rec.clock = 0;
        P.records[0] = rec;
    }
    // b == 0: x = 0 is the exact solution; no iteration runs
    if (ok && dot > 0.0) {
        uint32_t it = 0;
        for (;;) {
            // A: Ap = A p and the p.Ap partial (reference krnl_storeApk.cpp:63-76)
            double pap = 0.0;
            for (int t = sw.first; t < sw.end; t += sw.step) {
                double y;
                int row;
                tile_spmv<VT, CT>(t, P.tile_row, P.tile_k, P.nnz, P.rowptr, col, val, P.dsplit, p, lds, red, &y,
                                  &row);
                if (row >= 0) {
                    Ap[row] = y;
                    pap += p[row] * y;
                }
            }
            pap = block_sum(pap, red);
            if (tid == 0) part[L] = pap;
            if (!grid_barrier(P.bar, target, G, P.abortf, &sflag)) break;
            const double alpha = rz / grid_total(part, G, red);
            // B: r -= alpha Ap; z = J r; r.r and r.z partials (reference update_rk_jacobi.hpp:82-110)
            double rr_l = 0.0, rz_l = 0.0;
            for (int i = gtid; i < n; i += gthreads) {
                const double ri = -alpha * Ap[i] + r[i];
                r[i] = ri;
                const double zi = J[i] * ri;
                rr_l += ri * ri;
                rz_l += ri * zi;
            }
            rr_l = block_sum(rr_l, red);
            rz_l = block_sum(rz_l, red);
            if (tid == 0) {
                part[G + L] = rr_l;
                part[2 * G + L] = rz_l;
            }
            if (!grid_barrier(P.bar, target, G, P.abortf, &sflag)) break;
            rr = grid_total(part + G, G, red);
            const double rzn = grid_total(part + 2 * G, G, red);
            const double beta = rzn / rz;
            rz = rzn;
            ++K;
            if (L == 0 && tid == 0 && P.records) {
                Record rec;
                rec.id = 0; rec.iter = it; rec.vec_size = vec_size; rec.max_iter = P.max_rounds;
                rec.tols = tols; rec.res = rr; rec.rz = rz; rec.alpha = alpha; rec.beta = beta;
                rec.clock = wall_clock64();
                P.records[K] = rec;
            }
            // stop rule of control.hpp:83-88 with Task::meetTol / increase (tasks.hpp:43-47)
            bool more = false;
            if (!(rr < tols)) {
                ++it;
                more = it < P.max_rounds;
            }
            // C: x += alpha p (reference update_xk.hpp:53-67) with the p of this iteration, then
            // p = beta p + z (update_pk.hpp:52-72); skipped after the last iteration
            for (int i = gtid; i < n; i += gthreads) {
                const double pi = p[i];
                x[i] = alpha * pi + x[i];
                if (more) p[i] = beta * pi + J[i] * r[i];
            }
            if (!more) break;
            if (!grid_barrier(P.bar, target, G, P.abortf, &sflag)) break;
        }
    }
    if (L == 0 && tid == 0) {
        P.result[0] = (double)K;
        P.result[1] = rr;
        P.result[2] = dot;
        P.result[3] = rz0;
        P.result[4] = (double)__hip_atomic_load(P.abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// --------------------------------------------------------------------------
// multi-kernel JPCG: coherence by kernel boundaries.
//
//   cg_spmv      Ap = A p, p.Ap partial per tile           (one workgroup per tile)
//   cg_alpha     alpha = rz / sum of the partials           (one workgroup; only when tiles are many)
//   cg_update_r  r -= alpha Ap; r.r and r.z partials       (G2, grid-stride)
//   cg_update_xp beta = rz'/rz; x += alpha p; p = beta p + J r; record; stop rule
//
// Partials are reduced in one fixed order, either once by cg_alpha or by
// every workgroup of the vector kernels, so all workgroups hold bit-identical
// alpha / beta.  The loop state lives in CgState; each field is written by
// workgroup 0 of one kernel only and read by the others only in later
// kernels, so no kernel reads a field that is being written during its own
// launch.  Once `stop` is set the remaining enqueued launches return at once.

// Work vectors r, p, Ap are stored as PT: double, or float for the inner
// operator of a refinement (P.w32), which then also keeps the Jacobi factors
// J = 1/diag in PT.  Values are widened to double before any arithmetic and
// rounded to PT only when stored; dot products use the stored values and
// accumulate in double; x and the loop state stay double.  With PT = double
// every expression is the plain fp64 loop.
template <typename PT>
__device__ __forceinline__ PT* wv(double* p) { return reinterpret_cast<PT*>(p); }

template <typename PT>
__global__ __launch_bounds__(kBlock) void cg_setup(CgParams P) {
    __shared__ double red[kWaves];
    const int L = (int)blockIdx.x;
    const int stride = gridDim.x * kBlock;
    PT* __restrict__ r = wv<PT>(P.r);
    PT* __restrict__ p = wv<PT>(P.p);
    PT* __restrict__ J = wv<PT>(P.J);
    double dot = 0.0, rz = 0.0;
    for (int i = L * kBlock + threadIdx.x; i < P.n; i += stride) {
        const double bi = P.b[i];
        const double ri = (double)(PT)bi;
        const double Ji = (double)(PT)(1.0 / P.diag[i]);
        const double zi = Ji * ri;
        r[i] = (PT)ri;
        J[i] = (PT)Ji;
        p[i] = (PT)zi;
        P.x[i] = 0.0;
        dot += bi * bi;  // the stop rule's ||b|| from the exact right-hand side
        rz += ri * zi;
    }
    dot = block_sum(dot, red);
    rz = block_sum(rz, red);
    if (threadIdx.x == 0) {
        P.partB[L] = dot;
        P.partC[L] = rz;
    }
}

__global__ __launch_bounds__(kBlock) void cg_setup_finish(CgParams P) {
    __shared__ double red[kWaves];
    const double dot = grid_total(P.partB, P.G2, red);
    const double rz = grid_total(P.partC, P.G2, red);
    if (threadIdx.x == 0) {
        CgState* st = P.st;
        st->dot = dot;
        st->tols = dot * P.tol * P.tol;
        st->rr = 0.0;
        st->rz0 = rz;
        st->rz[0] = rz;
        st->rz[1] = 0.0;
        st->alpha = 0.0;
        st->k = 0xFFFFFFFFu;  // the first cg_spmv advances it to 0
        st->it = 0;
        st->stop = dot > 0.0 ? 0 : 1;  // b == 0: x = 0, no iteration
        if (P.host_stop && !(dot > 0.0)) {
            *P.host_stop = 1;
            __threadfence_system();
        }
        st->stop_k2 = 0;
        st->K = 0;
        if (P.records) {
            Record rec;
            rec.id = 0; rec.iter = 0; rec.vec_size = (uint32_t)((P.n + 3) / 4 * 4); rec.max_iter = P.max_rounds;
            rec.tols = st->tols; rec.res = dot; rec.rz = rz; rec.alpha = 0.0; rec.beta = 0.0; rec.clock = 0;
            P.records[0] = rec;
        }
    }
}

// CG-loop SpMV on row tiles: logical workgroup w runs tiles kRowTiles*w ..
// (consecutive rows), the p.Ap partial of tile t in partA[t] (the same
// partials, in the same order, whatever the layout).  The stop flag, the
// tiles' loads and this lane's p[row] are all issued before the workgroup
// waits on any of them.
template <typename VT, typename CT>
__global__ __launch_bounds__(kBlock) void cg_spmv(CgParams P) {
    __shared__ double lds[kTileNnz];
    __shared__ double red[kWaves];
    CgState* st = P.st;
    const int stop = st->stop;
    const int w = (int)blockIdx.x;
    const double* __restrict__ p = P.p;
    const CT* __restrict__ col = static_cast<const CT*>(P.col);
    const VT* __restrict__ val = static_cast<const VT*>(P.val);
    RowTile<VT, CT> T[kRowTiles];
    double prow[kRowTiles];
#pragma unroll
    for (int u = 0; u < kRowTiles; ++u) {
        const int t = w * kRowTiles + u;
        prow[u] = 0.0;
        if (t < P.tiles) {
            row_tile_load<VT, CT>(T[u], t, P.tile_row, P.tile_k, P.nnz, P.rowptr, col, val);
            if ((int)threadIdx.x < T[u].r1 - T[u].r0) prow[u] = p[T[u].r0 + threadIdx.x];
        }
    }
    if (stop) return;
#pragma unroll
    for (int u = 0; u < kRowTiles; ++u) {
        if (T[u].t < 0) break;
        double y;
        int row;
        row_tile_compute<VT, CT>(T[u], P.rowptr, col, val, P.dsplit, p, lds, red, &y, &row);
        double pap = 0.0;
        if (row >= 0) {
            P.Ap[row] = y;
            pap = prow[u] * y;
        }
        pap = block_sum(pap, red);
        if (threadIdx.x == 0) P.partA[T[u].t] = pap;
    }
    // k advances after the last load: a store ahead of the loads could alias
    // the tables and would turn their scalar loads into vector loads
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = st->k + 1;  // no other workgroup of this launch reads k
}

// CG-loop SpMV on column-offset pattern tiles: logical workgroup w runs tiles
// 2w and 2w+1 (consecutive rows); the p.Ap partial of tile t is partA[t]
template <typename VT>
__global__ __launch_bounds__(kBlock) void cg_spmv_pattern(CgParams P) {
    __shared__ double lds[kTileNnz];
    __shared__ double red[kWaves];
    __shared__ int loffs[32];
    CgState* st = P.st;
    // the stop flag, the tiles' loads and this lane's p[row] are all issued
    // before the workgroup waits on any of them
    const int stop = st->stop;
    const int w = (int)blockIdx.x;
    const double* __restrict__ p = P.p;
    const VT* __restrict__ val = static_cast<const VT*>(P.val);
    PatternTile<VT> T[kPatternTiles];
    double prow[kPatternTiles];
#pragma unroll
    for (int u = 0; u < kPatternTiles; ++u) {
        const int t = w * kPatternTiles + u;
        prow[u] = 0.0;
        if (t < P.tiles) {
            pattern_tile_load<VT>(T[u], t, P.tile_row, P.tile_k, P.nnz, P.rowptr, P.pmask, P.poffs, val);
            if ((int)threadIdx.x < T[u].r1 - T[u].r0) prow[u] = p[T[u].r0 + threadIdx.x];
        }
    }
    if (stop) return;
#pragma unroll
    for (int u = 0; u < kPatternTiles; ++u) {
        if (T[u].t < 0) break;
        double y;
        int row;
        pattern_tile_compute<VT>(T[u], P.dsplit, p, lds, loffs, &y, &row);
        double pap = 0.0;
        if (row >= 0) {
            P.Ap[row] = y;
            pap = prow[u] * y;
        }
        pap = block_sum(pap, red);
        if (threadIdx.x == 0) P.partA[T[u].t] = pap;
    }
    // k advances after the last load: a store ahead of the loads could alias
    // the tables and would turn their scalar loads into vector loads
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = st->k + 1;  // no other workgroup of this launch reads k
}

// second reduction level for many tiles: workgroup b sums a contiguous slice
// of the per-tile partials (fixed order) into partA2[b]
__global__ __launch_bounds__(kBlock) void cg_alpha(CgParams P) {
    __shared__ double red[kWaves];
    if (P.st->stop) return;
    const int b = blockIdx.x;
    const int t0 = (int)((int64_t)P.ntiles * b / P.nA2), t1 = (int)((int64_t)P.ntiles * (b + 1) / P.nA2);
    double s = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += kBlock) s += P.partA[t];
    s = block_sum(s, red);
    if (threadIdx.x == 0) P.partA2[b] = s;
}

// The p.Ap partials: the SpMV's, or cg_alpha's fold of them.
__device__ __forceinline__ const double* pap_parts(const CgParams& P) { return P.split_alpha ? P.partA2 : P.partA; }
__device__ __forceinline__ int pap_count(const CgParams& P) { return P.split_alpha ? P.nA2 : P.ntiles; }

// The vector kernels load their first row's operands before they sum the
// partials they need (alpha; beta and the stop rule), so those memory round
// trips overlap instead of following one another; the partial sets are then
// reduced together with one pair of barriers.  Sums and their order are
// those of one grid_total per set.
template <typename PT>
__global__ __launch_bounds__(kBlock) void cg_update_r(CgParams P) {
    __shared__ double red[2 * kWaves];
    CgState* st = P.st;
    const int L = (int)blockIdx.x;
    const int stop = st->stop;
    if (stop) {
        if (L == 0 && threadIdx.x == 0) st->stop_k2 = stop;
        return;
    }
    const int stride = gridDim.x * kBlock;
    const int i0 = L * kBlock + threadIdx.x;
    PT* __restrict__ r = wv<PT>(P.r);
    const PT* __restrict__ Ap = wv<PT>(P.Ap);
    const PT* __restrict__ J = wv<PT>(P.J);
    double a0 = 0.0, r0 = 0.0, J0 = 0.0;
    if (i0 < P.n) {
        a0 = (double)Ap[i0];
        r0 = (double)r[i0];
        J0 = (double)J[i0];
    }
    double pap[1] = {strided_part(pap_parts(P), pap_count(P))};
    block_sums<1>(pap, red);
    const double alpha = st->rz[st->k & 1] / pap[0];
    double acc[2] = {0.0, 0.0};  // r.r, r.z
    // every operand of a row is loaded before its store to r (J could alias r
    // for all the compiler knows, which would order its load after the store)
    // (contractions written out: the compiler may otherwise fuse a different
    // product in the peeled first row than in the loop)
    auto row = [&](int i, double a, double rv, double Jv) {
        const double ri = (double)(PT)__builtin_fma(-alpha, a, rv);
        r[i] = (PT)ri;
        const double zi = Jv * ri;
        acc[0] = __builtin_fma(ri, ri, acc[0]);
        acc[1] = __builtin_fma(ri, zi, acc[1]);
    };
    if (i0 < P.n) row(i0, a0, r0, J0);
    for (int i = i0 + stride; i < P.n; i += stride) {
        const double a = (double)Ap[i], rv = (double)r[i], Jv = (double)J[i];
        row(i, a, rv, Jv);
    }
    block_sums<2>(acc, red);
    if (threadIdx.x == 0) {
        P.partB[L] = acc[0];
        P.partC[L] = acc[1];
        if (L == 0) st->stop_k2 = 0;  // stored last, after every load of this workgroup
    }
}

template <typename PT>
__global__ __launch_bounds__(kBlock) void cg_update_xp(CgParams P) {
    __shared__ double red[3 * kWaves];
    CgState* st = P.st;
    if (st->stop_k2) return;
    const int L = (int)blockIdx.x;
    const uint32_t k = st->k;
    const double rz = st->rz[k & 1];
    const int stride = gridDim.x * kBlock;
    const int i0 = L * kBlock + threadIdx.x;
    PT* __restrict__ p = wv<PT>(P.p);
    const PT* __restrict__ r = wv<PT>(P.r);
    const PT* __restrict__ J = wv<PT>(P.J);
    double p0 = 0.0, x0 = 0.0, J0 = 0.0, r0 = 0.0;
    if (i0 < P.n) {
        p0 = (double)p[i0];
        x0 = P.x[i0];
        J0 = (double)J[i0];
        r0 = (double)r[i0];
    }
    const double* pa = pap_parts(P);
    const int na = pap_count(P);
    double sums[3] = {strided_head(pa, na), strided_head(P.partB, P.G2), strided_head(P.partC, P.G2)};
    sums[0] = strided_tail(pa, na, sums[0]);
    sums[1] = strided_tail(P.partB, P.G2, sums[1]);
    sums[2] = strided_tail(P.partC, P.G2, sums[2]);
    block_sums<3>(sums, red);
    const double alpha = rz / sums[0];
    const double rr = sums[1];
    const double rzn = sums[2];
    const double beta = rzn / rz;
    // every operand of a row is loaded before its first store (x could alias
    // J or r for all the compiler knows, which would order their loads after
    // it: two memory round trips per row instead of one)
    auto row = [&](int i, double pi, double xi, double Ji, double ri) {
        P.x[i] = __builtin_fma(alpha, pi, xi);
        p[i] = (PT)__builtin_fma(beta, pi, Ji * ri);
    };
    if (i0 < P.n) row(i0, p0, x0, J0, r0);
    for (int i = i0 + stride; i < P.n; i += stride) {
        const double pi = (double)p[i], xi = P.x[i], Ji = (double)J[i], ri = (double)r[i];
        row(i, pi, xi, Ji, ri);
    }
    if (L == 0 && threadIdx.x == 0) {
        st->rz[(k + 1) & 1] = rzn;
        st->rr = rr;
        st->K = (int)(k + 1);
        uint32_t it = st->it;
        if (P.records) {
            Record rec;
            rec.id = 0; rec.iter = it; rec.vec_size = (uint32_t)((P.n + 3) / 4 * 4); rec.max_iter = P.max_rounds;
            rec.tols = st->tols; rec.res = rr; rec.rz = rzn; rec.alpha = alpha; rec.beta = beta;
            rec.clock = wall_clock64();
            P.records[k + 1] = rec;
        }
        bool more = false;
        if (!(rr < st->tols)) {
            ++it;
            more = it < P.max_rounds;
        }
        st->it = it;
        if (!more) {
            st->stop = 1;
            if (P.host_stop) {
                *P.host_stop = 1;
                __threadfence_system();
            }
        }
    }
}

// --------------------------------------------------------------------------
// standalone kernels

template <typename VT, typename CT>
__global__ __launch_bounds__(kBlock) void spmv_tiles(int ntiles, int nnz, const int* __restrict__ tile_row,
                                                     const int* __restrict__ tile_k,
                                                     const int* __restrict__ rowptr, const CT* __restrict__ col,
                                                     const VT* __restrict__ val, const double* __restrict__ dsplit,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     const double* __restrict__ b) {
    __shared__ double lds[kTileNnz];
    __shared__ double red[kWaves];
    const int w = (int)blockIdx.x;
    RowTile<VT, CT> T[kRowTiles];
#pragma unroll
    for (int u = 0; u < kRowTiles; ++u) {
        const int t = w * kRowTiles + u;
        if (t < ntiles) row_tile_load<VT, CT>(T[u], t, tile_row, tile_k, nnz, rowptr, col, val);
    }
#pragma unroll
    for (int u = 0; u < kRowTiles; ++u) {
        if (T[u].t < 0) break;
        double v;
        int row;
        row_tile_compute<VT, CT>(T[u], rowptr, col, val, dsplit, x, lds, red, &v, &row);
        if (row >= 0) y[row] = b ? b[row] - v : v;
    }
}

// --------------------------------------------------------------------------
// Panel-swept SpMV.
//
// For operators whose columns scatter over the whole of x (random graphs),
// every gather of the tile kernel misses the XCD's L2 and costs a full line
// fetch from the MALL or HBM.  In the panel layout, row block b keeps its
// rows' running sums in LDS and walks the column panels (kPanelCols columns
// each) in order; all workgroups walk them in the same order, so the x
// entries being gathered at any moment lie in one 512 KB window that stays in
// L2.  Block b's entries of panel p are segment [seg[b*P+p], seg[b*P+p+1]),
// row-major with ascending columns.  A chunk of up to kPanelChunk entries is
// loaded (value, uint16 column offset, uint16 row word: row offset and, on a
// run's first entry, the run's length), the products are staged in LDS, and
// each row's run in the chunk is added to the row's sum in one of two bins.
// A run of fewer than kPanelRunScan entries is added left to right by the
// lane holding its first entry; with the length known, its loads of the run's
// products go out together (walking the run to find its end cost one LDS
// round trip per entry, and every chunk waited at its barrier for the longest
// run: 4M-row random graph, 746 us against 533 us with no gathers at all and
// 220 us without the run sums).  A longer run is summed by the wavefront
// holding its first entry in a fixed order (64 entries per step, an xor
// butterfly) and its total added to the row's sum.  A row's runs in later
// chunks and panels hold larger columns and are added after it, so a row
// whose runs are all short is summed left to right from dsplit[i]*x[i] (or
// 0), exactly as the tile kernel sums a short row.
constexpr int kPanelCols = kPanelColsHost;
constexpr int kPanelBlock = kPanelBlockHost;
constexpr int kPanelChunk = kPanelChunkHost;
static_assert(kPanelChunk % kPanelBlock == 0, "whole chunk entries per lane");

// acc: LDS of panel_lds_bytes(A.block_rows) bytes; on return acc[i] holds
// y[b*block_rows + i] for the block's rows (after a __syncthreads).
template <typename VT, typename XT = double>
__device__ __forceinline__ void panel_rows(const PanelDev& A, int n, int b, const double* __restrict__ dsplit,
                                           const XT* __restrict__ x, double* acc) {
#pragma clang fp contract(off)
    constexpr int J = kPanelChunk / kPanelBlock;
    double* prod = acc + A.block_rows;
    uint16_t* lr = reinterpret_cast<uint16_t*>(prod + kPanelChunk);
    const int tid = threadIdx.x;
    const int row0 = b * A.block_rows;
    const int rows = min(A.block_rows, n - row0);
    for (int i = tid; i < rows; i += kPanelBlock) acc[i] = dsplit ? dsplit[row0 + i] * (double)x[row0 + i] : 0.0;
    __syncthreads();
    const int* seg = A.seg + (int64_t)b * A.P;
    const VT* __restrict__ val = static_cast<const VT*>(A.val);
    const uint16_t* __restrict__ off = A.off;
    const uint16_t* __restrict__ lrow = A.lrow;
    for (int p = 0; p < A.P; ++p) {
        const int s1 = seg[p + 1];
        const XT* __restrict__ xp = x + (int64_t)p * kPanelCols;
        for (int c0 = seg[p]; c0 < s1; c0 += kPanelChunk) {
            const int cn = min(kPanelChunk, s1 - c0);
            // lanes past the chunk's end re-read its first entry (in bounds)
            VT v[J];
            int o[J];
            uint16_t w[J];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int k = tid + j * kPanelBlock;
                const int e = c0 + (k < cn ? k : 0);
                v[j] = ld_stream(val + e);
                o[j] = off[e];
                w[j] = lrow[e];
            }
            // every stream load of the chunk is issued before anything waits
            // for one (left to itself the scheduler held a value load behind
            // the first column load's return: one more HBM latency per chunk)
            __builtin_amdgcn_sched_barrier(0);
            double pr[J];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int k = tid + j * kPanelBlock;
                pr[j] = (double)v[j] * (double)xp[o[j]];
                prod[k] = pr[j];
                lr[k] = k < cn ? w[j] : (uint16_t)0xFFFF;
            }
            __syncthreads();
            // the lane holding a run's first entry adds a short run (its
            // length in the word): its own product, then the run's other
            // products from LDS, all loaded before the first add
            bool wide[J];
            int rw[J];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int k = tid + j * kPanelBlock;
                const int len = w[j] >> kPanelRowBits;
                rw[j] = w[j] & ((1 << kPanelRowBits) - 1);
                wide[j] = k < cn && len >= kPanelRunScan;
                if (k < cn && len != 0 && len < kPanelRunScan) {
                    const int r = rw[j];
                    double a = acc[r];
                    a += pr[j];
                    int q = k + 1;
                    const int end = k + len;
                    for (; q + 3 < end; q += 4) {
                        const double t0 = prod[q], t1 = prod[q + 1], t2 = prod[q + 2], t3 = prod[q + 3];
                        a += t0;
                        a += t1;
                        a += t2;
                        a += t3;
                    }
                    for (; q < end; ++q) a += prod[q];
                    acc[r] = a;
                }
            }
            // a long run (kPanelRunScan or more entries, a high-degree row):
            // the wavefront holding its first entry sums it, 64 entries per
            // step (lane l adds entries start+l, start+l+64, ... in order,
            // then the xor butterfly); the run ends at the first entry whose
            // row word differs (continuation words hold the bare row offset,
            // slots past the chunk 0xFFFF)
            const int lane = tid & 63;
#pragma unroll
            for (int j = 0; j < J; ++j) {
                uint64_t m = __ballot(wide[j]);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    const int k0r = (tid - lane + l) + j * kPanelBlock;  // the run's first entry
                    const int r = __shfl(rw[j], l, 64);
                    double part = 0.0;
                    for (int q = k0r + lane;; q += 64) {
                        const int qq = min(q, kPanelChunk - 1);
                        const bool in = q < cn && (q == k0r || lr[qq] == r);
                        if (in) part += prod[qq];
                        if (__ballot(!in) != 0) break;
                    }
                    part = wave_sum(part);
                    if (lane == l) acc[r] = acc[r] + part;
                }
            }
            __syncthreads();
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kPanelBlock) void spmv_panels(PanelDev A, int n, const double* __restrict__ dsplit,
                                                           const double* __restrict__ x, double* __restrict__ y,
                                                           const double* __restrict__ b) {
    extern __shared__ double plds[];
    const int blk = blockIdx.x;
    panel_rows<VT>(A, n, blk, dsplit, x, plds);
    const int row0 = blk * A.block_rows, rows = min(A.block_rows, n - row0);
    for (int i = threadIdx.x; i < rows; i += kPanelBlock) st_stream(y + row0 + i, b ? b[row0 + i] - plds[i] : plds[i]);
}

// CG-loop SpMV on the panel layout: Ap = A p and block blk's p.Ap partial in partA[blk]
template <typename VT, typename PT>
__global__ __launch_bounds__(kPanelBlock) void cg_spmv_panels(CgParams P) {
    extern __shared__ double plds[];
    __shared__ double red[kPanelBlock / 64];
    CgState* st = P.st;
    if (st->stop) return;
    const int blk = blockIdx.x;
    const PT* __restrict__ p = wv<PT>(P.p);
    PT* __restrict__ Ap = wv<PT>(P.Ap);
    panel_rows<VT, PT>(P.pan, P.n, blk, P.dsplit, p, plds);
    const int row0 = blk * P.pan.block_rows, rows = min(P.pan.block_rows, P.n - row0);
    double pap = 0.0;
    for (int i = threadIdx.x; i < rows; i += kPanelBlock) {
        const PT y = (PT)plds[i];  // with fp32 work vectors the partial uses the stored Ap
        st_stream(Ap + row0 + i, y);
        pap += (double)p[row0 + i] * (double)y;
    }
    pap = block_sum_w<kPanelBlock / 64>(pap, red);
    if (threadIdx.x == 0) P.partA[blk] = pap;
    // k advances after the last load: a store ahead of the loads could alias
    // the tables and would turn their scalar loads into vector loads
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = st->k + 1;  // no other workgroup of this launch reads k
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void spmv_pattern(int ntiles, int nnz, const int* __restrict__ tile_row,
                                                       const int* __restrict__ tile_k,
                                                       const int* __restrict__ rowptr, const uint32_t* __restrict__ pmask,
                                                       const int* __restrict__ poffs, const VT* __restrict__ val,
                                                       const double* __restrict__ dsplit, const double* __restrict__ x,
                                                       double* __restrict__ y, const double* __restrict__ b) {
    __shared__ double lds[kTileNnz];
    __shared__ int loffs[32];
    const int w = (int)blockIdx.x;
    PatternTile<VT> T[kPatternTiles];
#pragma unroll
    for (int u = 0; u < kPatternTiles; ++u) {
        const int t = w * kPatternTiles + u;
        if (t < ntiles) pattern_tile_load<VT>(T[u], t, tile_row, tile_k, nnz, rowptr, pmask, poffs, val);
    }
#pragma unroll
    for (int u = 0; u < kPatternTiles; ++u) {
        if (T[u].t < 0) break;
        double v;
        int row;
        pattern_tile_compute<VT>(T[u], dsplit, x, lds, loffs, &v, &row);
        if (row >= 0) y[row] = b ? b[row] - v : v;
    }
}

// --------------------------------------------------------------------------
// diagonal layout

// Rows i0 .. i0+R-1 of A times x on the diagonal layout (R = 2 for fp64
// values, 1 for fp32), C diagonals at a time (R * C = 16): the zero-padded
// values of each diagonal (one R-wide load: the arrays have an even leading
// dimension and i0 is a multiple of R) and x at the entries' columns (clamped
// into [0, n) for a diagonal a row does not have) are loaded together with
// the masks, then each row's products are added in ascending offset order,
// each rounded before it is added, left to right from dsplit[i]*x[i] (or 0)
// -- the arithmetic of the row tiles, row by row.  Value loads and the x
// reads of one offset are contiguous across the lanes of a wave.  A padding
// row (i0 + r == n when n is odd) reads in-bounds copies of row n-1's
// operands and has mask 0; callers drop its sum.  dia_rows_s reads x one
// element per row; dia_rows_v (R > 1, n >= R) reads a diagonal's R x values
// with one load.
template <int R>
struct DiaLane {
    static constexpr int C = 16 / R;  // diagonals per batch of loads
};
static_assert(dia_rows_per_lane(true, kDiaWideRows) == 1 && dia_rows_per_lane(false, kDiaWideRows) == 2 &&
                  dia_rows_per_lane(false, kDiaWideRows - 1) == 1,
              "the kernels are instantiated for these rows per lane");
static_assert(kDiaBlock == kBlock && kDiaWideBlock % 64 == 0 && kDiaWideBlock <= 1024,
              "one-row diagonal kernels run kBlock lanes, two-row ones whole wavefronts");
static_assert(kVectorBlock == kBlock, "vector kernels run kBlock lanes");

template <typename T, int R>
__device__ __forceinline__ void ld_rows(const T* p, T (&o)[R]) {
    if constexpr (R == 1) {
        o[0] = ld_stream(p);
    } else {
        typedef T V __attribute__((ext_vector_type(R)));
        const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
#pragma unroll
        for (int r = 0; r < R; ++r) o[r] = v[r];
    }
}

template <typename VT, int R, typename XT = double>
__device__ __forceinline__ void dia_rows_s(const DiaDev& D, int n, int i0, const double* __restrict__ dsplit,
                                         const XT* __restrict__ x, double (&s)[R]) {
#pragma clang fp contract(off)
    constexpr int C = DiaLane<R>::C;
    const VT* __restrict__ val = static_cast<const VT*>(D.val);
    const int* __restrict__ off = D.off;
    const int nd = D.ndia;
    const size_t ld = (size_t)D.ld;
    uint32_t m[R];
    ld_rows<uint32_t, R>(D.mask + i0, m);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + r < n ? i0 + r : n - 1;
        s[r] = dsplit ? dsplit[i] * (double)x[i] : 0.0;
    }
#pragma nounroll
    for (int d0 = 0; d0 < nd; d0 += C) {
        VT v[C][R];
        XT g[C][R];  // converted where it is used: a conversion here would wait for each load
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int d = d0 + u;
            if (d < nd) {
                ld_rows<VT, R>(val + (size_t)d * ld + i0, v[u]);
                const int o = off[d];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int c = i0 + r + o;
                    c = c < 0 ? 0 : (c >= n ? n - 1 : c);
                    g[u][r] = x[c];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int d = d0 + u;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (d < nd && ((m[r] >> d) & 1u)) s[r] += (double)v[u][r] * (double)g[u][r];
        }
    }
}

// The same rows with x read as one R-wide load per diagonal (R > 1, n >= R):
// the load's start is clamped into [0, n-R]; a column shifted out of [0, n)
// belongs to a diagonal the row does not have (mask 0).
template <typename VT, int R>
__device__ __forceinline__ void dia_rows_v(const DiaDev& D, int n, int i0, const double* __restrict__ dsplit,
                                           const double* __restrict__ x, double (&s)[R]) {
#pragma clang fp contract(off)
    constexpr int C = DiaLane<R>::C;
    typedef double XV __attribute__((ext_vector_type(R), aligned(8)));
    const VT* __restrict__ val = static_cast<const VT*>(D.val);
    const int* __restrict__ off = D.off;
    const int nd = D.ndia;
    const size_t ld = (size_t)D.ld;
    uint32_t m[R];
    ld_rows<uint32_t, R>(D.mask + i0, m);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = i0 + r < n ? i0 + r : n - 1;
        s[r] = dsplit ? dsplit[i] * x[i] : 0.0;
    }
#pragma nounroll
    for (int d0 = 0; d0 < nd; d0 += C) {
        VT v[C][R];
        double g[C][R];
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int d = d0 + u;
            if (d < nd) {
                ld_rows<VT, R>(val + (size_t)d * ld + i0, v[u]);
                const int c0 = i0 + off[d];
                const int cc = c0 < 0 ? 0 : (c0 > n - R ? n - R : c0);
                const XV xv = *reinterpret_cast<const XV*>(x + cc);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    int k = r + c0 - cc;
                    k = k < 0 ? 0 : (k > R - 1 ? R - 1 : k);
                    double t = xv[0];
#pragma unroll
                    for (int q = 1; q < R; ++q) t = k == q ? xv[q] : t;
                    g[u][r] = t;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int d = d0 + u;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (d < nd && ((m[r] >> d) & 1u)) s[r] += (double)v[u][r] * g[u][r];
        }
    }
}

// The kernels pick dia_rows_v or dia_rows_s themselves: behind a shared
// forceinline wrapper the compiler gave the fp32 kernels 84 VGPRs instead of
// 59 (occupancy 5 instead of 8).

// y = A x (or b - A x), R consecutive rows per lane, BS R rows per workgroup
template <typename VT, int R, int BS>
__global__ __launch_bounds__(BS) void spmv_dia(int n, DiaDev D, const double* __restrict__ dsplit,
                                               const double* __restrict__ x, double* __restrict__ y,
                                               const double* __restrict__ b) {
    const int i0 = ((int)blockIdx.x * BS + (int)threadIdx.x) * R;
    if (i0 >= n) return;
    double s[R];
    if constexpr (R > 1) {
        if (n >= R) dia_rows_v<VT, R>(D, n, i0, dsplit, x, s);
        else dia_rows_s<VT, R>(D, n, i0, dsplit, x, s);
    } else {
        dia_rows_s<VT, R>(D, n, i0, dsplit, x, s);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        // streamed: y would otherwise evict x lines other workgroups still
        // gather (160^3 fp64: 172.8 -> 165.9 us, same bits)
        if (i0 + r < n) st_stream(y + i0 + r, b ? b[i0 + r] - s[r] : s[r]);
}

// CG SpMV on the diagonal layout: BS R consecutive rows per workgroup, R per
// lane (every address follows from the block index, so nothing waits on a
// table); Ap is stored and each workgroup leaves one p.Ap partial, a lane's
// rows' products added in row order, an xor-butterfly within each wavefront,
// then the wavefronts in order -- a fixed order.  With fp32 work vectors (PT =
// float) p is widened as it is read, Ap is rounded as it is stored and the
// partial uses the stored value.
template <typename VT, typename PT, int R, int BS>
__global__ __launch_bounds__(BS) void cg_spmv_dia(CgParams P) {
    CgState* st = P.st;
    const int stop = st->stop;
    const int b = (int)blockIdx.x;
    if (stop) return;
    const PT* __restrict__ p = wv<PT>(P.p);
    PT* __restrict__ Ap = wv<PT>(P.Ap);
    double pap = 0.0;
    const int i0 = (b * BS + (int)threadIdx.x) * R;
    if (i0 < P.n) {
        double pi[R], y[R];
#pragma unroll
        for (int r = 0; r < R; ++r) pi[r] = (double)p[i0 + r < P.n ? i0 + r : P.n - 1];
        if constexpr (R > 1 && sizeof(PT) == 8) {
            if (P.n >= R) dia_rows_v<VT, R>(P.dia, P.n, i0, P.dsplit, P.p, y);
            else dia_rows_s<VT, R>(P.dia, P.n, i0, P.dsplit, P.p, y);
        } else {
            dia_rows_s<VT, R, PT>(P.dia, P.n, i0, P.dsplit, p, y);
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (i0 + r < P.n) {
                const PT a = (PT)y[r];
                if constexpr (sizeof(PT) == 8)
                    st_stream(Ap + i0 + r, a);  // see spmv_dia
                else
                    Ap[i0 + r] = a;  // fp32 work vectors: the plain store measured faster
                pap += pi[r] * (double)a;
            }
    }
    // one partial per workgroup rather than per wavefront: a quarter of the
    // partials for cg_alpha or the vector kernels to fold (160^3 mixed solve
    // with fp32 work vectors 43.7 ms against 44.0 ms, same box)
    __shared__ double red[BS / 64];
    pap = block_sum_w<BS / 64>(pap, red);
    if (threadIdx.x == 0) P.partA[b] = pap;
    // k advances after the last load: a store ahead of the loads could alias
    // the tables and would turn their scalar loads into vector loads
    if (blockIdx.x == 0 && threadIdx.x == 0) st->k = st->k + 1;  // no other workgroup of this launch reads k
}

// Diagonal values and masks from the CSR, one lane per row: entry k of row i
// lands in diagonal d with off[d] = col[k] - i, at dval[d * dia_ld(n) + i].
// dval and dmask are zeroed beforehand (a refresh of the values rewrites the
// same slots; the padding row keeps mask 0).
template <typename VT>
__global__ __launch_bounds__(kBlock) void fill_dia(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const VT* __restrict__ val, int ndia, const int* __restrict__ off,
                                                   VT* __restrict__ dval, uint32_t* __restrict__ dmask,
                                                   int* __restrict__ fail) {
    __shared__ int loff[kDiaMax];
    if ((int)threadIdx.x < kDiaMax) loff[threadIdx.x] = (int)threadIdx.x < ndia ? off[threadIdx.x] : INT32_MAX;
    __syncthreads();
    for (int i = blockIdx.x * kBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBlock) {
        uint32_t m = 0;
        int prev = -1;
        bool bad = ndia <= 0 && rowptr[i + 1] > rowptr[i];
        for (int k = rowptr[i]; k < rowptr[i + 1] && !bad; ++k) {
            const int dlt = col[k] - i;
            int lo = 0, hi = ndia - 1;
            while (lo < hi) {  // first slot holding an offset >= dlt
                const int mid = (lo + hi) >> 1;
                if (loff[mid] < dlt) lo = mid + 1;
                else hi = mid;
            }
            if (loff[lo] != dlt || lo <= prev) {
                bad = true;
                break;
            }
            prev = lo;
            m |= 1u << lo;
            dval[(size_t)lo * dia_ld(n) + i] = val[k];
        }
        dmask[i] = m;
        if (bad) atomicOr(fail, 1);
    }
}

// Column-offset pattern plan on the GPU, one workgroup per tile (the same
// plan as the host planner plan_patterns): lane t checks row t (at most 32
// entries, columns strictly ascending and inside [0, n)) and inserts its
// offsets col - row into a 64-slot LDS hash set; the tile's distinct offsets
// (at most 32) are ranked into ascending order, written to poffs[t*32 ..]
// (unused slots 0), and lane t builds row t's mask against them.  Any tile
// that cannot be held sets *fail.
__global__ __launch_bounds__(kBlock) void plan_pattern_tiles(int n, int ntiles, const int* __restrict__ tile_row,
                                                             const int* __restrict__ rowptr,
                                                             const int* __restrict__ col, uint32_t* __restrict__ pmask,
                                                             int* __restrict__ poffs, int unused,
                                                             int* __restrict__ fail) {
    constexpr int kSlots = 64;
    constexpr int kEmpty = INT32_MIN;  // no offset col - row reaches it for n < 2^31
    __shared__ int tab[kSlots];
    __shared__ int sorted[32];
    __shared__ int cnt, bad;
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int tid = threadIdx.x;
    const int r0 = tile_row[t], r1 = tile_row[t + 1];
    if (tid < kSlots) tab[tid] = kEmpty;
    if (tid == 0) {
        cnt = 0;
        bad = rowptr[r1] - rowptr[r0] > kTilePlanNnz ? 1 : 0;
    }
    __syncthreads();
    const bool mine = tid < r1 - r0;
    const int i = r0 + tid;
    const int a = mine ? rowptr[i] : 0, b = mine ? rowptr[i + 1] : 0;
    if (mine && !bad) {
        if (b - a > 32) atomicOr(&bad, 1);
        int prev = -1;
        for (int k = a; k < b && !bad; ++k) {
            const int c = col[k];
            if (c < 0 || c >= n || c <= prev) {
                atomicOr(&bad, 1);
                break;
            }
            prev = c;
            const int d = c - i;
            int h = (int)(((uint32_t)d * 2654435761u) >> 26);  // 6-bit multiplicative hash
            for (int probe = 0; probe < kSlots; ++probe) {
                const int old = atomicCAS(&tab[h], kEmpty, d);
                if (old == kEmpty) {
                    if (atomicAdd(&cnt, 1) >= 32) atomicOr(&bad, 1);
                    break;
                }
                if (old == d) break;
                h = (h + 1) & (kSlots - 1);
            }
        }
    }
    __syncthreads();
    if (bad) {
        if (tid == 0) atomicOr(fail, 1);
        return;
    }
    if (tid < kSlots) {
        const int v = tab[tid];
        if (v != kEmpty) {
            int rank = 0;
            for (int u = 0; u < kSlots; ++u) {
                const int w = tab[u];
                rank += (w != kEmpty && w < v) ? 1 : 0;
            }
            sorted[rank] = v;
        }
    }
    __syncthreads();
    const int m = cnt;
    if (tid < 32) poffs[t * 32 + tid] = tid < m ? sorted[tid] : unused;
    if (mine && pmask) {
        uint32_t mask = 0;
        for (int k = a; k < b; ++k) {
            const int d = col[k] - i;
            int lo = 0, hi = m - 1;
            while (lo < hi) {  // sorted[lo..hi] holds d
                const int mid = (lo + hi) >> 1;
                if (sorted[mid] < d) lo = mid + 1; else hi = mid;
            }
            mask |= 1u << lo;
        }
        pmask[i] = mask;
    }
}

// col16[k] = col[k] - first row of k's tile; *overflow = 1 if any difference leaves int16
__global__ __launch_bounds__(kBlock) void build_col16(int ntiles, const int* __restrict__ tile_row,
                                                      const int* __restrict__ rowptr, const int* __restrict__ col,
                                                      int16_t* __restrict__ col16, int* __restrict__ overflow) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int r0 = tile_row[t];
    const int k0 = rowptr[r0], k1 = rowptr[tile_row[t + 1]];
    int bad = 0;
    for (int k = k0 + threadIdx.x; k < k1; k += kBlock) {
        const int d = col[k] - r0;
        if (d < -32768 || d > 32767) bad = 1;
        col16[k] = (int16_t)d;
    }
    if (bad) atomicOr(overflow, 1);
}

// part[blk] = sum over this block's grid-stride slice of a[i]*b[i]
__global__ __launch_bounds__(kBlock) void dot_partials(int n, const double* __restrict__ a,
                                                       const double* __restrict__ b, double* __restrict__ part) {
    __shared__ double red[kWaves];
    double s = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) s += a[i] * b[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void sum_partials(int m, const double* __restrict__ part,
                                                       double* __restrict__ out) {
    __shared__ double red[kWaves];
    const double s = grid_total(part, m, red);
    if (threadIdx.x == 0) *out = s;
}

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void gather_values(int64_t nnz, const int64_t* __restrict__ perm,
                                                        const TI* __restrict__ src, TO* __restrict__ dst) {
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * kBlock)
        dst[k] = (TO)src[perm[k]];
}

// dst = src for n doubles: 16-byte non-temporal loads and stores, grid-stride
// (the achievable-bandwidth reference: every byte read once, written once)
__global__ __launch_bounds__(kBlock) void stream_copy(int64_t n, const double* __restrict__ src,
                                                      double* __restrict__ dst) {
    typedef unsigned int W __attribute__((ext_vector_type(4)));
    const int64_t n2 = n >> 1;
    const W* __restrict__ s = reinterpret_cast<const W*>(src);
    W* __restrict__ d = reinterpret_cast<W*>(dst);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n2; i += (int64_t)gridDim.x * kBlock)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[n - 1] = src[n - 1];
}

__global__ __launch_bounds__(kBlock) void fill_kernel(int64_t n, double v, double* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) x[i] = v;
}

// x = x + a*d  (iterative-refinement update)
__global__ __launch_bounds__(kBlock) void axpy_kernel(int n, double a, const double* __restrict__ d,
                                                      double* __restrict__ x) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) x[i] = a * d[i] + x[i];
}

// g[k] = -lam[row(k)] * x[col[k]] over one tile per workgroup
__global__ __launch_bounds__(kBlock) void sddmm_tiles(int ntiles, const int* __restrict__ tile_row,
                                                      const int* __restrict__ rowptr, const int* __restrict__ col,
                                                      const double* __restrict__ lam,
                                                      const double* __restrict__ x, double* __restrict__ g) {
    __shared__ int srow[kTileNnz];
    const int t = (int)blockIdx.x;
    if (t >= ntiles) return;
    const int r0 = tile_row[t], r1 = tile_row[t + 1];
    const int k0 = rowptr[r0], k1 = rowptr[r1];
    if (k1 - k0 > kTilePlanNnz) {
        const double l = -lam[r0];
        for (int k = k0 + threadIdx.x; k < k1; k += kBlock) g[k] = l * x[col[k]];
        return;
    }
    if ((int)threadIdx.x < r1 - r0) {
        const int i = r0 + threadIdx.x;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) srow[k - k0] = i;
    }
    __syncthreads();
    for (int k = k0 + threadIdx.x; k < k1; k += kBlock) g[k] = -lam[srow[k - k0]] * x[col[k]];
}

// --------------------------------------------------------------------------
// launchers

// Runs CALL with VT (value type) and CT (column type) bound to the operator's storage.
#define XJ_DISPATCH_VC(F32, C16, CALL)                         \
    do {                                                       \
        if (F32) {                                             \
            using VT = float;                                  \
            if (C16) { using CT = int16_t; CALL; }             \
            else { using CT = int; CALL; }                     \
        } else {                                               \
            using VT = double;                                 \
            if (C16) { using CT = int16_t; CALL; }             \
            else { using CT = int; CALL; }                     \
        }                                                      \
    } while (0)

static int g_coop_grid[4] = {0, 0, 0, 0};  // cached persistent grid cap per (value, column) type (one device)

// Largest grid the persistent kernel is launched with: one workgroup per CU
// below the co-resident maximum.  A grid of exactly cus * per workgroups
// needs every CU filled to its limit; the fp32-value instantiation (64
// VGPRs, 8 waves per SIMD) then asks for every wave slot on the chip, and a
// single slot held by any other dispatch leaves a workgroup unplaced while
// the placed ones wait for it at the first grid barrier.  One workgroup of
// slack per CU keeps every workgroup placeable.
template <typename VT, typename CT>
static hipError_t coop_grid(int* G) {
    const int slot = (sizeof(VT) == 4 ? 1 : 0) + (sizeof(CT) == 2 ? 2 : 0);
    if (g_coop_grid[slot] > 0) {
        *G = g_coop_grid[slot];
        return hipSuccess;
    }
    int dev = 0, cus = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, jpcg_persistent<VT, CT>, kBlock, 0);
    if (e != hipSuccess) return e;
    if (per < 1) return hipErrorCooperativeLaunchTooLarge;
    g_coop_grid[slot] = cus * (per > 1 ? per - 1 : 1);
    *G = g_coop_grid[slot];
    return hipSuccess;
}

hipError_t persistent_grid(bool f32, bool c16, int* G) {
    hipError_t e = hipSuccess;
    XJ_DISPATCH_VC(f32, c16, e = (coop_grid<VT, CT>(G)));
    return e;
}

hipError_t launch_jpcg(const JpcgParams& P0, bool f32, bool c16, int G, hipStream_t s) {
    JpcgParams P = P0;
    void* args[] = {&P};
    hipError_t e = hipSuccess;
    XJ_DISPATCH_VC(f32, c16,
                   e = hipLaunchCooperativeKernel((const void*)jpcg_persistent<VT, CT>, dim3(G), dim3(kBlock), args,
                                                  0, s));
    return e;
}

hipError_t launch_cg_setup(const CgParams& P, hipStream_t s) {
    if (P.w32 && !P.dia_on && !P.pan_on) return hipErrorInvalidValue;
    if (P.w32)
        cg_setup<float><<<P.G2, kBlock, 0, s>>>(P);
    else
        cg_setup<double><<<P.G2, kBlock, 0, s>>>(P);
    cg_setup_finish<<<1, kBlock, 0, s>>>(P);
    return hipGetLastError();
}

hipError_t launch_cg_iteration(const CgParams& P, bool f32, bool c16, hipStream_t s) {
    if (P.pan_on) {
        const size_t lds = (size_t)panel_lds_bytes(P.pan.block_rows);
        if (P.w32) {
            if (f32)
                cg_spmv_panels<float, float><<<P.pan.nb, kPanelBlock, lds, s>>>(P);
            else
                cg_spmv_panels<double, float><<<P.pan.nb, kPanelBlock, lds, s>>>(P);
        } else {
            if (f32)
                cg_spmv_panels<float, double><<<P.pan.nb, kPanelBlock, lds, s>>>(P);
            else
                cg_spmv_panels<double, double><<<P.pan.nb, kPanelBlock, lds, s>>>(P);
        }
    } else if (P.dia_on) {
        const int g = dia_groups(P.n, f32);
        const bool wide = dia_rows_per_lane(f32, P.n) == 2;  // fp64 values only
        constexpr int WB = kDiaWideBlock, NB = kDiaBlock;
        const bool wb = dia_block(f32, P.n) == WB;
        if (P.w32) {
            if (f32)
                cg_spmv_dia<float, float, 1, NB><<<g, NB, 0, s>>>(P);
            else if (wide && wb)
                cg_spmv_dia<double, float, 2, WB><<<g, WB, 0, s>>>(P);
            else if (wide)
                cg_spmv_dia<double, float, 2, NB><<<g, NB, 0, s>>>(P);
            else
                cg_spmv_dia<double, float, 1, NB><<<g, NB, 0, s>>>(P);
        } else {
            if (f32)
                cg_spmv_dia<float, double, 1, NB><<<g, NB, 0, s>>>(P);
            else if (wide && wb)
                cg_spmv_dia<double, double, 2, WB><<<g, WB, 0, s>>>(P);
            else if (wide)
                cg_spmv_dia<double, double, 2, NB><<<g, NB, 0, s>>>(P);
            else
                cg_spmv_dia<double, double, 1, NB><<<g, NB, 0, s>>>(P);
        }
    } else if (P.pat_on) {
        if (f32)
            cg_spmv_pattern<float><<<pattern_groups(P.tiles), kBlock, 0, s>>>(P);
        else
            cg_spmv_pattern<double><<<pattern_groups(P.tiles), kBlock, 0, s>>>(P);
    } else {
        XJ_DISPATCH_VC(f32, c16, (cg_spmv<VT, CT><<<tile_groups(P.tiles), kBlock, 0, s>>>(P)));
    }
    return launch_cg_vector_steps(P, s);
}

hipError_t launch_cg_vector_steps(const CgParams& P, hipStream_t s) {
    if (P.w32 && !P.dia_on && !P.pan_on) return hipErrorInvalidValue;
    if (P.split_alpha) cg_alpha<<<P.nA2, kBlock, 0, s>>>(P);
    if (P.w32) {
        cg_update_r<float><<<P.G2, kBlock, 0, s>>>(P);
        cg_update_xp<float><<<P.G2, kBlock, 0, s>>>(P);
    } else {
        cg_update_r<double><<<P.G2, kBlock, 0, s>>>(P);
        cg_update_xp<double><<<P.G2, kBlock, 0, s>>>(P);
    }
    return hipGetLastError();
}

hipError_t launch_spmv(const CsrDev& A, const double* x, double* y, const double* b, hipStream_t s) {
    if (A.pan_on) {
        const size_t lds = (size_t)panel_lds_bytes(A.pan.block_rows);
        if (A.val_f32)
            spmv_panels<float><<<A.pan.nb, kPanelBlock, lds, s>>>(A.pan, A.n, A.dsplit, x, y, b);
        else
            spmv_panels<double><<<A.pan.nb, kPanelBlock, lds, s>>>(A.pan, A.n, A.dsplit, x, y, b);
        return hipGetLastError();
    }
    if (A.dia_on) {
        const int g = dia_groups(A.n, A.val_f32);
        if (g <= 0) return hipSuccess;
        if (A.val_f32)
            spmv_dia<float, 1, kDiaBlock><<<g, kDiaBlock, 0, s>>>(A.n, A.dia, A.dsplit, x, y, b);
        else if (dia_rows_per_lane(false, A.n) == 2 && dia_block(false, A.n) == kDiaWideBlock)
            spmv_dia<double, 2, kDiaWideBlock><<<g, kDiaWideBlock, 0, s>>>(A.n, A.dia, A.dsplit, x, y, b);
        else if (dia_rows_per_lane(false, A.n) == 2)
            spmv_dia<double, 2, kDiaBlock><<<g, kDiaBlock, 0, s>>>(A.n, A.dia, A.dsplit, x, y, b);
        else
            spmv_dia<double, 1, kDiaBlock><<<g, kDiaBlock, 0, s>>>(A.n, A.dia, A.dsplit, x, y, b);
        return hipGetLastError();
    }
    if (A.ntiles <= 0) return hipSuccess;
    if (A.pat_on) {
        if (A.val_f32)
            spmv_pattern<float><<<pattern_groups(A.ntiles), kBlock, 0, s>>>(A.ntiles, (int)A.nnz, A.tile_row, A.tile_k,
                                                                            A.rowptr, A.pmask,
                                                            A.poffs, static_cast<const float*>(A.val), A.dsplit, x, y, b);
        else
            spmv_pattern<double><<<pattern_groups(A.ntiles), kBlock, 0, s>>>(A.ntiles, (int)A.nnz, A.tile_row, A.tile_k,
                                                                             A.rowptr, A.pmask,
                                                             A.poffs, static_cast<const double*>(A.val), A.dsplit, x, y,
                                                             b);
        return hipGetLastError();
    }
    const void* col = A.col16 ? (const void*)A.col16 : (const void*)A.col;
    XJ_DISPATCH_VC(A.val_f32, A.col16 != nullptr,
                   (spmv_tiles<VT, CT><<<tile_groups(A.ntiles), kBlock, 0, s>>>(A.ntiles, (int)A.nnz, A.tile_row,
                                                                                A.tile_k,
                                                                   A.rowptr,
                                                                   static_cast<const CT*>(col),
                                                                   static_cast<const VT*>(A.val), A.dsplit, x, y,
                                                                   b)));
    return hipGetLastError();
}

hipError_t launch_plan_patterns(const CsrDev& A, uint32_t* pmask, int* poffs, int* fail, hipStream_t s) {
    if (A.ntiles <= 0) return hipSuccess;
    plan_pattern_tiles<<<A.ntiles, kBlock, 0, s>>>(A.n, A.ntiles, A.tile_row, A.rowptr, A.col, pmask, poffs, 0, fail);
    return hipGetLastError();
}

hipError_t launch_tile_offsets(const CsrDev& A, int* toffs, int* fail, hipStream_t s) {
    if (A.ntiles <= 0) return hipSuccess;
    plan_pattern_tiles<<<A.ntiles, kBlock, 0, s>>>(A.n, A.ntiles, A.tile_row, A.rowptr, A.col, nullptr, toffs,
                                                   INT32_MIN, fail);
    return hipGetLastError();
}

hipError_t launch_fill(int64_t n, double v, double* x, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    fill_kernel<<<(int)g, kBlock, 0, s>>>(n, v, x);
    return hipGetLastError();
}

hipError_t launch_fill_dia(const CsrDev& A, int ndia, const int* off, void* dval, uint32_t* dmask, int* fail,
                           hipStream_t s) {
    if (A.n <= 0) return hipSuccess;
    int g = (A.n + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (A.val_f32)
        fill_dia<float><<<g, kBlock, 0, s>>>(A.n, A.rowptr, A.col, static_cast<const float*>(A.val), ndia, off,
                                             static_cast<float*>(dval), dmask, fail);
    else
        fill_dia<double><<<g, kBlock, 0, s>>>(A.n, A.rowptr, A.col, static_cast<const double*>(A.val), ndia, off,
                                              static_cast<double*>(dval), dmask, fail);
    return hipGetLastError();
}

hipError_t launch_build_col16(const CsrDev& A, int16_t* col16, int* overflow, hipStream_t s) {
    if (A.ntiles <= 0) return hipSuccess;
    build_col16<<<A.ntiles, kBlock, 0, s>>>(A.ntiles, A.tile_row, A.rowptr, A.col, col16, overflow);
    return hipGetLastError();
}

int dot_blocks(int n) {
    int g = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    return g;
}

hipError_t launch_dot(int n, const double* a, const double* b, double* part, double* out, hipStream_t s) {
    const int g = dot_blocks(n);
    dot_partials<<<g, kBlock, 0, s>>>(n, a, b, part);
    sum_partials<<<1, kBlock, 0, s>>>(g, part, out);
    return hipGetLastError();
}

static int grid_for(int64_t m) {
    int64_t g = (m + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

hipError_t launch_gather_f64(int64_t nnz, const int64_t* perm, const double* src, double* dst, hipStream_t s) {
    gather_values<double, double><<<grid_for(nnz), kBlock, 0, s>>>(nnz, perm, src, dst);
    return hipGetLastError();
}

hipError_t launch_gather_f32(int64_t nnz, const int64_t* perm, const double* src, float* dst, hipStream_t s) {
    gather_values<float, double><<<grid_for(nnz), kBlock, 0, s>>>(nnz, perm, src, dst);
    return hipGetLastError();
}

hipError_t launch_gather_vals(int64_t nnz, const int64_t* perm, const void* src, void* dst, bool f32, hipStream_t s) {
    if (nnz <= 0) return hipSuccess;
    if (f32)
        gather_values<float, float><<<grid_for(nnz), kBlock, 0, s>>>(nnz, perm, static_cast<const float*>(src),
                                                                     static_cast<float*>(dst));
    else
        gather_values<double, double><<<grid_for(nnz), kBlock, 0, s>>>(nnz, perm, static_cast<const double*>(src),
                                                                       static_cast<double*>(dst));
    return hipGetLastError();
}

hipError_t launch_stream_copy(int64_t n, const double* src, double* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t g = (n / 2 + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    stream_copy<<<(unsigned)g, kBlock, 0, s>>>(n, src, dst);
    return hipGetLastError();
}

hipError_t launch_axpy(int n, double a, const double* d, double* x, hipStream_t s) {
    axpy_kernel<<<grid_for(n), kBlock, 0, s>>>(n, a, d, x);
    return hipGetLastError();
}

hipError_t launch_sddmm(const CsrDev& A, const double* lam, const double* x, double* g, hipStream_t s) {
    if (A.ntiles <= 0) return hipSuccess;
    sddmm_tiles<<<A.ntiles, kBlock, 0, s>>>(A.ntiles, A.tile_row, A.rowptr, A.col, lam, x, g);
    return hipGetLastError();
}

}  // namespace xj
