// Device-side building blocks shared by the HIP translation units:
// wavefront / workgroup reductions, the XCD-aware block remap, the grid
// barrier used by the persistent kernels, and the CSR tile SpMV.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_consts.hpp"

namespace xj {

constexpr int kBlock = 256;                // threads per workgroup (4 wavefronts of 64)
constexpr int kWaves = kBlock / 64;
constexpr int kTileNnz = kTileNnzHost;     // max stored entries per short-row tile (16 KB of fp64 products in LDS)
constexpr int kTileRows = kTileRowsHost;   // max rows per tile: one lane reduces one row
constexpr int kWaveRowMin = kWaveRowMinHost;  // longer rows of a short-row tile: summed by their wavefront
static_assert(kTileRows == kBlock, "one lane per tile row");
static_assert(kTileNnz % kBlock == 0, "whole loads per lane");


// 64-byte per-iteration record, field order of the reference's CGSolverInstr
// (pcg/hw/include/impl/cgInstr.hpp:117-132), little-endian.
struct alignas(64) Record {
    uint32_t id, iter, vec_size, max_iter;
    double tols, res, rz, alpha, beta;
    uint64_t clock;
};
static_assert(sizeof(Record) == 64, "record layout");

// Block index -> logical index so that each XCD (workgroups are dealt to the
// 8 XCDs round-robin) receives one contiguous range of logical blocks.  Only
// the persistent loop's tile sweep uses it.  The launch-per-step kernels take
// blocks in launch order: measured on MI355X, the contiguous ranges put eight
// distant fronts of the matrix streams on HBM at once (the 160^3 diagonal SpMV
// ran 3% slower) and leave an XCD idle once its own range of a batch has
// converged (the batched iteration ran 29% slower); no kernel measured gained
// from it.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, i = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// xor-butterfly: every lane ends with the same, order-fixed sum
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Wavefront bin of a row-binned SpMV.  Each lane of the wavefront may own a
// row whose products sit in lds[a, b); the lanes with wide == true (rows longer
// than kWaveRowMin) are served one after another in lane order by the whole
// wavefront: lane l adds entries a+l, a+l+64, ... left to right, the 64
// partials are folded by the xor butterfly, and the owner adds the total to
// its s.  The order depends only on the row's length, so the sum is
// deterministic.  Called by all 64 lanes of the wavefront together.
__device__ __forceinline__ void wave_rows_sum(bool wide, int a, int b, const double* lds, double& s) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    uint64_t m = __ballot(wide);
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const int ra = __shfl(a, l, 64), rb = __shfl(b, l, 64);
        double p = 0.0;
        int j = ra + lane;
        for (; j + 64 < rb; j += 128) {
            const double q0 = lds[j], q1 = lds[j + 64];
            p += q0;
            p += q1;
        }
        if (j < rb) p += lds[j];
        p = wave_sum(p);
        if (lane == l) s += p;
    }
}

// Sum over a workgroup of NW wavefronts; every thread returns the identical
// value.  red must hold NW doubles of LDS.  Contains two __syncthreads.
template <int NW>
__device__ __forceinline__ double block_sum_w(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) s += red[i];
    __syncthreads();
    return s;
}
__device__ __forceinline__ double block_sum(double v, double* red) { return block_sum_w<kWaves>(v, red); }

// K sums over a workgroup with one pair of __syncthreads: value j is summed
// exactly as block_sum sums it alone (wavefront butterfly, then wavefronts
// 0..kWaves-1 in order).  red must hold K * kWaves doubles of LDS.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        v[j] = wave_sum(v[j]);
        if (lane == 0) red[j * kWaves + w] = v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double s = red[j * kWaves];
#pragma unroll
        for (int i = 1; i < kWaves; ++i) s += red[j * kWaves + i];
        v[j] = s;
    }
    __syncthreads();
}

// This thread's share of part[0..G), G >= 1: 0.0 + part[threadIdx.x] +
// part[threadIdx.x + kBlock] + ... in that order, split in two so that
// several sets' first loads can be in flight together.  The head load uses a
// clamped index and no branch (a load inside a branch is waited for before
// the branch ends); the tail issues four loads before adding them in order.
__device__ __forceinline__ double strided_head(const double* part, int G) {
    const int t = (int)threadIdx.x;
    const double v = part[t < G ? t : G - 1];
    return 0.0 + (t < G ? v : 0.0);
}
__device__ __forceinline__ double strided_tail(const double* part, int G, double s) {
    for (int i = (int)threadIdx.x + kBlock; i < G; i += 4 * kBlock) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = i + u * kBlock;
            v[u] = part[j < G ? j : G - 1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * kBlock < G) s += v[u];
    }
    return s;
}
__device__ __forceinline__ double strided_part(const double* part, int G) {
    return strided_tail(part, G, strided_head(part, G));
}

// Sum of part[0..G) in a fixed order, identical in every workgroup.
__device__ __forceinline__ double grid_total(const double* part, int G, double* red) {
    return block_sum(strided_part(part, G), red);
}

// Grid-wide barrier for cooperatively launched kernels (all workgroups
// co-resident).  `bar` is a zeroed 64-bit arrival counter; `target` is the
// caller's running count of expected arrivals.
//
// On a multi-XCD part an agent-scope release writes back the XCD's L2 and an
// acquire invalidates it, so only lane 0 of each workgroup fences, once on
// each side, and the spin uses relaxed loads (an acquire load per poll would
// invalidate L2 under the workgroups still computing).  The __syncthreads
// before the release orders the other wavefronts' stores ahead of it.
//
// A wait longer than kBarrierTimeout ticks of the 100 MHz wall clock sets
// *abortf and every workgroup leaves the barrier; returns false once the
// abort flag is set.
constexpr uint64_t kBarrierTimeout = 500000000ull;  // 5 s

__device__ __forceinline__ bool grid_barrier(unsigned long long* bar, unsigned long long& target, int G,
                                             int* abortf, int* sflag) {
    __syncthreads();
    target += (unsigned long long)G;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t0 = wall_clock64();
        int ab = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(2);
            if (__hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ab = 1; break; }
            if (wall_clock64() - t0 > kBarrierTimeout) {
                __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ab = 1;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *sflag = ab;
    }
    __syncthreads();
    return *sflag == 0;
}

template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

// A streamed output (y = A x, Ap, the vector updates' r, x, p): written once
// and read again only by a later kernel, so it is stored without allocating
// in L2, where it would evict lines the launch still reuses (x of an SpMV).
// The stored value is the same; only the cache policy differs.
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}

// ---------------------------------------------------------------------------
// CSR tile SpMV.
//
// Rows are cut into tiles on the host: a tile is either a run of consecutive
// rows with at most kTileRows rows and kTilePlanNnz stored entries, or one row
// holding more than kTilePlanNnz entries.  For a short-row tile, lanes stream
// the tile's values and column indices contiguously with 16-byte
// non-temporal loads, gather x, and park the products in LDS; then the rows
// are summed in two bins: lane t sums row t's products left to right when the
// row has at most kWaveRowMin entries, and the lane's wavefront sums a longer
// row together (wave_rows_sum).  A row longer than a tile is summed by the
// whole workgroup with a strided loop and a fixed-order workgroup reduction.
//
// VT is the stored value type (double, or float for the mixed-precision
// inner operator).  With dsplit != nullptr the matrix holds no diagonal and
// dsplit[i]*x[i] starts row i's sum.  Value and int32 column arrays must be
// 16-byte aligned.
//
// On return, lanes with *row >= 0 hold y[*row] in *y.

// Column index of entry k in tile row-base r0: int32 columns are stored as
// is; int16 columns are stored as col - r0 (the compact index stream, used
// when every tile's columns lie within 32767 of its first row).
__device__ __forceinline__ int col_at(const int* __restrict__ col, int k, int) { return ld_stream(col + k); }
__device__ __forceinline__ int col_at(const int16_t* __restrict__ col, int k, int r0) {
    return r0 + (int)ld_stream(col + k);
}
__device__ __forceinline__ int col_of(int c, int) { return c; }
__device__ __forceinline__ int col_of(int16_t c, int r0) { return r0 + (int)c; }

// E consecutive elements of T as one non-temporal load of E*sizeof(T) bytes
// (p must be aligned to that size).
template <int BYTES> struct LoadWord;
template <> struct LoadWord<4> { typedef uint32_t type; };
template <> struct LoadWord<8> { typedef uint64_t type; };
template <> struct LoadWord<16> { typedef unsigned int type __attribute__((ext_vector_type(4))); };

template <typename T, int E>
__device__ __forceinline__ void load_run(const T* __restrict__ p, T (&out)[E]) {
    typedef typename LoadWord<E * sizeof(T)>::type W;
    const W w = __builtin_nontemporal_load(reinterpret_cast<const W*>(p));
    __builtin_memcpy(out, &w, sizeof(W));
}

// What one lane needs of a row tile before its compute starts, loaded up
// front so a workgroup can have a later tile's loads (and the CG loop its
// stop flag) in flight together.  k0 = rowptr[r0] and k1 = rowptr[r1] come
// from the per-tile table tile_k, so they load together with r0 and r1.
template <typename VT, typename CT>
struct RowTile {
    static constexpr int E = 16 / (int)sizeof(VT);
    static constexpr int kGroups = kTileNnz / (kBlock * E);
    int t = -1, r0 = 0, r1 = 0, k0 = 0, k1 = 0;
    VT v[kGroups][E];
    CT c[kGroups][E];
    __device__ bool long_row() const { return k1 - k0 > kTilePlanNnz; }
};

template <typename VT, typename CT>
__device__ __forceinline__ void row_tile_load(RowTile<VT, CT>& T, int t, const int* __restrict__ tile_row,
                                              const int* __restrict__ tile_k, int nnz,
                                              const int* __restrict__ rowptr, const CT* __restrict__ col,
                                              const VT* __restrict__ val) {
    constexpr int E = RowTile<VT, CT>::E, kGroups = RowTile<VT, CT>::kGroups;
    const int tid = threadIdx.x;
    T.t = t;
    T.r0 = tile_row[t];
    T.r1 = tile_row[t + 1];
    T.k0 = tile_k[t];
    T.k1 = tile_k[t + 1];
    const int k0 = T.k0, k1 = T.k1;
    if (k1 - k0 > kTilePlanNnz) return;  // one long row: its compute streams it
    if (k1 > k0) {
        // Each lane loads runs of E = 16/sizeof(VT) consecutive entries (one
        // 16-byte value load, one matching column load), kGroups runs spaced a
        // workgroup-span apart, so every wave-wide load is one contiguous
        // block.  The range starts at k0 rounded down to a multiple of E (the
        // plan leaves room for the extra entries); entries outside [k0, k1)
        // are dropped.  Every E-aligned run that meets [k0, k1) lies inside
        // the arrays unless the tile ends in a final partial run of the whole
        // array (nnz not a multiple of E); only that tile loads entry by
        // entry.  Lanes with nothing to load re-read the tile's first run.
        const int k0a = k0 & ~(E - 1);
        if (((k1 + E - 1) & ~(E - 1)) <= nnz) {
#pragma unroll
            for (int g = 0; g < kGroups; ++g) {
                int base = k0a + (g * kBlock + tid) * E;
                if (base >= k1) base = k0a;
                load_run<VT, E>(val + base, T.v[g]);
                load_run<CT, E>(col + base, T.c[g]);
            }
        } else {
#pragma unroll
            for (int g = 0; g < kGroups; ++g) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int kk = k0a + (g * kBlock + tid) * E + e;
                    const int k = kk >= k0 && kk < k1 ? kk : k0;
                    T.v[g][e] = ld_stream(val + k);
                    T.c[g][e] = ld_stream(col + k);
                }
            }
        }
    }
}

// Compute of a loaded row tile: gathers (all of a lane's in flight
// together), products in LDS, binned row sums (lane per row up to
// kWaveRowMin entries, wavefront per row above); a row longer than a tile is
// summed by the whole workgroup with a strided loop and a fixed-order
// workgroup reduction.  On return lanes with *row >= 0 hold y[*row] in *y;
// ends with a __syncthreads so the workgroup's next tile can reuse lds / red.
template <typename VT, typename CT>
__device__ __forceinline__ void row_tile_compute(const RowTile<VT, CT>& T, const int* __restrict__ rowptr,
                                                 const CT* __restrict__ col,
                                                 const VT* __restrict__ val, const double* __restrict__ dsplit,
                                                 const double* __restrict__ x, double* lds, double* red, double* y,
                                                 int* row) {
#pragma clang fp contract(off)
    constexpr int E = RowTile<VT, CT>::E, kGroups = RowTile<VT, CT>::kGroups;
    const int tid = threadIdx.x;
    const int r0 = T.r0, r1 = T.r1, k0 = T.k0, k1 = T.k1;
    if (T.long_row()) {
        double s = 0.0;
        for (int k = k0 + tid; k < k1; k += kBlock) s += (double)ld_stream(val + k) * x[col_at(col, k, r0)];
        s = block_sum(s, red);
        if (dsplit) s = dsplit[r0] * x[r0] + s;
        *y = s;
        *row = (tid == 0) ? r0 : -1;
        return;
    }
    if (k1 > k0) {
        // Entries outside [k0, k1) belong to neighbouring tiles: their int16
        // deltas are relative to another tile's first row, so decoding them
        // here could index past x.  They gather x[r0] instead (and are dropped).
        const int k0a = k0 & ~(E - 1);
        double gx[kGroups][E];
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = k0a + (g * kBlock + tid) * E + e;
                gx[g][e] = x[(k >= k0 && k < k1) ? col_of(T.c[g][e], r0) : r0];
            }
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const int base = k0a + (g * kBlock + tid) * E;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = base + e;
                if (k >= k0 && k < k1) lds[k - k0] = (double)T.v[g][e] * gx[g][e];
            }
        }
    }
    __syncthreads();
    const bool mine = tid < r1 - r0;
    // a lane's row bounds load here, not with the tile: holding them across
    // the gathers costs more (measured) than the load
    const int a = mine ? rowptr[r0 + tid] - k0 : 0, b = mine ? rowptr[r0 + tid + 1] - k0 : 0;
    double s = (mine && dsplit) ? dsplit[r0 + tid] * x[r0 + tid] : 0.0;
    const bool wide = b - a > kWaveRowMin;
    if (mine && !wide) {
        // left-to-right sum; the LDS reads of a group of four issue together
        int j = a;
        for (; j + 4 <= b; j += 4) {
            const double q0 = lds[j], q1 = lds[j + 1], q2 = lds[j + 2], q3 = lds[j + 3];
            s += q0;
            s += q1;
            s += q2;
            s += q3;
        }
        for (; j < b; ++j) s += lds[j];
    }
    wave_rows_sum(wide, a, b, lds, s);
    *y = s;
    *row = mine ? r0 + tid : -1;
    __syncthreads();
}

// One row tile, loaded and computed (the persistent kernel's tile loop).
template <typename VT, typename CT>
__device__ __forceinline__ void tile_spmv(int t, const int* __restrict__ tile_row, const int* __restrict__ tile_k,
                                          int nnz, const int* __restrict__ rowptr, const CT* __restrict__ col,
                                          const VT* __restrict__ val, const double* __restrict__ dsplit,
                                          const double* __restrict__ x, double* lds, double* red, double* y,
                                          int* row) {
    RowTile<VT, CT> T;
    row_tile_load<VT, CT>(T, t, tile_row, tile_k, nnz, rowptr, col, val);
    row_tile_compute<VT, CT>(T, rowptr, col, val, dsplit, x, lds, red, y, row);
}

// Row tiles per workgroup in the standalone and CG-loop SpMV kernels (with
// more than one, later tiles' loads are issued before the first compute).
constexpr int kRowTiles = kRowTilesHost;

// Column-offset pattern tiles.  When every tile of the plan uses at most 32
// distinct column offsets (col - row), as stencil operators do, the operator
// keeps, instead of a column index per entry, each tile's ascending offsets
// (poffs[t*32 + b]) and one 32-bit mask per row (bit b set: the row has an
// entry at column row + poffs[t*32 + b]).  With strictly ascending columns in
// every row, a row's stored entries are its set bits in ascending order.
// Lane t first expands row t's mask into the tile's column indices in LDS;
// then the tile runs as in tile_spmv (16-byte value loads, all gathers of a
// lane in flight together, products in LDS, lane-per-row sums left to
// right), reading each entry's column from its LDS slot before the product
// overwrites it.  Same products and sums as tile_spmv, with no column stream.
// What one lane needs of a pattern tile before the tile's compute starts
// (all of it loaded up front, so that a workgroup can have the loads of its
// next tile in flight while it computes the current one).
template <typename VT>
struct PatternTile {
    static constexpr int E = 16 / (int)sizeof(VT);
    static constexpr int kGroups = kTileNnz / (kBlock * E);
    int t = -1, r0 = 0, r1 = 0, k0 = 0, k1 = 0;
    int a = 0, b = 0;      // this lane's row: entries [a, b) relative to k0
    uint32_t m = 0;        // this lane's row mask
    int off = 0;           // lanes < 32: the tile's offset #lane
    VT v[kGroups][E];      // this lane's share of the value stream
};

template <typename VT>
__device__ __forceinline__ void pattern_tile_load(PatternTile<VT>& T, int t, const int* __restrict__ tile_row,
                                                  const int* __restrict__ tile_k, int nnz,
                                                  const int* __restrict__ rowptr, const uint32_t* __restrict__ pmask,
                                                  const int* __restrict__ poffs, const VT* __restrict__ val) {
    constexpr int E = PatternTile<VT>::E, kGroups = PatternTile<VT>::kGroups;
    const int tid = threadIdx.x;
    T.t = t;
    T.r0 = tile_row[t];
    T.r1 = tile_row[t + 1];
    T.k0 = tile_k[t];
    T.k1 = tile_k[t + 1];
    const int i = T.r0 + tid;
    const bool mine = tid < T.r1 - T.r0;
    T.a = mine ? rowptr[i] - T.k0 : 0;
    T.b = mine ? rowptr[i + 1] - T.k0 : 0;
    T.m = mine ? pmask[i] : 0u;
    T.off = tid < 32 ? poffs[t * 32 + tid] : 0;
    const int k0 = T.k0, k1 = T.k1;
    const int k0a = k0 & ~(E - 1);
    if (k1 > k0) {
        if (((k1 + E - 1) & ~(E - 1)) <= nnz) {
#pragma unroll
            for (int g = 0; g < kGroups; ++g) {
                int base = k0a + (g * kBlock + tid) * E;
                if (base >= k1) base = k0a;
                load_run<VT, E>(val + base, T.v[g]);
            }
        } else {
#pragma unroll
            for (int g = 0; g < kGroups; ++g)
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int kk = k0a + (g * kBlock + tid) * E + e;
                    T.v[g][e] = ld_stream(val + (kk >= k0 && kk < k1 ? kk : k0));
                }
        }
    }
}

// Compute of a loaded pattern tile; on return lanes with *row >= 0 hold
// y[*row] in *y.  Starts and ends with a __syncthreads, so consecutive tiles
// of a workgroup can reuse lds and loffs.
template <typename VT>
__device__ __forceinline__ void pattern_tile_compute(const PatternTile<VT>& T, const double* __restrict__ dsplit,
                                                     const double* __restrict__ x, double* lds, int* loffs, double* y,
                                                     int* row) {
#pragma clang fp contract(off)
    constexpr int E = PatternTile<VT>::E, kGroups = PatternTile<VT>::kGroups;
    const int tid = threadIdx.x;
    const int k0 = T.k0, k1 = T.k1, r0 = T.r0;
    const int i = r0 + tid;
    const bool mine = tid < T.r1 - r0;
    __syncthreads();  // the previous tile of this workgroup is done with lds / loffs
    if (tid < 32) loffs[tid] = T.off;
    __syncthreads();
    // columns are parked as doubles (exact below 2^53) in the slots the
    // products reuse, so the column read and the product write of an entry
    // are accesses of one type the compiler keeps in order
    uint32_t m = T.m;
    for (int j = T.a; m; ++j) {
        lds[j] = (double)(i + loffs[__builtin_ctz(m)]);
        m &= m - 1;
    }
    __syncthreads();
    if (k1 > k0) {
        const int k0a = k0 & ~(E - 1);
        double gx[kGroups][E];
#pragma unroll
        for (int g = 0; g < kGroups; ++g)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = k0a + (g * kBlock + tid) * E + e;
                gx[g][e] = x[(k >= k0 && k < k1) ? (int)lds[k - k0] : r0];
            }
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const int base = k0a + (g * kBlock + tid) * E;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int k = base + e;
                if (k >= k0 && k < k1) lds[k - k0] = (double)T.v[g][e] * gx[g][e];
            }
        }
    }
    __syncthreads();
    if (mine) {
        double s = dsplit ? dsplit[i] * x[i] : 0.0;
        int j = T.a;
        const int b = T.b;
        for (; j + 4 <= b; j += 4) {
            const double q0 = lds[j], q1 = lds[j + 1], q2 = lds[j + 2], q3 = lds[j + 3];
            s += q0;
            s += q1;
            s += q2;
            s += q3;
        }
        for (; j < b; ++j) s += lds[j];
        *y = s;
        *row = i;
    } else {
        *row = -1;
    }
}

// Pattern tiles per workgroup: the loads of both are issued before the first
// tile's compute (measured: 2 beats 1 and 4 for fp64 and fp32 values).
constexpr int kPatternTiles = kPatternTilesHost;

// Tiles of logical workgroup L in a grid of G for a tile loop: the logical
// workgroups of one XCD (a contiguous range under xcd_remap) own one
// contiguous eighth of the tiles and deal them out round-robin, so the XCD's
// resident workgroups sweep its band of rows together and the gathered
// x-vector window stays in that XCD's L2.  Tiles of L are
// first, first + step, ... below end.
struct TileSweep {
    int first, step, end;
};
__device__ __forceinline__ TileSweep tile_sweep(int L, int G, int ntiles) {
    const int groups = G < 8 ? G : 8;
    const int q = G / groups, r = G % groups;  // the first r groups hold q+1 workgroups
    int grp, l, size;
    if (L < r * (q + 1)) {
        grp = L / (q + 1);
        l = L % (q + 1);
        size = q + 1;
    } else {
        grp = r + (L - r * (q + 1)) / q;
        l = (L - r * (q + 1)) % q;
        size = q;
    }
    TileSweep s;
    const int gbeg = (int)((int64_t)ntiles * grp / groups);
    s.end = (int)((int64_t)ntiles * (grp + 1) / groups);
    s.first = gbeg + l;
    s.step = size;
    return s;
}

}  // namespace xj
