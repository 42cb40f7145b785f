// Host-side view of the device data and the kernel launchers (jpcg.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xj {

struct Record;

// The panel-swept copy of an operator (see panel_rows in jpcg.hip): row
// block b (block_rows rows) stores its entries of column panel p (65536
// columns) as segment [seg[b*P+p], seg[b*P+p+1]), row-major with ascending
// columns; val in the operator's value type, off = column - 65536*p,
// lrow = row - block_rows*b.
struct PanelDev {
    int nb = 0;          // row blocks
    int block_rows = 0;  // rows per block (the last block may be shorter)
    int P = 0;           // column panels
    const int* seg = nullptr;
    const void* val = nullptr;
    const uint16_t* off = nullptr;
    const uint16_t* lrow = nullptr;
};

// A CSR operator resident in HBM plus its tile plan.
//   rowptr[n+1], col[nnz]  int32
//   col16[nnz]             optional compact column stream: col - first row of the entry's tile
//   val[nnz]               fp64, or fp32 when val_f32
//   dsplit[n]              optional fp64 diagonal kept outside the CSR
//                          (the CSR then holds off-diagonal entries only)
//   tile_row[ntiles+1]     first row of each tile
// Diagonal layout (stencil-like operators with few distinct offsets): one
// value array per global column offset col - row, ascending, with 0 where a
// row has no entry there, and a presence mask per row.
constexpr int kDiaMax = 32;
// Rows per lane of the diagonal kernels: two consecutive rows for fp64 values
// (one 16-byte load per diagonal) from kDiaWideRows rows up, one row for fp32
// values and for smaller fp64 operators (two rows per lane would leave a
// small operator's grid at a workgroup or two per CU).
constexpr int kDiaWideRows = 5 << 16;  // 327,680 rows: the measured crossover
constexpr int dia_rows_per_lane(bool f32, int n) { return f32 || n < kDiaWideRows ? 1 : 2; }
// Leading dimension of the diagonal value arrays and length of the mask
// array: n rounded up to even, so a lane's two rows are one aligned 16-byte
// load; the padding row's mask is 0 and its values are 0.
constexpr int dia_ld(int n) { return (n + 1) & ~1; }
// workgroups of the diagonal kernels (kDiaBlock lanes each) and the CG
// SpMV's p.Ap partials (one per workgroup)
// (kDiaWideBlock lanes from kDiaWideBlockRows rows with two rows per lane:
// 1,024 rows per workgroup against 512, 27-point SpMV 100^3 48.5 / 51.5 us,
// 128^3 89.5 / 95.0, 160^3 168 / 172.5; 81^3 40.3 / 31.5, too few workgroups.
// One row per lane keeps 256: the 160^3 fp32-value SpMV took 119.6-120.6 us
// with 256 or 128 lanes, 121.5-121.8 with 512, 122.5-123.2 with 1024; 128
// lanes below 2^18 rows slowed the 256^2 and 400^2 solves by 6-10%, twice
// the p.Ap partials per iteration)
constexpr int kDiaBlock = 256;
constexpr int kDiaWideBlock = 512;
constexpr int kDiaWideBlockRows = 3 << 18;  // 786,432
constexpr int dia_block(bool f32, int n) {
    return dia_rows_per_lane(f32, n) == 2 && n >= kDiaWideBlockRows ? kDiaWideBlock : kDiaBlock;
}
constexpr int dia_groups(int n, bool f32) {
    const int rows = dia_block(f32, n) * dia_rows_per_lane(f32, n);
    return (n + rows - 1) / rows;
}
constexpr int dia_cg_partials(int n, bool f32) { return dia_groups(n, f32); }
// Workgroups of the CG vector kernels (256 lanes, grid-stride; one r.r / r.z
// partial each, and every workgroup of the x/p update re-sums all of them):
// one row per lane up to 256 workgroups, so a small system's update is one
// memory round trip per lane (256^2 Laplacian, with the load-order change of
// the vector kernels: 13.0 us per iteration against 14.6 with n/1024 = 64
// workgroups); beyond that n/1024, up to 1024 (more
// workgroups cost more redundant partial sums than they save: 512^2 diffusion
// 64.0 ms with 1024 against 56.8 with 256; 1024 measured best on the 160^3
// stencil, fp64 solve 74.6 ms; 2048: 75.4, 4096: 79.9, 512: 75.8).
constexpr int kVectorBlock = 256;
constexpr int kMaxVectorGroups = 1024;
constexpr int vector_groups(int n) {
    const int per_lane = (n + kVectorBlock - 1) / kVectorBlock;
    const int small = per_lane < 256 ? per_lane : 256;
    const int large = (n + 1023) / 1024;
    const int g = small > large ? small : large;
    return g < 1 ? 1 : (g > kMaxVectorGroups ? kMaxVectorGroups : g);
}
static_assert(vector_groups(65536) == 256 && vector_groups(262144) == 256 && vector_groups(4096000) == 1024,
              "vector grid rule");
struct DiaDev {
    int ndia = 0;
    int ld = 0;                      // dia_ld(n)
    const int* off = nullptr;        // [ndia] offsets, ascending
    const void* val = nullptr;       // [ndia][ld] in the operator's value type
    const uint32_t* mask = nullptr;  // [ld]: bit d set when row i has an entry at off[d]
};

struct CsrDev {
    int n = 0;
    int64_t nnz = 0;
    const int* rowptr = nullptr;
    const int* col = nullptr;
    const int16_t* col16 = nullptr;
    const void* val = nullptr;
    bool val_f32 = false;
    const double* dsplit = nullptr;
    const int* tile_row = nullptr;
    const int* tile_k = nullptr;      // rowptr[tile_row[t]] for t = 0..ntiles
    int ntiles = 0;
    PanelDev pan;           // panel-swept copy, used by the SpMV kernels when pan_on
    bool pan_on = false;
    const uint32_t* pmask = nullptr;  // column-offset pattern tiles (tile_spmv_pattern): mask per row
    const int* poffs = nullptr;       // 32 offsets per tile
    bool pat_on = false;
    DiaDev dia;                       // diagonal layout, used by the SpMV kernels when dia_on
    bool dia_on = false;
};

struct JpcgParams {
    int n = 0, ntiles = 0, nnz = 0;
    const int* rowptr = nullptr;
    const void* col = nullptr;           // int32, or int16 deltas when the operator is compact
    const void* val = nullptr;
    const double* dsplit = nullptr;
    const int* tile_row = nullptr;
    const int* tile_k = nullptr;
    const double* b = nullptr;
    const double* diag = nullptr;  // Jacobi diagonal supplied by the caller
    double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *J = nullptr;
    double* part = nullptr;              // 5*G doubles
    unsigned long long* bar = nullptr;   // zeroed before launch
    int* abortf = nullptr;               // zeroed before launch
    Record* records = nullptr;           // max_rounds+1 records, or null
    double* result = nullptr;            // [K, rr_last, dot, rz0, abort]
    double tol = 0.0;
    uint32_t max_rounds = 0;             // maxIter + 1, as the reference host programs it
};

// Device state of the multi-kernel CG loop (one per operator).
struct CgState {
    double dot, tols, rr, rz0;
    double rz[2];      // rz of iteration k lives in rz[k & 1]
    double alpha;      // alpha of the running iteration (written by cg_alpha when the tile partials are many)
    uint32_t k;        // index of the running iteration; advanced by workgroup 0 of the SpMV kernel
    uint32_t it;       // the reference Task's iteration counter (tasks.hpp:30-47)
    int stop;          // set by workgroup 0 of the update kernel when the loop ends
    int stop_k2;       // snapshot of `stop` taken by the residual kernel
    int K;             // iterations executed
    int pad;
};

// Arguments of the multi-kernel loop: SpMV (one workgroup per tile or per row
// block of the panel layout or workgroup of the diagonal layout, with one
// p.Ap partial each), optionally a
// reduction of those partials, then two vector kernels (G2 workgroups,
// grid-stride).  ntiles counts the SpMV workgroups, hence the partials.
struct CgParams {
    int n = 0, ntiles = 0, nnz = 0, G2 = 0;
    int tiles = 0;                       // row tiles of the plan (ntiles counts the partials)
    int pan_on = 0;                      // 1: the SpMV runs on the panel layout `pan`
    PanelDev pan;
    int pat_on = 0;                      // 1: the SpMV runs on column-offset pattern tiles
    const uint32_t* pmask = nullptr;
    const int* poffs = nullptr;
    int dia_on = 0;                      // 1: the SpMV runs on the diagonal layout `dia`
    DiaDev dia;
    int split_alpha = 0;                 // 1: cg_alpha first folds the ntiles partials into nA2
    int nA2 = 0;
    double* partA2 = nullptr;            // nA2 doubles
    const int* rowptr = nullptr;
    const void* col = nullptr;
    const void* val = nullptr;
    const double* dsplit = nullptr;
    const int* tile_row = nullptr;
    const int* tile_k = nullptr;
    const double* b = nullptr;
    const double* diag = nullptr;
    double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr, *J = nullptr;
    int w32 = 0;                         // 1: r, p, Ap hold fp32 (diagonal layout only); x, J, dots fp64
    double *partA = nullptr, *partB = nullptr, *partC = nullptr;  // ntiles, G2, G2 doubles
    CgState* st = nullptr;
    Record* records = nullptr;
    double tol = 0.0;
    uint32_t max_rounds = 0;
    // pinned host word the kernel that ends the loop also sets to 1 (may be
    // null): the host reads it after a chunk's completion event instead of
    // copying the device flag, which would put a copy between chunks
    int* host_stop = nullptr;
};

// tiles above which the SpMV's p.Ap partials are first folded by cg_alpha
// into kAlphaBlocks partials instead of being summed by every workgroup of
// the vector kernels
constexpr int kMaxRedundantPartials = 2048;
constexpr int kAlphaBlocks = 128;

hipError_t launch_cg_setup(const CgParams& P, hipStream_t s);
hipError_t launch_cg_iteration(const CgParams& P, bool f32, bool c16, hipStream_t s);
// the steps of an iteration after its SpMV: the optional partial fold, the
// residual update and the x / p update (for SpMV kernels launched elsewhere)
hipError_t launch_cg_vector_steps(const CgParams& P, hipStream_t s);

hipError_t persistent_grid(bool f32, bool c16, int* G);
hipError_t launch_jpcg(const JpcgParams& P, bool f32, bool c16, int G, hipStream_t s);
hipError_t launch_spmv(const CsrDev& A, const double* x, double* y, const double* b, hipStream_t s);
hipError_t launch_build_col16(const CsrDev& A, int16_t* col16, int* overflow, hipStream_t s);
hipError_t launch_plan_patterns(const CsrDev& A, uint32_t* pmask, int* poffs, int* fail, hipStream_t s);
// per-tile ascending offsets as for the pattern plan, unused slots = INT32_MIN, no masks
hipError_t launch_tile_offsets(const CsrDev& A, int* toffs, int* fail, hipStream_t s);
// diagonal values and masks from the CSR (dval zeroed beforehand); *fail when
// a row's offset is not in `off` or its offsets do not ascend
hipError_t launch_fill_dia(const CsrDev& A, int ndia, const int* off, void* dval, uint32_t* dmask, int* fail,
                           hipStream_t s);
int dot_blocks(int n);
hipError_t launch_dot(int n, const double* a, const double* b, double* part, double* out, hipStream_t s);
hipError_t launch_gather_f64(int64_t nnz, const int64_t* perm, const double* src, double* dst, hipStream_t s);
hipError_t launch_gather_f32(int64_t nnz, const int64_t* perm, const double* src, float* dst, hipStream_t s);
// dst[k] = src[perm[k]] in the operator's value type (fp32 when f32)
hipError_t launch_gather_vals(int64_t nnz, const int64_t* perm, const void* src, void* dst, bool f32, hipStream_t s);
hipError_t launch_axpy(int n, double a, const double* d, double* x, hipStream_t s);
hipError_t launch_fill(int64_t n, double v, double* x, hipStream_t s);
hipError_t launch_stream_copy(int64_t n, const double* src, double* dst, hipStream_t s);
hipError_t launch_sddmm(const CsrDev& A, const double* lam, const double* x, double* g, hipStream_t s);

}  // namespace xj
