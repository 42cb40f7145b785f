// Batched JPCG: kernel arguments and launchers (batched.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace xj {

// Diagonal layout of one system: its rows [r0, r1), nd <= kBDiaMax column
// offsets (ascending), and where its nd value arrays of r1 - r0 entries start.
constexpr int kBDiaMax = 8;
struct BDiaSys {
    int off[kBDiaMax];
    int r0, r1, nd, pad;
    long long base;
    long long pad2;
};
static_assert(sizeof(BDiaSys) == 64, "system record layout");

struct BatchParams {
    int N = 0, S = 0, T = 0;             // rows, systems, tiles
    const int* rowptr = nullptr;         // [N+1]
    const int* col = nullptr;            // [nnz] global column indices
    const void* val = nullptr;           // [nnz] fp64 or fp32
    const double* dsplit = nullptr;      // [N] split diagonal, or null
    const int* tile_row = nullptr;       // [T+1] first row of each tile
    const int* tile_sys = nullptr;       // [T] system of each tile
    const int* sys_tile = nullptr;       // [S+1] first tile of each system
    const double* b = nullptr;           // [N][R]
    const double* diag = nullptr;        // [N] Jacobi diagonal
    double *x = nullptr, *r = nullptr, *p = nullptr, *Ap = nullptr;  // [N][R]
    double* J = nullptr;                 // [N]
    double *partA = nullptr, *partB = nullptr, *partC = nullptr;     // [T][R]
    double *dot = nullptr, *tols = nullptr, *rz = nullptr, *rr = nullptr;   // [S][R]
    double *alpha = nullptr, *beta = nullptr;                        // [S][R]
    int *active = nullptr, *upd = nullptr, *iters = nullptr, *it = nullptr;  // [S][R]
    int *sys_active = nullptr, *sys_upd = nullptr;                   // [S]
    int* any_active = nullptr;           // [1]
    const int* sys_row = nullptr;        // [S+1] first row of each system
    int dia_on = 0;                      // 1: the SpMM runs on the per-system diagonal layout
    const BDiaSys* dsys = nullptr;       // [S]
    const void* dval = nullptr;          // per system nd*(r1-r0) values of the value type, zero where absent
    const uint8_t* dmask = nullptr;      // [N]: bit d set when the row has an entry at its system's off[d]
    double tol = 0.0;
    uint32_t max_rounds = 0;
};

// State of the persistent batched loop (one cooperative launch for all
// iterations): a zeroed grid-barrier arrival counter, a zeroed abort flag,
// result[0] = iterations run, result[1] = 1 when a barrier timed out.
struct BatchLoop {
    unsigned long long* bar = nullptr;
    int* abortf = nullptr;
    long long* result = nullptr;
    uint32_t rounds = 0;
};

// w32: the work vectors r, p, Ap (P.r, P.p, P.Ap) hold fp32 values (x,
// arithmetic and dot products stay fp64)
hipError_t batch_launch_setup(const BatchParams& P, int R, bool w32, int block, hipStream_t s);
hipError_t batch_launch_iteration(const BatchParams& P, int R, bool f32, bool w32, int block, hipStream_t s);
hipError_t batch_launch_spmm(const BatchParams& P, int R, bool f32, int block, const double* x, double* y,
                             const double* b, hipStream_t s);
// the iteration loop as one cooperative launch of G workgroups of `block`
// threads (G at most batch_persistent_grid's)
hipError_t batch_persistent_grid(int R, bool f32, bool w32, bool dia, int block, int* G);
hipError_t batch_launch_persistent(const BatchParams& P, const BatchLoop& L, int R, bool f32, bool w32, int block,
                                   int G, hipStream_t s);

hipError_t batch_launch_norms(const BatchParams& P, int R, int block, const double* v, double* part, double* out,
                              hipStream_t s);
hipError_t batch_launch_mask(const BatchParams& P, int R, int block, const int* keep, double* v, hipStream_t s);
hipError_t batch_launch_finish(int nq, const double* dot, const double* rr, const int* iters, int* out_iters,
                               double* out_res, hipStream_t s);
hipError_t batch_launch_axpy(size_t m, const double* d, double* x, hipStream_t s);
// per tile: its distinct column offsets in 16 slots (INT32_MIN = unused); *fail when
// a tile has more than kBDiaMax offsets, a column lies outside its system or a
// row's columns do not strictly ascend
hipError_t batch_launch_tile_offsets(const BatchParams& P, int* toffs, int* fail, hipStream_t s);
// diagonal values and masks (dval zeroed beforehand); *fail when an entry's
// offset is not its system's or a row's offsets do not ascend
hipError_t batch_launch_fill_dia(const BatchParams& P, bool f32, void* dval, uint8_t* dmask, int* fail, hipStream_t s);

}  // namespace xj
