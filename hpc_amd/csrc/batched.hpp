// Batched JPCG: kernel arguments and launchers (batched.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace xj {

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
    double tol = 0.0;
    uint32_t max_rounds = 0;
};

hipError_t batch_launch_setup(const BatchParams& P, int R, int block, hipStream_t s);
hipError_t batch_launch_iteration(const BatchParams& P, int R, bool f32, int block, hipStream_t s);
hipError_t batch_launch_spmm(const BatchParams& P, int R, bool f32, int block, const double* x, double* y,
                             const double* b, hipStream_t s);
hipError_t batch_launch_norms(const BatchParams& P, int R, int block, const double* v, double* part, double* out,
                              hipStream_t s);
hipError_t batch_launch_mask(const BatchParams& P, int R, int block, const int* keep, double* v, hipStream_t s);
hipError_t batch_launch_finish(int nq, const double* dot, const double* rr, const int* iters, int* out_iters,
                               double* out_res, hipStream_t s);
hipError_t batch_launch_axpy(size_t m, const double* d, double* x, hipStream_t s);

}  // namespace xj
