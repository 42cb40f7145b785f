/*
 * C integration program for the pcg.h ABI, linked through the dlopen loader
 * (libxjpcg_loader.a), reproducing the three-solve sequence and inputs of the
 * reference's pcg/sw/examples/c/example.c:36-176:
 *   n = 65536 tridiagonal, diagonal 2.2 (also the Jacobi diagonal), off-diagonals +1.0;
 *   b[i] = (rand() % n - n/2) / 97.0 from glibc rand() without srand;
 *   tol 1e-16, maxIter 100;
 *   1) COO, XJPCG_MODE_DEFAULT
 *   2) same COO values, XJPCG_MODE_KEEP_NZ_LAYOUT with NULL index arrays, fresh b
 *   3) lower-triangle CSC with Fortran indices, XJPCG_MODE_DEFAULT | XJPCG_MODE_FORTRAN_INDEX, fresh b
 *
 * usage: example_tridiag <out-prefix>     writes <out-prefix>.{b,x}{1,2,3}.bin and prints one
 *                                         "SOLVE k status iter res" line per solve
 *        example_tridiag --loader-error   checks the loader's error path (library unreachable)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pcg.h"

static void make_coo(uint32_t n, uint32_t* row, uint32_t* col, double* val, double* diag) {
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (i > 0) { row[k] = i; col[k] = i - 1; val[k++] = 1.0; }
        if (i + 1 < n) { row[k] = i; col[k] = i + 1; val[k++] = 1.0; }
        row[k] = i; col[k] = i; diag[i] = 2.2; val[k++] = 2.2;
    }
}

/* lower triangle by columns, Fortran base: column j holds (j,j)=2.2 then (j+1,j)=1.0 */
static void make_csc_lower(uint32_t n, int64_t* row_idx, int64_t* col_ptr, double* val, double* diag) {
    int64_t k = 0;
    for (uint32_t j = 0; j < n; ++j) {
        col_ptr[j] = k + 1;
        row_idx[k] = (int64_t)j + 1; val[k++] = 2.2;
        if (j + 1 < n) { row_idx[k] = (int64_t)j + 2; val[k++] = 1.0; }
        diag[j] = 2.2;
    }
    col_ptr[n] = k + 1;
}

static void make_b(uint32_t n, double* b) {
    for (uint32_t i = 0; i < n; ++i) {
        int v = rand() % (int)n - (int)(n / 2);
        b[i] = v / 97.0;
    }
}

static void dump(const char* prefix, const char* tag, int k, const double* a, uint32_t n) {
    char path[4096];
    snprintf(path, sizeof path, "%s.%s%d.bin", prefix, tag, k);
    FILE* f = fopen(path, "wb");
    if (!f) { perror(path); exit(2); }
    fwrite(a, sizeof(double), n, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <out-prefix> | --loader-error\n", argv[0]);
        return 1;
    }
    if (strcmp(argv[1], "--loader-error") == 0) {
        XJPCG_Handle_t* h = NULL;
        XJPCG_Status_t s = xJPCG_createHandle(&h, NULL);
        const char* msg = xJPCG_getLastMessage(NULL);
        printf("LOADER %d %d %s\n", (int)s, (int)xJPCG_peekAtLastStatus(h), xJPCG_getErrorString(s));
        printf("MESSAGE %s", msg ? msg : "(null)\n");
        return s == XJPCG_STATUS_DYNAMIC_LOADING_ERROR ? 0 : 3;
    }
    const char* prefix = argv[1];
    const uint32_t n = 65536, nnz = 3 * n - 2, max_iter = 100;
    const double tol = 1e-16;
    uint32_t* row = malloc(sizeof(uint32_t) * nnz);
    uint32_t* col = malloc(sizeof(uint32_t) * nnz);
    double* val = malloc(sizeof(double) * nnz);
    double* diag = malloc(sizeof(double) * n);
    double* b = malloc(sizeof(double) * n);
    double* x = calloc(n, sizeof(double));
    const uint32_t half = (nnz + n) / 2;
    int64_t* row_idx = malloc(sizeof(int64_t) * half);
    int64_t* col_ptr = malloc(sizeof(int64_t) * (n + 1));
    double* hval = malloc(sizeof(double) * half);

    XJPCG_Handle_t* h = NULL;
    XJPCG_Status_t s = xJPCG_createHandle(&h, NULL);
    printf("CREATE %d %s\n", (int)s, xJPCG_getLastMessage(h));
    if (s != XJPCG_STATUS_SUCCESS) return 4;

    uint32_t it = 0;
    double res = 0;
    make_coo(n, row, col, val, diag);
    make_b(n, b);
    s = xJPCG_cooSolver(h, n, nnz, row, col, val, diag, b, x, max_iter, tol, &it, &res, XJPCG_MODE_DEFAULT);
    printf("SOLVE 1 %d %u %.17g\n", (int)s, it, res);
    printf("MSG 1 %s\n", xJPCG_getLastMessage(h));
    dump(prefix, "b", 1, b, n);
    dump(prefix, "x", 1, x, n);

    make_coo(n, row, col, val, diag);
    make_b(n, b);
    s = xJPCG_cooSolver(h, n, nnz, NULL, NULL, val, diag, b, x, max_iter, tol, &it, &res, XJPCG_MODE_KEEP_NZ_LAYOUT);
    printf("SOLVE 2 %d %u %.17g\n", (int)s, it, res);
    printf("MSG 2 %s\n", xJPCG_getLastMessage(h));
    dump(prefix, "b", 2, b, n);
    dump(prefix, "x", 2, x, n);

    make_csc_lower(n, row_idx, col_ptr, hval, diag);
    make_b(n, b);
    int64_t it64 = 0;
    s = xJPCG_cscSymSolver(h, n, half, row_idx, col_ptr, hval, diag, b, x, max_iter, tol, &it64, &res,
                           (XJPCG_Mode_t)(XJPCG_MODE_DEFAULT | XJPCG_MODE_FORTRAN_INDEX));
    printf("SOLVE 3 %d %lld %.17g\n", (int)s, (long long)it64, res);
    printf("MSG 3 %s\n", xJPCG_getLastMessage(h));
    dump(prefix, "b", 3, b, n);
    dump(prefix, "x", 3, x, n);

    XJPCG_Metric_t m;
    s = xJPCG_getMetrics(h, &m);
    printf("METRICS %d %g %g %g %g\n", (int)s, m.m_objInit, m.m_matProc, m.m_vecProc, m.m_solver);
    printf("DESTROY %d\n", (int)xJPCG_destroyHandle(h));
    free(row); free(col); free(val); free(diag); free(b); free(x);
    free(row_idx); free(col_ptr); free(hval);
    return 0;
}
