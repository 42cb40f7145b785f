/*
 * CPU oracle for the xJPCG solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's JPCG semantics.  Used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker
 * and CPU baseline; the product library (hpc_amd/) never links or calls it.
 * Compiled with -ffp-contract=off so every a*b+c rounds twice, like the
 * reference's separate multiply and add stages.
 *
 * The pin of each function is listed in jpcg_oracle.h.
 */
#include "jpcg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* COO -> CSR.
 * Follows L2/sparse/include/sw/fp64/matrix_params.hpp:51-79 (loadCoo): the
 * index base (0 = C, 1 = Fortran) is subtracted from every row and column
 * index, values are addressed through the identity permutation, and
 * duplicates are kept so the SpMV accumulates them
 * (L1/sparse/include/hw/xf_sparse/fp64/spmv.hpp:445-476).
 * Rows are bucketed stably: entries of a row keep their input order.
 * perm[k] = input position of CSR slot k (the KEEP_NZ_LAYOUT refill map,
 * L2/sparse/include/sw/fp64/signature.hpp:427-469).
 * Returns 0, or -1 for an unsupported base / out-of-range index. */
int orc_coo_to_csr(uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, const double* val, int base,
                   int32_t* rowptr, int32_t* col_out, double* val_out, int64_t* perm) {
    if (base != 0 && base != 1) return -1;
    uint32_t off = (uint32_t)base;
    for (uint32_t i = 0; i <= n; ++i) rowptr[i] = 0;
    for (uint32_t k = 0; k < nnz; ++k) {
        uint32_t r = row[k] - off, c = col[k] - off;
        if (r >= n || c >= n) return -1;
        rowptr[r + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
    memcpy(pos, rowptr, sizeof(int32_t) * n);
    for (uint32_t k = 0; k < nnz; ++k) {
        uint32_t r = row[k] - off;
        int32_t p = pos[r]++;
        col_out[p] = (int32_t)(col[k] - off);
        if (val_out) val_out[p] = val[k];
        if (perm) perm[p] = k;
    }
    free(pos);
    return 0;
}

/* Half-stored symmetric CSC -> full COO.
 * Follows matrix_params.hpp:81-134 (loadCscSym) for the indices and
 * gen_signature.hpp:88-112 (getCooDatFromCscSym) for the values: columns j in
 * order, their entries k in order, emit (i, j) and, off the diagonal, (j, i)
 * with the same value.  Fortran base requires col_ptr[0] >= 1.  The caller
 * sizes the outputs for 2*nnz_half - n entries (pcgImp.hpp:87); returns the
 * number written, or -1 when the expansion does not produce exactly that many
 * (the reference raises an internal error there). */
int64_t orc_cscsym_to_coo(int64_t n, int64_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                          const double* data, int base, uint32_t* row, uint32_t* col, double* val) {
    if (base != 0 && base != 1) return -1;
    int64_t off = base;
    if (base == 1 && col_ptr[0] < 1) return -1;
    int64_t full = 2 * nnz_half - n;
    int64_t idx = 0;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t k = col_ptr[j] - off; k < col_ptr[j + 1] - off; ++k) {
            int64_t i = row_idx[k] - off;
            if (idx >= full) return -1;
            row[idx] = (uint32_t)i;
            col[idx] = (uint32_t)j;
            val[idx] = data[k];
            ++idx;
            if (i != j) {
                if (idx >= full) return -1;
                row[idx] = (uint32_t)j;
                col[idx] = (uint32_t)i;
                val[idx] = data[k];
                ++idx;
            }
        }
    }
    return idx == full ? idx : -1;
}

/* y = A x, each row summed left to right in fp64.  The reference's SpMV
 * (spmv.hpp:135-181 lane sums, :445-476 row accumulation) multiplies every
 * stored entry, including duplicates, and sums per row; its summation order
 * depends on the channel partition, so this restatement fixes one order. */
void orc_spmv(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* x, double* y) {
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
        y[i] = s;
    }
}

void orc_spmv_omp(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* x,
                  double* y, int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int32_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
        y[i] = s;
    }
    (void)threads;
}

/* Fixed-order dot product of the reference's vector units:
 * products in words of 4 (CG_vecParEntries=4, pcg/hw/Makefile:103) are
 * summed as (p0+p1)+(p2+p3) (BinarySum, L1/blas/.../helpers/utils/utils.hpp:39-50,
 * preProcess in helpers/funcs/sum.hpp); the word sums are taken in groups of
 * 8 (AdderDelay<double>, utils.hpp:98-101), zero padded, each group summed as
 * a binary tree and the groups accumulated left to right from 0 (postProcess,
 * sum.hpp).  n is padded with zeros to a multiple of 4. */
double orc_dot_fixed(int64_t n, const double* a, const double* b) {
    int64_t words = (n + 3) / 4;
    int64_t groups = (words + 7) / 8;
    double total = 0.0;
    for (int64_t g = 0; g < groups; ++g) {
        double w[8];
        for (int q = 0; q < 8; ++q) {
            int64_t wi = g * 8 + q;
            double p[4];
            for (int e = 0; e < 4; ++e) {
                int64_t i = wi * 4 + e;
                p[e] = (wi < words && i < n) ? a[i] * b[i] : 0.0;
            }
            w[q] = (wi < words) ? (p[0] + p[1]) + (p[2] + p[3]) : 0.0;
        }
        double s = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
        total += s;
    }
    return total;
}

/* Vector set-up, pcg/sw/include/impl/cgVector.hpp:55-114 (loadVec + init):
 * dimAligned = ceil(n/4)*4, padding b = 0 and diag = 1; r0 = b, J = 1/diag,
 * z0 = J*r0, p0 = z0; dot = sum b^2 and rz = sum r*z accumulated left to
 * right over dimAligned entries. */
void orc_cg_init(uint32_t n, const double* b, const double* diag, double* dot, double* rz, uint32_t* dim_aligned) {
    uint32_t da = (n + 3) / 4 * 4;
    double d = 0.0, q = 0.0;
    for (uint32_t i = 0; i < da; ++i) {
        double bi = i < n ? b[i] : 0.0;
        double di = i < n ? diag[i] : 1.0;
        double j = 1.0 / di;
        double z = j * bi;
        d += bi * bi;
        q += bi * z;
    }
    *dot = d;
    *rz = q;
    if (dim_aligned) *dim_aligned = da;
}

static void put_le(uint8_t* p, uint64_t v, int bytes) {
    for (int i = 0; i < bytes; ++i) p[i] = (uint8_t)(v >> (8 * i));
}
static uint64_t get_le(const uint8_t* p, int bytes) {
    uint64_t v = 0;
    for (int i = bytes - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
static uint64_t dbits(double d) {
    uint64_t u;
    memcpy(&u, &d, 8);
    return u;
}
static double bitsd(uint64_t u) {
    double d;
    memcpy(&d, &u, 8);
    return d;
}

/* 64-byte little-endian record, field order of CGSolverInstr::encode
 * (pcg/hw/include/impl/cgInstr.hpp:117-132) with MemInstr's byte packing
 * (L1/hpc/include/memInstr.hpp:40-83). */
void orc_record_encode(const orc_record_t* r, uint8_t out[64]) {
    memset(out, 0, 64);
    put_le(out + 0, r->id, 4);
    put_le(out + 4, r->iter, 4);
    put_le(out + 8, r->vec_size, 4);
    put_le(out + 12, r->max_iter, 4);
    put_le(out + 16, dbits(r->tols), 8);
    put_le(out + 24, dbits(r->res), 8);
    put_le(out + 32, dbits(r->rz), 8);
    put_le(out + 40, dbits(r->alpha), 8);
    put_le(out + 48, dbits(r->beta), 8);
    put_le(out + 56, r->clock, 8);
}

void orc_record_decode(const uint8_t in[64], orc_record_t* r) {
    r->id = (uint32_t)get_le(in + 0, 4);
    r->iter = (uint32_t)get_le(in + 4, 4);
    r->vec_size = (uint32_t)get_le(in + 8, 4);
    r->max_iter = (uint32_t)get_le(in + 12, 4);
    r->tols = bitsd(get_le(in + 16, 8));
    r->res = bitsd(get_le(in + 24, 8));
    r->rz = bitsd(get_le(in + 32, 8));
    r->alpha = bitsd(get_le(in + 40, 8));
    r->beta = bitsd(get_le(in + 48, 8));
    r->clock = get_le(in + 56, 8);
}

static double dot_par(int64_t n, const double* a, const double* b, int threads) {
    if (threads <= 1) return orc_dot_fixed(n, a, b);
    /* CPU-baseline variant: fixed-order dot per contiguous chunk of 32-entry
     * groups, chunks added in thread order */
    double part[256];
    int t = threads > 256 ? 256 : threads;
    int64_t groups = (n + 31) / 32;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(t)
#endif
    for (int q = 0; q < t; ++q) {
        int64_t g0 = groups * q / t, g1 = groups * (q + 1) / t;
        int64_t lo = g0 * 32, hi = g1 * 32 < n ? g1 * 32 : n;
        part[q] = hi > lo ? orc_dot_fixed(hi - lo, a + lo, b + lo) : 0.0;
    }
    double s = 0.0;
    for (int q = 0; q < t; ++q) s += part[q];
    return s;
}

/* Jacobi-preconditioned CG, the reference's device loop restated.
 *
 *  set-up      cgVector.hpp:101-114 (orc_cg_init); record 0 = {MaxIter =
 *              maxIter+1, Tols = dot*tol^2, Res = dot, RZ = rz, VecSize =
 *              dimAligned} (cgVector.hpp:149-163, pcgImp.hpp:157-158, :215-225)
 *  iteration   Ap = A p; pAp = p.Ap; alpha = rz/pAp   (krnl_storeApk.cpp:63-76)
 *              x = alpha*p + x                          (update_xk.hpp:26-67, axpy.hpp:52-76)
 *              r = (-alpha)*Ap + r; z = J*r; res = r.r; rz' = r.z; beta = rz'/rz; rz = rz'
 *                                                      (update_rk_jacobi.hpp:29-110)
 *              p = beta*p + z                           (update_pk.hpp:26-72)
 *  control     record k = {Iter = k-1, Res = res, RZ, Alpha, Beta}; stop when
 *              res < Tols, else continue while ++iter < MaxIter
 *              (control.hpp:65-91, tasks.hpp:43-47), so at most maxIter+1 iterations run
 *  results     iter = index of the first empty record scanning 1..maxIter+1, capped at
 *              maxIter; res = Res of the last written record (pcgImp.hpp:171-180);
 *              reported residual sqrt(res/dot) (pcg.cpp:159)
 *
 * x receives the final iterate (n entries).  records, if not NULL, receives
 * (maxIter+2)*64 bytes; rr_hist, if not NULL, receives the res value of each
 * executed iteration (maxIter+1 entries).  threads > 1 selects the OpenMP
 * CPU-baseline variant (parallel loops, chunked dot products).
 * Returns the number of iterations executed (K), or -1 on allocation failure. */
int orc_jpcg(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val, const double* diag,
             const double* b, uint32_t max_iter, double tol, double* x, uint32_t* iter, double* res,
             uint8_t* records, double* rr_hist, int threads) {
    uint32_t da;
    double dot, rz;
    orc_cg_init((uint32_t)n, b, diag, &dot, &rz, &da);
    double* buf = (double*)calloc((size_t)da * 6, sizeof(double));
    if (!buf) return -1;
    double *xv = buf, *r = buf + da, *J = buf + 2 * (size_t)da, *z = buf + 3 * (size_t)da;
    double *p = buf + 4 * (size_t)da, *Ap = buf + 5 * (size_t)da;
    int nt = threads > 0 ? threads : 1;
    for (uint32_t i = 0; i < da; ++i) {
        double bi = i < (uint32_t)n ? b[i] : 0.0;
        double di = i < (uint32_t)n ? diag[i] : 1.0;
        r[i] = bi;
        J[i] = 1.0 / di;
        z[i] = J[i] * r[i];
        p[i] = z[i];
    }
    const uint32_t max_rounds = max_iter + 1;
    const double tols = dot * tol * tol;
    if (records) {
        memset(records, 0, (size_t)(max_iter + 2) * 64);
        orc_record_t r0 = {0, 0, da, max_rounds, tols, dot, rz, 0.0, 0.0, 0};
        orc_record_encode(&r0, records);
    }
    uint32_t it = 0; /* Task::m_iter */
    int executed = 0;
    double last_res = 0.0;
    for (;;) {
        if (nt > 1)
            orc_spmv_omp(n, rowptr, col, val, p, Ap, nt);
        else
            orc_spmv(n, rowptr, col, val, p, Ap);
        double pAp = dot_par(da, p, Ap, nt);
        double alpha = rz / pAp;
        double nalpha = -alpha;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
#endif
        for (int64_t i = 0; i < (int64_t)da; ++i) {
            xv[i] = alpha * p[i] + xv[i];
            r[i] = nalpha * Ap[i] + r[i];
            z[i] = J[i] * r[i];
        }
        double rr = dot_par(da, r, r, nt);
        double rzn = dot_par(da, r, z, nt);
        double beta = rzn / rz;
        rz = rzn;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
#endif
        for (int64_t i = 0; i < (int64_t)da; ++i) p[i] = beta * p[i] + z[i];
        ++executed;
        last_res = rr;
        if (rr_hist) rr_hist[executed - 1] = rr;
        if (records) {
            orc_record_t rk = {0, it, da, max_rounds, tols, rr, rz, alpha, beta, 0};
            orc_record_encode(&rk, records + (size_t)executed * 64);
        }
        if (!(rr < tols)) {
            ++it;
            if (it < max_rounds) continue;
        }
        break;
    }
    memcpy(x, xv, sizeof(double) * (size_t)n);
    *iter = (uint32_t)executed < max_iter ? (uint32_t)executed : max_iter;
    *res = sqrt(last_res / dot);
    free(buf);
    return executed;
}
