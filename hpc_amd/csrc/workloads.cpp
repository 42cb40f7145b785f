// Seeded synthetic workload generators (declared in include/xjpcg_workloads.h).
// Host code only; nothing here touches the GPU.
#include "xjpcg_workloads.h"
#include "xjpcg_rng.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

int host_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return (int)std::min(t, 16u);
}

// run f(lo, hi) over [0, n) split into contiguous chunks, one per thread
template <typename F>
void parallel_ranges(int64_t n, F f) {
    int nt = host_threads();
    if (n < 65536 || nt == 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([=] { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

struct Stencil {
    int kind, nx, ny, nz;
    double c[3];
    // neighbour offsets in ascending column order and their values
    int nb;
    int dx[27], dy[27], dz[27];
    double v[27];

    bool init(int k, int x, int y, int z, const double* coef) {
        kind = k; nx = x; ny = y; nz = z;
        if (nx <= 0 || ny <= 0 || nz <= 0) return false;
        nb = 0;
        auto add = [&](int a, int b, int cc, double val) {
            dx[nb] = a; dy[nb] = b; dz[nb] = cc; v[nb] = val; ++nb;
        };
        if (kind == XJPCG_STENCIL_2D5) {
            if (nz != 1) return false;
            double kx = coef[0], ky = coef[1];
            add(0, -1, 0, -ky);
            add(-1, 0, 0, -kx);
            add(0, 0, 0, 2.0 * (kx + ky));
            add(1, 0, 0, -kx);
            add(0, 1, 0, -ky);
        } else if (kind == XJPCG_STENCIL_3D7) {
            double kx = coef[0], ky = coef[1], kz = coef[2];
            add(0, 0, -1, -kz);
            add(0, -1, 0, -ky);
            add(-1, 0, 0, -kx);
            add(0, 0, 0, 2.0 * (kx + ky + kz));
            add(1, 0, 0, -kx);
            add(0, 1, 0, -ky);
            add(0, 0, 1, -kz);
        } else if (kind == XJPCG_STENCIL_3D27) {
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b)
                    for (int cc = -1; cc <= 1; ++cc)
                        add(cc, b, a, (a == 0 && b == 0 && cc == 0) ? coef[0] : coef[1]);
        } else {
            return false;
        }
        return true;
    }
    int64_t n() const { return (int64_t)nx * ny * nz; }
    double diag() const {
        for (int k = 0; k < nb; ++k)
            if (dx[k] == 0 && dy[k] == 0 && dz[k] == 0) return v[k];
        return 0.0;
    }
    int row_count(int x, int y, int z) const {
        int c = 0;
        for (int k = 0; k < nb; ++k) {
            int a = x + dx[k], b = y + dy[k], cc = z + dz[k];
            c += (a >= 0 && a < nx && b >= 0 && b < ny && cc >= 0 && cc < nz);
        }
        return c;
    }
};

// number of grid points along one axis of length m whose +-1 neighbour exists: used for closed-form nnz
int64_t stencil_nnz_closed(const Stencil& s) {
    int64_t total = 0;
    for (int k = 0; k < s.nb; ++k) {
        int64_t cx = s.nx - std::abs(s.dx[k]);
        int64_t cy = s.ny - std::abs(s.dy[k]);
        int64_t cz = s.nz - std::abs(s.dz[k]);
        if (cx > 0 && cy > 0 && cz > 0) total += cx * cy * cz;
    }
    return total;
}

}  // namespace

extern "C" int64_t xjpcg_stencil_nnz(int kind, int nx, int ny, int nz) {
    Stencil s;
    double coef[3] = {1, 1, 1};
    if (!s.init(kind, nx, ny, nz, coef)) return -1;
    return stencil_nnz_closed(s);
}

extern "C" int xjpcg_stencil_fill(int kind, int nx, int ny, int nz, const double* coef, int32_t* rowptr,
                                  int32_t* col, double* val, double* diag) {
    Stencil s;
    if (coef == nullptr || !s.init(kind, nx, ny, nz, coef)) return -1;
    const int64_t n = s.n();
    const int64_t nnz = stencil_nnz_closed(s);
    if (n >= (int64_t)INT32_MAX || nnz >= (int64_t)INT32_MAX) return -1;
    // row pointer: closed-form count per row, prefix-summed sequentially
    std::vector<int32_t> rp_local;
    int32_t* rp = rowptr;
    if (rp == nullptr) {
        rp_local.resize(n + 1);
        rp = rp_local.data();
    }
    rp[0] = 0;
    {
        int64_t acc = 0, i = 0;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    acc += s.row_count(x, y, z);
                    rp[++i] = (int32_t)acc;
                }
    }
    if (col != nullptr || val != nullptr) {
        parallel_ranges(n, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
                int64_t p = rp[i];
                for (int k = 0; k < s.nb; ++k) {
                    int a = x + s.dx[k], b = y + s.dy[k], cc = z + s.dz[k];
                    if (a < 0 || a >= nx || b < 0 || b >= ny || cc < 0 || cc >= nz) continue;
                    if (col) col[p] = (int32_t)(((int64_t)cc * ny + b) * nx + a);
                    if (val) val[p] = s.v[k];
                    ++p;
                }
            }
        });
    }
    if (diag != nullptr) {
        double d = s.diag();
        for (int64_t i = 0; i < n; ++i) diag[i] = d;
    }
    return 0;
}

extern "C" void xjpcg_fill_uniform(uint64_t seed, uint64_t stream, int64_t n, double lo, double hi, double* out) {
    parallel_ranges(n, [&](int64_t a, int64_t b) {
        for (int64_t i = a; i < b; ++i) out[i] = xjpcg_uniform(seed, stream, (uint64_t)i, lo, hi);
    });
}

extern "C" void xjpcg_host_csr_matvec(int32_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                                      const double* x, double* y) {
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            double s = 0.0;
            for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * x[col[k]];
            y[i] = s;
        }
    });
}

// ---------------------------------------------------------------------------
// random graph Laplacian

struct xjpcg_graph {
    int32_t n = 0;
    std::vector<int32_t> rowptr, col;
    std::vector<double> val, diag;
};

extern "C" xjpcg_graph_t* xjpcg_graph_create(int32_t n, double mu, double sigma, int32_t cap, uint64_t seed,
                                             double wlo, double whi, double shift) {
    if (n <= 0 || cap <= 0) return nullptr;
    auto* g = new xjpcg_graph();
    g->n = n;
    // 1) degrees: Box-Muller normal from streams 0 and 1
    std::vector<int32_t> deg(n);
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            double u1 = xjpcg_u01(xjpcg_draw(seed, 0, i));
            double u2 = xjpcg_u01(xjpcg_draw(seed, 1, i));
            double z = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * M_PI * u2);
            double d = std::nearbyint(std::exp(mu + sigma * z));
            if (d < 1) d = 1;
            if (d > cap) d = cap;
            deg[i] = (int32_t)d;
        }
    });
    // 2) stubs, shuffled with Fisher-Yates over stream 2
    int64_t S = 0;
    for (int32_t i = 0; i < n; ++i) S += deg[i];
    std::vector<int32_t> stub(S);
    {
        int64_t p = 0;
        for (int32_t i = 0; i < n; ++i)
            for (int32_t k = 0; k < deg[i]; ++k) stub[p++] = i;
    }
    for (int64_t i = S - 1; i > 0; --i) {
        uint64_t j = xjpcg_draw(seed, 2, (uint64_t)i) % (uint64_t)(i + 1);
        std::swap(stub[i], stub[j]);
    }
    // 3) pair consecutive stubs; drop self loops; count both directions per row
    const int64_t E = S / 2;
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t e = 0; e < E; ++e) {
        int32_t u = stub[2 * e], v = stub[2 * e + 1];
        if (u == v) continue;
        cnt[u + 1]++;
        cnt[v + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int32_t> adj(cnt[n]);
    {
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (int64_t e = 0; e < E; ++e) {
            int32_t u = stub[2 * e], v = stub[2 * e + 1];
            if (u == v) continue;
            adj[pos[u]++] = v;
            adj[pos[v]++] = u;
        }
    }
    std::vector<int32_t>().swap(stub);
    // 4) per row: sort, drop repeats, insert the diagonal
    std::vector<int32_t> rowlen(n);
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            int32_t* b = adj.data() + cnt[i];
            int32_t* e = adj.data() + cnt[i + 1];
            std::sort(b, e);
            rowlen[i] = (int32_t)(std::unique(b, e) - b) + 1;
        }
    });
    g->rowptr.resize((size_t)n + 1);
    g->rowptr[0] = 0;
    for (int32_t i = 0; i < n; ++i) g->rowptr[i + 1] = g->rowptr[i] + rowlen[i];
    const int64_t nnz = g->rowptr[n];
    g->col.resize(nnz);
    g->val.resize(nnz);
    g->diag.resize(n);
    const uint64_t un = (uint64_t)n;
    parallel_ranges(n, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            const int32_t* b = adj.data() + cnt[i];
            int32_t m = rowlen[i] - 1;
            int64_t p = g->rowptr[i];
            int64_t dpos = -1;
            double wsum = 0.0;
            bool placed = false;
            for (int32_t k = 0; k < m; ++k) {
                int32_t j = b[k];
                if (!placed && j > i) {
                    dpos = p;
                    g->col[p++] = (int32_t)i;
                    placed = true;
                }
                uint64_t lo_ = (uint64_t)std::min<int64_t>(i, j), hi_ = (uint64_t)std::max<int64_t>(i, j);
                float w = (float)xjpcg_uniform(seed, 3, lo_ * un + hi_, wlo, whi);
                wsum += (double)w;
                g->col[p] = j;
                g->val[p] = -(double)w;
                ++p;
            }
            if (!placed) {
                dpos = p;
                g->col[p++] = (int32_t)i;
            }
            double d = wsum + shift;
            g->val[dpos] = d;
            g->diag[i] = d;
        }
    });
    return g;
}

extern "C" int64_t xjpcg_graph_nnz(const xjpcg_graph_t* g) { return g ? (int64_t)g->col.size() : -1; }

extern "C" void xjpcg_graph_copy(const xjpcg_graph_t* g, int32_t* rowptr, int32_t* col, double* val, double* diag) {
    if (!g) return;
    if (rowptr) std::memcpy(rowptr, g->rowptr.data(), sizeof(int32_t) * g->rowptr.size());
    if (col) std::memcpy(col, g->col.data(), sizeof(int32_t) * g->col.size());
    if (val) std::memcpy(val, g->val.data(), sizeof(double) * g->val.size());
    if (diag) std::memcpy(diag, g->diag.data(), sizeof(double) * g->diag.size());
}

extern "C" void xjpcg_graph_free(xjpcg_graph_t* g) { delete g; }
