// C++ host engine core (see operator.hpp).
#include "operator.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "kernels_consts.hpp"
#include "staging.hpp"

namespace xj {


// Workgroups of the persistent loop: every grid barrier costs each workgroup
// an L2 write-back and invalidate on its XCD, so a small operator runs best on
// far fewer workgroups than fit on the chip.  Per-iteration times measured on
// MI355X put the best grid near 0.1 * sqrt(nnz) (256^2 Laplacian: 64
// workgroups, 25 us against 288 us on all 2048; 128^3 27-point: 512), a
// multiple of 8 so each XCD gets the same share, at most gmax (co-resident).
static int persistent_size(int gmax, int64_t nnz) {
    int g = (int)(0.1 * std::sqrt((double)std::max<int64_t>(nnz, 1)));
    g = (g + 7) / 8 * 8;
    return std::max(1, std::min(gmax, std::max(8, g)));
}

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) throw alloc_failed(m);
    throw internal(m);
}

std::vector<int32_t> plan_tiles(int32_t n, const int32_t* rp) {
    std::vector<int32_t> t;
    t.reserve((size_t)n / 32 + 2);
    t.push_back(0);
    int32_t r = 0;
    while (r < n) {
        const int32_t start = r;
        if (rp[r + 1] - rp[r] > kTilePlanNnz) {
            r += 1;  // one long row
        } else {
            while (r < n && r - start < kTileRowsHost && rp[r + 1] - rp[start] <= kTilePlanNnz) ++r;
        }
        t.push_back(r);
    }
    return t;
}

bool plan_patterns(int32_t n, const int32_t* rp, const int32_t* col, const std::vector<int32_t>& tiles,
                   std::vector<uint32_t>& mask, std::vector<int32_t>& offs) {
    const int nt = (int)tiles.size() - 1;
    mask.assign((size_t)n, 0u);
    offs.assign((size_t)nt * 32, 0);
    std::vector<int32_t> ds;
    ds.reserve(32);
    // offset of the u-th entry of the previous row and its bit: stencil rows
    // repeat their neighbour's pattern, so most entries skip the search
    int32_t last_d[32], last_b[32];
    for (int t = 0; t < nt; ++t) {
        const int32_t r0 = tiles[(size_t)t], r1 = tiles[(size_t)t + 1];
        if (rp[r1] - rp[r0] > kTilePlanNnz) return false;
        ds.clear();
        int nlast = 0;
        for (int32_t i = r0; i < r1; ++i) {
            const int len = rp[i + 1] - rp[i];
            if (len > 32) return false;
            int32_t prev = -1;
            for (int u = 0; u < len; ++u) {
                const int32_t c = col[rp[i] + u];
                if (c < 0 || c >= n || c <= prev) return false;
                prev = c;
                const int32_t d = c - i;
                if (!(u < nlast && last_d[u] == d) && std::find(ds.begin(), ds.end(), d) == ds.end()) {
                    if (ds.size() == 32) return false;
                    ds.push_back(d);
                }
                last_d[u] = d;
            }
            nlast = len;
        }
        std::sort(ds.begin(), ds.end());
        for (size_t b = 0; b < ds.size(); ++b) offs[(size_t)t * 32 + b] = ds[b];
        nlast = 0;
        for (int32_t i = r0; i < r1; ++i) {
            const int len = rp[i + 1] - rp[i];
            uint32_t m = 0;
            for (int u = 0; u < len; ++u) {
                const int32_t d = col[rp[i] + u] - i;
                int b;
                if (u < nlast && last_d[u] == d) {
                    b = last_b[u];
                } else {
                    b = (int)(std::lower_bound(ds.begin(), ds.end(), d) - ds.begin());
                }
                last_d[u] = d;
                last_b[u] = b;
                m |= 1u << b;
            }
            nlast = len;
            mask[(size_t)i] = m;
        }
    }
    return true;
}

int panel_block_rows(int32_t n) {
    // largest power of two up to kPanelRowsMaxHost that still gives >= 512 blocks
    int br = kPanelRowsMaxHost;
    while (br > 256 && (int64_t)n < (int64_t)br * 512) br >>= 1;
    return br;
}

int64_t panel_segments(int32_t n, int block_rows) {
    const int64_t nb = ((int64_t)n + block_rows - 1) / block_rows;
    const int64_t P = ((int64_t)n + kPanelColsHost - 1) / kPanelColsHost;
    return nb * P;
}

PanelPlan plan_panels(int32_t n, const int32_t* rp, const int32_t* col, int block_rows) {
    PanelPlan pl;
    pl.block_rows = block_rows;
    pl.nb = (int)(((int64_t)n + block_rows - 1) / block_rows);
    pl.P = (int)(((int64_t)n + kPanelColsHost - 1) / kPanelColsHost);
    const int64_t nseg = (int64_t)pl.nb * pl.P;
    const int64_t nnz = rp[n];
    pl.seg.assign((size_t)nseg + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        const int64_t base = (int64_t)(i / block_rows) * pl.P;
        int last = 0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int32_t c = col[k];
            if (c < 0 || c >= n) throw invalid("column index out of range.");
            const int p = c / kPanelColsHost;
            if (p < last) pl.ordered = false;
            last = p;
            ++pl.seg[(size_t)(base + p + 1)];
        }
    }
    for (int64_t q = 0; q < nseg; ++q) pl.seg[(size_t)q + 1] += pl.seg[(size_t)q];
    std::vector<int32_t> cur(pl.seg.begin(), pl.seg.end() - 1);
    pl.off.resize((size_t)nnz);
    pl.lrow.resize((size_t)nnz);
    pl.perm.resize((size_t)nnz);
    for (int32_t i = 0; i < n; ++i) {
        const int b = i / block_rows;
        const int64_t base = (int64_t)b * pl.P;
        const uint16_t lr = (uint16_t)(i - b * block_rows);
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const int32_t c = col[k];
            const int p = c / kPanelColsHost;
            const int32_t q = cur[(size_t)(base + p)]++;
            pl.off[(size_t)q] = (uint16_t)(c - p * kPanelColsHost);
            pl.lrow[(size_t)q] = lr;
            pl.perm[(size_t)q] = k;
        }
    }
    // run codes: the kernels stage a segment kPanelChunkHost entries at a
    // time; the first entry of each row's run inside such a chunk carries the
    // run's length (capped at kPanelRunScan: a wavefront sums the run and
    // scans for its end) in the bits above the row offset, the other entries
    // carry 0
    for (int64_t q = 0; q < nseg; ++q) {
        const int64_t s1 = pl.seg[(size_t)q + 1];
        for (int64_t c0 = pl.seg[(size_t)q]; c0 < s1; c0 += kPanelChunkHost) {
            const int64_t c1 = std::min<int64_t>(c0 + kPanelChunkHost, s1);
            for (int64_t k = c0; k < c1;) {
                int64_t e = k + 1;
                while (e < c1 && pl.lrow[(size_t)e] == pl.lrow[(size_t)k]) ++e;
                const int64_t len = std::min<int64_t>(e - k, kPanelRunScan);
                pl.lrow[(size_t)k] = (uint16_t)(pl.lrow[(size_t)k] | (len << kPanelRowBits));
                k = e;
            }
        }
    }
    return pl;
}

Operator::Operator(int32_t n, int64_t nnz, const int32_t* rowptr, const int32_t* col, const void* val, bool val_f32,
                   const double* dsplit, const int32_t* rowptr_host, hipStream_t s) {
    if (n <= 0) throw invalid("Wrong dimension size.");
    if (nnz < 0 || nnz >= (int64_t)INT32_MAX) throw invalid("number of stored entries must be in [0, 2^31-1).");
    if (rowptr == nullptr || (nnz > 0 && (col == nullptr || val == nullptr)))
        throw invalid("Matrix is nullptr.");
    if ((reinterpret_cast<uintptr_t>(col) | reinterpret_cast<uintptr_t>(val)) & 15u)
        throw invalid("column and value arrays must be 16-byte aligned.");
    std::vector<int32_t> rp_local;
    const int32_t* rp = rowptr_host;
    if (rp == nullptr) {
        rp_local.resize((size_t)n + 1);
        hip_check(hipMemcpyAsync(rp_local.data(), rowptr, sizeof(int32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, s),
                  "copy row pointers");
        hip_check(hipStreamSynchronize(s), "copy row pointers");
        rp = rp_local.data();
    }
    if (rp[0] != 0 || (int64_t)rp[n] != nnz) throw invalid("row pointers do not span [0, nnz].");
    for (int32_t i = 0; i < n; ++i)
        if (rp[i + 1] < rp[i]) throw invalid("row pointers must not decrease.");
    std::vector<int32_t> t = plan_tiles(n, rp);
    tile_plan_ = t;
    std::vector<int32_t> tk(t.size());
    for (size_t q = 0; q < t.size(); ++q) tk[q] = rp[t[q]];
    tiles_.ensure(t.size());
    tilek_.ensure(tk.size());
    hip_check(hipMemcpyAsync(tiles_.ptr, t.data(), sizeof(int32_t) * t.size(), hipMemcpyHostToDevice, s),
              "upload tile plan");
    hip_check(hipMemcpyAsync(tilek_.ptr, tk.data(), sizeof(int32_t) * tk.size(), hipMemcpyHostToDevice, s),
              "upload tile entry offsets");
    A_.n = n;
    A_.nnz = nnz;
    A_.rowptr = rowptr;
    A_.col = col;
    A_.val = val;
    A_.val_f32 = val_f32;
    A_.dsplit = dsplit;
    A_.tile_row = tiles_.ptr;
    A_.tile_k = tilek_.ptr;
    A_.ntiles = (int)t.size() - 1;
    G2_ = vector_groups(n);
    chunk_ = chunk_for(n);
    partA_.ensure(std::max(1, A_.ntiles));
    partA2_.ensure(kAlphaBlocks);
    partB_.ensure(G2_);
    partC_.ensure(G2_);
    state_.ensure(1);
    r_.ensure(n);
    p_.ensure(n);
    Ap_.ensure(n);
    J_.ensure(n);
    result_.ensure(8);
    bar_.ensure(1);
    abort_.ensure(1);
    hip_check(hipStreamCreateWithFlags(&work_, hipStreamNonBlocking), "operator stream");
    hip_check(hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming), "event");
    hip_check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming), "event");
    hip_check(hipHostMalloc((void**)&host_flags_, 2 * sizeof(int), hipHostMallocDefault), "pinned flags");
    hip_check(hipHostMalloc((void**)&host_result_, 16 * sizeof(double), hipHostMallocDefault), "pinned result");
    static_assert(sizeof(CgState) <= 16 * sizeof(double), "state fits the pinned result buffer");
    // compact int16 column stream when every tile's columns fit
    if (nnz > 0) {
        col16_.ensure((size_t)nnz);
        hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset flag");
        hip_check(launch_build_col16(A_, col16_.ptr, abort_.ptr, s), "build int16 columns");
        int overflow = 1;
        hip_check(hipMemcpyAsync(&overflow, abort_.ptr, sizeof(int), hipMemcpyDeviceToHost, s), "int16 check");
        hip_check(hipStreamSynchronize(s), "int16 check");
        if (overflow) {
            col16_.release();
        } else {
            has_col16_ = true;
            A_.col16 = col16_.ptr;
        }
    }
    // Layouts for the standalone SpMV, timed on this operator: column-offset
    // pattern tiles when every tile has at most 32 distinct offsets
    // (stencils), the diagonal layout when the whole matrix has at most 32
    // (and its value arrays stay within twice the stored entries), and, when
    // the columns are too spread for the int16 stream and x is far larger than
    // an XCD's L2, the panel layout.  Every layout sums each row left to
    // right, so the SpMV result does not depend on the choice.
    //
    // The solves' layout follows from the matrix instead, so that operators
    // built from the same matrix give bit-identical iterates whatever the
    // timings: the diagonals when they hold at most 25% padding slots (at
    // every size: below kDiaWideRows the fp64 kernel takes one row per lane,
    // so a small operator still fills the chip -- 512^2 diffusion 45.0 ms
    // against 48.8 ms on the tiles, 27-point 64^3 4.0 against 5.6 ms), else
    // the panels when their rule holds (no int16 column stream, at least
    // kPanelAutoRows rows, columns ascending in panel order), else the tiles
    // or the pattern tiles (whichever SpMV is faster; both give the same
    // per-tile p.Ap partials, so the same iterates).  The diagonal and panel
    // layouts group the p.Ap partials their own way (per workgroup, per row
    // block).
    if (nnz > 0) {
        const float t_tiles = time_layout(kTiles, s);
        float best = t_tiles, t_pat = 0.0f;
        if (build_patterns(s)) {
            t_pat = time_layout(kPatterns, s);
            if (t_pat > 0.0f && t_pat < best) {
                best = t_pat;
                auto_layout_ = kPatterns;
            }
        }
        if (build_diagonals(s)) {
            const float t_dia = time_layout(kDiagonals, s);
            if (t_dia > 0.0f && t_dia < best) {
                best = t_dia;
                auto_layout_ = kDiagonals;
            }
        }
        // panels by rule (columns too spread for the int16 stream, x beyond
        // an XCD's L2, every row's columns in ascending panel order), so the
        // solve layout does not depend on a timing; the standalone SpMV still
        // takes them only when they time faster
        bool panels_ok = false;
        if (!has_col16_ && n >= kPanelAutoRows && build_panels(s)) {
            panels_ok = true;
            const float t_pan = time_layout(kPanels, s);
            if (t_pan > 0.0f && t_pan < best) {
                best = t_pan;
                auto_layout_ = kPanels;
            }
        }
        auto_solve_ = (t_pat > 0.0f && t_pat < t_tiles) ? kPatterns : kTiles;
        if (panels_ok) auto_solve_ = kPanels;
        if (has_dia_ && 4 * (int64_t)A_.dia.ndia * n <= 5 * nnz) auto_solve_ = kDiagonals;
        const auto used = [&](int L) { return auto_layout_ == L || auto_solve_ == L; };
        if (!used(kPanels)) free_panels();
        if (!used(kPatterns)) free_patterns();
        if (!used(kDiagonals)) free_diagonals();
        A_.pan_on = auto_layout_ == kPanels;
        A_.pat_on = auto_layout_ == kPatterns;
        A_.dia_on = auto_layout_ == kDiagonals;
        solve_layout_ = auto_solve_;
    }
    hip_check(persistent_grid(val_f32, A_.col16 != nullptr, &G_), "persistent kernel occupancy");
    G_ = persistent_size(G_, A_.nnz);
    part_.ensure((size_t)5 * G_ + 4096);
    hip_check(hipStreamSynchronize(s), "operator set-up");
}

void Operator::copy_pattern_to_host(hipStream_t s, std::vector<int32_t>& rp, std::vector<int32_t>& col) const {
    rp.resize((size_t)A_.n + 1);
    col.resize((size_t)A_.nnz);
    hip_check(hipMemcpyAsync(rp.data(), A_.rowptr, sizeof(int32_t) * rp.size(), hipMemcpyDeviceToHost, s),
              "copy row pointers");
    if (!col.empty())
        hip_check(hipMemcpyAsync(col.data(), A_.col, sizeof(int32_t) * col.size(), hipMemcpyDeviceToHost, s),
                  "copy columns");
    hip_check(hipStreamSynchronize(s), "copy pattern");
}

// plan on the GPU (plan_pattern_tiles); no copy of the pattern to the host
bool Operator::build_patterns(hipStream_t s) {
    if (has_pat_) return true;
    if (pat_tried_) return false;
    pat_tried_ = true;
    try {
        pmask_.ensure((size_t)A_.n);
        poffs_.ensure((size_t)std::max(1, A_.ntiles) * 32);
    } catch (const Error& e) {
        if (e.status != XJPCG_STATUS_ALLOC_FAILED) throw;
        free_patterns();
        pat_tried_ = true;
        return false;
    }
    hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset flag");
    hip_check(launch_plan_patterns(A_, pmask_.ptr, poffs_.ptr, abort_.ptr, s), "pattern plan");
    int fail = 1;
    hip_check(hipMemcpyAsync(&fail, abort_.ptr, sizeof(int), hipMemcpyDeviceToHost, s), "pattern plan");
    hip_check(hipStreamSynchronize(s), "pattern plan");
    if (fail) {
        free_patterns();
        pat_tried_ = true;
        return false;
    }
    A_.pmask = pmask_.ptr;
    A_.poffs = poffs_.ptr;
    has_pat_ = true;
    return true;
}

void Operator::free_patterns() {
    pmask_.release();
    poffs_.release();
    A_.pmask = nullptr;
    A_.poffs = nullptr;
    A_.pat_on = false;
    has_pat_ = false;
    pat_tried_ = false;
}

// Offsets from the per-tile plan (on the GPU), their union on the host, then
// the value arrays and masks on the GPU.  Refused (false) for more than 32
// global offsets, rows whose columns do not ascend, or value arrays beyond
// twice the stored entries.
bool Operator::build_diagonals(hipStream_t s) {
    if (has_dia_) return true;
    if (dia_tried_ || A_.nnz <= 0 || A_.ntiles <= 0) return false;
    dia_tried_ = true;
    std::vector<int32_t> offs;
    {
        DevBuf<int32_t> toffs;
        try {
            toffs.ensure((size_t)A_.ntiles * 32);
        } catch (const Error& e) {
            if (e.status != XJPCG_STATUS_ALLOC_FAILED) throw;
            return false;
        }
        hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset flag");
        hip_check(launch_tile_offsets(A_, toffs.ptr, abort_.ptr, s), "tile offsets");
        std::vector<int32_t> h((size_t)A_.ntiles * 32);
        int fail = 1;
        hip_check(hipMemcpyAsync(h.data(), toffs.ptr, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost, s),
                  "tile offsets");
        hip_check(hipMemcpyAsync(&fail, abort_.ptr, sizeof(int), hipMemcpyDeviceToHost, s), "tile offsets");
        hip_check(hipStreamSynchronize(s), "tile offsets");
        if (fail) return false;
        // distinct offsets, ascending, kept in a short sorted list: sorting
        // every tile's copy (1.45M values on the 160^3 stencil) took 21 ms
        for (int32_t v : h) {
            if (v == INT32_MIN) continue;
            const auto it = std::lower_bound(offs.begin(), offs.end(), v);
            if (it != offs.end() && *it == v) continue;
            if ((int)offs.size() > kDiaMax) break;  // already more than the layout holds
            offs.insert(it, v);
        }
    }
    const int ndia = (int)offs.size();
    if (ndia == 0 || ndia > kDiaMax || (int64_t)ndia * A_.n > 2 * A_.nnz) return false;
    const int nblk = dia_cg_partials(A_.n, A_.val_f32);  // the CG kernel's p.Ap partials
    const int ld = dia_ld(A_.n);
    const size_t vbytes = (size_t)ndia * ld * (A_.val_f32 ? 4 : 8);
    try {
        doff_.ensure(kDiaMax + 8);  // padded: lanes never read past the table
        dmask_.ensure((size_t)ld);
        dval_.ensure((vbytes + 7) / 8);
        partA_.ensure((size_t)std::max({A_.ntiles, nblk, A_.pan.nb}));
    } catch (const Error& e) {
        if (e.status != XJPCG_STATUS_ALLOC_FAILED) throw;
        free_diagonals();
        dia_tried_ = true;
        return false;
    }
    offs.resize(kDiaMax + 8, 0);
    hip_check(hipMemcpyAsync(doff_.ptr, offs.data(), sizeof(int32_t) * offs.size(), hipMemcpyHostToDevice, s),
              "upload offsets");
    hip_check(hipMemsetAsync(dval_.ptr, 0, vbytes, s), "clear diagonals");
    hip_check(hipMemsetAsync(dmask_.ptr, 0, sizeof(uint32_t) * ld, s), "clear masks");
    hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset flag");
    hip_check(launch_fill_dia(A_, ndia, doff_.ptr, dval_.ptr, dmask_.ptr, abort_.ptr, s), "fill diagonals");
    int fail = 1;
    hip_check(hipMemcpyAsync(&fail, abort_.ptr, sizeof(int), hipMemcpyDeviceToHost, s), "fill diagonals");
    hip_check(hipStreamSynchronize(s), "fill diagonals");  // host tables end with this scope
    if (fail) {
        free_diagonals();
        dia_tried_ = true;
        return false;
    }
    A_.dia.ndia = ndia;
    A_.dia.ld = ld;
    A_.dia.off = doff_.ptr;
    A_.dia.val = dval_.ptr;
    A_.dia.mask = dmask_.ptr;
    has_dia_ = true;
    dia_stale_ = false;
    return true;
}

void Operator::free_diagonals() {
    doff_.release();
    dmask_.release();
    dval_.release();
    A_.dia = DiaDev();
    A_.dia_on = false;
    has_dia_ = false;
    dia_tried_ = false;
}

void Operator::refresh_copies(hipStream_t s) const {
    refresh_panels(s);
    if (!(A_.dia_on || solve_layout_ == kDiagonals) || !dia_stale_) return;
    // same pattern: the fill rewrites every stored slot, the zeros stay
    hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset flag");
    hip_check(launch_fill_dia(A_, A_.dia.ndia, A_.dia.off, const_cast<void*>(A_.dia.val), dmask_.ptr, abort_.ptr, s),
              "diagonal values");
    dia_stale_ = false;
}

bool Operator::build_panels(hipStream_t s, const int32_t* rp_host, const int32_t* col_host) {
    if (has_pan_) return true;
    const int32_t n = A_.n;
    const int br = panel_block_rows(n);
    if (panel_segments(n, br) > kPanelMaxSegments) return false;
    std::vector<int32_t> rp_l, col_l;
    if (!rp_host || !col_host) copy_pattern_to_host(s, rp_l, col_l);
    const int32_t* rp = rp_host && col_host ? rp_host : rp_l.data();
    const int32_t* col = rp_host && col_host ? col_host : col_l.data();
    PanelPlan pl = plan_panels(n, rp, col, br);
    if (!pl.ordered) return false;
    try {
        pseg_.ensure(pl.seg.size());
        poff_.ensure(pl.off.size());
        plrow_.ensure(pl.lrow.size());
        pperm_.ensure(pl.perm.size());
        pval_.ensure(A_.val_f32 ? ((size_t)A_.nnz + 1) / 2 : (size_t)A_.nnz);
        partA_.ensure((size_t)std::max(A_.ntiles, pl.nb));
    } catch (const Error& e) {
        if (e.status != XJPCG_STATUS_ALLOC_FAILED) throw;
        free_panels();  // the panel layout is optional: keep the tiles
        return false;
    }
    const HostPiece up[4] = {{pseg_.ptr, pl.seg.data(), sizeof(int32_t) * pl.seg.size()},
                             {poff_.ptr, pl.off.data(), sizeof(uint16_t) * pl.off.size()},
                             {plrow_.ptr, pl.lrow.data(), sizeof(uint16_t) * pl.lrow.size()},
                             {pperm_.ptr, pl.perm.data(), sizeof(int64_t) * pl.perm.size()}};
    staged_upload(up, 4, s);
    A_.pan.nb = pl.nb;
    A_.pan.block_rows = pl.block_rows;
    A_.pan.P = pl.P;
    A_.pan.seg = pseg_.ptr;
    A_.pan.val = pval_.ptr;
    A_.pan.off = poff_.ptr;
    A_.pan.lrow = plrow_.ptr;
    has_pan_ = true;
    pan_stale_ = true;
    hip_check(hipStreamSynchronize(s), "upload panel layout");  // host plan buffers end with this scope
    return true;
}

void Operator::free_panels() {
    pseg_.release();
    poff_.release();
    plrow_.release();
    pperm_.release();
    pval_.release();
    A_.pan = PanelDev();
    A_.pan_on = false;
    has_pan_ = false;
}

void Operator::refresh_panels(hipStream_t s) const {
    if (!(A_.pan_on || solve_layout_ == kPanels) || !pan_stale_) return;
    hip_check(launch_gather_vals(A_.nnz, pperm_.ptr, A_.val, pval_.ptr, A_.val_f32, s), "panel values");
    pan_stale_ = false;
}

// time 5 SpMV launches on `layout` (after 2 warm-up launches); leaves that layout switched on
float Operator::time_layout(int layout, hipStream_t s) {
    A_.pan_on = layout == kPanels;
    A_.pat_on = layout == kPatterns;
    A_.dia_on = layout == kDiagonals;
    refresh_copies(s);
    hip_check(hipMemsetAsync(r_.ptr, 0, sizeof(double) * (size_t)A_.n, s), "clear timing input");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "event");
    hipError_t e = hipEventCreate(&e1);
    float ms = 0.0f;
    if (e == hipSuccess) {
        for (int i = 0; i < 2 && e == hipSuccess; ++i) e = launch_spmv(A_, r_.ptr, Ap_.ptr, nullptr, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int i = 0; i < 5 && e == hipSuccess; ++i) e = launch_spmv(A_, r_.ptr, Ap_.ptr, nullptr, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    hip_check(e, "time SpMV layouts");
    return ms;
}

int64_t Operator::spmv_bytes(int layout) const {
    const int L = layout == kLayoutAuto ? this->layout() : layout;
    const int64_t n = A_.n, nnz = A_.nnz;
    const int64_t vb = A_.val_f32 ? 4 : 8;
    const int64_t vec = 16 * n + (A_.dsplit ? 8 * n : 0);  // x and y once, the split diagonal once
    switch (L) {
        case kTiles: {
            // values, the column stream in use, both row pointers of each row
            const int64_t cb = A_.col16 ? 2 : 4;
            return (vb + cb) * nnz + 4 * (n + 1) + vec;
        }
        case kPatterns:
            if (!has_pat_) return -1;
            // values, row pointers, one mask per row, 32 offsets per tile
            return vb * nnz + 4 * (n + 1) + 4 * n + 128 * (int64_t)A_.ntiles + vec;
        case kDiagonals:
            if (!has_dia_) return -1;
            // every diagonal's zero-padded value array, one mask per row
            return vb * (int64_t)A_.dia.ndia * A_.dia.ld + 4 * (int64_t)A_.dia.ld + vec;
        case kPanels:
            if (!has_pan_) return -1;
            // value, uint16 column and uint16 row offset per entry, the segment table
            return (vb + 4) * nnz + 4 * ((int64_t)A_.pan.nb * A_.pan.P + 1) + vec;
        default:
            return -1;
    }
}

void Operator::set_layout(int layout, hipStream_t s) {
    int want = layout == kLayoutAuto ? auto_layout_ : layout;
    if (want != kTiles && want != kPanels && want != kPatterns && want != kDiagonals)
        throw invalid("layout must be 0 (auto), 1 (tiles), 2 (panels), 3 (column-offset patterns) or 4 "
                      "(diagonals).");
    if (want == kPanels && !has_pan_ && !build_panels(s))
        throw invalid("the panel layout needs each row's columns in ascending order and fewer than 2^26 "
                      "(row block, column panel) segments.");
    if (want == kPatterns && !has_pat_ && !build_patterns(s))
        throw invalid("column-offset patterns need every row tile to use at most 32 distinct offsets (col - row), "
                      "rows of at most 32 entries and strictly ascending columns.");
    if (want == kDiagonals && !has_dia_ && !build_diagonals(s))
        throw invalid("the diagonal layout needs at most 32 distinct offsets (col - row) over the whole matrix, "
                      "strictly ascending columns in every row and diagonals no longer than twice the stored "
                      "entries.");
    const bool pan = want == kPanels, pat = want == kPatterns, dia = want == kDiagonals;
    const int solve = layout == kLayoutAuto ? auto_solve_ : want;
    if (solve == kPanels && !has_pan_ && !build_panels(s)) throw internal("panel layout lost.");
    if (solve == kPatterns && !has_pat_ && !build_patterns(s)) throw internal("pattern layout lost.");
    if (solve == kDiagonals && !has_dia_ && !build_diagonals(s)) throw internal("diagonal layout lost.");
    if (pan == A_.pan_on && pat == A_.pat_on && dia == A_.dia_on && solve == solve_layout_) return;
    A_.pan_on = pan;
    A_.pat_on = pat;
    A_.dia_on = dia;
    solve_layout_ = solve;
    drop_graph();
}

Operator::~Operator() {
    drop_graph();
    if (host_result_) (void)hipHostFree(host_result_);
    if (host_flags_) (void)hipHostFree(host_flags_);
    for (hipEvent_t e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (work_) (void)hipStreamDestroy(work_);
}

void Operator::set_compact(bool on) {
    const int16_t* want = (on && has_col16_) ? col16_.ptr : nullptr;
    if (want == A_.col16) return;
    A_.col16 = want;
    drop_graph();
    hip_check(persistent_grid(A_.val_f32, A_.col16 != nullptr, &G_), "persistent kernel occupancy");
    G_ = persistent_size(G_, A_.nnz);
    part_.ensure((size_t)5 * G_ + 4096);
}

void Operator::spmv(const double* x, double* y, hipStream_t s) const {
    refresh_copies(s);
    hip_check(launch_spmv(A_, x, y, nullptr, s), "spmv launch");
}

double Operator::dot(const double* a, const double* b, int n, hipStream_t s) {
    hip_check(launch_dot(n, a, b, part_.ptr, result_.ptr, s), "dot launch");
    hip_check(hipMemcpyAsync(host_result_, result_.ptr, sizeof(double), hipMemcpyDeviceToHost, s), "dot result");
    hip_check(hipStreamSynchronize(s), "dot");
    return host_result_[0];
}

double Operator::residual(const double* b, const double* x, double* r, hipStream_t s) {
    refresh_copies(s);
    hip_check(launch_spmv(A_, x, r, b, s), "residual launch");
    return dot(r, r, A_.n, s);
}

void Operator::sddmm(const double* lam, const double* x, double* g, hipStream_t s) const {
    hip_check(launch_sddmm(A_, lam, x, g, s), "sddmm launch");
}

SolveResult Operator::solve(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                            void* records, hipStream_t s) {
    if (diag == nullptr || b == nullptr || x == nullptr) throw invalid("Vector is nullptr.");
    if (max_iter == 0xFFFFFFFFu) throw invalid("maxIter too large.");
    refresh_copies(s);
    int p = path_;
    if (p == kAuto) p = kMultiKernel;
    // the persistent kernel runs the tile SpMV with fp64 vectors only
    if (!coop_ok_ || solve_layout_ == kPanels || vector_bits() == 32) p = kMultiKernel;
    last_path_ = p;
    if (p == kPersistent) return solve_persistent(diag, b, x, max_iter, tol, records, s);
    return solve_multi(diag, b, x, max_iter, tol, records, s);
}

static SolveResult finish_result(int64_t K, double rr, double dot, uint32_t max_iter) {
    SolveResult R;
    R.executed = K;
    R.rr = rr;
    R.dot = dot;
    // result scan of pcgImp.hpp:171-180: count capped at maxIter
    R.iter = R.executed < (int64_t)max_iter ? (uint32_t)R.executed : max_iter;
    R.res = R.dot > 0.0 ? std::sqrt(R.rr / R.dot) : 0.0;
    return R;
}

CgParams Operator::cg_params(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                             void* records) const {
    CgParams P;
    P.n = A_.n;
    // the solve layout (which may differ from the standalone SpMV's)
    P.pan_on = solve_layout_ == kPanels ? 1 : 0;
    P.pan = A_.pan;
    P.pat_on = solve_layout_ == kPatterns ? 1 : 0;
    P.pmask = A_.pmask;
    P.poffs = A_.poffs;
    P.dia_on = solve_layout_ == kDiagonals ? 1 : 0;
    P.dia = A_.dia;
    P.tiles = A_.ntiles;
    // p.Ap partials: one per panel row block, per workgroup of the diagonal
    // layout, or per row tile
    P.ntiles = P.pan_on ? A_.pan.nb : P.dia_on ? dia_cg_partials(A_.n, A_.val_f32) : A_.ntiles;
    P.nnz = (int)A_.nnz;
    P.G2 = G2_;
    P.split_alpha = P.ntiles > kMaxRedundantPartials ? 1 : 0;
    P.nA2 = kAlphaBlocks;
    P.partA2 = partA2_.ptr;
    P.rowptr = A_.rowptr;
    P.col = col_stream();
    P.val = A_.val;
    P.dsplit = A_.dsplit;
    P.tile_row = A_.tile_row;
    P.tile_k = A_.tile_k;
    P.b = b;
    P.diag = diag;
    P.x = x;
    P.r = r_.ptr;
    P.p = p_.ptr;
    P.Ap = Ap_.ptr;
    P.J = J_.ptr;
    P.w32 = vector_bits() == 32 ? 1 : 0;
    P.partA = partA_.ptr;
    P.partB = partB_.ptr;
    P.partC = partC_.ptr;
    P.st = state_.ptr;
    P.records = static_cast<Record*>(records);
    P.tol = tol;
    P.max_rounds = max_iter + 1u;  // pcgImp.hpp:158 programs maxIter+1 rounds
    return P;
}

void Operator::set_vector_bits(int bits) {
    if (bits != 32 && bits != 64) throw invalid("vector bits must be 32 or 64.");
    if (bits == 32 && solve_layout_ != kDiagonals && solve_layout_ != kPanels)
        throw invalid("fp32 work vectors need the diagonal or the panel solve layout.");
    w32_ = bits == 32;
    drop_graph();
}

void Operator::drop_graph() {
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    gexec_ = nullptr;
}

void Operator::build_graph(const CgParams& P) {
    drop_graph();
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(work_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        graph_ok_ = false;
        return;
    }
    hipError_t e = hipSuccess;
    for (int c = 0; c < chunk_ && e == hipSuccess; ++c) e = launch_cg_iteration(P, A_.val_f32, A_.col16, work_);
    hipError_t e2 = hipStreamEndCapture(work_, &g);
    if (e == hipSuccess && e2 == hipSuccess && g) e = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess || e2 != hipSuccess) {
        (void)hipGetLastError();
        gexec_ = nullptr;
        graph_ok_ = false;
        return;
    }
    gkey_ = P;
}

SolveResult Operator::solve_multi(const double* diag, const double* b, double* x, uint32_t max_iter, double tol,
                                  void* records, hipStream_t s) {
    CgParams P = cg_params(diag, b, x, max_iter, tol, records);
    // the loop's end is signalled to the host in pinned memory (no copy in the
    // stream between chunks); the previous solve has finished, so the word
    // can be cleared from the host
    volatile int* host_stop = host_flags_;
    *host_stop = 0;
    P.host_stop = host_flags_;
    // the work stream follows everything already queued on the caller's stream
    hip_check(hipEventRecord(ev_join_, s), "event record");
    hip_check(hipStreamWaitEvent(work_, ev_join_, 0), "stream wait");
    if (records) hip_check(hipMemsetAsync(records, 0, (size_t)(max_iter + 2) * 64, work_), "clear records");
    hip_check(launch_cg_setup(P, work_), "CG set-up launch");
    // The iteration kernels read b, diag and tol only through the set-up, so a
    // graph captured for (x, records, max_rounds, values) replays for any b.
    if (graph_ok_ && (!gexec_ || gkey_.x != P.x || gkey_.records != P.records || gkey_.max_rounds != P.max_rounds ||
                      gkey_.val != P.val || gkey_.col != P.col || gkey_.pan_on != P.pan_on ||
                      gkey_.pat_on != P.pat_on || gkey_.dia_on != P.dia_on || gkey_.w32 != P.w32 ||
                      gkey_.host_stop != P.host_stop))
        build_graph(P);
    // Enqueue iterations in chunks, with an event after each.  The host waits
    // for chunk c-1's event while chunk c is queued and then reads the pinned
    // stop word, so the GPU never idles and at most one chunk of launches
    // returns at once after convergence.
    uint64_t enq = 0;
    int slot = 0;
    bool pending = false;
    while (enq < P.max_rounds) {
        const uint64_t left = P.max_rounds - enq;
        if (gexec_ && left >= (uint64_t)chunk_) {
            hip_check(hipGraphLaunch(gexec_, work_), "graph launch");
            enq += chunk_;
        } else {
            const uint32_t c = (uint32_t)std::min<uint64_t>(left, chunk_);
            for (uint32_t j = 0; j < c; ++j)
                hip_check(launch_cg_iteration(P, A_.val_f32, A_.col16, work_), "CG iteration launch");
            enq += c;
        }
        hip_check(hipEventRecord(ev_[slot], work_), "event record");
        if (pending) {
            hip_check(hipEventSynchronize(ev_[slot ^ 1]), "event sync");
            if (*host_stop) break;
        }
        pending = true;
        slot ^= 1;
    }
    hip_check(hipMemcpyAsync(host_result_, state_.ptr, sizeof(CgState), hipMemcpyDeviceToHost, work_), "read state");
    hip_check(hipEventRecord(ev_join_, work_), "event record");
    hip_check(hipStreamWaitEvent(s, ev_join_, 0), "stream wait");
    hip_check(hipStreamSynchronize(work_), "JPCG solve");
    CgState hs;
    std::memcpy(&hs, host_result_, sizeof(CgState));
    return finish_result(hs.K, hs.rr, hs.dot, max_iter);
}

SolveResult Operator::solve_persistent(const double* diag, const double* b, double* x, uint32_t max_iter,
                                       double tol, void* records, hipStream_t s) {
    JpcgParams P;
    P.n = A_.n;
    P.ntiles = A_.ntiles;
    P.nnz = (int)A_.nnz;
    P.rowptr = A_.rowptr;
    P.col = col_stream();
    P.val = A_.val;
    P.dsplit = A_.dsplit;
    P.tile_row = A_.tile_row;
    P.tile_k = A_.tile_k;
    P.b = b;
    P.diag = diag;
    P.x = x;
    P.r = r_.ptr;
    P.p = p_.ptr;
    P.Ap = Ap_.ptr;
    P.J = J_.ptr;
    P.part = part_.ptr;
    P.bar = bar_.ptr;
    P.abortf = abort_.ptr;
    P.records = static_cast<Record*>(records);
    P.result = result_.ptr;
    P.tol = tol;
    P.max_rounds = max_iter + 1u;  // pcgImp.hpp:158 programs maxIter+1 rounds
    hip_check(hipMemsetAsync(bar_.ptr, 0, sizeof(unsigned long long), s), "reset barrier");
    hip_check(hipMemsetAsync(abort_.ptr, 0, sizeof(int), s), "reset abort flag");
    if (records) hip_check(hipMemsetAsync(records, 0, (size_t)(max_iter + 2) * 64, s), "clear records");
    hipError_t e = launch_jpcg(P, A_.val_f32, A_.col16, G_, s);
    if (e != hipSuccess) {
        // Some runtimes refuse the cooperative launch (grid above what they will
        // make co-resident, or no cooperative support on the queue).  Clear the
        // error, try half the grid once, then use the multi-kernel loop for good.
        (void)hipGetLastError();
        if (G_ > 1) {
            G_ /= 2;
            e = launch_jpcg(P, A_.val_f32, A_.col16, G_, s);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            coop_ok_ = false;
            last_path_ = kMultiKernel;
            return solve_multi(diag, b, x, max_iter, tol, records, s);
        }
    }
    hip_check(hipMemcpyAsync(host_result_, result_.ptr, 5 * sizeof(double), hipMemcpyDeviceToHost, s),
              "read solve result");
    hip_check(hipStreamSynchronize(s), "JPCG solve");
    if (host_result_[4] != 0.0) throw internal("grid barrier timed out; solve aborted.");
    return finish_result((int64_t)host_result_[0], host_result_[1], host_result_[2], max_iter);
}

}  // namespace xj
