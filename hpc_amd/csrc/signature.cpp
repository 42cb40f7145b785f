// Signature-format encoder and decoder (format described in signature.hpp).
//
// Encoder steps and the reference code whose output they reproduce byte for
// byte (L2/sparse/include/sw/fp64/signature.hpp):
//   order      entries by (row, column, input position)        matrix_params.hpp:219-246, 275-301 (stable sorts)
//   row blocks rows [r0, r0+4096) from the first row present     signature.hpp:92-133 gen_rbs
//   partitions columns [c0, c0+4096), c0 a present column // 4*4 signature.hpp:135-168 genPars4Rb
//   padding    whole 4-lane words per row, rows to 32 entries    signature.hpp:182-231 pad_par
//   channels   ~1/16 of a partition each, cut at row ends        signature.hpp:251-316 gen_chPars
//   block info per-channel row span and entry count              signature.hpp:318-358 update_rbParams
//   streams    row/column index words and value words            signature.hpp:360-421 gen_nnzStore
#include "signature.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <fstream>
#include <thread>

#include "ingest.hpp"
#include "operator.hpp"

namespace xj {
namespace sig {

namespace {

// ---------------------------------------------------------------- bytes

inline void put32(std::vector<uint8_t>& b, size_t off, uint32_t v) { std::memcpy(&b[off], &v, 4); }
inline void put16(std::vector<uint8_t>& b, size_t off, uint16_t v) { std::memcpy(&b[off], &v, 2); }
inline uint32_t get32(const std::vector<uint8_t>& b, size_t off) {
    uint32_t v;
    std::memcpy(&v, &b[off], 4);
    return v;
}
inline uint16_t get16(const std::vector<uint8_t>& b, size_t off) {
    uint16_t v;
    std::memcpy(&v, &b[off], 2);
    return v;
}
inline size_t add_word(std::vector<uint8_t>& b) {
    const size_t off = b.size();
    b.resize(off + kWordBytes, 0);
    return off;
}
inline uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

Error malformed(const std::string& m) { return invalid("malformed signature files: " + m); }

// Word offsets inside one row block's 8 words of rbParam.dat and one
// partition's 8 words of parParam.dat (matrix_params.hpp:355-400, 476-520).
constexpr size_t kRbBytes = 8 * kWordBytes;
constexpr size_t kParBytes = 8 * kWordBytes;
inline size_t rb_at(uint32_t r) { return kWordBytes + (size_t)r * kRbBytes; }
inline size_t par_at(uint32_t p) { return kWordBytes + (size_t)p * kParBytes; }

// Words a channel segment of `cnt` padded entries occupies in its stream.
inline uint64_t segment_words(uint32_t cnt) {
    return cnt / kLanes + div_up(cnt, kColWordSpan) + div_up(cnt, kRowWordSpan);
}

// ---------------------------------------------------------------- encoder

// Entries in (row, column, input position) order; src = value index.
struct Ordered {
    std::vector<uint32_t> row, col, src;
};

// Two stable counting passes (column, then row) over the validated COO.
void order_coo(Ordered& o, uint32_t m, uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col, int base) {
    if (base != 0 && base != 1)
        throw spm_not_supported("from getCooDatFromCscSym in gen_signature.hpp, unsupported storage mode.");
    if ((int64_t)nnz >= (int64_t)INT32_MAX) throw invalid("number of non-zeros must be below 2^31.");
    const uint32_t off = (uint32_t)base;
    for (uint32_t k = 0; k < nnz; ++k)
        if (row[k] - off >= m || col[k] - off >= n)
            throw invalid("matrix index out of range at entry " + std::to_string(k) + ".");
    std::vector<uint32_t> by_col(nnz);
    {
        std::vector<uint32_t> pos((size_t)n + 1, 0);
        for (uint32_t k = 0; k < nnz; ++k) pos[col[k] - off + 1]++;
        for (uint32_t j = 0; j < n; ++j) pos[j + 1] += pos[j];
        for (uint32_t k = 0; k < nnz; ++k) by_col[pos[col[k] - off]++] = k;
    }
    o.src.resize(nnz);
    {
        std::vector<uint32_t> pos((size_t)m + 1, 0);
        for (uint32_t k = 0; k < nnz; ++k) pos[row[k] - off + 1]++;
        for (uint32_t i = 0; i < m; ++i) pos[i + 1] += pos[i];
        for (uint32_t k : by_col) o.src[pos[row[k] - off]++] = k;
    }
    o.row.resize(nnz);
    o.col.resize(nnz);
    for (uint32_t p = 0; p < nnz; ++p) {
        o.row[p] = row[o.src[p]] - off;
        o.col[p] = col[o.src[p]] - off;
    }
}

// One padded partition and its split over the channels.
struct Part {
    std::vector<uint32_t> row, col, src;  // padded slots, rows ascending
    uint32_t base_blk = 0, col_blks = 0, rows = 0;
    std::array<uint32_t, kChannels> beg{}, end{}, min_row{}, nrows{}, ncols{}, blk_off{};
};

void push_slot(Part& P, uint32_t r, uint32_t c, uint32_t s) {
    P.row.push_back(r);
    P.col.push_back(c);
    P.src.push_back(s);
}

// Rows of the partition (members index into `o`, in order) laid out as whole
// 4-lane words; a lane whose column has no entry, and the tail that brings a
// row to a multiple of 32 slots, are padding.  A word starts at the column
// of the next unplaced entry rounded down to 4.
void pad_partition(const Ordered& o, const std::vector<uint32_t>& mem, Part& P) {
    size_t a = 0;
    while (a < mem.size()) {
        const uint32_t r = o.row[mem[a]];
        size_t b = a;
        while (b < mem.size() && o.row[mem[b]] == r) ++b;
        uint32_t lane = 0, base = 0, placed = 0;
        for (size_t k = a; k < b;) {
            const uint32_t c = o.col[mem[k]];
            if (lane == 0) base = c - c % kLanes;
            if (c == base + lane) {
                push_slot(P, r, c, o.src[mem[k]]);
                ++k;
            } else {
                push_slot(P, r, base + lane, kPadSlot);
            }
            ++placed;
            lane = (lane + 1) % kLanes;
        }
        for (; placed % kRowQuantum != 0; ++placed, lane = (lane + 1) % kLanes) push_slot(P, r, base + lane, kPadSlot);
        a = b;
    }
}

// Partition summary and the channel cut: channel c takes the next
// total/16 slots, extended to the end of the row it stops in; the last
// channel takes the rest.
void split_channels(Part& P) {
    const uint32_t tot = (uint32_t)P.row.size();
    uint32_t cmin = UINT32_MAX, cmax = 0;
    for (uint32_t c : P.col) {
        cmin = std::min(cmin, c);
        cmax = std::max(cmax, c);
    }
    P.base_blk = cmin / kLanes;
    P.col_blks = div_up(cmax - cmin + 1, kLanes);
    P.rows = P.row.back() - P.row.front() + 1;
    const uint32_t per = tot / kChannels;
    uint32_t s = 0;
    for (uint32_t ch = 0; ch < kChannels; ++ch) {
        uint32_t e = (ch == kChannels - 1 || s + per >= tot) ? tot : s + per;
        while (e > 0 && e < tot && P.row[e] == P.row[e - 1]) ++e;
        P.beg[ch] = s;
        P.end[ch] = e;
        if (e > s) {
            uint32_t lo = UINT32_MAX, hi = 0;
            for (uint32_t k = s; k < e; ++k) {
                lo = std::min(lo, P.col[k]);
                hi = std::max(hi, P.col[k]);
            }
            P.min_row[ch] = P.row[s];
            P.nrows[ch] = P.row[e - 1] - P.row[s] + 1;
            P.ncols[ch] = hi - lo + 1;
            P.blk_off[ch] = lo / kLanes - P.base_blk;
        } else {
            // an empty channel's column offset is 0 - base, kept as the 16-bit wrap
            P.min_row[ch] = 0;
            P.nrows[ch] = 0;
            P.ncols[ch] = 0;
            P.blk_off[ch] = 0u - P.base_blk;
        }
        s = e;
    }
}

void write_value_word(std::vector<uint8_t>& st, const double* val, const uint32_t* src) {
    const size_t off = add_word(st);
    for (uint32_t j = 0; j < kLanes; ++j) {
        const double v = src[j] == kPadSlot ? 0.0 : val[src[j]];
        std::memcpy(&st[off + 8 * j], &v, 8);
    }
}

// Channel ch's words for one partition: a row-index word every 512 slots
// (row of slot i + 32 k minus the block's channel start row, k < 16), a
// column word every 64 slots (4-column block of slot i + 4 k minus the
// channel's first block), one value word per 4 slots.  Index lanes past the
// segment's end keep the value of the previous index word of the segment.
void emit_segment(Image& im, uint32_t ch, const Part& P, uint32_t row0, uint32_t blk0, const double* val) {
    std::vector<uint8_t>& st = im.nnz_val[ch];
    std::vector<uint32_t>& slots = im.slot_src[ch];
    const uint32_t b = P.beg[ch], cnt = P.end[ch] - P.beg[ch];
    uint32_t ridx[kIdxPerWord] = {}, cidx[kIdxPerWord] = {};
    for (uint32_t i = 0; i < cnt; i += kLanes) {
        if (i % kRowWordSpan == 0) {
            for (uint32_t k = 0; k < kIdxPerWord; ++k)
                if (i + k * kRowQuantum < cnt) ridx[k] = P.row[b + i + k * kRowQuantum] - row0;
            const size_t off = add_word(st);
            for (uint32_t k = 0; k < kIdxPerWord; ++k) put16(st, off + 2 * k, (uint16_t)ridx[k]);
        }
        if (i % kColWordSpan == 0) {
            for (uint32_t k = 0; k < kIdxPerWord; ++k)
                if (i + k * kLanes < cnt) cidx[k] = P.col[b + i + k * kLanes] / kLanes - blk0;
            const size_t off = add_word(st);
            for (uint32_t k = 0; k < kIdxPerWord; ++k) put16(st, off + 2 * k, (uint16_t)cidx[k]);
        }
        write_value_word(st, val, &P.src[b + i]);
        slots.insert(slots.end(), &P.src[b + i], &P.src[b + i] + kLanes);
    }
}

void encode_ordered(Image& im, uint32_t m, uint32_t n, uint32_t nnz, const Ordered& o, const double* val,
                    uint64_t src_len) {
    if (nnz == 0) throw spm_invalid("the matrix has no entries to partition.");
    im = Image{};
    im.m = m;
    im.n = n;
    im.nnz = nnz;
    im.m_pad = m;
    im.n_pad = div_up(n, kLanes) * kLanes;
    im.src_len = src_len;
    im.has_slots = true;
    add_word(im.rb_param);
    add_word(im.par_param);
    for (auto& st : im.nnz_val) add_word(st);
    std::array<uint32_t, kChannels> row_words{}, col_words{}, val_words{};
    uint64_t nnz_pad = 0;
    uint32_t n_rb = 0, n_par = 0;

    size_t s = 0;
    while (s < nnz) {
        // row block [r0, r0 + 4096)
        const uint32_t r0 = o.row[s];
        const uint64_t rlim = (uint64_t)r0 + kBlockRows;
        const size_t e =
            std::partition_point(o.row.begin() + s, o.row.end(), [&](uint32_t r) { return r < rlim; }) - o.row.begin();
        std::vector<uint32_t> cols(o.col.begin() + s, o.col.begin() + e);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        std::vector<uint32_t> starts;
        for (size_t q = 0; q < cols.size();) {
            const uint32_t c0 = cols[q] - cols[q] % kLanes;
            starts.push_back(c0);
            const uint64_t clim = (uint64_t)c0 + kBlockCols;
            while (q < cols.size() && cols[q] < clim) ++q;
        }
        const uint32_t npars = (uint32_t)starts.size();
        std::vector<std::vector<uint32_t>> members(npars);
        for (size_t k = s; k < e; ++k) {
            const size_t p = std::upper_bound(starts.begin(), starts.end(), o.col[k]) - starts.begin() - 1;
            members[p].push_back((uint32_t)k);
        }
        std::vector<Part> parts(npars);
        for (uint32_t p = 0; p < npars; ++p) {
            pad_partition(o, members[p], parts[p]);
            split_channels(parts[p]);
        }

        // block-level channel spans: the start row only moves below r0 for a
        // partition channel starting there, so it stays r0
        std::array<uint32_t, kChannels> ch_first{}, ch_rows{}, ch_cnt{};
        uint32_t blk_cnt = 0;
        for (uint32_t ch = 0; ch < kChannels; ++ch) {
            uint32_t lo = r0, hi = r0, cnt = 0;
            for (const Part& P : parts) {
                const uint32_t c = P.end[ch] - P.beg[ch];
                if (lo > P.min_row[ch] && c != 0) lo = P.min_row[ch];
                hi = std::max(hi, P.min_row[ch] + P.nrows[ch]);
                cnt += c;
            }
            ch_first[ch] = lo - r0;
            ch_rows[ch] = hi - lo;
            ch_cnt[ch] = cnt;
            blk_cnt += cnt;
        }
        uint32_t cmin = UINT32_MAX, cmax = 0;
        for (size_t k = s; k < e; ++k) {
            cmin = std::min(cmin, o.col[k]);
            cmax = std::max(cmax, o.col[k]);
        }
        const size_t rb = add_word(im.rb_param);
        for (int w = 1; w < 8; ++w) add_word(im.rb_param);
        put32(im.rb_param, rb + 0, r0);
        put32(im.rb_param, rb + 4, cmin);
        put32(im.rb_param, rb + 8, cmax - cmin + 1);
        put32(im.rb_param, rb + 12, npars);
        put32(im.rb_param, rb + kWordBytes, o.row[e - 1] - r0 + 1);
        put32(im.rb_param, rb + kWordBytes + 4, blk_cnt);
        for (uint32_t ch = 0; ch < kChannels; ++ch) {
            put16(im.rb_param, rb + 2 * kWordBytes + 2 * ch, (uint16_t)ch_first[ch]);
            put16(im.rb_param, rb + 3 * kWordBytes + 2 * ch, (uint16_t)ch_rows[ch]);
            put32(im.rb_param, rb + 4 * kWordBytes + 4 * ch, ch_cnt[ch]);
        }

        for (const Part& P : parts) {
            const size_t pw = add_word(im.par_param);
            for (int w = 1; w < 8; ++w) add_word(im.par_param);
            for (uint32_t ch = 0; ch < kChannels; ++ch) {
                put32(im.par_param, pw + 4 * ch, P.ncols[ch]);
                put32(im.par_param, pw + 2 * kWordBytes + 4 * ch, P.end[ch] - P.beg[ch]);
                put16(im.par_param, pw + 5 * kWordBytes + 2 * ch, (uint16_t)P.blk_off[ch]);
                put16(im.par_param, pw + 6 * kWordBytes + 2 * ch, (uint16_t)(P.ncols[ch] / kLanes));
            }
            const uint32_t tot = (uint32_t)P.row.size();
            put32(im.par_param, pw + 4 * kWordBytes + 0, P.base_blk);
            put32(im.par_param, pw + 4 * kWordBytes + 4, P.col_blks);
            put32(im.par_param, pw + 4 * kWordBytes + 8, P.rows);
            put32(im.par_param, pw + 4 * kWordBytes + 12, tot);
            nnz_pad += tot;
            for (uint32_t ch = 0; ch < kChannels; ++ch) {
                const uint32_t cnt = P.end[ch] - P.beg[ch];
                const uint32_t row0 = (uint16_t)ch_first[ch] + r0;
                const uint32_t blk0 = (uint16_t)P.blk_off[ch] + P.base_blk;
                emit_segment(im, ch, P, row0, blk0, val);
                row_words[ch] += div_up(cnt, kRowWordSpan);
                col_words[ch] += div_up(cnt, kColWordSpan);
                val_words[ch] += cnt / kLanes;
            }
        }
        n_par += npars;
        ++n_rb;
        s = e;
    }
    if (nnz_pad > UINT32_MAX) throw spm_invalid("the padded matrix has more than 2^32 entries.");
    put32(im.rb_param, 0, m);
    put32(im.rb_param, 4, n_rb);
    put32(im.par_param, 0, n_par);
    for (uint32_t ch = 0; ch < kChannels; ++ch) {
        put32(im.nnz_val[ch], 0, row_words[ch] + col_words[ch] + val_words[ch]);
        put32(im.nnz_val[ch], 4, row_words[ch]);
        put32(im.nnz_val[ch], 8, col_words[ch]);
        put32(im.nnz_val[ch], 12, val_words[ch]);
    }
    im.nnz_pad = (uint32_t)nnz_pad;
}

// ---------------------------------------------------------------- decoder

// Checked view of the parameter files: per row block its first row and
// partition range, per partition its first 4-column block.
struct Layout {
    uint32_t n_rb = 0, n_par = 0;
    std::vector<uint32_t> rb_row0, rb_par0;  // rb_par0 has n_rb + 1 entries
    std::vector<uint32_t> par_blk;
};

void check_layout(const Image& im, Layout& L) {
    if (im.rb_param.size() < kWordBytes || im.rb_param.size() % kWordBytes)
        throw malformed("rbParam.dat has " + std::to_string(im.rb_param.size()) + " bytes.");
    if (im.par_param.size() < kWordBytes || im.par_param.size() % kWordBytes)
        throw malformed("parParam.dat has " + std::to_string(im.par_param.size()) + " bytes.");
    L.n_rb = get32(im.rb_param, 4);
    L.n_par = get32(im.par_param, 0);
    if (im.rb_param.size() != rb_at(L.n_rb))
        throw malformed("rbParam.dat describes " + std::to_string(L.n_rb) + " row blocks in " +
                        std::to_string(im.rb_param.size()) + " bytes.");
    if (im.par_param.size() != par_at(L.n_par))
        throw malformed("parParam.dat describes " + std::to_string(L.n_par) + " partitions in " +
                        std::to_string(im.par_param.size()) + " bytes.");
    if (get32(im.rb_param, 0) != im.m) throw malformed("rbParam.dat's row count differs from info.dat's.");
    L.rb_row0.resize(L.n_rb);
    L.rb_par0.assign((size_t)L.n_rb + 1, 0);
    uint64_t pars = 0;
    for (uint32_t r = 0; r < L.n_rb; ++r) {
        L.rb_row0[r] = get32(im.rb_param, rb_at(r));
        L.rb_par0[r] = (uint32_t)pars;
        pars += get32(im.rb_param, rb_at(r) + 12);
        if (pars > L.n_par) break;
    }
    if (pars != L.n_par) throw malformed("row blocks list a partition count different from parParam.dat's.");
    L.rb_par0[L.n_rb] = L.n_par;
    L.par_blk.resize(L.n_par);
    for (uint32_t p = 0; p < L.n_par; ++p) L.par_blk[p] = get32(im.par_param, par_at(p) + 4 * kWordBytes);
}

inline uint32_t seg_count(const Image& im, uint32_t p, uint32_t ch) {
    return get32(im.par_param, par_at(p) + 2 * kWordBytes + 4 * ch);
}
inline uint32_t seg_row0(const Image& im, const Layout& L, uint32_t r, uint32_t ch) {
    return L.rb_row0[r] + get16(im.rb_param, rb_at(r) + 2 * kWordBytes + 2 * ch);
}
inline uint32_t seg_blk0(const Image& im, const Layout& L, uint32_t p, uint32_t ch) {
    return L.par_blk[p] + get16(im.par_param, par_at(p) + 5 * kWordBytes + 2 * ch);
}

// Walk channel ch's stream: visit(row0, blk0, count, word offset of the
// segment) per partition, after checking the stream holds exactly the
// segments the parameter files describe and that its header agrees.
template <typename Visit>
void walk_channel(const Image& im, const Layout& L, uint32_t ch, Visit visit) {
    const std::vector<uint8_t>& st = im.nnz_val[ch];
    const std::string name = "nnzVal_" + std::to_string(ch) + ".dat";
    if (st.size() < kWordBytes || st.size() % kWordBytes)
        throw malformed(name + " has " + std::to_string(st.size()) + " bytes.");
    uint64_t pos = 1, rw = 0, cw = 0, vw = 0;
    const uint64_t words = st.size() / kWordBytes;
    for (uint32_t r = 0; r < L.n_rb; ++r) {
        for (uint32_t p = L.rb_par0[r]; p < L.rb_par0[r + 1]; ++p) {
            const uint32_t cnt = seg_count(im, p, ch);
            if (cnt % kLanes) throw malformed("partition " + std::to_string(p) + " gives channel " + std::to_string(ch) +
                                              " a count that is not a multiple of 4.");
            const uint64_t w = segment_words(cnt);
            if (pos + w > words) throw malformed(name + " is shorter than its partitions.");
            visit(seg_row0(im, L, r, ch), seg_blk0(im, L, p, ch), cnt, pos);
            pos += w;
            rw += div_up(cnt, kRowWordSpan);
            cw += div_up(cnt, kColWordSpan);
            vw += cnt / kLanes;
        }
    }
    if (pos != words) throw malformed(name + " is longer than its partitions.");
    if (get32(st, 0) != rw + cw + vw || get32(st, 4) != rw || get32(st, 8) != cw || get32(st, 12) != vw)
        throw malformed(name + "'s header word disagrees with its partitions.");
}

// Entries of one segment; emit(row, col, value) for each non-zero value.
template <typename Emit>
void read_segment(const std::vector<uint8_t>& st, uint64_t word, uint32_t row0, uint32_t blk0, uint32_t cnt,
                  Emit emit) {
    size_t at = (size_t)word * kWordBytes;
    size_t rword = 0, cword = 0;
    for (uint32_t i = 0; i < cnt; i += kLanes) {
        if (i % kRowWordSpan == 0) {
            rword = at;
            at += kWordBytes;
        }
        if (i % kColWordSpan == 0) {
            cword = at;
            at += kWordBytes;
        }
        const uint32_t r = row0 + get16(st, rword + 2 * ((i % kRowWordSpan) / kRowQuantum));
        const uint32_t c0 = (blk0 + get16(st, cword + 2 * ((i % kColWordSpan) / kLanes))) * kLanes;
        for (uint32_t j = 0; j < kLanes; ++j) {
            double v;
            std::memcpy(&v, &st[at + 8 * j], 8);
            if (v != 0.0) emit(r, c0 + j, v);
        }
        at += kWordBytes;
    }
}

// Run f(ch) for every channel, on threads when the streams are large; the
// first exception (lowest channel) is rethrown.
template <typename F>
void for_channels(const Image& im, F f) {
    size_t bytes = 0;
    for (const auto& st : im.nnz_val) bytes += st.size();
    if (bytes < ((size_t)8 << 20)) {
        for (uint32_t ch = 0; ch < kChannels; ++ch) f(ch);
        return;
    }
    std::array<std::exception_ptr, kChannels> err{};
    std::vector<std::thread> th;
    for (uint32_t ch = 0; ch < kChannels; ++ch)
        th.emplace_back([&, ch] {
            try {
                f(ch);
            } catch (...) {
                err[ch] = std::current_exception();
            }
        });
    for (auto& t : th) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// ---------------------------------------------------------------- files

void write_file(const std::string& path, const void* data, size_t bytes) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw invalid("cannot open " + path + " for writing.");
    f.write(static_cast<const char*>(data), (std::streamsize)bytes);
    f.close();
    if (!f) throw invalid("cannot write " + path + ".");
}

void read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw invalid("cannot open " + path + ".");
    const std::streamoff len = f.tellg();
    f.seekg(0);
    out.resize((size_t)len);
    if (len > 0) f.read(reinterpret_cast<char*>(out.data()), len);
    if (!f) throw invalid("cannot read " + path + ".");
}

std::string channel_file(const std::string& dir, uint32_t ch) {
    return dir + "/nnzVal_" + std::to_string(ch) + ".dat";
}

}  // namespace

// ---------------------------------------------------------------- public

void encode_coo(Image& im, uint32_t m, uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col,
                const double* val, int base) {
    Ordered o;
    order_coo(o, m, n, nnz, row, col, base);
    encode_ordered(im, m, n, nnz, o, val, nnz);
}

uint32_t encode_cscsym(Image& im, uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                       const double* val, int base) {
    Pattern P;
    pattern_from_cscsym(P, n, nnz_half, row_idx, col_ptr, base);
    Ordered o;
    o.row.resize((size_t)P.nnz);
    o.col.resize((size_t)P.nnz);
    o.src.resize((size_t)P.nnz);
    for (uint32_t i = 0; i < n; ++i)
        for (int32_t k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
            o.row[k] = i;
            o.col[k] = (uint32_t)P.col[k];
            o.src[k] = (uint32_t)P.perm[k];
        }
    encode_ordered(im, n, n, (uint32_t)P.nnz, o, val, nnz_half);
    return (uint32_t)P.nnz;
}

uint64_t source_length(const Image& im) { return im.has_slots ? im.src_len : 0; }

void update_values(Image& im, const double* val, uint64_t len) {
    if (!im.has_slots) throw invalid("the signature was loaded from files and has no value map to update.");
    if (len != im.src_len)
        throw spm_invalid("Invalid dimension and/or number of nnzs used for updateing matrix.");
    Layout L;
    check_layout(im, L);
    for_channels(im, [&](uint32_t ch) {
        std::vector<uint8_t>& st = im.nnz_val[ch];
        const uint32_t* src = im.slot_src[ch].data();
        walk_channel(im, L, ch, [&](uint32_t, uint32_t, uint32_t cnt, uint64_t word) {
            uint64_t w = word;
            for (uint32_t i = 0; i < cnt; i += kLanes) {
                if (i % kRowWordSpan == 0) ++w;
                if (i % kColWordSpan == 0) ++w;
                for (uint32_t j = 0; j < kLanes; ++j) {
                    const double v = src[j] == kPadSlot ? 0.0 : val[src[j]];
                    std::memcpy(&st[(size_t)w * kWordBytes + 8 * j], &v, 8);
                }
                src += kLanes;
                ++w;
            }
        });
    });
}

void store(const Image& im, const std::string& dir) {
    write_file(dir + "/rbParam.dat", im.rb_param.data(), im.rb_param.size());
    write_file(dir + "/parParam.dat", im.par_param.data(), im.par_param.size());
    for (uint32_t ch = 0; ch < kChannels; ++ch)
        write_file(channel_file(dir, ch), im.nnz_val[ch].data(), im.nnz_val[ch].size());
    const uint32_t info[6] = {im.m, im.n, im.nnz, im.m_pad, im.n_pad, im.nnz_pad};
    write_file(dir + "/info.dat", info, sizeof(info));
}

void load(Image& im, const std::string& dir) {
    im = Image{};
    std::vector<uint8_t> info;
    read_file(dir + "/info.dat", info);
    if (info.size() != 24) throw malformed("info.dat has " + std::to_string(info.size()) + " bytes, not 24.");
    uint32_t v[6];
    std::memcpy(v, info.data(), sizeof(v));
    im.m = v[0];
    im.n = v[1];
    im.nnz = v[2];
    im.m_pad = v[3];
    im.n_pad = v[4];
    im.nnz_pad = v[5];
    read_file(dir + "/rbParam.dat", im.rb_param);
    read_file(dir + "/parParam.dat", im.par_param);
    for (uint32_t ch = 0; ch < kChannels; ++ch) read_file(channel_file(dir, ch), im.nnz_val[ch]);
    Layout L;
    check_layout(im, L);
    for (uint32_t ch = 0; ch < kChannels; ++ch) walk_channel(im, L, ch, [](uint32_t, uint32_t, uint32_t, uint64_t) {});
}

void decode(const Image& im, Coo& out) {
    Layout L;
    check_layout(im, L);
    std::array<Coo, kChannels> part;
    for_channels(im, [&](uint32_t ch) {
        Coo& c = part[ch];
        const std::vector<uint8_t>& st = im.nnz_val[ch];
        walk_channel(im, L, ch, [&](uint32_t row0, uint32_t blk0, uint32_t cnt, uint64_t word) {
            read_segment(st, word, row0, blk0, cnt, [&](uint32_t r, uint32_t col, double v) {
                if (r >= im.m || col >= im.n)
                    throw malformed("channel " + std::to_string(ch) + " holds entry (" + std::to_string(r) + ", " +
                                    std::to_string(col) + ") outside the " + std::to_string(im.m) + " x " +
                                    std::to_string(im.n) + " matrix.");
                c.row.push_back(r);
                c.col.push_back(col);
                c.val.push_back(v);
            });
        });
    });
    size_t tot = 0;
    for (const Coo& c : part) tot += c.row.size();
    out.row.clear();
    out.col.clear();
    out.val.clear();
    out.row.reserve(tot);
    out.col.reserve(tot);
    out.val.reserve(tot);
    for (const Coo& c : part) {
        out.row.insert(out.row.end(), c.row.begin(), c.row.end());
        out.col.insert(out.col.end(), c.col.begin(), c.col.end());
        out.val.insert(out.val.end(), c.val.begin(), c.val.end());
    }
}

uint64_t count_entries(const Image& im) {
    Layout L;
    check_layout(im, L);
    std::array<uint64_t, kChannels> cnt{};
    for_channels(im, [&](uint32_t ch) {
        const std::vector<uint8_t>& st = im.nnz_val[ch];
        walk_channel(im, L, ch, [&](uint32_t row0, uint32_t blk0, uint32_t c, uint64_t word) {
            read_segment(st, word, row0, blk0, c, [&](uint32_t, uint32_t, double) { ++cnt[ch]; });
        });
    });
    uint64_t t = 0;
    for (uint64_t c : cnt) t += c;
    return t;
}

std::vector<Segment> segments(const Image& im, uint64_t* value_words) {
    Layout L;
    check_layout(im, L);
    std::vector<Segment> out;
    uint64_t base = 0, vw = 0;
    for (uint32_t ch = 0; ch < kChannels; ++ch) {
        walk_channel(im, L, ch, [&](uint32_t row0, uint32_t blk0, uint32_t cnt, uint64_t word) {
            if (cnt == 0) return;
            if (vw + cnt / kLanes > UINT32_MAX) throw malformed("more than 2^32 value words.");
            out.push_back(Segment{base + word, (uint32_t)vw, cnt, row0, blk0});
            vw += cnt / kLanes;
        });
        base += im.nnz_val[ch].size() / kWordBytes;
    }
    if (value_words) *value_words = vw;
    return out;
}

Summary summarize(const Image& im) {
    Layout L;
    check_layout(im, L);
    Summary s;
    s.row_blocks = L.n_rb;
    s.partitions = L.n_par;
    for (uint32_t ch = 0; ch < kChannels; ++ch) s.words[ch] = (uint32_t)(im.nnz_val[ch].size() / kWordBytes);
    return s;
}

}  // namespace sig
}  // namespace xj
