// Signature streams decoded on the GPU (the host only builds the per-segment
// table): one thread per value word finds its segment (one binary search on
// the value-word prefix per block, then a short walk per thread), locates its
// row-index and column words from the fixed interleave (a row word per 128
// value words, a column word per 16), and turns its four lanes into (key = row*n + col, position, value) triples for
// the stable radix sort that builds the CSR.  Zero lanes (padding, and stored
// zeros, which the format cannot tell apart) are dropped via count + scan.
//
// Word positions inside a segment, for value word q (entries 4q..4q+3):
//   value word      q + (q/16 + 1) + (q/128 + 1)
//   its row word    137 * (q/128)                  (128 value + 8 column + 1 row words per 512 entries)
//   its column word 17 * (q/16) + q/128 + 1
// Row of the word: row0 + row word lane (q % 128) / 8; first column:
// 4 * (blk0 + column word lane q % 16).
#include "kernels.hpp"
#include "signature.hpp"

namespace xj {
namespace sig {

namespace {

__device__ __forceinline__ uint32_t find_segment(const Segment* __restrict__ seg, uint32_t nseg, uint32_t w) {
    uint32_t lo = 0, hi = nseg;  // last segment with vw0 <= w
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg[mid].vw0 <= w) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Segment of value word w (w >= w0), walking forward from s0 = segment of w0:
// the block searches once for its first word, and segments are long (a
// padded row is 8 value words at least), so a thread takes a step or two.
__device__ __forceinline__ uint32_t walk_segment(const Segment* __restrict__ seg, uint32_t nseg, uint32_t s0,
                                                 uint32_t w) {
    uint32_t s = s0;
    while (s + 1 < nseg && seg[s + 1].vw0 <= w) ++s;
    return s;
}

struct WordView {
    uint32_t row, col0;
    double v[4];
};

__device__ __forceinline__ WordView view_word(const Segment& S, const uint8_t* __restrict__ words, uint32_t w) {
    const uint32_t q = w - S.vw0;
    const uint64_t vpos = S.word0 + q + (q >> 4) + 1 + (q >> 7) + 1;
    const uint64_t rpos = S.word0 + 137ull * (q >> 7);
    const uint64_t cpos = S.word0 + 17ull * (q >> 4) + (q >> 7) + 1;
    const uint16_t* rw = reinterpret_cast<const uint16_t*>(words + rpos * kWordBytes);
    const uint16_t* cw = reinterpret_cast<const uint16_t*>(words + cpos * kWordBytes);
    const double2* vw = reinterpret_cast<const double2*>(words + vpos * kWordBytes);
    WordView out;
    out.row = S.row0 + rw[(q & 127) >> 3];
    out.col0 = (S.blk0 + cw[q & 15]) * kLanes;
    const double2 a = vw[0], b = vw[1];
    out.v[0] = a.x;
    out.v[1] = a.y;
    out.v[2] = b.x;
    out.v[3] = b.y;
    return out;
}

__global__ __launch_bounds__(kBlock) void sig_count(uint64_t vw, const Segment* __restrict__ seg, uint32_t nseg,
                                                    const uint8_t* __restrict__ words, uint32_t* __restrict__ cnt) {
    for (uint64_t w0 = (uint64_t)blockIdx.x * kBlock; w0 < vw; w0 += (uint64_t)gridDim.x * kBlock) {
        const uint32_t s0 = find_segment(seg, nseg, (uint32_t)w0);
        const uint64_t w = w0 + threadIdx.x;
        if (w >= vw) continue;
        const Segment S = seg[walk_segment(seg, nseg, s0, (uint32_t)w)];
        const uint32_t q = (uint32_t)w - S.vw0;
        const uint64_t vpos = S.word0 + q + (q >> 4) + 1 + (q >> 7) + 1;
        const double2* p = reinterpret_cast<const double2*>(words + vpos * kWordBytes);
        const double2 a = p[0], b = p[1];
        cnt[w] = (a.x != 0.0) + (a.y != 0.0) + (b.x != 0.0) + (b.y != 0.0);
    }
}

__global__ __launch_bounds__(kBlock) void sig_emit(uint64_t vw, const Segment* __restrict__ seg, uint32_t nseg,
                                                   const uint8_t* __restrict__ words, const uint32_t* __restrict__ pos,
                                                   uint32_t m, uint32_t n, uint64_t* __restrict__ key,
                                                   int32_t* __restrict__ idx, double* __restrict__ val,
                                                   int32_t* __restrict__ bad) {
    for (uint64_t w0 = (uint64_t)blockIdx.x * kBlock; w0 < vw; w0 += (uint64_t)gridDim.x * kBlock) {
        const uint32_t s0 = find_segment(seg, nseg, (uint32_t)w0);
        const uint64_t w = w0 + threadIdx.x;
        if (w >= vw) continue;
        const WordView V = view_word(seg[walk_segment(seg, nseg, s0, (uint32_t)w)], words, (uint32_t)w);
        uint32_t p = pos[w];
        for (int j = 0; j < 4; ++j) {
            if (V.v[j] == 0.0) continue;
            const uint32_t c = V.col0 + j;
            const bool ok = V.row < m && c < n;
            if (!ok) atomicOr(bad, 1);
            key[p] = ok ? (uint64_t)V.row * n + c : 0;
            idx[p] = (int32_t)p;
            val[p] = V.v[j];
            ++p;
        }
    }
}

int grid_for_words(uint64_t vw) {
    uint64_t g = (vw + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

}  // namespace

hipError_t decode_count(uint64_t vw, const Segment* seg, uint32_t nseg, const uint8_t* words, uint32_t* cnt,
                        hipStream_t s) {
    if (vw == 0 || nseg == 0) return hipSuccess;
    sig_count<<<grid_for_words(vw), kBlock, 0, s>>>(vw, seg, nseg, words, cnt);
    return hipGetLastError();
}

hipError_t decode_emit(uint64_t vw, const Segment* seg, uint32_t nseg, const uint8_t* words, const uint32_t* pos,
                       uint32_t m, uint32_t n, uint64_t* key, int32_t* idx, double* val, int32_t* bad, hipStream_t s) {
    if (vw == 0 || nseg == 0) return hipSuccess;
    sig_emit<<<grid_for_words(vw), kBlock, 0, s>>>(vw, seg, nseg, words, pos, m, n, key, idx, val, bad);
    return hipGetLastError();
}

}  // namespace sig
}  // namespace xj
