// The partitioned matrix format of the reference's SpMV engine ("signature"
// files: info.dat, rbParam.dat, parParam.dat, nnzVal_<c>.dat), read and
// written so that matrices prepared for the reference load here and matrices
// built here can be handed to the reference's tools.
//
// Format (reference L2/sparse/include/sw/fp64/matrix_params.hpp:329-688,
// signature.hpp:92-421, src/sw/c++/gen_signature.cpp:58-81), for the PCG
// build's geometry (pcg/sw/include/impl/pcgImp.hpp: 4 values per word, 8-word
// accumulation, 16 channels, 4096-row blocks, 4096-column partitions,
// 256-bit words):
//   * row blocks: rows sorted, greedily cut into spans of at most 4096 rows
//     starting at the first row present;
//   * partitions: inside a row block, columns cut into spans of at most 4096
//     starting at the first present column rounded down to a multiple of 4;
//   * each row of a partition becomes whole 4-lane words (lane k holds column
//     base+k, missing lanes are padding), padded to a multiple of 32 entries;
//   * each padded partition is split over 16 channels at row boundaries,
//     about 1/16 of its entries each;
//   * channel c's stream: a header word, then per partition (row block order)
//     a row-index word every 512 entries (16 rows, one per 32 entries), a
//     column word every 64 entries (16 4-column blocks, one per word) and one
//     value word (4 doubles, padding = 0) per 4 entries.
// The encoder keeps, per padded slot, the index of its value in the caller's
// array, so new values with the same pattern rewrite only the value words
// (the reference's update_sig, signature.hpp:427-469).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace xj {
namespace sig {

constexpr uint32_t kLanes = 4;            // values per word (parEntries)
constexpr uint32_t kRowQuantum = 32;      // row padding quantum (parEntries x accLatency)
constexpr uint32_t kChannels = 16;        // HBM channels of the reference build
constexpr uint32_t kBlockRows = 4096;     // rows per row block (maxRows)
constexpr uint32_t kBlockCols = 4096;     // columns per partition (maxCols)
constexpr uint32_t kWordBytes = 32;       // 256-bit word
constexpr uint32_t kIdxPerWord = kWordBytes / 2;              // uint16 indices per word
constexpr uint32_t kRowWordSpan = kIdxPerWord * kRowQuantum;  // entries per row-index word (512)
constexpr uint32_t kColWordSpan = kIdxPerWord * kLanes;       // entries per column word (64)
constexpr uint32_t kPadSlot = 0xffffffffu;                    // slot with no caller value

struct Image {
    // info.dat: m, n, nnz as given, m, n rounded up to 4, padded entry count
    uint32_t m = 0, n = 0, nnz = 0, m_pad = 0, n_pad = 0, nnz_pad = 0;
    std::vector<uint8_t> rb_param, par_param;
    std::array<std::vector<uint8_t>, kChannels> nnz_val;
    // encoder only: per channel, per padded slot in stream order, the index of
    // its value in the caller's array or kPadSlot
    std::array<std::vector<uint32_t>, kChannels> slot_src;
    bool has_slots = false;
    uint64_t src_len = 0;  // length of the caller's value array
};

// Encode a COO matrix (indices with base 0 or 1, duplicates kept as separate
// entries) with m rows and n columns.  val[k] is the value of entry k.
// Throws xj::Error (INTERNAL_ERROR, like the reference partitioner's
// SpmException) on bad input.
void encode_coo(Image& im, uint32_t m, uint32_t n, uint32_t nnz, const uint32_t* row, const uint32_t* col,
                const double* val, int base);
// Encode a half-stored symmetric CSC (diagonal included), expanded to the
// full matrix in the reference's order (gen_signature.hpp:88-112,
// matrix_params.hpp:82-134).  Returns the expanded entry count.
uint32_t encode_cscsym(Image& im, uint32_t n, uint32_t nnz_half, const int64_t* row_idx, const int64_t* col_ptr,
                       const double* val, int base);
// New values for the encoded pattern: val has the length of the encoding
// call's value array (nnz for COO, nnz_half for CSC-sym).
void update_values(Image& im, const double* val, uint64_t len);
// Length of the value array update_values expects.
uint64_t source_length(const Image& im);

// Files: dir/info.dat, dir/rbParam.dat, dir/parParam.dat, dir/nnzVal_<c>.dat.
void store(const Image& im, const std::string& dir);
void load(Image& im, const std::string& dir);

// Decoded entries in stream order (channel, then partition, then slot).
// Zero values are dropped: the format does not tell stored zeros from
// padding (the reference's decoder, tests/fp64/spmv/python/signature.py
// create_spm, drops them the same way).  Checks the structure and every
// index against m and n.
struct Coo {
    std::vector<uint32_t> row, col;
    std::vector<double> val;
};
void decode(const Image& im, Coo& out);
// Entry count decode() returns, without building it.
uint64_t count_entries(const Image& im);

// Non-empty channel segments with the 16 streams laid end to end (channel 0
// first, each with its header word): first word, index of its first value
// word among all value words, padded entry count, start row, start 4-column
// block.  Checks the structure as load() does; *value_words = total.
struct Segment {
    uint64_t word0;
    uint32_t vw0, cnt, row0, blk0;
};
std::vector<Segment> segments(const Image& im, uint64_t* value_words);

// GPU decode (sig_decode.hip) over the streams in device memory `words`
// (32-byte words) and the segment table `seg` (nseg entries, vw0 ascending):
//   count: cnt[w] = non-zero values in value word w (w < vw)
//   emit:  from pos = exclusive scan of cnt, key = row*n + col, idx = its
//          position, val = value; *bad |= 1 for an entry outside m x n
hipError_t decode_count(uint64_t vw, const Segment* seg, uint32_t nseg, const uint8_t* words, uint32_t* cnt,
                        hipStream_t s);
hipError_t decode_emit(uint64_t vw, const Segment* seg, uint32_t nseg, const uint8_t* words, const uint32_t* pos,
                       uint32_t m, uint32_t n, uint64_t* key, int32_t* idx, double* val, int32_t* bad, hipStream_t s);

// Partition-level summary: row blocks, partitions, total words per channel.
struct Summary {
    uint32_t row_blocks = 0, partitions = 0;
    std::array<uint32_t, kChannels> words{};
};
Summary summarize(const Image& im);

}  // namespace sig
}  // namespace xj
