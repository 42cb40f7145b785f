// Host-to-device uploads through pinned staging chunks.
//
// A copy from pageable host memory runs at about half the rate of one from
// pinned memory (30 against 58 GB/s on an MI355X host link).  Large uploads
// (matrix streams, COO indices and values, signature streams) therefore go
// through a process-wide pool of pinned chunks: worker threads copy the
// source into their chunks and queue each chunk's DMA on the caller's stream,
// two chunks per worker, so host copies and DMAs overlap.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace xj {

struct HostPiece {
    void* dst;        // device address
    const void* src;  // host address (pageable or pinned)
    size_t bytes;
};

// Copies every piece to the device in the order of `s`: work queued on `s`
// afterwards sees the data.  Returns once the sources may be reused (their
// bytes are in pinned chunks or already on the device); the DMAs may still be
// in flight.  Uploads below kStageMinBytes in total are plain
// hipMemcpyAsync calls.  Calls are serialised on one lock; the pool (up to
// kStageWorkers x 2 chunks of kStageChunkBytes per device, allocated as they
// are first needed) is kept until the process ends.  Throws like hip_check.
void staged_upload(const HostPiece* pieces, size_t count, hipStream_t s);
inline void staged_upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
    const HostPiece p{dst, src, bytes};
    staged_upload(&p, 1, s);
}

constexpr size_t kStageChunkBytes = size_t(8) << 20;
constexpr int kStageWorkers = 8;
constexpr size_t kStageMinBytes = size_t(4) << 20;

}  // namespace xj
