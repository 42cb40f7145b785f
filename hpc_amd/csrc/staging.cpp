// Pinned staging pool for large host-to-device uploads (see staging.hpp).
#include "staging.hpp"

#include <algorithm>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "operator.hpp"

namespace xj {
namespace {

constexpr int kSlotsPerWorker = 2;

// One pinned chunk and the event recorded after the DMA that last read it:
// a worker waits on it before it overwrites the chunk.
struct Slot {
    void* host = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct DevicePool {
    Slot slot[kStageWorkers * kSlotsPerWorker];
};

std::mutex g_mu;
std::map<int, DevicePool> g_pools;  // per device: events belong to the device they were created on

void ready(Slot& sl) {
    if (sl.pending) hip_check(hipEventSynchronize(sl.done), "staging chunk reuse");
    sl.pending = false;
    if (!sl.done) hip_check(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "staging event");
    if (!sl.host) {
        void* h = nullptr;
        hip_check(hipHostMalloc(&h, kStageChunkBytes, hipHostMallocPortable), "pinned staging chunk");
        sl.host = h;
    }
}

struct Chunk {
    char* dst;
    const char* src;
    size_t bytes;
};

}  // namespace

void staged_upload(const HostPiece* pieces, size_t count, hipStream_t s) {
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) total += pieces[i].bytes;
    if (total == 0) return;
    if (total < kStageMinBytes) {
        for (size_t i = 0; i < count; ++i)
            if (pieces[i].bytes)
                hip_check(hipMemcpyAsync(pieces[i].dst, pieces[i].src, pieces[i].bytes, hipMemcpyHostToDevice, s),
                          "upload");
        return;
    }
    std::vector<Chunk> chunks;
    chunks.reserve(total / kStageChunkBytes + count);
    for (size_t i = 0; i < count; ++i)
        for (size_t o = 0; o < pieces[i].bytes; o += kStageChunkBytes)
            chunks.push_back({static_cast<char*>(pieces[i].dst) + o, static_cast<const char*>(pieces[i].src) + o,
                              std::min(kStageChunkBytes, pieces[i].bytes - o)});

    int dev = 0;
    hip_check(hipGetDevice(&dev), "current device");
    std::lock_guard<std::mutex> lock(g_mu);
    DevicePool& pool = g_pools[dev];
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int W = (int)std::min<size_t>({(size_t)kStageWorkers, (size_t)hw, chunks.size()});

    // worker w takes chunks w, w+W, ... into its own slots in turn; the DMAs
    // of all workers go to `s`, so later work on `s` sees every chunk
    std::vector<std::exception_ptr> err(W);
    auto work = [&](int w) {
        try {
            hip_check(hipSetDevice(dev), "staging worker device");
            int j = 0;
            for (size_t c = (size_t)w; c < chunks.size(); c += (size_t)W, ++j) {
                Slot& sl = pool.slot[w * kSlotsPerWorker + j % kSlotsPerWorker];
                ready(sl);
                std::memcpy(sl.host, chunks[c].src, chunks[c].bytes);
                hip_check(hipMemcpyAsync(chunks[c].dst, sl.host, chunks[c].bytes, hipMemcpyHostToDevice, s),
                          "staged upload");
                hip_check(hipEventRecord(sl.done, s), "staging event");
                sl.pending = true;
            }
        } catch (...) {
            err[w] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    th.reserve(W > 0 ? W - 1 : 0);
    for (int w = 1; w < W; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    for (int w = 0; w < W; ++w)
        if (err[w]) std::rethrow_exception(err[w]);
}

}  // namespace xj
