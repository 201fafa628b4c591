// State of the serving loop: requests, the pixel arena that admission draws
// from, and the schedule of sample chunks in the device view ring.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pt_core.h"
#include "launch.hpp"

namespace pt {

// First-fit allocator of accumulation pixels: free ranges by offset.
class PixelArena {
public:
    void reset(uint32_t capacity) {
        free_.clear();
        if (capacity) free_[0] = capacity;
        capacity_ = capacity;
        used_ = 0;
    }
    bool alloc(uint32_t n, uint32_t& offset) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            offset = it->first;
            const uint32_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[offset + n] = rest;
            used_ += n;
            return true;
        }
        return false;
    }
    // Returns [offset, offset+n) and merges it with free neighbours.
    void release(uint32_t offset, uint32_t n) {
        used_ -= n;
        auto next = free_.lower_bound(offset);
        if (next != free_.end() && offset + n == next->first) {
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == offset) {
                prev->second += n;
                return;
            }
        }
        free_[offset] = n;
    }
    uint32_t capacity() const { return capacity_; }
    uint64_t used() const { return used_; }

private:
    std::map<uint32_t, uint32_t> free_;
    uint32_t capacity_ = 0;
    uint64_t used_ = 0;
};

struct ServerRequest {
    uint64_t id = 0;
    ViewRecord rec{};                // camera and settings; per-chunk fields set when scheduled
    uint32_t depth = 1;
    float* hdr = nullptr;
    uint32_t* rgba = nullptr;
    uint64_t total = 0;              // local samples
    uint64_t scheduled = 0;          // local samples handed to the device so far
    uint32_t pixels = 0;
    uint32_t accum_offset = 0;
    uint64_t first_virtual = 0;      // view-ring index of its first and last chunk
    uint64_t last_virtual = 0;
    int64_t launched_at = -1;        // iterations run when every sample had been launched
    uint32_t invalid_slot = 0xFFFFFFFFu; // counter of its non-finite or negative pixels (postprocess to readback)
    uint32_t state = PT_REQUEST_QUEUED;
    std::string message;             // why it failed
    double t_submit = 0, t_admit = 0, t_done = 0;
};

struct ServerChunk {
    uint64_t virt;                   // view-ring index
    uint64_t end;                    // global sample end
    std::shared_ptr<ServerRequest> req;
};

struct ServerState {
    pt_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    std::thread thread, readback;
    hipStream_t rb_stream = nullptr;

    std::mutex mu;
    std::condition_variable cv_work;     // server thread: new requests, freed pixels, stop
    std::condition_variable cv_done;     // waiters
    std::condition_variable cv_rb;       // readback thread
    bool stop = false;                   // drain and exit
    bool loop_exited = false;
    std::string error;                   // first failure of the server thread

    std::deque<std::shared_ptr<ServerRequest>> pending;
    std::vector<std::shared_ptr<ServerRequest>> active;     // admitted, not yet postprocessed
    std::unordered_map<uint64_t, std::shared_ptr<ServerRequest>> all;
    std::deque<std::pair<std::shared_ptr<ServerRequest>, hipEvent_t>> rb_queue;
    std::vector<hipEvent_t> free_events;
    std::vector<uint32_t> free_slots;    // invalid-pixel counters not held by a request
    PixelArena arena;
    uint64_t next_id = 1;

    // schedule (server thread only)
    std::deque<ServerChunk> chunks;      // appended chunks not yet seen fully launched
    uint64_t total_samples = 0;          // global end of scheduled samples
    uint64_t view_hi = 0;                // next view-ring index
    uint64_t launched_seen = 0;          // from the newest counter snapshot
    int64_t iterations_seen = 0;
    uint64_t graph_launches = 0;
    uint64_t ext_rays = 0, shadow_rays = 0;
    bool header_dirty = false;

    // statistics (under mu)
    uint64_t submitted = 0, done = 0, failed = 0, samples_done = 0;
    double busy_ms = 0, busy_since = -1;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

}  // namespace pt
