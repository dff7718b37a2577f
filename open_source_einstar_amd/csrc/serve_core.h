// Device-independent core of the serving loop: the HBM arena that admission is decided on, and the
// FIFO admission rule. Pure C++ so the logic is tested on the CPU with a fake device
// (tests/cpp/serve_core_test.cpp); sgm_server.hip drives it with HIP streams.
#pragma once

#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <optional>
#include <unordered_map>
#include <vector>

namespace sgm::serve {

// First-fit allocator over [0, capacity) in granule units; freed ranges coalesce with neighbours.
class Arena {
public:
    explicit Arena(size_t capacity = 0, size_t granule = size_t(1) << 21) : granule_(granule ? granule : 1) {
        capacity_ = capacity / granule_ * granule_;
        if (capacity_) free_[0] = capacity_;
        free_bytes_ = capacity_;
    }
    [[nodiscard]] size_t capacity() const { return capacity_; }
    [[nodiscard]] size_t free_bytes() const { return free_bytes_; }
    [[nodiscard]] size_t round(size_t bytes) const { return (bytes + granule_ - 1) / granule_ * granule_; }
    [[nodiscard]] size_t largest_free() const {
        size_t m = 0;
        for (const auto& [o, s] : free_) m = s > m ? s : m;
        return m;
    }
    // Offset of a block of at least `bytes`, or nullopt when no free range is large enough.
    std::optional<size_t> alloc(size_t bytes) {
        const size_t need = round(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const size_t off = it->first, rest = it->second - need;
            free_.erase(it);
            if (rest) free_[off + need] = rest;
            used_[off] = need;
            free_bytes_ -= need;
            return off;
        }
        return std::nullopt;
    }
    // Returns a block; false for an offset that is not allocated.
    bool release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return false;
        size_t start = off, size = u->second;
        used_.erase(u);
        free_bytes_ += size;
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                size += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = size;
        return true;
    }
    [[nodiscard]] size_t live_blocks() const { return used_.size(); }

private:
    size_t granule_, capacity_ = 0, free_bytes_ = 0;
    std::map<size_t, size_t> free_;            // offset -> size, disjoint, coalesced
    std::unordered_map<size_t, size_t> used_;  // offset -> size
};

// Requests wait in arrival order; the head is admitted when a slot is free and its footprint fits
// the arena. Nothing overtakes the head, so a large frame is never starved by smaller ones.
struct Pending {
    uint64_t ticket;
    size_t bytes;
};

struct Admission {
    uint64_t ticket;
    int slot;
    size_t offset;
};

class AdmissionQueue {
public:
    AdmissionQueue(int slots, Arena arena)
        : arena_(std::move(arena)), slot_busy_(static_cast<size_t>(slots), false), slot_retired_(static_cast<size_t>(slots), false) {}

    // A request that could never fit the arena is refused at submission.
    [[nodiscard]] bool can_ever_fit(size_t bytes) const { return arena_.round(bytes) <= arena_.capacity(); }
    void push(uint64_t ticket, size_t bytes) { q_.push_back({ticket, bytes}); }

    // Admits from the head while a slot and arena space are available.
    std::vector<Admission> admit() {
        std::vector<Admission> out;
        while (!q_.empty()) {
            int slot = -1;
            for (size_t i = 0; i < slot_busy_.size(); ++i)
                if (!slot_busy_[i]) {
                    slot = static_cast<int>(i);
                    break;
                }
            if (slot < 0) break;
            const auto off = arena_.alloc(q_.front().bytes);
            if (!off) break;
            slot_busy_[static_cast<size_t>(slot)] = true;
            out.push_back({q_.front().ticket, slot, *off});
            q_.pop_front();
        }
        return out;
    }

    // A finished request gives back its slot and memory.
    void finish(int slot, size_t offset) {
        slot_busy_[static_cast<size_t>(slot)] = false;
        arena_.release(offset);
    }

    // A request whose device work could not be confirmed finished: its slot is taken out of service
    // and its arena range is never handed out again (the device may still write it).
    void retire(int slot) {
        if (!slot_retired_[static_cast<size_t>(slot)]) ++retired_;
        slot_retired_[static_cast<size_t>(slot)] = true;
        slot_busy_[static_cast<size_t>(slot)] = true;
    }
    [[nodiscard]] int live_slots() const { return static_cast<int>(slot_busy_.size()) - retired_; }
    // Every waiting request, in arrival order, removed from the queue (when no slot is left to run them).
    std::vector<uint64_t> drop_all() {
        std::vector<uint64_t> out;
        for (const auto& p : q_) out.push_back(p.ticket);
        q_.clear();
        return out;
    }

    [[nodiscard]] size_t queued() const { return q_.size(); }
    [[nodiscard]] int busy_slots() const {
        int n = 0;
        for (size_t i = 0; i < slot_busy_.size(); ++i) n += (slot_busy_[i] && !slot_retired_[i]) ? 1 : 0;
        return n;
    }
    [[nodiscard]] const Arena& arena() const { return arena_; }

private:
    Arena arena_;
    std::vector<bool> slot_busy_;
    std::vector<bool> slot_retired_;
    int retired_ = 0;
    std::deque<Pending> q_;
};

// How a request ends once its work is enqueued (enqueue_ok false: some enqueue failed part-way).
// The streams it used are always drained first (drain() returns true when every wait succeeded),
// whatever the enqueue status, so no copy or kernel of the request can still touch its arena range
// or the caller's buffers afterwards. Its memory returns to the arena only when the drain succeeded.
struct Outcome {
    bool ok;       // the request's results are complete and valid
    bool reclaim;  // its slot and arena range may be reused
};
template <class Drain>
Outcome finish_request(bool enqueue_ok, Drain&& drain) {
    const bool drained = drain();
    return {enqueue_ok && drained, drained};
}

}  // namespace sgm::serve
