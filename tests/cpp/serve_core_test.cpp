// CPU test of the serving loop's admission logic (open_source_einstar_amd/csrc/serve_core.h) with a
// fake device: random request sizes, random completion order; checks that admitted requests never
// overlap in the arena, never exceed the slots, are admitted in arrival order, and that everything
// submitted completes; that a slot whose drain failed is retired with its memory; that a request's
// streams are drained before anything is released. Exit status 0 = pass.
#include <cstdio>
#include <random>
#include <set>
#include <vector>

#include "serve_core.h"

using namespace sgm::serve;

int main() {
    int failures = 0;
    auto expect = [&](bool ok, const char* what) {
        std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
        if (!ok) ++failures;
    };

    {  // arena: first fit, split, coalesce
        Arena a(10 << 20, 1 << 20);
        auto x = a.alloc(3 << 20), y = a.alloc(1), z = a.alloc(4 << 20);
        expect(x && y && z && *x == 0 && *y == (3u << 20) && *z == (4u << 20), "first-fit offsets");
        expect(a.free_bytes() == (2u << 20) && !a.alloc(3 << 20), "no fit beyond capacity");
        a.release(*y);
        a.release(*x);
        expect(a.largest_free() == (4u << 20), "neighbours coalesce");
        a.release(*z);
        expect(a.largest_free() == (10u << 20) && a.live_blocks() == 0, "all free again");
        expect(!a.release(12345), "unknown block refused");
    }

    std::mt19937_64 rng(7);
    const int slots = 4;
    const size_t cap = 64ull << 20;
    AdmissionQueue q(slots, Arena(cap, 1 << 20));
    expect(!q.can_ever_fit(cap + 1) && q.can_ever_fit(cap), "requests larger than the arena are refused");
    struct Live {
        uint64_t ticket;
        int slot;
        size_t off, bytes;
    };
    std::vector<Live> live;
    std::vector<size_t> sizes;
    uint64_t next = 0, expected_next_admit = 0;
    int completed = 0;
    bool overlap = false, order = true, slot_ok = true;
    const int total = 5000;
    while (completed < total) {
        // Arrivals.
        const int arrivals = static_cast<int>(rng() % 3);
        for (int i = 0; i < arrivals && next < static_cast<uint64_t>(total); ++i) {
            const size_t b = (rng() % 4 == 0) ? (24ull << 20) + rng() % (16ull << 20) : (1ull << 20) + rng() % (6ull << 20);
            sizes.push_back(b);
            q.push(next++, b);
        }
        for (const auto& ad : q.admit()) {
            order &= ad.ticket == expected_next_admit++;
            for (const auto& l : live) {
                slot_ok &= l.slot != ad.slot;
                const size_t b = sizes[ad.ticket];
                overlap |= ad.offset < l.off + l.bytes && l.off < ad.offset + b;
            }
            live.push_back({ad.ticket, ad.slot, ad.offset, sizes[ad.ticket]});
        }
        slot_ok &= static_cast<int>(live.size()) <= slots && q.busy_slots() == static_cast<int>(live.size());
        // Completions in random order.
        if (!live.empty() && (rng() % 2 == 0 || next == static_cast<uint64_t>(total))) {
            const size_t k = rng() % live.size();
            q.finish(live[k].slot, live[k].off);
            live.erase(live.begin() + static_cast<long>(k));
            ++completed;
        }
    }
    {  // a retired slot keeps its arena range and is never admitted to again
        AdmissionQueue r(2, Arena(8 << 20, 1 << 20));
        r.push(1, 3 << 20);
        r.push(2, 3 << 20);
        auto a = r.admit();
        expect(a.size() == 2, "two slots admit two requests");
        r.retire(a[0].slot);
        r.finish(a[1].slot, a[1].offset);
        expect(r.live_slots() == 1 && r.busy_slots() == 0 && r.arena().free_bytes() == (5u << 20),
               "a retired slot's range stays allocated");
        r.push(3, 5 << 20);
        r.push(4, 1 << 20);
        auto b = r.admit();
        expect(b.size() == 1 && b[0].slot == a[1].slot && b[0].offset != a[0].offset, "only the live slot admits");
        r.retire(b[0].slot);
        expect(r.live_slots() == 0 && r.admit().empty(), "no live slot: nothing admitted");
        const auto dropped = r.drop_all();
        expect(dropped.size() == 1 && dropped[0] == 4 && r.queued() == 0, "waiting requests are dropped in order");
    }
    {  // finish_request drains first, always, and reclaims only after a good drain
        int drains = 0;
        auto ok = finish_request(true, [&] { ++drains; return true; });
        auto enq_fail = finish_request(false, [&] { ++drains; return true; });
        auto drain_fail = finish_request(true, [&] { ++drains; return false; });
        auto both = finish_request(false, [&] { ++drains; return false; });
        expect(drains == 4, "the streams are drained even when the enqueue failed");
        expect(ok.ok && ok.reclaim, "good request: valid, memory reused");
        expect(!enq_fail.ok && enq_fail.reclaim, "failed enqueue, drained: error status, memory reused");
        expect(!drain_fail.ok && !drain_fail.reclaim, "failed drain: error status, memory never reused");
        expect(!both.ok && !both.reclaim, "both failed: error, memory never reused");
    }
    expect(order, "admission in arrival order");
    expect(!overlap, "admitted blocks never overlap");
    expect(slot_ok, "at most one request per slot, slots bounded");
    expect(q.queued() == 0 && q.arena().free_bytes() == q.arena().capacity() && q.arena().live_blocks() == 0,
           "every request completed and returned its memory");
    std::printf("%s (%d failures)\n", failures ? "serve_core test FAILED" : "serve_core test ok", failures);
    return failures ? 1 : 0;
}
