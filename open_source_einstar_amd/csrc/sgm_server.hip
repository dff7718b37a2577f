// Serving loop of the gfx950 SGM library (sgm_server_* in include/sgm_hip.h).
//
// Requests (host image pairs of any size up to the configured maximum, each with its own disparity
// range) are admitted in arrival order when a slot (a worker thread and its copy stream) is free and
// the request's device footprint fits the HBM arena reserved at start (serve_core.h). Requests that
// arrive while others run are admitted as soon as slots and memory free up, so several frames are on
// the device at once. A request's copies run on its slot's stream; its kernels run on a compute
// stream, handed over by events (upload done -> kernels -> results copied back). With shared compute
// streams (the default: one), requests' kernels run one request after another in admission order
// while other requests' uploads and result copies overlap them on the copy engines; with one compute
// stream per slot, the requests' kernels share the GPU. Each request runs the same kernels as
// sgm_compute on its own scratch inside the arena, so its result does not depend on what else is in
// flight.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "serve_core.h"
#include "sgm_hip.h"
#include "sgm_internal.h"

using namespace sgm;

namespace sgm {
// Shared with sgm_api.hip.
int validate_params_ext(const sgm_params* p, std::string& err);
int check_row_lds_ext(int w, std::string& err);
SgmConsts make_consts_ext(const sgm_params& p);
hipError_t enqueue_pair_ext(const PairDesc& d, const SgmConsts& c, bool padded, hipStream_t s, hipEvent_t* ev5,
                            int path_group);
void set_geometry_ext(PairDesc& d, int w, int h, int min_d, int nd, int ndirs, int es);
size_t census_words_ext(int w, int h, int min_d, int nd);
void set_thread_error(const std::string& e);
}  // namespace sgm

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Byte offsets of one request's buffers inside its arena block. raw holds the two raw frames of a
// request that is rectified on the device (raw_px pixels each).
struct Footprint {
    size_t raw, img, census, cvol, lvol, disp, median, conf, dint, status, rdisp, best, second, total;
};

// wide: u16 path costs and step penalties (SgmConsts::wide_costs).
Footprint footprint(int w, int h, int min_d, int nd, int ndirs, size_t raw_px, bool wide) {
    const size_t npx = (size_t)w * h;
    const size_t dp = (size_t)dpl_for(nd) * kGroup;
    Footprint f{};
    size_t o = 0;
    f.raw = o;
    o = align_up(o + 2 * raw_px);
    f.img = o;
    o = align_up(o + 2 * npx);
    f.census = o;
    o = align_up(o + 2 * census_words_ext(w, h, min_d, nd) * 8 + npx * edge_bytes_per_px(wide));  // census words, then penalties
    f.cvol = o;  // matching costs as bytes: the path kernel's cost-volume mode, where the volume fits a view
    if (npx * dp <= kMaxViewBytes) o = align_up(o + npx * dp);
    f.lvol = o;
    o = align_up(o + (size_t)ndirs * npx * dp * (wide ? 2 : 1));
    f.disp = o;
    o = align_up(o + npx * 4);
    f.median = o;
    o = align_up(o + npx * 4);
    f.conf = o;
    o = align_up(o + npx * 4);
    f.dint = o;
    o = align_up(o + npx * 4);
    f.status = o;
    o = align_up(o + npx);
    f.rdisp = o;
    o = align_up(o + npx * 4);
    f.best = o;
    o = align_up(o + npx * 2);
    f.second = o;
    o = align_up(o + npx * 2);
    f.total = o;
    return f;
}

// A resident rectification map set: both cameras' (x, y) maps interleaved, left then right.
struct MapSet {
    float* d_maps = nullptr;
    int map_w = 0, map_h = 0, raw_w = 0, raw_h = 0;
};

struct Request {
    uint64_t ticket = 0, tag = 0;
    int w = 0, h = 0, min_d = 0, nd = 0;
    MapSet rect;                 // d_maps == null: the images are already rectified
    uint8_t* pinned = nullptr;   // left rows then right rows (raw frames when rectified) in a pooled
    size_t pinned_class = 0;     // pinned buffer of this size class, or
    std::vector<uint8_t> img;    // the same in heap memory when the pool was empty, or
    const uint8_t* borrowed[2] = {nullptr, nullptr};  // the caller's own buffers, read in place
    int64_t borrowed_stride = 0;
    sgm_outputs out{};         // caller's host buffers
    int64_t t_submit = 0, t_admit = 0;
    size_t bytes = 0, offset = 0;
};

}  // namespace

struct sgm_server {
    sgm_params p{};
    SgmConsts consts{};
    int ndirs = 4, max_w = 0, max_h = 0, max_nd = 0;
    uint8_t* d_arena = nullptr;
    std::mutex mu;
    std::condition_variable cv_done;
    serve::AdmissionQueue aq{1, serve::Arena(0)};
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<sgm_completion> done;
    uint64_t next_ticket = 1;
    int64_t completed = 0, failed = 0;
    bool stopping = false;
    std::string err;

    // Pinned input buffers by size class (bytes): 2 * max_w * max_h at start, plus one class per
    // raw frame size that a map set brings in.
    std::map<size_t, std::vector<uint8_t*>> pinned_free;
    std::vector<uint8_t*> pinned_all;
    std::vector<MapSet> map_sets;  // id k is map_sets[k - 1]

    // Takes a pooled pinned buffer of at least `bytes` (null when every such buffer is in use).
    uint8_t* take_pinned(size_t bytes, size_t& cls) {  // with mu held
        for (auto it = pinned_free.lower_bound(bytes); it != pinned_free.end(); ++it)
            if (!it->second.empty()) {
                uint8_t* b = it->second.back();
                it->second.pop_back();
                cls = it->first;
                return b;
            }
        return nullptr;
    }
    hipError_t add_pinned_class(size_t bytes, int count) {  // with mu held
        auto& v = pinned_free[bytes];
        for (int i = 0; i < count; ++i) {
            uint8_t* b = nullptr;
            const hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&b), bytes, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            pinned_all.push_back(b);
            v.push_back(b);
        }
        return hipSuccess;
    }

    struct Slot {
        hipStream_t stream = nullptr;  // the request's copies
        hipEvent_t ev[6] = {};  // census .. median as sgm_compute, [5] = rectification start
        hipEvent_t ev_up = nullptr, ev_done = nullptr;  // inputs on the device; kernels finished
        std::thread th;
        std::condition_variable cv;
        bool has_job = false;
        uint64_t ticket = 0;
    };
    std::vector<Slot> slots;
    // Shared compute streams, each with a lock that keeps one request's launches together.
    struct Compute {
        hipStream_t stream = nullptr;
        std::mutex mu;
    };
    std::vector<std::unique_ptr<Compute>> compute;
    bool broken = false;  // a stream could not be drained: nothing new is admitted

    // Lanes per path scanline by request shape (width, height, disparities): fixed by the config, or
    // chosen by timing the path kernel at each width the lane layout allows on the first requests of
    // the shape (two each, the fastest minimum wins). Every width gives the same results. 32 lanes
    // need a multiple of 8 disparities per lane of the 16-lane layout, 64 lanes a multiple of 16:
    // otherwise each lane would move its path costs in 2-byte pieces.
    int fixed_group = 0;
    static constexpr int kWidths[3] = {16, 32, 64};
    struct LaneChoice {
        float best_ms[3] = {0, 0, 0};
        int samples[3] = {0, 0, 0};
        int allowed = 1;  // widths kWidths[0 .. allowed - 1] are timed
        int chosen = 0;
    };
    std::map<std::tuple<int, int, int>, LaneChoice> lanes;
    int pick_group(int w, int h, int nd) {  // with mu held
        if (fixed_group) return fixed_group;
        LaneChoice& c = lanes[{w, h, nd}];
        if (c.chosen) return c.chosen;
        const int dpl = dpl_for_pair(nd, w, h);
        c.allowed = dpl % 16 == 0 ? 3 : (dpl % 8 == 0 ? 2 : 1);
        if (c.allowed == 1) return c.chosen = 16;
        int i = 0;  // the least sampled width
        for (int k = 1; k < c.allowed; ++k)
            if (c.samples[k] < c.samples[i]) i = k;
        return kWidths[i];
    }
    void record_group(int w, int h, int nd, int group, float paths_ms) {  // with mu held
        LaneChoice& c = lanes[{w, h, nd}];
        if (c.chosen || paths_ms <= 0) return;
        const int i = group == 64 ? 2 : (group == 32 ? 1 : 0);
        if (i >= c.allowed) return;
        c.best_ms[i] = c.samples[i] ? std::min(c.best_ms[i], paths_ms) : paths_ms;
        ++c.samples[i];
        for (int k = 0; k < c.allowed; ++k)
            if (c.samples[k] < 2) return;
        int best = 0;
        for (int k = 1; k < c.allowed; ++k)
            if (c.best_ms[k] < c.best_ms[best]) best = k;
        c.chosen = kWidths[best];
    }

    void dispatch(const std::vector<serve::Admission>& adm) {  // with mu held
        const int64_t t = now_ns();
        for (const auto& a : adm) {
            Request& r = reqs[a.ticket];
            r.offset = a.offset;
            r.t_admit = t;
            Slot& s = slots[(size_t)a.slot];
            s.ticket = a.ticket;
            s.has_job = true;
            s.cv.notify_one();
        }
    }

    void worker(int si);
};

void sgm_server::worker(int si) {
    Slot& s = slots[(size_t)si];
    for (;;) {
        Request r;
        {
            std::unique_lock lk(mu);
            s.cv.wait(lk, [&] { return s.has_job || stopping; });
            if (!s.has_job) return;
            r = std::move(reqs.at(s.ticket));  // the worker owns the request until it completes
        }
        const size_t raw_px = r.rect.d_maps ? (size_t)r.rect.raw_w * r.rect.raw_h : 0;
        const Footprint f = footprint(r.w, r.h, r.min_d, r.nd, ndirs, raw_px, consts.wide_costs != 0);
        uint8_t* base = d_arena + r.offset;
        const size_t npx = (size_t)r.w * r.h;
        PairDesc d{};
        d.left = base + f.img;
        d.right = base + f.img + npx;
        d.pitch = r.w;
        d.census_l = reinterpret_cast<uint64_t*>(base + f.census);
        d.census_r = d.census_l + census_words_ext(r.w, r.h, r.min_d, r.nd);
        d.edge = reinterpret_cast<uint32_t*>(d.census_r + census_words_ext(r.w, r.h, r.min_d, r.nd));
        d.lvol = base + f.lvol;
        d.cvol = npx * (size_t)dpl_for(r.nd) * kGroup <= kMaxViewBytes ? base + f.cvol : nullptr;
        set_geometry_ext(d, r.w, r.h, r.min_d, r.nd, ndirs, consts.wide_costs ? 2 : 1);
        d.disp = reinterpret_cast<float*>(base + f.disp);
        d.median = r.out.disparity_median ? reinterpret_cast<float*>(base + f.median) : nullptr;
        d.conf = r.out.confidence ? reinterpret_cast<float*>(base + f.conf) : nullptr;
        d.dint = r.out.disparity_int ? reinterpret_cast<int32_t*>(base + f.dint) : nullptr;
        d.status = r.out.status ? base + f.status : nullptr;
        d.rdisp = r.out.right_disparity ? reinterpret_cast<int32_t*>(base + f.rdisp) : nullptr;
        d.best = r.out.best_cost ? reinterpret_cast<uint16_t*>(base + f.best) : nullptr;
        d.second = r.out.second_cost ? reinterpret_cast<uint16_t*>(base + f.second) : nullptr;

        struct Copy {
            void* host;
            size_t dev_off;
            size_t bytes;
        };
        const Copy copies[] = {
            {r.out.disparity, f.disp, npx * 4},       {r.out.disparity_median, f.median, npx * 4},
            {r.out.confidence, f.conf, npx * 4},      {r.out.disparity_int, f.dint, npx * 4},
            {r.out.status, f.status, npx},            {r.out.right_disparity, f.rdisp, npx * 4},
            {r.out.best_cost, f.best, npx * 2},       {r.out.second_cost, f.second, npx * 2},
        };
        // Inputs from the pooled pinned buffer (or heap); results straight into the caller's buffers
        // (a direct DMA when they are pinned, e.g. from sgm_host_alloc).
        const uint8_t* src = r.pinned ? r.pinned : r.img.data();
        // Both inputs (raw frames when rectified on the device) to `dst`, rows of in_w bytes.
        auto upload = [&](uint8_t* dst, int in_w, int in_h) {
            const size_t n = (size_t)in_w * in_h;
            if (!r.borrowed[0]) return hipMemcpyAsync(dst, src, 2 * n, hipMemcpyHostToDevice, s.stream);
            hipError_t ue = hipMemcpy2DAsync(dst, in_w, r.borrowed[0], r.borrowed_stride, in_w, in_h, hipMemcpyHostToDevice, s.stream);
            if (ue == hipSuccess)
                ue = hipMemcpy2DAsync(dst + n, in_w, r.borrowed[1], r.borrowed_stride, in_w, in_h, hipMemcpyHostToDevice, s.stream);
            return ue;
        };
        // Copies on the slot's stream; kernels on the compute stream (the slot's own stream when
        // there are no shared ones), handed over by events.
        int group = 16;
        {
            std::lock_guard lk(mu);
            group = pick_group(r.w, r.h, r.nd);
        }
        Compute* cs = compute.empty() ? nullptr : compute[(size_t)si % compute.size()].get();
        hipStream_t ks = cs ? cs->stream : s.stream;
        hipError_t e;
        RectifyDesc rd{};
        if (r.rect.d_maps) {
            e = upload(base + f.raw, r.rect.raw_w, r.rect.raw_h);
            const size_t nmap = (size_t)r.rect.map_w * r.rect.map_h;
            rd.raw[0] = base + f.raw;
            rd.raw[1] = base + f.raw + raw_px;
            rd.map[0] = r.rect.d_maps;
            rd.map[1] = r.rect.d_maps + 2 * nmap;
            rd.out[0] = base + f.img;
            rd.out[1] = base + f.img + npx;
            rd.raw_pitch = r.rect.raw_w;
            rd.out_pitch = r.w;
            rd.raw_w = r.rect.raw_w;
            rd.raw_h = r.rect.raw_h;
            rd.map_w = r.rect.map_w;
            rd.out_w = r.w;
            rd.out_h = r.h;
        } else {
            e = upload(base + f.img, r.w, r.h);
        }
        bool compute_enqueued = false;
        if (e == hipSuccess && cs) e = hipEventRecord(s.ev_up, s.stream);
        if (e == hipSuccess) {
            std::unique_lock<std::mutex> klk;
            if (cs) klk = std::unique_lock<std::mutex>(cs->mu);
            if (cs) e = hipStreamWaitEvent(ks, s.ev_up, 0);
            compute_enqueued = cs != nullptr;
            if (e == hipSuccess && r.rect.d_maps) {
                e = hipEventRecord(s.ev[5], ks);
                if (e == hipSuccess) e = launch_rectify(rd, ks);
            }
            if (e == hipSuccess) e = enqueue_pair_ext(d, consts, d.dp != d.nd, ks, s.ev, group);
            if (e == hipSuccess && cs) e = hipEventRecord(s.ev_done, ks);
        }
        if (e == hipSuccess && cs) e = hipStreamWaitEvent(s.stream, s.ev_done, 0);
        for (const Copy& c : copies) {
            if (!c.host || e != hipSuccess) continue;
            e = hipMemcpyAsync(c.host, base + c.dev_off, c.bytes, hipMemcpyDeviceToHost, s.stream);
        }
        // Drain before anything is released: the slot's stream always, and the compute stream too
        // when the hand-off back to the slot's stream may not have been enqueued.
        const serve::Outcome out = serve::finish_request(e == hipSuccess, [&] {
            bool ok = hipStreamSynchronize(s.stream) == hipSuccess;
            if (e != hipSuccess && compute_enqueued) ok = hipStreamSynchronize(ks) == hipSuccess && ok;
            return ok;
        });
        float gpu_ms = 0, paths_ms = 0;
        if (out.ok) {
            (void)hipEventElapsedTime(&gpu_ms, r.rect.d_maps ? s.ev[5] : s.ev[0], s.ev[4]);
            (void)hipEventElapsedTime(&paths_ms, s.ev[1], s.ev[2]);
        }
        sgm_completion cpl{};
        cpl.ticket = r.ticket;
        cpl.user_tag = r.tag;
        cpl.status = out.ok ? SGM_OK : SGM_ERR_IO;
        cpl.t_submit_ns = r.t_submit;
        cpl.t_admit_ns = r.t_admit;
        cpl.t_done_ns = now_ns();
        cpl.gpu_ms = gpu_ms;
        {
            std::lock_guard lk(mu);
            if (!out.ok) {
                err = std::string("request ") + std::to_string(r.ticket) + ": " +
                      (e != hipSuccess ? hipGetErrorString(e) : "the device did not finish its work");
                ++failed;
            }
            ++completed;
            if (out.ok) record_group(r.w, r.h, r.nd, group, paths_ms);
            if (r.pinned) pinned_free[r.pinned_class].push_back(r.pinned);
            reqs.erase(r.ticket);
            s.has_job = false;
            done.push_back(cpl);
            if (out.reclaim) {
                aq.finish(si, r.offset);
            } else {
                // The device may still write this request's range: keep it, and the slot, out of use.
                broken = true;
                aq.retire(si);
                if (aq.live_slots() == 0) {
                    for (uint64_t t : aq.drop_all()) {
                        Request& q = reqs.at(t);
                        sgm_completion dc{};
                        dc.ticket = t;
                        dc.user_tag = q.tag;
                        dc.status = SGM_ERR_IO;
                        dc.t_submit_ns = q.t_submit;
                        dc.t_done_ns = now_ns();
                        if (q.pinned) pinned_free[q.pinned_class].push_back(q.pinned);
                        reqs.erase(t);
                        ++failed;
                        ++completed;
                        done.push_back(dc);
                    }
                }
            }
            dispatch(aq.admit());
        }
        cv_done.notify_all();
        if (!out.reclaim) return;  // a retired slot's worker ends
    }
}

extern "C" {

int64_t sgm_server_now_ns(void) { return now_ns(); }

int sgm_server_stop(sgm_server* sv) {
    if (!sv) return SGM_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard lk(sv->mu);
        sv->stopping = true;
        for (auto& s : sv->slots) s.cv.notify_all();
    }
    for (auto& s : sv->slots)
        if (s.th.joinable()) s.th.join();
    for (auto& c : sv->compute)
        if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        for (hipEvent_t e : s.ev)
            if (e) (void)hipEventDestroy(e);
        if (s.ev_up) (void)hipEventDestroy(s.ev_up);
        if (s.ev_done) (void)hipEventDestroy(s.ev_done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (auto& c : sv->compute)
        if (c->stream) (void)hipStreamDestroy(c->stream);
    for (uint8_t* b : sv->pinned_all) (void)hipHostFree(b);
    for (const MapSet& m : sv->map_sets) (void)hipFree(m.d_maps);
    if (sv->d_arena) (void)hipFree(sv->d_arena);
    delete sv;
    return SGM_OK;
}

int sgm_server_start(const sgm_params* params, const sgm_server_config* cfg, sgm_server** out) {
    if (!out || !cfg) {
        set_thread_error("out and cfg must not be NULL");
        return SGM_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::string err;
    if (int rc = validate_params_ext(params, err); rc != SGM_OK) {
        set_thread_error(err);
        return rc;
    }
    sgm_params pmax = *params;
    pmax.num_disparities = cfg->max_disparities;
    if (validate_params_ext(&pmax, err) != SGM_OK || cfg->max_width < 1 || cfg->max_width > kMaxWidth || cfg->max_height < 1 ||
        cfg->slots < 1 || cfg->slots > 64 || cfg->compute_streams < -1 || cfg->compute_streams > cfg->slots ||
        (cfg->path_group != 0 && cfg->path_group != 16 && cfg->path_group != 32 && cfg->path_group != 64)) {
        set_thread_error("server config: max sizes, max_disparities (1..512), slots (1..64), compute_streams "
                         "(-1..slots) or path_group (0, 16, 32, 64) out of range" +
                         (err.empty() ? std::string() : ": " + err));
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (int rc = check_row_lds_ext(cfg->max_width, err); rc != SGM_OK) {
        set_thread_error("server config: " + err);
        return rc;
    }
    auto* sv = new sgm_server;
    sv->p = *params;
    sv->consts = make_consts_ext(*params);
    sv->ndirs = params->eight_paths ? 8 : 4;
    sv->max_w = cfg->max_width;
    sv->max_h = cfg->max_height;
    sv->max_nd = cfg->max_disparities;
    sv->fixed_group = cfg->path_group;
    size_t free_b = 0, total_b = 0;
    hipError_t e = hipMemGetInfo(&free_b, &total_b);
    size_t budget = cfg->hbm_budget_bytes > 0 ? (size_t)cfg->hbm_budget_bytes : std::min(free_b / 2, (size_t)64 << 30);
    if (e == hipSuccess && budget > free_b / 10 * 9) budget = free_b / 10 * 9;
    if (e == hipSuccess) e = hipMalloc(&sv->d_arena, budget);
    if (e != hipSuccess) {
        set_thread_error(std::string("sgm_server_start: arena: ") + hipGetErrorString(e));
        delete sv;
        return SGM_ERR_IO;
    }
    sv->aq = serve::AdmissionQueue(cfg->slots, serve::Arena(budget));
    sv->slots = std::vector<sgm_server::Slot>((size_t)cfg->slots);
    const size_t npx = (size_t)cfg->max_width * cfg->max_height;
    for (auto& s : sv->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        // Timing only (a request completes by a stream synchronisation): no system-scope fence.
        for (int i = 0; e == hipSuccess && i < 6; ++i) e = hipEventCreateWithFlags(&s.ev[i], hipEventDisableSystemFence);
        // Hand-offs between the copy and compute streams of one device.
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev_up, hipEventDisableTiming | hipEventDisableSystemFence);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ev_done, hipEventDisableTiming | hipEventDisableSystemFence);
    }
    const int ncs = cfg->compute_streams == 0 ? 1 : (cfg->compute_streams < 0 ? 0 : cfg->compute_streams);
    for (int i = 0; e == hipSuccess && i < ncs; ++i) {
        sv->compute.push_back(std::make_unique<sgm_server::Compute>());
        e = hipStreamCreateWithFlags(&sv->compute.back()->stream, hipStreamNonBlocking);
    }
    if (e == hipSuccess) e = sv->add_pinned_class(2 * npx, 2 * cfg->slots + 2);
    if (e != hipSuccess) {
        set_thread_error(std::string("sgm_server_start: slots: ") + hipGetErrorString(e));
        sgm_server_stop(sv);
        return SGM_ERR_IO;
    }
    for (size_t i = 0; i < sv->slots.size(); ++i) sv->slots[i].th = std::thread(&sgm_server::worker, sv, (int)i);
    *out = sv;
    return SGM_OK;
}

int sgm_server_submit(sgm_server* sv, const sgm_request* rq, uint64_t* ticket) {
    if (!sv || !rq) return SGM_ERR_INVALID_ARGUMENT;
    std::string err;
    sgm_params pp = sv->p;
    pp.min_disparity = rq->min_disparity;
    pp.num_disparities = rq->num_disparities;
    MapSet ms;
    bool bad_rect = rq->rectification < 0;
    if (rq->rectification > 0) {
        std::lock_guard lk(sv->mu);
        if ((size_t)rq->rectification > sv->map_sets.size()) bad_rect = true;
        else ms = sv->map_sets[(size_t)rq->rectification - 1];
        bad_rect = bad_rect || rq->width != ms.map_w / 2 || rq->height != ms.map_h / 2;
    }
    const int in_w = ms.d_maps ? ms.raw_w : rq->width, in_h = ms.d_maps ? ms.raw_h : rq->height;
    if (validate_params_ext(&pp, err) != SGM_OK || bad_rect || rq->width < 1 || rq->height < 1 || rq->width > sv->max_w ||
        rq->height > sv->max_h || rq->num_disparities > sv->max_nd || !rq->left || !rq->right || rq->stride < in_w ||
        !rq->out.disparity || rq->out.sum_volume || rq->out.cost_volume || rq->out.census_left || rq->out.census_right ||
        census_words_ext(rq->width, rq->height, rq->min_disparity, rq->num_disparities) * 8 > kMaxViewBytes) {
        std::lock_guard lk(sv->mu);
        sv->err = "sgm_server_submit: size, disparity range, rectification, pointers or outputs not servable" +
                  (err.empty() ? "" : ": " + err);
        return SGM_ERR_INVALID_ARGUMENT;
    }
    Request r;
    r.rect = ms;
    r.tag = rq->user_tag;
    r.w = rq->width;
    r.h = rq->height;
    r.min_d = rq->min_disparity;
    r.nd = rq->num_disparities;
    r.out = rq->out;
    const size_t in_px = (size_t)in_w * in_h;
    r.bytes = footprint(r.w, r.h, r.min_d, r.nd, sv->ndirs, ms.d_maps ? in_px : 0, sv->consts.wide_costs != 0).total;
    r.t_submit = now_ns();
    if (rq->borrow_inputs) {
        r.borrowed[0] = rq->left;
        r.borrowed[1] = rq->right;
        r.borrowed_stride = rq->stride;
    } else {
        {
            std::lock_guard lk(sv->mu);
            r.pinned = sv->take_pinned(2 * in_px, r.pinned_class);
        }
        uint8_t* dst = r.pinned;
        if (!dst) {
            r.img.resize(2 * in_px);
            dst = r.img.data();
        }
        for (int y = 0; y < in_h; ++y) {
            std::memcpy(dst + (size_t)y * in_w, rq->left + (ptrdiff_t)y * rq->stride, (size_t)in_w);
            std::memcpy(dst + ((size_t)in_h + y) * in_w, rq->right + (ptrdiff_t)y * rq->stride, (size_t)in_w);
        }
    }
    std::lock_guard lk(sv->mu);
    if (sv->stopping || sv->broken || !sv->aq.can_ever_fit(r.bytes)) {
        if (r.pinned) sv->pinned_free[r.pinned_class].push_back(r.pinned);
        if (sv->stopping) {
            sv->err = "sgm_server_submit: server is stopping";
            return SGM_ERR_DISCONNECTED;
        }
        if (sv->broken) {
            sv->err = "sgm_server_submit: a request's device work could not be confirmed finished; the server "
                      "takes no new requests (stop it)";
            return SGM_ERR_IO;
        }
        sv->err = "sgm_server_submit: request needs " + std::to_string(r.bytes) + " bytes, more than the HBM arena";
        return SGM_ERR_BUSY;
    }
    r.ticket = sv->next_ticket++;
    if (ticket) *ticket = r.ticket;
    sv->aq.push(r.ticket, r.bytes);
    sv->reqs.emplace(r.ticket, std::move(r));
    sv->dispatch(sv->aq.admit());
    return SGM_OK;
}

int sgm_server_add_rectification(sgm_server* sv, const float* left_map_x, const float* left_map_y,
                                 const float* right_map_x, const float* right_map_y, int map_w, int map_h, int raw_w,
                                 int raw_h, int32_t* id) {
    if (!sv || !id) return SGM_ERR_INVALID_ARGUMENT;
    auto fail = [&](int code, const std::string& m) {
        std::lock_guard lk(sv->mu);
        sv->err = "sgm_server_add_rectification: " + m;
        return code;
    };
    if (!left_map_x || !left_map_y || !right_map_x || !right_map_y) return fail(SGM_ERR_INVALID_ARGUMENT, "null map");
    if (map_w < 2 || map_h < 2 || map_w / 2 > sv->max_w || map_h / 2 > sv->max_h)
        return fail(SGM_ERR_INVALID_ARGUMENT, "remap resolution must be twice the stereo resolution (at most the server's maximum)");
    if (raw_w < 1 || raw_h < 1 || raw_w >= (1 << 24) || raw_h >= (1 << 24))
        return fail(SGM_ERR_INVALID_ARGUMENT, "raw size must be in [1, 2^24)");
    MapSet m;
    m.map_w = map_w;
    m.map_h = map_h;
    m.raw_w = raw_w;
    m.raw_h = raw_h;
    const size_t n = (size_t)map_w * map_h;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&m.d_maps), 16 * n);
    std::vector<float> packed(e == hipSuccess ? 2 * n : 0);
    const float* mx[2] = {left_map_x, right_map_x};
    const float* my[2] = {left_map_y, right_map_y};
    for (int side = 0; side < 2 && e == hipSuccess; ++side) {
        for (size_t i = 0; i < n; ++i) {
            packed[2 * i] = mx[side][i];
            packed[2 * i + 1] = my[side][i];
        }
        e = hipMemcpy(m.d_maps + 2 * n * side, packed.data(), 8 * n, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (m.d_maps) (void)hipFree(m.d_maps);
        return fail(SGM_ERR_IO, std::string("maps: ") + hipGetErrorString(e));
    }
    std::lock_guard lk(sv->mu);
    const size_t raw_bytes = 2 * (size_t)raw_w * raw_h;
    if (sv->pinned_free.lower_bound(raw_bytes) == sv->pinned_free.end()) {
        // A size class for this set's raw frames; if pinned memory runs out, submissions of such
        // frames fall back to heap staging.
        (void)sv->add_pinned_class(raw_bytes, 2 * (int)sv->slots.size() + 2);
    }
    sv->map_sets.push_back(m);
    *id = (int32_t)sv->map_sets.size();
    return SGM_OK;
}

int sgm_server_poll(sgm_server* sv, sgm_completion* out, int max, int timeout_us, int* n) {
    if (!sv || !out || max < 1 || !n) return SGM_ERR_INVALID_ARGUMENT;
    std::unique_lock lk(sv->mu);
    if (sv->done.empty() && timeout_us != 0) {
        auto pred = [&] { return !sv->done.empty(); };
        if (timeout_us < 0) sv->cv_done.wait(lk, pred);
        else sv->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), pred);
    }
    int k = 0;
    while (k < max && !sv->done.empty()) {
        out[k++] = sv->done.front();
        sv->done.pop_front();
    }
    *n = k;
    return SGM_OK;
}

int sgm_server_get_stats(sgm_server* sv, sgm_server_stats* st) {
    if (!sv || !st) return SGM_ERR_INVALID_ARGUMENT;
    std::lock_guard lk(sv->mu);
    st->queued = (int32_t)sv->aq.queued();
    st->in_flight = sv->aq.busy_slots();
    st->completed = sv->completed;
    st->failed = sv->failed;
    st->arena_bytes = (int64_t)sv->aq.arena().capacity();
    st->arena_free_bytes = (int64_t)sv->aq.arena().free_bytes();
    return SGM_OK;
}

const char* sgm_server_last_error(sgm_server* sv) {
    if (!sv) return sgm_last_error(nullptr);
    std::lock_guard lk(sv->mu);
    thread_local std::string copy;
    copy = sv->err;
    return copy.c_str();
}

}  // extern "C"
