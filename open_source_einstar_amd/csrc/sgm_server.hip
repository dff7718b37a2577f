// Serving loop of the gfx950 SGM library (sgm_server_* in include/sgm_hip.h).
//
// Requests (host image pairs of any size up to the configured maximum, each with its own disparity
// range) are admitted in arrival order when a slot (a stream and its worker thread) is free and the
// request's device footprint fits the HBM arena reserved at start (serve_core.h). Requests that
// arrive while others run are admitted as soon as slots and memory free up, so several frames are on
// the device at once on separate streams. Each request runs the same kernels as sgm_compute on its
// own scratch inside the arena, so its result does not depend on what else is in flight.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "serve_core.h"
#include "sgm_hip.h"
#include "sgm_internal.h"

using namespace sgm;

namespace sgm {
// Shared with sgm_api.hip.
int validate_params_ext(const sgm_params* p, std::string& err);
SgmConsts make_consts_ext(const sgm_params& p);
hipError_t enqueue_pair_ext(const PairDesc& d, const SgmConsts& c, bool padded, hipStream_t s, hipEvent_t* ev5);
void set_geometry_ext(PairDesc& d, int w, int h, int min_d, int nd, int ndirs);
size_t census_words_ext(int w, int h, int min_d, int nd);
void set_thread_error(const std::string& e);
}  // namespace sgm

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Byte offsets of one request's buffers inside its arena block.
struct Footprint {
    size_t img, census, cvol, lvol, disp, median, conf, dint, status, rdisp, best, second, total;
};

Footprint footprint(int w, int h, int min_d, int nd, int ndirs) {
    const size_t npx = (size_t)w * h;
    const size_t dp = (size_t)dpl_for(nd) * kGroup;
    Footprint f{};
    size_t o = 0;
    f.img = o;
    o = align_up(o + 2 * npx);
    f.census = o;
    o = align_up(o + 2 * census_words_ext(w, h, min_d, nd) * 8);
    f.cvol = o;  // matching costs as bytes: the path kernel's cost-volume mode
    o = align_up(o + npx * dp);
    f.lvol = o;
    o = align_up(o + (size_t)ndirs * npx * dp);
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

struct Request {
    uint64_t ticket = 0, tag = 0;
    int w = 0, h = 0, min_d = 0, nd = 0;
    uint8_t* pinned = nullptr;  // left rows then right rows (pitch w) in a pooled pinned buffer, or
    std::vector<uint8_t> img;   // the same in heap memory when the pool was empty
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

    std::vector<uint8_t*> pinned_free;  // pinned input buffers (2 * max_w * max_h bytes each)
    std::vector<uint8_t*> pinned_all;

    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t ev[5] = {};
        std::thread th;
        std::condition_variable cv;
        bool has_job = false;
        uint64_t ticket = 0;
    };
    std::vector<Slot> slots;

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
        const Footprint f = footprint(r.w, r.h, r.min_d, r.nd, ndirs);
        uint8_t* base = d_arena + r.offset;
        const size_t npx = (size_t)r.w * r.h;
        PairDesc d{};
        d.left = base + f.img;
        d.right = base + f.img + npx;
        d.pitch = r.w;
        d.census_l = reinterpret_cast<uint64_t*>(base + f.census);
        d.census_r = d.census_l + census_words_ext(r.w, r.h, r.min_d, r.nd);
        d.lvol = base + f.lvol;
        d.cvol = base + f.cvol;
        set_geometry_ext(d, r.w, r.h, r.min_d, r.nd, ndirs);
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
        hipError_t e = hipMemcpyAsync(base + f.img, src, 2 * npx, hipMemcpyHostToDevice, s.stream);
        if (e == hipSuccess) e = enqueue_pair_ext(d, consts, d.dp != d.nd, s.stream, s.ev);
        for (const Copy& c : copies) {
            if (!c.host || e != hipSuccess) continue;
            e = hipMemcpyAsync(c.host, base + c.dev_off, c.bytes, hipMemcpyDeviceToHost, s.stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
        float gpu_ms = 0;
        if (e == hipSuccess) (void)hipEventElapsedTime(&gpu_ms, s.ev[0], s.ev[4]);
        sgm_completion cpl{};
        cpl.ticket = r.ticket;
        cpl.user_tag = r.tag;
        cpl.status = e == hipSuccess ? SGM_OK : SGM_ERR_IO;
        cpl.t_submit_ns = r.t_submit;
        cpl.t_admit_ns = r.t_admit;
        cpl.t_done_ns = now_ns();
        cpl.gpu_ms = gpu_ms;
        {
            std::lock_guard lk(mu);
            if (e != hipSuccess) {
                err = std::string("request ") + std::to_string(r.ticket) + ": " + hipGetErrorString(e);
                ++failed;
            }
            ++completed;
            if (r.pinned) pinned_free.push_back(r.pinned);
            reqs.erase(r.ticket);
            s.has_job = false;
            aq.finish(si, r.offset);
            done.push_back(cpl);
            dispatch(aq.admit());
        }
        cv_done.notify_all();
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
    for (auto& s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        for (hipEvent_t e : s.ev)
            if (e) (void)hipEventDestroy(e);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (uint8_t* b : sv->pinned_all) (void)hipHostFree(b);
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
        cfg->slots < 1 || cfg->slots > 64) {
        set_thread_error("server config: max sizes, max_disparities (1..512) or slots (1..64) out of range" +
                         (err.empty() ? std::string() : ": " + err));
        return SGM_ERR_INVALID_ARGUMENT;
    }
    auto* sv = new sgm_server;
    sv->p = *params;
    sv->consts = make_consts_ext(*params);
    sv->ndirs = params->eight_paths ? 8 : 4;
    sv->max_w = cfg->max_width;
    sv->max_h = cfg->max_height;
    sv->max_nd = cfg->max_disparities;
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
        for (int i = 0; e == hipSuccess && i < 5; ++i) e = hipEventCreate(&s.ev[i]);
    }
    for (int i = 0; e == hipSuccess && i < 2 * cfg->slots + 2; ++i) {
        uint8_t* b = nullptr;
        e = hipHostMalloc(reinterpret_cast<void**>(&b), 2 * npx, hipHostMallocDefault);
        if (e == hipSuccess) {
            sv->pinned_all.push_back(b);
            sv->pinned_free.push_back(b);
        }
    }
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
    if (validate_params_ext(&pp, err) != SGM_OK || rq->width < 1 || rq->height < 1 || rq->width > sv->max_w ||
        rq->height > sv->max_h || rq->num_disparities > sv->max_nd || !rq->left || !rq->right || rq->stride < rq->width ||
        !rq->out.disparity || rq->out.sum_volume || rq->out.cost_volume || rq->out.census_left || rq->out.census_right) {
        std::lock_guard lk(sv->mu);
        sv->err = "sgm_server_submit: size, disparity range, pointers or outputs not servable" + (err.empty() ? "" : ": " + err);
        return SGM_ERR_INVALID_ARGUMENT;
    }
    Request r;
    r.tag = rq->user_tag;
    r.w = rq->width;
    r.h = rq->height;
    r.min_d = rq->min_disparity;
    r.nd = rq->num_disparities;
    r.out = rq->out;
    r.bytes = footprint(r.w, r.h, r.min_d, r.nd, sv->ndirs).total;
    r.t_submit = now_ns();
    {
        std::lock_guard lk(sv->mu);
        if (!sv->pinned_free.empty()) {
            r.pinned = sv->pinned_free.back();
            sv->pinned_free.pop_back();
        }
    }
    uint8_t* dst = r.pinned;
    if (!dst) {
        r.img.resize(2 * (size_t)r.w * r.h);
        dst = r.img.data();
    }
    for (int y = 0; y < r.h; ++y) {
        std::memcpy(dst + (size_t)y * r.w, rq->left + (ptrdiff_t)y * rq->stride, (size_t)r.w);
        std::memcpy(dst + ((size_t)r.h + y) * r.w, rq->right + (ptrdiff_t)y * rq->stride, (size_t)r.w);
    }
    std::lock_guard lk(sv->mu);
    if (sv->stopping || !sv->aq.can_ever_fit(r.bytes)) {
        if (r.pinned) sv->pinned_free.push_back(r.pinned);
        if (sv->stopping) {
            sv->err = "sgm_server_submit: server is stopping";
            return SGM_ERR_DISCONNECTED;
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
