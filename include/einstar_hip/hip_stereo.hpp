#pragma once

// einstar::depth_hip::HipStereo -- C++ replacement of einstar::depth_metal::MetalStereo
// (libs/depth_metal/include/einstar/depth_metal/metal_stereo.hpp:67-107) for its stereo matching,
// on the gfx950 library behind include/sgm_hip.h. Header-only: it only calls the C ABI.
//
// The value types restate the reference's (C++20 here; the reference's need C++23 <expected>):
//   ImageView / Image            libs/core/include/einstar/core/image.hpp:16-102 (stride in elements)
//   Errc / Error / Result        libs/core/include/einstar/core/error.hpp:8-31
//   SgmParams ... StereoParams   libs/depth/include/einstar/depth/stereo.hpp:26-88
//   StereoResult, StereoTimings  stereo.hpp:90-93, metal_stereo.hpp:61-65
//
// Scope: create / set_rectification / compute / compute_raw / last_timings as MetalStereo has them,
// running depth::compute_disparity with every stage of StereoParams (pyramid, SGM, ZNCC refinement,
// detail refinement, occlusion and spike filters, speckle removal). compute_full is
// compute_disparity's second form. compute_sgm additionally returns every per-pixel output of the
// SGM stage (integer disparity, status bits, right disparity, best/second costs, median3) when SGM
// runs at the configured size (pyramid_levels == 0). set_point_params / compute_frame_raw produce the
// point frame (points, normals, weights) and pack_depth the recording layout; MetalStereo's
// marker-blob search (compute_frame's FrameRequest) is not part of this library.

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "sgm_hip.h"

namespace einstar::depth_hip {

enum class Errc { ok = 0, not_found, io, timeout, protocol, invalid_argument, blocked, unsupported, busy, disconnected };

struct Error {
    Errc code = Errc::ok;
    std::string message;
};

// The subset of std::expected<T, Error> the reference's callers use.
template <typename T>
class Result {
public:
    Result(T v) : ok_(true), value_(std::move(v)) {}
    Result(Error e) : ok_(false), error_(std::move(e)) {}
    [[nodiscard]] bool has_value() const { return ok_; }
    explicit operator bool() const { return ok_; }
    T& operator*() { return value_; }
    const T& operator*() const { return value_; }
    T* operator->() { return &value_; }
    const T* operator->() const { return &value_; }
    T& value() { return value_; }
    [[nodiscard]] const Error& error() const { return error_; }

private:
    bool ok_;
    T value_{};
    Error error_;
};

template <>
class Result<void> {
public:
    Result() = default;
    Result(Error e) : ok_(false), error_(std::move(e)) {}
    [[nodiscard]] bool has_value() const { return ok_; }
    explicit operator bool() const { return ok_; }
    [[nodiscard]] const Error& error() const { return error_; }

private:
    bool ok_ = true;
    Error error_;
};

inline Error make_error(Errc code, std::string message) { return Error{code, std::move(message)}; }

template <typename T>
struct ImageView {
    T* data = nullptr;
    int width = 0;
    int height = 0;
    std::ptrdiff_t stride = 0;  // elements

    constexpr ImageView() = default;
    constexpr ImageView(T* d, int w, int h, std::ptrdiff_t s) : data(d), width(w), height(h), stride(s) {}
    constexpr ImageView(T* d, int w, int h) : ImageView(d, w, h, w) {}
    template <typename U>
        requires std::is_same_v<const U, T>
    constexpr ImageView(const ImageView<U>& o) : data(o.data), width(o.width), height(o.height), stride(o.stride) {}
    [[nodiscard]] constexpr T& operator()(int x, int y) const { return data[y * stride + x]; }
    [[nodiscard]] constexpr T* row(int y) const { return data + y * stride; }
};

template <typename T>
class Image {
public:
    Image() = default;
    Image(int w, int h, T fill = T{}) : width_(w), height_(h), pixels_(static_cast<std::size_t>(w) * static_cast<std::size_t>(h), fill) {}
    [[nodiscard]] int width() const { return width_; }
    [[nodiscard]] int height() const { return height_; }
    [[nodiscard]] std::size_t size() const { return pixels_.size(); }
    [[nodiscard]] T* data() { return pixels_.data(); }
    [[nodiscard]] const T* data() const { return pixels_.data(); }
    T& operator()(int x, int y) { return pixels_[static_cast<std::size_t>(y) * static_cast<std::size_t>(width_) + static_cast<std::size_t>(x)]; }
    const T& operator()(int x, int y) const {
        return pixels_[static_cast<std::size_t>(y) * static_cast<std::size_t>(width_) + static_cast<std::size_t>(x)];
    }
    [[nodiscard]] ImageView<T> view() { return {pixels_.data(), width_, height_}; }
    [[nodiscard]] ImageView<const T> view() const { return {pixels_.data(), width_, height_}; }

private:
    int width_ = 0;
    int height_ = 0;
    std::vector<T> pixels_;
};

using ImageU8 = Image<std::uint8_t>;
using ImageU16 = Image<std::uint16_t>;
using ImageI32 = Image<std::int32_t>;
using ImageF32 = Image<float>;

inline constexpr float kInvalidDisparity = SGM_INVALID_DISPARITY;
[[nodiscard]] constexpr bool valid_disparity(float d) { return d > -1.0e8f; }

struct SgmParams {
    int min_disparity = 0;
    int num_disparities = 64;
    int census_radius_x = 2;
    int census_radius_y = 2;
    int p1 = 8;
    int p2 = 96;
    bool adaptive_p2 = true;
    bool eight_paths = false;
    bool slant_steps = true;
    float uniqueness = 0.9f;
    int lr_max_diff = 1;
};
struct RefineParams {
    int search_radius = 2;
    int zncc_radius = 3;
    float min_zncc = 0.5f;
    bool reject_edge_winners = true;
};
struct DetailParams {
    bool enabled = true;
    int radius = 5;
    float search = 1.0f;
    float step = 0.5f;
    float min_zncc = 0.5f;
};
struct FilterParams {
    bool occlusion = true;
    float occlusion_tolerance = 1.0f;
    bool spikes = true;
    float spike_threshold = 1.0f;
};
struct SpeckleParams {
    int max_region_size = 200;
    float max_diff = 1.0f;
};
struct StereoParams {
    int pyramid_levels = 2;
    SgmParams sgm;
    RefineParams refine;
    DetailParams detail;
    FilterParams filter;
    SpeckleParams speckle;
    bool subpixel = true;
};

struct StereoResult {
    ImageF32 disparity;
    ImageF32 confidence;
};

// calib::RemapTable (libs/calib/include/einstar/calib/rectify.hpp:40-43): for each rectified pixel,
// the source pixel in the raw (distorted) frame.
struct RemapTable {
    int width = 0, height = 0;
    ImageF32 map_x, map_y;
};

struct StereoTimings {
    double gpu_ms = 0;
    double speckle_ms = 0;
    double total_ms = 0;
};

// Every per-pixel output of the SGM stage.
struct SgmResult {
    ImageF32 disparity;         // sgm_disparity (stereo.cpp:361-463)
    ImageF32 disparity_median;  // median3 of it (stereo.cpp:129-147)
    ImageI32 disparity_int;     // min_disparity + best_k, SGM_NO_MATCH for an empty window
    ImageU8 status;             // SGM_STATUS_* bits
    ImageI32 right_disparity;   // right WTA, SGM_NO_MATCH where none
    ImageU16 best_cost;
    ImageU16 second_cost;
};

// depth::RectifiedGeometry and the point limits MetalStereo::set_point_params takes
// (point_image.hpp:12-26, 43-49; metal_stereo.hpp:83-85).
struct RectifiedGeometry {
    double f = 0, cx = 0, cy = 0, baseline = 0, cx_offset = 0;
};
struct PointImageParams {
    float min_depth = 150.0f;
    float max_depth = 700.0f;
    float max_depth_jump = 4.0f;
    float edge_weight = 0.25f;
    float weight_reference_depth = 0.0f;
};

// A point frame as MetalFrameData exposes it (points_xyzw, normals_xyzw, weights), on the host.
struct FrameResult {
    int width = 0, height = 0;
    std::vector<float> points_xyzw;   // 4 per pixel, zero where invalid
    std::vector<float> normals_xyzw;  // 4 per pixel, w = 1 where a normal exists
    std::vector<float> weights;
};

inline sgm_params to_c(const StereoParams& p) {
    sgm_params c{};
    c.min_disparity = p.sgm.min_disparity;
    c.num_disparities = p.sgm.num_disparities;
    c.census_radius_x = p.sgm.census_radius_x;
    c.census_radius_y = p.sgm.census_radius_y;
    c.p1 = p.sgm.p1;
    c.p2 = p.sgm.p2;
    c.adaptive_p2 = p.sgm.adaptive_p2 ? 1 : 0;
    c.eight_paths = p.sgm.eight_paths ? 1 : 0;
    c.slant_steps = p.sgm.slant_steps ? 1 : 0;
    c.uniqueness = p.sgm.uniqueness;
    c.lr_max_diff = p.sgm.lr_max_diff;
    c.subpixel = p.subpixel ? 1 : 0;
    return c;
}

inline sgm_stereo_params to_c_stereo(const StereoParams& p) {
    sgm_stereo_params c{};
    c.sgm = to_c(p);
    c.pyramid_levels = p.pyramid_levels;
    c.refine_search_radius = p.refine.search_radius;
    c.refine_zncc_radius = p.refine.zncc_radius;
    c.refine_min_zncc = p.refine.min_zncc;
    c.refine_reject_edge_winners = p.refine.reject_edge_winners ? 1 : 0;
    c.detail_enabled = p.detail.enabled ? 1 : 0;
    c.detail_radius = p.detail.radius;
    c.detail_search = p.detail.search;
    c.detail_step = p.detail.step;
    c.detail_min_zncc = p.detail.min_zncc;
    c.filter_occlusion = p.filter.occlusion ? 1 : 0;
    c.filter_occlusion_tolerance = p.filter.occlusion_tolerance;
    c.filter_spikes = p.filter.spikes ? 1 : 0;
    c.filter_spike_threshold = p.filter.spike_threshold;
    c.speckle_max_region_size = p.speckle.max_region_size;
    c.speckle_max_diff = p.speckle.max_diff;
    return c;
}

// True when depth::compute_disparity with these parameters is the SGM stage alone.
inline bool sgm_only(const StereoParams& p) {
    return p.pyramid_levels == 0 && !p.detail.enabled && !p.filter.occlusion && !p.filter.spikes &&
           p.speckle.max_region_size <= 1;
}

class HipStereo {
public:
    // MetalStereo::create (metal_stereo.hpp:70-71) without the Metal context argument: the current
    // HIP device is used. `width`/`height`: the resolution the disparity is computed at (the finest
    // level); SGM runs at (width >> pyramid_levels) x (height >> pyramid_levels).
    static Result<std::unique_ptr<HipStereo>> create(const StereoParams& params, int width, int height) {
        sgm_context* ctx = nullptr;
        int rc;
        if (sgm_only(params)) {
            const sgm_params c = to_c(params);
            rc = sgm_create(&c, width, height, &ctx);
        } else {
            const sgm_stereo_params c = to_c_stereo(params);
            rc = sgm_create_stereo(&c, width, height, &ctx);
        }
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(nullptr));
        return std::unique_ptr<HipStereo>(new HipStereo(ctx, params, width, height));
    }
    ~HipStereo() { sgm_destroy(ctx_); }
    HipStereo(const HipStereo&) = delete;
    HipStereo& operator=(const HipStereo&) = delete;

    // MetalStereo::compute (metal_stereo.hpp:77-78): depth::compute_disparity (stereo.cpp:496-536) on
    // rectified images at the configured size, every stage of the StereoParams.
    Result<StereoResult> compute(ImageView<const std::uint8_t> left, ImageView<const std::uint8_t> right) {
        return compute_full(left, right, {}, {});
    }

    // depth::compute_disparity's second form (stereo.hpp:96-98): with the full-resolution images the
    // pair was box-downsampled from, for the detail refinement (used when they are twice the size).
    Result<StereoResult> compute_full(ImageView<const std::uint8_t> left, ImageView<const std::uint8_t> right,
                                      ImageView<const std::uint8_t> full_left, ImageView<const std::uint8_t> full_right) {
        if (auto r = check(left, right); !r) return r.error();
        if ((full_left.data == nullptr) != (full_right.data == nullptr))
            return make_error(Errc::invalid_argument, "full-resolution images must be given together");
        if (full_left.data && (full_left.stride != full_right.stride || full_left.width != full_right.width ||
                               full_left.height != full_right.height))
            return make_error(Errc::invalid_argument, "full-resolution images must share a size and stride");
        StereoResult res;
        res.disparity = ImageF32(w_, h_);
        res.confidence = ImageF32(w_, h_);
        const int rc = sgm_compute_disparity(ctx_, left.data, right.data, left.stride, w_, h_, full_left.data,
                                             full_right.data, full_left.stride, full_left.width, full_left.height,
                                             res.disparity.data(), res.confidence.data());
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(ctx_));
        read_timings();
        return res;
    }

    // The SGM stage's full output set (sgm_disparity plus median3 and the integer intermediates) at the
    // configured size, for configurations that run SGM at that size (pyramid_levels == 0).
    Result<SgmResult> compute_sgm(ImageView<const std::uint8_t> left, ImageView<const std::uint8_t> right) {
        if (params_.pyramid_levels != 0)
            return make_error(Errc::unsupported, "compute_sgm needs pyramid_levels 0 (SGM at the configured size)");
        if (auto r = check(left, right); !r) return r.error();
        SgmResult res;
        res.disparity = ImageF32(w_, h_);
        res.disparity_median = ImageF32(w_, h_);
        res.disparity_int = ImageI32(w_, h_);
        res.status = ImageU8(w_, h_);
        res.right_disparity = ImageI32(w_, h_);
        res.best_cost = ImageU16(w_, h_);
        res.second_cost = ImageU16(w_, h_);
        sgm_outputs o{};
        o.disparity = res.disparity.data();
        o.disparity_median = res.disparity_median.data();
        o.disparity_int = res.disparity_int.data();
        o.status = res.status.data();
        o.right_disparity = res.right_disparity.data();
        o.best_cost = res.best_cost.data();
        o.second_cost = res.second_cost.data();
        if (auto r = run(left, right, o); !r) return r.error();
        return res;
    }

    // MetalStereo::set_rectification (metal_stereo.hpp:74-75): raw full-resolution frames + remap
    // tables at twice the stereo resolution; the tables stay resident on the device.
    Result<void> set_rectification(const RemapTable& left, const RemapTable& right, int raw_width, int raw_height) {
        if (left.width / 2 != w_ || left.height / 2 != h_ || right.width != left.width || right.height != left.height)
            return make_error(Errc::invalid_argument, "remap resolution must be twice the stereo resolution");
        const int rc = sgm_set_rectification(ctx_, left.map_x.data(), left.map_y.data(), right.map_x.data(),
                                             right.map_y.data(), left.width, left.height, raw_width, raw_height);
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(ctx_));
        raw_w_ = raw_width;
        raw_h_ = raw_height;
        return {};
    }

    // MetalStereo::compute_raw (metal_stereo.hpp:79-81): rectify (to half of the remap resolution)
    // and run stereo; optionally returns the rectified pair.
    Result<StereoResult> compute_raw(ImageView<const std::uint8_t> raw_left, ImageView<const std::uint8_t> raw_right,
                                     ImageU8* rect_left = nullptr, ImageU8* rect_right = nullptr) {
        if (raw_w_ == 0) return make_error(Errc::invalid_argument, "set_rectification() not called");
        if (raw_left.width != raw_w_ || raw_left.height != raw_h_ || raw_right.width != raw_w_ || raw_right.height != raw_h_)
            return make_error(Errc::invalid_argument, "raw size mismatch");
        if (raw_left.stride != raw_right.stride)
            return make_error(Errc::invalid_argument, "left and right images must share a stride");
        StereoResult res;
        res.disparity = ImageF32(w_, h_);
        res.confidence = ImageF32(w_, h_);
        if (rect_left) *rect_left = ImageU8(w_, h_);
        if (rect_right) *rect_right = ImageU8(w_, h_);
        const int rc = sgm_compute_disparity_raw(ctx_, raw_left.data, raw_right.data, raw_left.stride, raw_w_, raw_h_,
                                                 res.disparity.data(), res.confidence.data(),
                                                 rect_left ? rect_left->data() : nullptr,
                                                 rect_right ? rect_right->data() : nullptr);
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(ctx_));
        read_timings();
        return res;
    }

    // MetalStereo::set_point_params (metal_stereo.hpp:83-85).
    Result<void> set_point_params(const RectifiedGeometry& g, float min_depth_mm, float max_depth_mm,
                                  float max_depth_jump_mm = 4.0f) {
        const sgm_rectified_geometry cg{g.f, g.cx, g.cy, g.baseline, g.cx_offset};
        sgm_point_params pp{};
        sgm_point_params_default(&pp);
        pp.min_depth = min_depth_mm;
        pp.max_depth = max_depth_mm;
        pp.max_depth_jump = max_depth_jump_mm;
        const int rc = sgm_set_point_params(ctx_, &cg, &pp);
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(ctx_));
        return {};
    }

    // MetalStereo::compute_frame_raw (metal_stereo.hpp:86-90): raw pair -> rectified -> stereo ->
    // points, normals, weights; the frame comes back to the host (the reference keeps it on the GPU).
    Result<FrameResult> compute_frame_raw(ImageView<const std::uint8_t> raw_left, ImageView<const std::uint8_t> raw_right,
                                          ImageU8* rect_left = nullptr, ImageU8* rect_right = nullptr) {
        if (raw_left.stride != raw_right.stride)
            return make_error(Errc::invalid_argument, "left and right images must share a stride");
        if (raw_left.width != raw_right.width || raw_left.height != raw_right.height ||
            (raw_w_ && (raw_left.width != raw_w_ || raw_left.height != raw_h_)))
            return make_error(Errc::invalid_argument, "raw size mismatch");
        FrameResult f;
        f.width = w_;
        f.height = h_;
        const auto n = static_cast<std::size_t>(w_) * static_cast<std::size_t>(h_);
        f.points_xyzw.resize(4 * n);
        f.normals_xyzw.resize(4 * n);
        f.weights.resize(n);
        if (rect_left) *rect_left = ImageU8(w_, h_);
        if (rect_right) *rect_right = ImageU8(w_, h_);
        const sgm_frame out{f.points_xyzw.data(), f.normals_xyzw.data(), f.weights.data()};
        const int rc = sgm_compute_frame_raw(ctx_, raw_left.data, raw_right.data, raw_left.stride, raw_left.width,
                                             raw_left.height, &out, rect_left ? rect_left->data() : nullptr,
                                             rect_right ? rect_right->data() : nullptr);
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(ctx_));
        read_timings();
        return f;
    }

    // MetalStereo::last_timings (metal_stereo.hpp:97).
    [[nodiscard]] const StereoTimings& last_timings() const { return timings_; }

private:
    HipStereo(sgm_context* ctx, const StereoParams& p, int w, int h) : ctx_(ctx), params_(p), w_(w), h_(h) {}

    // MetalStereo::compute's size check (metal_stereo.cpp:235-236).
    Result<void> check(ImageView<const std::uint8_t> l, ImageView<const std::uint8_t> r) const {
        if (l.width != w_ || l.height != h_ || r.width != w_ || r.height != h_)
            return make_error(Errc::invalid_argument, "image size does not match the configured stereo size");
        if (l.stride != r.stride)
            return make_error(Errc::invalid_argument, "left and right images must share a stride");
        return {};
    }

    Result<void> run(ImageView<const std::uint8_t> l, ImageView<const std::uint8_t> r, const sgm_outputs& o) {
        const int rc = sgm_compute(ctx_, l.data, r.data, l.stride, w_, h_, &o);
        if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), sgm_last_error(ctx_));
        read_timings();
        return {};
    }

    void read_timings() {
        sgm_timings t{};
        sgm_last_timings(ctx_, &t);
        timings_.gpu_ms = t.gpu_ms;
        timings_.speckle_ms = 0;
        timings_.total_ms = t.total_ms;
    }

    sgm_context* ctx_;
    StereoParams params_;
    int w_, h_;
    int raw_w_ = 0, raw_h_ = 0;
    StereoTimings timings_;
};

// depth_metal::DepthPacker::pack (libs/depth_metal/include/einstar/depth_metal/depth_packer.hpp:17-38)
// for a host frame: the session file's uncompressed depth layout, 3 * width * height bytes, packed on
// the GPU (sgm_pack_depth_device) through temporary device buffers.
inline Result<std::vector<std::uint8_t>> pack_depth(const FrameResult& f) {
    const auto n = static_cast<std::size_t>(f.width) * static_cast<std::size_t>(f.height);
    if (n == 0 || f.points_xyzw.size() != 4 * n || (!f.weights.empty() && f.weights.size() != n))
        return make_error(Errc::invalid_argument, "frame arrays do not match its size");
    void* dp = nullptr;
    void* dw = nullptr;
    void* dout = nullptr;
    std::vector<std::uint8_t> bytes(3 * n);
    int rc = sgm_dev_malloc(16 * n, &dp);
    if (rc == SGM_OK && !f.weights.empty()) rc = sgm_dev_malloc(4 * n, &dw);
    if (rc == SGM_OK) rc = sgm_dev_malloc(3 * n, &dout);
    if (rc == SGM_OK) rc = sgm_dev_upload(dp, f.points_xyzw.data(), 16 * n);
    if (rc == SGM_OK && dw) rc = sgm_dev_upload(dw, f.weights.data(), 4 * n);
    if (rc == SGM_OK)
        rc = sgm_pack_depth_device(static_cast<const float*>(dp), static_cast<const float*>(dw), f.width, f.height,
                                   static_cast<std::uint8_t*>(dout), nullptr);
    if (rc == SGM_OK) rc = sgm_dev_download(bytes.data(), dout, 3 * n);
    const std::string err = rc == SGM_OK ? std::string() : sgm_last_error(nullptr);
    for (void* p : {dp, dw, dout})
        if (p) sgm_dev_free(p);
    if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), err);
    return bytes;
}

// depth::remove_speckles (libs/depth/include/einstar/depth/stereo.hpp:117-118) on the GPU, in place:
// regions (4-neighbourhood, |d_a - d_b| <= max_diff) of fewer than max_region_size pixels become
// kInvalidDisparity. Through temporary device buffers (sgm_remove_speckles_device).
inline Result<void> remove_speckles(ImageF32& disparity, const SpeckleParams& p) {
    const auto n = disparity.size();
    if (n == 0) return {};
    void* dd = nullptr;
    void* dw = nullptr;
    const std::size_t work = sgm_speckle_workspace_bytes(disparity.width(), disparity.height());
    int rc = sgm_dev_malloc(4 * n, &dd);
    if (rc == SGM_OK && p.max_region_size > 1) rc = sgm_dev_malloc(work, &dw);
    if (rc == SGM_OK) rc = sgm_dev_upload(dd, disparity.data(), 4 * n);
    if (rc == SGM_OK)
        rc = sgm_remove_speckles_device(static_cast<const float*>(dd), static_cast<float*>(dd), disparity.width(),
                                        disparity.height(), p.max_region_size, p.max_diff, dw, nullptr);
    if (rc == SGM_OK) rc = sgm_dev_download(disparity.data(), dd, 4 * n);
    const std::string err = rc == SGM_OK ? std::string() : sgm_last_error(nullptr);
    for (void* q : {dd, dw})
        if (q) sgm_dev_free(q);
    if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), err);
    return {};
}

// depth::filter_disparity (libs/depth/include/einstar/depth/stereo.hpp:124-125) on the GPU, in place:
// the occlusion filter, then the spike filter. Through temporary device buffers
// (sgm_filter_disparity_device).
inline Result<void> filter_disparity(ImageF32& disparity, const FilterParams& p) {
    const auto n = disparity.size();
    if (n == 0) return {};
    void* dd = nullptr;
    void* dw = nullptr;
    int rc = sgm_dev_malloc(4 * n, &dd);
    if (rc == SGM_OK && p.spikes) rc = sgm_dev_malloc(4 * n, &dw);
    if (rc == SGM_OK) rc = sgm_dev_upload(dd, disparity.data(), 4 * n);
    if (rc == SGM_OK)
        rc = sgm_filter_disparity_device(static_cast<float*>(dd), disparity.width(), disparity.height(), p.occlusion ? 1 : 0,
                                         p.occlusion_tolerance, p.spikes ? 1 : 0, p.spike_threshold,
                                         static_cast<float*>(dw), nullptr);
    if (rc == SGM_OK) rc = sgm_dev_download(disparity.data(), dd, 4 * n);
    const std::string err = rc == SGM_OK ? std::string() : sgm_last_error(nullptr);
    for (void* q : {dd, dw})
        if (q) sgm_dev_free(q);
    if (rc != SGM_OK) return make_error(static_cast<Errc>(rc), err);
    return {};
}

}  // namespace einstar::depth_hip
