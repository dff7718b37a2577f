// C++ self-test of einstar::depth_hip::HipStereo (include/einstar_hip/hip_stereo.hpp) on the GPU,
// against the CPU oracle: the drop-in class's outputs, error codes and timings, the way
// tests/depth/test_metal_stereo.cpp exercises MetalStereo. Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "einstar_hip/hip_stereo.hpp"
#include "sgm_oracle.h"

using namespace einstar::depth_hip;

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
    std::fflush(stdout);
    if (!ok) ++failures;
}

// A textured pair: left random blobs, right = left shifted by `shift` px with fresh columns.
void make_pair(int w, int h, int shift, ImageU8& l, ImageU8& r) {
    std::mt19937 rng(1234);
    std::uniform_int_distribution<int> u(0, 255);
    l = ImageU8(w, h);
    r = ImageU8(w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) l(x, y) = static_cast<std::uint8_t>(u(rng));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) r(x, y) = x + shift < w ? l(x + shift, y) : static_cast<std::uint8_t>(u(rng));
}

orc_params oracle_params(const StereoParams& p) {
    orc_params o{};
    o.min_disparity = p.sgm.min_disparity;
    o.num_disparities = p.sgm.num_disparities;
    o.census_radius_x = p.sgm.census_radius_x;
    o.census_radius_y = p.sgm.census_radius_y;
    o.p1 = p.sgm.p1;
    o.p2 = p.sgm.p2;
    o.adaptive_p2 = p.sgm.adaptive_p2;
    o.eight_paths = p.sgm.eight_paths;
    o.slant_steps = p.sgm.slant_steps;
    o.uniqueness = p.sgm.uniqueness;
    o.lr_max_diff = p.sgm.lr_max_diff;
    o.subpixel = p.subpixel;
    return o;
}

template <typename T>
bool same_bytes(const Image<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), b.size() * sizeof(T)) == 0;
}

}  // namespace

int main() {
    const int w = 200, h = 120;
    ImageU8 l, r;
    make_pair(w, h, 9, l, r);

    StereoParams sp;
    sp.pyramid_levels = 0;
    sp.filter.occlusion = false;
    sp.filter.spikes = false;
    sp.speckle.max_region_size = 0;
    sp.detail.enabled = false;  // compute_disparity is then sgm_disparity itself
    sp.sgm.num_disparities = 32;
    sp.sgm.census_radius_x = 4;
    sp.sgm.census_radius_y = 3;
    sp.sgm.p1 = 10;
    sp.sgm.p2 = 120;
    sp.sgm.eight_paths = true;

    auto st = HipStereo::create(sp, w, h);
    expect(st.has_value(), "create");
    if (!st) {
        std::printf("%s\n", st.error().message.c_str());
        return 1;
    }

    // Oracle on the same pair.
    const auto n = static_cast<std::size_t>(w * h);
    std::vector<float> disp(n), med(n);
    std::vector<std::int32_t> dint(n), rdisp(n);
    std::vector<std::uint8_t> status(n);
    std::vector<std::uint16_t> best(n), second(n);
    orc_outputs oo{};
    oo.disparity = disp.data();
    oo.disparity_median = med.data();
    oo.disparity_int = dint.data();
    oo.status = status.data();
    oo.right_disparity = rdisp.data();
    oo.best_cost = best.data();
    oo.second_cost = second.data();
    const orc_params op = oracle_params(sp);
    expect(orc_sgm(l.data(), r.data(), w, w, h, &op, &oo, 0) == 0, "oracle");

    auto res = (*st)->compute_sgm(l.view(), r.view());
    expect(res.has_value(), "compute_sgm");
    if (res) {
        expect(same_bytes(res->disparity, disp), "compute_sgm disparity bit-exact");
        expect(same_bytes(res->disparity_median, med), "compute_sgm median3 bit-exact");
        expect(same_bytes(res->disparity_int, dint), "compute_sgm integer disparity bit-exact");
        expect(same_bytes(res->status, status), "compute_sgm status bit-exact");
        expect(same_bytes(res->right_disparity, rdisp), "compute_sgm right disparity bit-exact");
        expect(same_bytes(res->best_cost, best) && same_bytes(res->second_cost, second), "compute_sgm costs bit-exact");
        int good = 0, valid = 0;
        for (int y = 10; y < h - 10; ++y)
            for (int x = 50; x < w - 10; ++x)
                if (valid_disparity(res->disparity(x, y))) {
                    ++valid;
                    good += std::abs(res->disparity(x, y) - 9.0f) < 0.5f;
                }
        expect(valid > 0 && good > 0.97 * valid, "shifted texture recovered (> 97% within 0.5 px)");
    }

    auto full = (*st)->compute(l.view(), r.view());
    expect(full.has_value(), "compute (SGM-only configuration)");
    if (full) {
        expect(same_bytes(full->disparity, disp), "compute disparity == sgm_disparity");
        bool conf_ok = true;
        for (std::size_t i = 0; i < n; ++i) conf_ok &= full->confidence.data()[i] == (valid_disparity(disp[i]) ? 1.0f : 0.0f);
        expect(conf_ok, "compute confidence = validity (stereo.cpp:515-519)");
        expect((*st)->last_timings().gpu_ms > 0 && (*st)->last_timings().total_ms >= (*st)->last_timings().gpu_ms,
               "last_timings");
    }

    // A strided view of a wider buffer gives the same result.
    {
        const int pitch = w + 24;
        std::vector<std::uint8_t> bl(static_cast<std::size_t>(pitch * h)), br(bl.size());
        for (int y = 0; y < h; ++y) {
            std::memcpy(&bl[static_cast<std::size_t>(y * pitch)], l.view().row(y), static_cast<std::size_t>(w));
            std::memcpy(&br[static_cast<std::size_t>(y * pitch)], r.view().row(y), static_cast<std::size_t>(w));
        }
        auto s2 = (*st)->compute_sgm(ImageView<const std::uint8_t>(bl.data(), w, h, pitch),
                                     ImageView<const std::uint8_t>(br.data(), w, h, pitch));
        expect(s2.has_value() && same_bytes(s2->disparity, disp), "strided input views");
    }

    // Fused rectification, as tests/depth/test_metal_stereo.cpp:70-103 checks it: raw frames at twice
    // the stereo resolution through slightly rotated, scaled remap tables; the rectified pair is
    // bit-exact with remap + downsample2 and the disparity with sgm_disparity on that pair.
    {
        const int mw = 2 * w, mh = 2 * h, rw = 2 * w + 30, rh = 2 * h + 20;
        ImageU8 raw_l, raw_r;
        make_pair(rw, rh, 18, raw_l, raw_r);
        RemapTable ml{mw, mh, ImageF32(mw, mh), ImageF32(mw, mh)}, mr = ml;
        for (int y = 0; y < mh; ++y)
            for (int x = 0; x < mw; ++x) {
                const double u = x - mw / 2.0, v = y - mh / 2.0;
                ml.map_x(x, y) = static_cast<float>(1.02 * (u * 0.9998 - v * 0.02) + rw / 2.0);
                ml.map_y(x, y) = static_cast<float>(1.02 * (u * 0.02 + v * 0.9998) + rh / 2.0);
                mr.map_x(x, y) = static_cast<float>(0.99 * u + rw / 2.0 + 0.37);
                mr.map_y(x, y) = static_cast<float>(0.99 * v + rh / 2.0 - 0.61);
            }
        auto none = (*st)->compute_raw(raw_l.view(), raw_r.view());
        expect(!none && none.error().message == "set_rectification() not called", "compute_raw before set_rectification");
        RemapTable wrong{mw + 2, mh, ImageF32(mw + 2, mh), ImageF32(mw + 2, mh)};
        auto bad_map = (*st)->set_rectification(wrong, wrong, rw, rh);
        expect(!bad_map && bad_map.error().code == Errc::invalid_argument, "remap resolution != 2x stereo -> invalid_argument");
        expect((*st)->set_rectification(ml, mr, rw, rh).has_value(), "set_rectification");
        ImageU8 gl, gr;
        auto rr = (*st)->compute_raw(raw_l.view(), raw_r.view(), &gl, &gr);
        expect(rr.has_value(), "compute_raw");
        std::vector<std::uint8_t> hl(n), hr(n);
        orc_rectify_half(raw_l.data(), rw, rh, rw, ml.map_x.data(), ml.map_y.data(), mw, mh, hl.data(), 0);
        orc_rectify_half(raw_r.data(), rw, rh, rw, mr.map_x.data(), mr.map_y.data(), mw, mh, hr.data(), 0);
        expect(same_bytes(gl, hl) && same_bytes(gr, hr), "rectified pair bit-exact");
        std::vector<float> rd(n);
        orc_outputs ro{};
        ro.disparity = rd.data();
        expect(orc_sgm(hl.data(), hr.data(), w, w, h, &op, &ro, 0) == 0, "oracle on the rectified pair");
        if (rr) expect(same_bytes(rr->disparity, rd), "compute_raw disparity bit-exact");
        ImageU8 short_raw(rw, rh - 1);
        auto mism = (*st)->compute_raw(short_raw.view(), short_raw.view());
        expect(!mism && mism.error().message == "raw size mismatch", "raw size mismatch");
    }

    // Point frames (compute_frame_raw) and packing against disparity_to_points and the session layout,
    // through the SGM-only context with the rectification above.
    {
        ImageU8 raw_l, raw_r;
        const int rw = 2 * w + 30, rh = 2 * h + 20;
        make_pair(rw, rh, 18, raw_l, raw_r);
        auto none = (*st)->compute_frame_raw(raw_l.view(), raw_r.view());
        expect(!none && none.error().message == "set_point_params() not called", "compute_frame_raw before set_point_params");
        const RectifiedGeometry g{420.0, w / 2.0, h / 2.0, 60.0, 0.0};
        expect((*st)->set_point_params(g, 50.0f, 20000.0f).has_value(), "set_point_params");
        auto fr = (*st)->compute_frame_raw(raw_l.view(), raw_r.view());
        auto dr = (*st)->compute_raw(raw_l.view(), raw_r.view());
        expect(fr.has_value() && dr.has_value(), "compute_frame_raw");
        if (fr && dr) {
            std::vector<float> pts(3 * n), nrm(3 * n), wts(n);
            const orc_geometry og{g.f, g.cx, g.cy, g.baseline, g.cx_offset};
            const orc_point_params opp{50.0f, 20000.0f, 4.0f, 0.25f, 0.0f};
            orc_disparity_to_points(dr->disparity.data(), dr->confidence.data(), w, h, &og, &opp, pts.data(), nrm.data(),
                                    wts.data(), 0);
            bool same = true;
            int valid = 0;
            for (std::size_t i = 0; i < n; ++i) {
                for (int c = 0; c < 3; ++c)
                    same &= fr->points_xyzw[4 * i + c] == pts[3 * i + c] && fr->normals_xyzw[4 * i + c] == nrm[3 * i + c];
                same &= fr->weights[i] == wts[i];
                valid += pts[3 * i + 2] > 0;
            }
            expect(same && valid > static_cast<int>(n / 4), "compute_frame_raw == disparity_to_points of compute_raw");
            auto packed = pack_depth(*fr);
            std::vector<std::uint8_t> want(3 * n);
            orc_pack_depth(pts.data(), wts.data(), w, h, want.data());
            expect(packed.has_value() && *packed == want, "pack_depth == the session layout");
        }
    }

    // Error behaviour of MetalStereo (metal_stereo.cpp:115-120, 235-236).
    ImageU8 small(w - 1, h);
    auto bad = (*st)->compute(small.view(), r.view());
    expect(!bad && bad.error().code == Errc::invalid_argument, "size mismatch -> invalid_argument");
    // The reference defaults: two pyramid levels and every stage on (stereo.hpp:42-89), against the
    // oracle's compute_disparity, without and with the full-resolution pair (detail refinement).
    StereoParams dflt;
    dflt.sgm = sp.sgm;
    dflt.sgm.num_disparities = 16;  // at the quarter-resolution SGM level
    auto st2 = HipStereo::create(dflt, w, h);
    expect(st2.has_value(), "create with reference defaults");
    if (st2) {
        ImageU8 fl(2 * w, 2 * h), fr(2 * w, 2 * h);  // pixel-replicated: downsample2 gives back (l, r)
        for (int y = 0; y < 2 * h; ++y)
            for (int x = 0; x < 2 * w; ++x) {
                fl(x, y) = l(x / 2, y / 2);
                fr(x, y) = r(x / 2, y / 2);
            }
        const orc_params op2 = oracle_params(dflt);
        orc_stereo_params os{};
        os.pyramid_levels = dflt.pyramid_levels;
        os.refine_search_radius = dflt.refine.search_radius;
        os.refine_zncc_radius = dflt.refine.zncc_radius;
        os.refine_min_zncc = dflt.refine.min_zncc;
        os.refine_reject_edge_winners = dflt.refine.reject_edge_winners;
        os.detail_enabled = dflt.detail.enabled;
        os.detail_radius = dflt.detail.radius;
        os.detail_search = dflt.detail.search;
        os.detail_step = dflt.detail.step;
        os.detail_min_zncc = dflt.detail.min_zncc;
        os.filter_occlusion = dflt.filter.occlusion;
        os.filter_occlusion_tolerance = dflt.filter.occlusion_tolerance;
        os.filter_spikes = dflt.filter.spikes;
        os.filter_spike_threshold = dflt.filter.spike_threshold;
        os.speckle_max_region_size = dflt.speckle.max_region_size;
        os.speckle_max_diff = dflt.speckle.max_diff;
        std::vector<float> d1(n), c1(n), d2(n), c2(n);
        expect(orc_compute_disparity(l.data(), r.data(), w, w, h, nullptr, nullptr, 0, 0, 0, &op2, &os, d1.data(),
                                     c1.data(), 0) == 0 &&
                   orc_compute_disparity(l.data(), r.data(), w, w, h, fl.data(), fr.data(), 2 * w, 2 * w, 2 * h, &op2,
                                         &os, d2.data(), c2.data(), 0) == 0,
               "oracle compute_disparity");
        auto u = (*st2)->compute(l.view(), r.view());
        expect(u.has_value() && same_bytes(u->disparity, d1) && same_bytes(u->confidence, c1),
               "compute == compute_disparity (pyramid, refinement, filters, speckles)");
        auto v = (*st2)->compute_full(l.view(), r.view(), fl.view(), fr.view());
        expect(v.has_value() && same_bytes(v->disparity, d2) && same_bytes(v->confidence, c2),
               "compute_full == compute_disparity with the detail refinement");
        int valid = 0;
        for (std::size_t i = 0; i < n; ++i) valid += valid_disparity(d2[i]);
        expect(valid > static_cast<int>(n / 2), "most pixels valid after every stage");
        auto s3 = (*st2)->compute_sgm(l.view(), r.view());
        expect(!s3 && s3.error().code == Errc::unsupported, "compute_sgm with pyramid levels -> unsupported");
    }
    // The post-filters as building blocks (stereo.hpp:117-118, 124-125) on a map of noisy 5x5 blocks with
    // holes, against the oracle's remove_speckles and filter_disparity.
    {
        const int mw = 301, mh = 203;
        ImageF32 map(mw, mh);
        std::uint32_t rng = 12345u;
        auto next = [&rng] { return (rng = rng * 1664525u + 1013904223u) >> 8; };
        std::vector<float> block((mw / 5 + 1) * (mh / 5 + 1));
        for (auto& b : block) b = 5.0f + static_cast<float>(next() % 5000) / 100.0f;
        for (int y = 0; y < mh; ++y)
            for (int x = 0; x < mw; ++x)
                map(x, y) = next() % 16 == 0 ? kInvalidDisparity
                                             : block[(y / 5) * (mw / 5 + 1) + x / 5] + static_cast<float>(next() % 60) / 100.0f;
        orc_stereo_params os{};
        os.filter_occlusion = 1;
        os.filter_occlusion_tolerance = 1.0f;
        os.filter_spikes = 1;
        os.filter_spike_threshold = 1.0f;
        os.speckle_max_region_size = 40;
        os.speckle_max_diff = 1.0f;
        std::vector<float> want(map.data(), map.data() + map.size());
        orc_filter_disparity(want.data(), mw, mh, &os, 0);
        ImageF32 got = map;
        auto fres = filter_disparity(got, FilterParams{});
        expect(fres.has_value() && same_bytes(got, want), "filter_disparity == the oracle's");
        expect(orc_remove_speckles(want.data(), mw, mh, &os) == 0, "oracle remove_speckles");
        auto sres = remove_speckles(got, SpeckleParams{40, 1.0f});
        expect(sres.has_value() && same_bytes(got, want), "remove_speckles == the oracle's");
    }
    StereoParams big = sp;
    big.sgm.num_disparities = 600;
    auto st3 = HipStereo::create(big, w, h);
    expect(!st3 && st3.error().code == Errc::invalid_argument, "num_disparities > 512 -> invalid_argument");
    StereoParams heavy = sp;
    heavy.sgm.p2 = 250;
    auto st4 = HipStereo::create(heavy, w, h);
    expect(!st4 && st4.error().code == Errc::invalid_argument, "census cells + P2 > 255 -> invalid_argument");

    std::printf("%s (%d failures)\n", failures ? "selftest FAILED" : "selftest ok", failures);
    return failures ? 1 : 0;
}
