// gfx950 kernels of the stages compute_disparity runs around SGM (libs/depth/src/stereo.cpp:98-326,
// 348-359, 465-536): the image pyramid, slanted-window ZNCC refinement to each finer level, detail
// refinement on the full-resolution pair, the occlusion and spike filters, and speckle removal by
// connected components.
//
// The ZNCC scores accumulate in double in the reference's order (window rows outer, columns inner)
// and the float sample positions keep its operation order; with -ffp-contract=off and IEEE double
// add/mul/div/sqrt every score, and so every winner and subpixel offset, equals the CPU's. The sums
// of products use fused multiply-adds: a sample is a float (24-bit significand) and a left pixel a
// byte, so ib * ib (<= 48 bits) and ia * ib (<= 32 bits) are exact in double, and rounding once after
// the add gives what rounding the exact product and then the sum gives.
#include <climits>

#include "sgm_internal.h"

namespace sgm {
namespace {

constexpr float kInvalid = -1.0e9f;
constexpr int kTw = 16, kTh = 16;  // 2-D blocks of 256 threads
#define RK_LDS __attribute__((address_space(3)))

__device__ __forceinline__ bool valid_d(float d) { return d > -1.0e8f; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }

// subpixel_offset, stereo.cpp:92-96.
__device__ __forceinline__ float subpixel_offset(float cm, float c0, float cp) {
    const float denom = cm - 2.0f * c0 + cp;
    if (denom <= 0.0f) return 0.0f;
    const float v = 0.5f * (cm - cp) / denom;
    return clampf(v, -0.5f, 0.5f);
}

// The n/2-th smallest valid value of the in-image 3x3 neighbourhood and the count n (median3,
// stereo.cpp:129-147, and the spike filter's median, :306-322).
__device__ __forceinline__ float median9(const float* in, int w, int h, int x, int y, int& n_out) {
    float buf[9];
    int n = 0;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = x + dx, yy = y + dy;
            float q = kInvalid;
            if (xx >= 0 && yy >= 0 && xx < w && yy < h) q = in[(int64_t)yy * w + xx];
            const bool ok = q > -1.0e8f;
            buf[dy * 3 + dx + 4] = ok ? q : __builtin_inff();  // +inf sorts after every valid value
            n += ok ? 1 : 0;
        }
    // A 25-comparator sorting network (it sorts every 0/1 input, so every input).
    constexpr int kNet[25][2] = {{0, 3}, {1, 7}, {2, 5}, {4, 8}, {0, 7}, {2, 4}, {3, 8}, {5, 6}, {0, 2},
                                 {1, 3}, {4, 5}, {7, 8}, {1, 4}, {3, 6}, {5, 7}, {0, 1}, {2, 4}, {3, 5},
                                 {6, 8}, {2, 3}, {4, 5}, {6, 7}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
    for (int c = 0; c < 25; ++c) {
        const float lo = fminf(buf[kNet[c][0]], buf[kNet[c][1]]), hi = fmaxf(buf[kNet[c][0]], buf[kNet[c][1]]);
        buf[kNet[c][0]] = lo;
        buf[kNet[c][1]] = hi;
    }
    float sel = buf[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) sel = (n / 2) == i ? buf[i] : sel;
    n_out = n;
    return sel;
}

// ------------------------------------------------------------------------------- pyramid
// downsample2, stereo.cpp:348-359; both images in one launch (grid z).

__global__ __launch_bounds__(kTw* kTh) void downsample2_kernel(const uint8_t* s0, const uint8_t* s1, int64_t spitch,
                                                               uint8_t* d0, uint8_t* d1, int64_t dpitch, int ow, int oh) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= ow || y >= oh) return;
    const uint8_t* s = blockIdx.z ? s1 : s0;
    const uint8_t* r0 = s + (int64_t)(2 * y) * spitch + 2 * x;
    const uint8_t* r1 = r0 + spitch;
    (blockIdx.z ? d1 : d0)[(int64_t)y * dpitch + x] = (uint8_t)(((uint32_t)r0[0] + r0[1] + r1[0] + r1[1] + 2) >> 2);
}

__global__ __launch_bounds__(kTw* kTh) void median3_kernel(const float* in, float* out, int w, int h) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    float m = kInvalid;
    if (valid_d(in[(int64_t)y * w + x])) {
        int n;
        m = median9(in, w, h, x, y, n);
    }
    out[(int64_t)y * w + x] = m;
}

// ------------------------------------------------------------------------------- refine_level
// stereo.cpp:151-207. One thread per fine pixel. The left window's sums do not depend on the
// candidate and are formed once, in the order the reference forms them for every candidate.

// zncc_slanted (stereo.cpp:101-125) with the left sums given.
__device__ float zncc_slanted(const uint8_t* a, const uint8_t* b, int64_t pitch, int bw, int x, int y, float d, float gx,
                              float gy, int r, double sa, double saa) {
    double sb = 0, sbb = 0, sab = 0;
    const int n = (2 * r + 1) * (2 * r + 1);
    for (int v = -r; v <= r; ++v) {
        const uint8_t* ra = a + (int64_t)(y + v) * pitch;
        const uint8_t* rb = b + (int64_t)(y + v) * pitch;
        for (int u = -r; u <= r; ++u) {
            const float xb = (float)(x + u) - (d + gx * (float)u + gy * (float)v);
            const float fl = __builtin_floorf(xb);
            // floor(xb) in [0, bw - 2]  <=>  0 <= xb < bw - 1 (bw < 2^24; NaN fails)
            if (!(xb >= 0.0f && xb < (float)(bw - 1))) return -2.0f;
            const int x0 = (int)fl;
            const float f = xb - fl;
            const double ib = (float)rb[x0] * (1.0f - f) + (float)rb[x0 + 1] * f;
            const double ia = ra[x + u];
            sb += ib;
            sbb = __builtin_fma(ib, ib, sbb);
            sab = __builtin_fma(ia, ib, sab);
        }
    }
    const double va = saa - sa * sa / n;
    const double vb = sbb - sb * sb / n;
    if (va < 1e-6 || vb < 1e-6) return -1.0f;
    return (float)((sab - sa * sb / n) / __builtin_sqrt(va * vb));
}

// Fixed window widths: the samples of a window row are unrolled, so their loads are issued together
// and waited for once per row; a window that leaves the right image clears a flag instead of leaving
// the loop (every load is clamped into the row and the whole score is discarded), so lanes do not
// diverge. Rows stay a loop (fully unrolled windows keep every load in flight and run out of
// registers). The sums run in the same order as above.
//
// Within a window row the sample position xb strictly increases with the column: each column step
// adds 1 - gx >= 0.1 (|gx| <= 0.9) while the float roundings of the position's terms change it by a
// few ulps of values below 2^15 (< 0.01). So every sample of the row lies in [0, bw - 1) exactly when
// the first is >= 0 and the last is < bw - 1, and only those two are tested; finite slopes and
// disparities leave no NaN to test for.

// The two taps rb[x0], rb[x0 + 1] of a bilinear sample as one 16-bit load (unaligned where x0 is odd).
typedef uint16_t u16_unaligned __attribute__((aligned(1)));
__device__ __forceinline__ void taps(const uint8_t* p, float& a, float& b) {
    const uint32_t v = *reinterpret_cast<const u16_unaligned*>(p);
    a = (float)(v & 0xFFu);
    b = (float)(v >> 8);
}

// The left pixel and the fraction of a sample position, for the windowed kernels: floor(xb) clamped
// into [lo, hi], and xb - floor(xb), wherever xb >= 0, which holds for every sample of a window that
// is scored (the others' scores are discarded, and their left pixel only has to stay inside
// [lo, hi]). For xb >= 0 and integers 0 <= lo <= hi, truncating xb clamped into [lo, hi] gives
// floor(xb) clamped, and v_fract_f32 returns xb - floor(xb), exact for 0 <= xb < 2^24.
__device__ __forceinline__ int left_pixel(float xb, float lo, float hi, float& f) {
    f = __builtin_amdgcn_fractf(xb);
    return (int)__builtin_amdgcn_fmed3f(xb, lo, hi);
}

// The bilinear sample at xb of row rb, its left pixel clamped into [0, bw - 2].
__device__ __forceinline__ double sample(const uint8_t* rb, int bw, float xb) {
    float f;
    const int x0 = left_pixel(xb, 0.0f, (float)(bw - 2), f);
    __builtin_assume(x0 >= 0);
    float a, b;
    taps(rb + x0, a, b);
    return a * (1.0f - f) + b * f;
}

__device__ __forceinline__ float zncc_finish(double sa, double saa, double sb, double sbb, double sab, double n) {
    const double va = saa - sa * sa / n;
    const double vb = sbb - sb * sb / n;
    if (va < 1e-6 || vb < 1e-6) return -1.0f;
    return (float)((sab - sa * sb / n) / __builtin_sqrt(va * vb));
}

// Right-image samples of the fixed-radius refine kernel come from global memory, or, when the block's
// samples fit, from an LDS copy of the rows they read: entry c - c0 of a copied row holds pixel c as a
// float, for the columns c0 .. c1 the block can reach. The left pixel is then clamped into
// [c0, c1 - 1] instead of [0, bw - 2]: every sample inside the image reads the same two pixels (a
// byte converted to float is exact), and the others (whose whole window score is discarded) stay
// inside the copy.
struct RightRows {
    const uint8_t* b;  // global image: the window's first row, rows `pitch` apart
    int64_t pitch;
    uint32_t t;        // LDS copy: byte address of the window's first row minus 4 c0 (rows kRightCols apart)
    float lo, hi;      // clamp of the left pixel: [0, bw - 2] or [c0, c1 - 1]
};
constexpr int kRightCols = 256;

// The bilinear sample at xb of row `row` (counted from the window's first).
template <bool TILE>
__device__ __forceinline__ double sample_row(const RightRows& rr, int row, float xb) {
    float f;
    const int x0 = left_pixel(xb, rr.lo, rr.hi, f);
    __builtin_assume(x0 >= 0);
    if constexpr (TILE) {
        const RK_LDS float* t = (const RK_LDS float*)(uintptr_t)(rr.t + (uint32_t)(row * kRightCols * 4) + 4u * (uint32_t)x0);
        return t[0] * (1.0f - f) + t[1] * f;
    } else {
        float a, b;
        taps(rr.b + (int64_t)row * rr.pitch + x0, a, b);
        return a * (1.0f - f) + b * f;
    }
}

// zncc_slanted (stereo.cpp:101-125) for radius R; gxu[u + R] = gx * u. The left window comes from the
// block's LDS tile (ta: the pixel, rows TS values apart), already converted to double; rr's rows
// start at the window's first row.
template <int R, int TS, bool TILE>
__device__ float zncc_slanted_w(const RK_LDS double* ta, const RightRows& rr, int bw, int x, float d, const float* gxu,
                                float gy, double sa, double saa) {
    constexpr int N = 2 * R + 1;
    double sb = 0, sbb = 0, sab = 0;
    bool ok = true;
#pragma unroll 1
    for (int v = -R; v <= R; ++v) {
        const RK_LDS double* ra = ta + v * TS;
        const float gyv = gy * (float)v;
#pragma unroll
        for (int u = -R; u <= R; ++u) {
            const float xb = (float)(x + u) - ((d + gxu[u + R]) + gyv);
            if (u == -R) ok = ok && xb >= 0.0f;
            if (u == R) ok = ok && xb < (float)(bw - 1);
            const double ib = sample_row<TILE>(rr, v + R, xb);
            const double ia = ra[u];
            sb += ib;
            sbb = __builtin_fma(ib, ib, sbb);
            sab = __builtin_fma(ia, ib, sab);
        }
    }
    if (!ok) return -2.0f;
    return zncc_finish(sa, saa, sb, sbb, sab, (double)(N * N));
}

// RR > 0: a fixed radius; RR == 0: any radius (R.radius) through the loop form above.
template <int RR>
__global__ __launch_bounds__(kTw* kTh) void refine_kernel(RefineDesc R) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    const int tid = (int)(threadIdx.y * kTw + threadIdx.x);
    // Fixed radius: the block's left pixels with an RR halo, as doubles (0 outside the image, where no
    // scored window reaches), and, when they fit, the right-image columns its samples reach.
    constexpr int TW = kTw + 2 * RR, TH = kTh + 2 * RR;
    __shared__ double tile[RR > 0 ? TW * TH : 1];
    __shared__ float rtile[RR > 0 ? TH * kRightCols : 1];
    __shared__ int rlim[2];
    const int bx = (int)blockIdx.x * kTw - RR, by = (int)blockIdx.y * kTh - RR;
    if constexpr (RR > 0) {
        for (int i = tid; i < TW * TH; i += kTw * kTh) {
            const int ly = i / TW, lx = i - ly * TW;
            const int gx = bx + lx, gy = by + ly;
            tile[i] = (gx >= 0 && gy >= 0 && gx < R.w && gy < R.h) ? (double)R.left[(int64_t)gy * R.pitch + gx] : 0.0;
        }
        if (tid == 0) {
            rlim[0] = INT_MAX;
            rlim[1] = INT_MIN;
        }
    }
    const bool inside = x < R.w && y < R.h;
    const int r = RR > 0 ? RR : R.radius;
    const float kStep = 0.5f;
    const int nd = min(127, (int)((float)(2 * R.search_radius) / kStep) + 1);
    bool work = false;
    float d0 = 0.0f, gx = 0.0f, gy = 0.0f;
    if (inside && y >= r && y < R.h - r && x >= r && x < R.w - r) {
        const int cy = min(y / 2, R.ch - 1), cx = min(x / 2, R.cw - 1);
        const float* C = R.coarse;
        d0 = C[(int64_t)cy * R.cw + cx];
        if (valid_d(d0)) {
            work = true;
            // The slope lambda, stereo.cpp:159-170.
            if (cx > 0 && cx + 1 < R.cw) {
                const float l = C[(int64_t)cy * R.cw + cx - 1], rr = C[(int64_t)cy * R.cw + cx + 1];
                if (valid_d(l) && valid_d(rr)) {
                    const float g = 0.5f * (rr - l);
                    if (fabsf(rr - d0) < 2 && fabsf(d0 - l) < 2) gx = g;
                }
            }
            if (cy > 0 && cy + 1 < R.ch) {
                const float u = C[(int64_t)(cy - 1) * R.cw + cx], dn = C[(int64_t)(cy + 1) * R.cw + cx];
                if (valid_d(u) && valid_d(dn)) {
                    const float g = 0.5f * (dn - u);
                    if (fabsf(dn - d0) < 2 && fabsf(d0 - u) < 2) gy = g;
                }
            }
            gx = clampf(gx, -0.9f, 0.9f);
            gy = clampf(gy, -0.9f, 0.9f);
        }
    }
    const float lo = 2.0f * d0 - (float)R.search_radius;
    bool tiled = false;
    RightRows rr{R.right + (int64_t)(y - r) * R.pitch, R.pitch, 0u, 0.0f, (float)(R.w - 2)};
    if constexpr (RR > 0) {
        __syncthreads();
        if (work) {
            // Columns of this pixel's samples, with a margin of two beyond the slopes' reach.
            const float m = 1.8f * (float)RR + 2.0f;
            atomicMin(&rlim[0], (int)__builtin_floorf((float)(x - RR) - (lo + kStep * (float)(nd - 1)) - m));
            atomicMax(&rlim[1], (int)__builtin_floorf((float)(x + RR) - lo + m) + 1);
        }
        __syncthreads();
        const int c0 = min(max(rlim[0], 0), R.w - 2), c1 = min(max(rlim[1], c0 + 1), R.w - 1);
        tiled = rlim[0] <= rlim[1] && c1 - c0 + 1 <= kRightCols;
        if (tiled) {
            const int nc = c1 - c0 + 1;
            for (int i = tid; i < TH * nc; i += kTw * kTh) {
                const int ly = i / nc, c = c0 + (i - ly * nc), gy = by + ly;
                rtile[ly * kRightCols + (c - c0)] = (gy >= 0 && gy < R.h) ? (float)R.right[(int64_t)gy * R.pitch + c] : 0.0f;
            }
            // (Unsigned arithmetic: the offset by -4 c0 wraps, the sampled addresses stay in the copy.)
            rr.t = (uint32_t)(uintptr_t)((const RK_LDS float*)rtile + ((int)threadIdx.y + RR - r) * kRightCols) -
                   4u * (uint32_t)c0;
            rr.lo = (float)c0;
            rr.hi = (float)(c1 - 1);
        }
        __syncthreads();
    }
    if (!inside) return;
    const int64_t o = (int64_t)y * R.w + x;
    float out_d = kInvalid, out_c = 0.0f;
    if (work) {
        const RK_LDS double* tc = (const RK_LDS double*)tile + (threadIdx.y + RR) * TW + threadIdx.x + RR;
        double sa = 0, saa = 0;
        if constexpr (RR > 0) {
            for (int v = -RR; v <= RR; ++v)
#pragma unroll
                for (int u = -RR; u <= RR; ++u) {
                    const double ia = tc[v * TW + u];
                    sa += ia;
                    saa += ia * ia;
                }
        } else {
            for (int v = -r; v <= r; ++v) {
                const uint8_t* ra = R.left + (int64_t)(y + v) * R.pitch;
                for (int u = -r; u <= r; ++u) {
                    const double ia = ra[x + u];
                    sa += ia;
                    saa += ia * ia;
                }
            }
        }
        float gxu[RR > 0 ? 2 * RR + 1 : 1];
        if constexpr (RR > 0) {
#pragma unroll
            for (int u = -RR; u <= RR; ++u) gxu[u + RR] = gx * (float)u;
        }
        // Only the winner and its two neighbours' scores matter afterwards.
        float best = -2.0f, prev = -3.0f, best_prev = -3.0f, best_next = -3.0f;
        int best_k = -1;
        for (int k = 0; k < nd; ++k) {
            const float d = lo + kStep * (float)k;
            float s;
            if constexpr (RR > 0) {
                if (tiled) s = zncc_slanted_w<RR, TW, true>(tc, rr, R.w, x, d, gxu, gy, sa, saa);
                else s = zncc_slanted_w<RR, TW, false>(tc, rr, R.w, x, d, gxu, gy, sa, saa);
            } else {
                s = zncc_slanted(R.left, R.right, R.pitch, R.w, x, y, d, gx, gy, r, sa, saa);
            }
            if (s > best) {
                best = s;
                best_k = k;
                best_prev = prev;
            } else if (k == best_k + 1) {
                best_next = s;
            }
            prev = s;
        }
        bool keep = !(best_k < 0 || best < R.min_zncc);
        if (keep && R.reject_edge && (best_k == 0 || best_k == nd - 1)) keep = false;
        if (keep) {
            float d = lo + kStep * (float)best_k;
            if (R.subpixel && best_k > 0 && best_k + 1 < nd && best_prev > -2.0f && best_next > -2.0f)
                d += kStep * subpixel_offset(1.0f - best_prev, 1.0f - best, 1.0f - best_next);
            out_d = d;
            out_c = best;
        }
    }
    R.disp[o] = out_d;
    R.conf[o] = out_c;
}

// ------------------------------------------------------------------------------- refine_detail
// stereo.cpp:211-284. One thread per half-resolution pixel; pixels it does not change keep the
// first estimate (the output starts as a copy).

__device__ float zncc_window(const uint8_t* a, const uint8_t* b, int64_t pitch, int bw, int x0, int y0, int n, float cx,
                             float cy, float D, float gx, float gy, double sa, double saa) {
    double sb = 0, sbb = 0, sab = 0;
    for (int j = 0; j < n; ++j) {
        const int Y = y0 + j;
        const uint8_t* ra = a + (int64_t)Y * pitch;
        const uint8_t* rb = b + (int64_t)Y * pitch;
        const float dy = gy * ((float)Y - cy);
        for (int i = 0; i < n; ++i) {
            const int X = x0 + i;
            const float xb = (float)X - (D + gx * ((float)X - cx) + dy);
            const float fl = __builtin_floorf(xb);
            if (!(xb >= 0.0f && xb < (float)(bw - 1))) return -2.0f;
            const int xi = (int)fl;
            const float f = xb - fl;
            const double ib = (float)rb[xi] * (1.0f - f) + (float)rb[xi + 1] * f;
            const double ia = ra[X];
            sb += ib;
            sbb = __builtin_fma(ib, ib, sbb);
            sab = __builtin_fma(ia, ib, sab);
        }
    }
    const double m = (double)n * n;
    const double va = saa - sa * sa / m;
    const double vb = sbb - sb * sb / m;
    if (va < 1e-6 || vb < 1e-6) return -1.0f;
    return (float)((sab - sa * sb / m) / __builtin_sqrt(va * vb));
}

__device__ __forceinline__ float detail_grad(float l, float c, float r) {
    if (!valid_d(l) || !valid_d(r) || fabsf(r - c) >= 1.0f || fabsf(c - l) >= 1.0f) return 0.0f;
    return clampf(0.5f * (r - l), -0.9f, 0.9f);
}

// zncc_window (stereo.cpp:211-239) for an N x N window; gxx[i] = gx * (x0 + i - cx). The left window
// comes from the block's LDS tile (ta: the window's first pixel, rows TS values apart) as doubles.
template <int N, int TS>
__device__ float zncc_window_w(const RK_LDS double* ta, const uint8_t* b, int64_t pitch, int bw, int x0, int y0, float cy,
                               float D, const float* gxx, float gy, double sa, double saa) {
    double sb = 0, sbb = 0, sab = 0;
    bool ok = true;
#pragma unroll 1
    for (int j = 0; j < N; ++j) {
        const int Y = y0 + j;
        const RK_LDS double* ra = ta + j * TS;
        const uint8_t* rb = b + (int64_t)Y * pitch;
        const float dy = gy * ((float)Y - cy);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const float xb = (float)(x0 + i) - ((D + gxx[i]) + dy);
            if (i == 0) ok = ok && xb >= 0.0f;
            if (i == N - 1) ok = ok && xb < (float)(bw - 1);
            const double ib = sample(rb, bw, xb);
            const double ia = ra[i];
            sb += ib;
            sbb = __builtin_fma(ib, ib, sbb);
            sab = __builtin_fma(ia, ib, sab);
        }
    }
    if (!ok) return -2.0f;
    return zncc_finish(sa, saa, sb, sbb, sab, (double)N * N);
}

// NW > 0: a fixed window (2 * radius); NW == 0: any radius through the loop form above.
template <int NW>
__global__ __launch_bounds__(kTw* kTh) void detail_kernel(DetailDesc D) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    // Fixed window: the full-resolution left pixels of the block's windows (thread (tx, ty) starts at
    // tile (2 tx, 2 ty)), as doubles (0 outside the image, where no scored window reaches).
    constexpr int TW = 2 * kTw - 2 + NW, TH = 2 * kTh - 2 + NW;
    __shared__ double tile[NW > 0 ? TW * TH : 1];
    if constexpr (NW > 0) {
        const int bx = 2 * (int)blockIdx.x * kTw - NW / 2 + 1, by = 2 * (int)blockIdx.y * kTh - NW / 2 + 1;
        for (int i = (int)(threadIdx.y * kTw + threadIdx.x); i < TW * TH; i += kTw * kTh) {
            const int ly = i / TW, lx = i - ly * TW;
            const int gx = bx + lx, gy = by + ly;
            tile[i] = (gx >= 0 && gy >= 0 && gx < D.W && gy < D.H) ? (double)D.fl[(int64_t)gy * D.fpitch + gx] : 0.0;
        }
        __syncthreads();
    }
    const RK_LDS double* tw = (const RK_LDS double*)tile + 2 * threadIdx.y * TW + 2 * threadIdx.x;
    if (x >= D.w || y >= D.h) return;
    const float* F = D.first;
    const int w = D.w;
    const float d = F[(int64_t)y * w + x];
    float out = d;
    const int n = NW > 0 ? NW : 2 * D.radius, K = D.K, nc = 2 * K + 1;
    const int y0 = 2 * y - n / 2 + 1, x0 = 2 * x - n / 2 + 1;
    if (y >= 1 && y + 1 < D.h && y0 >= 0 && y0 + n <= D.H && x >= 1 && x + 1 < w && valid_d(d) && x0 >= 0 &&
        x0 + n <= D.W) {
        const float gx = detail_grad(F[(int64_t)y * w + x - 1], d, F[(int64_t)y * w + x + 1]);
        const float gy = detail_grad(F[(int64_t)(y - 1) * w + x], d, F[(int64_t)(y + 1) * w + x]);
        const float cx = 2.0f * (float)x + 0.5f, cy = 2.0f * (float)y + 0.5f;
        const float D0 = 2.0f * d;
        double sa = 0, saa = 0;
        if constexpr (NW > 0) {
            for (int j = 0; j < NW; ++j)
#pragma unroll
                for (int i = 0; i < NW; ++i) {
                    const double ia = tw[j * TW + i];
                    sa += ia;
                    saa += ia * ia;
                }
        } else {
            for (int j = 0; j < n; ++j) {
                const uint8_t* ra = D.fl + (int64_t)(y0 + j) * D.fpitch;
                for (int i = 0; i < n; ++i) {
                    const double ia = ra[x0 + i];
                    sa += ia;
                    saa += ia * ia;
                }
            }
        }
        float gxx[NW > 0 ? NW : 1];
        if constexpr (NW > 0) {
#pragma unroll
            for (int i = 0; i < NW; ++i) gxx[i] = gx * ((float)(x0 + i) - cx);
        }
        float best = -2.0f, prev = -3.0f, best_prev = -3.0f, best_next = -3.0f;
        int best_k = -1;
        for (int k = 0; k < nc; ++k) {
            const float Dk = D0 + D.step * (float)(k - K);
            float s;
            if constexpr (NW > 0)
                s = zncc_window_w<NW, TW>(tw, D.fr, D.fpitch, D.W, x0, y0, cy, Dk, gxx, gy, sa, saa);
            else
                s = zncc_window(D.fl, D.fr, D.fpitch, D.W, x0, y0, n, cx, cy, Dk, gx, gy, sa, saa);
            if (s > best) {
                best = s;
                best_k = k;
                best_prev = prev;
            } else if (k == best_k + 1) {
                best_next = s;
            }
            prev = s;
        }
        if (!(best_k <= 0 || best_k + 1 >= nc || best < D.min_zncc || best_prev <= -2.0f || best_next <= -2.0f)) {
            float Dv = D0 + D.step * (float)(best_k - K);
            if (D.subpixel) Dv += D.step * subpixel_offset(1.0f - best_prev, 1.0f - best, 1.0f - best_next);
            out = 0.5f * Dv;
        }
    }
    D.out[(int64_t)y * w + x] = out;
}

// ------------------------------------------------------------------------------- filters
// Occlusion (stereo.cpp:288-304): one block per row; the nearest claimant of each right pixel is an
// LDS max over order-preserving integer keys of the claimants' disparities.

__device__ __forceinline__ uint32_t order_key(float f) {
    const uint32_t b = __builtin_bit_cast(uint32_t, f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_value(uint32_t k) {
    return __builtin_bit_cast(float, (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ __launch_bounds__(kBlock) void occlusion_kernel(float* disp, int w, float tol) {
    extern __shared__ uint32_t nearest[];
    float* row = disp + (int64_t)blockIdx.x * w;
    const uint32_t none = order_key(kInvalid);
    for (int x = (int)threadIdx.x; x < w; x += kBlock) nearest[x] = none;
    __syncthreads();
    for (int x = (int)threadIdx.x; x < w; x += kBlock) {
        const float d = row[x];
        if (!valid_d(d)) continue;
        const long xr = lroundf((float)x - d);
        if (xr >= 0 && xr < w) atomicMax(&nearest[xr], order_key(d));
    }
    __syncthreads();
    for (int x = (int)threadIdx.x; x < w; x += kBlock) {
        const float d = row[x];
        if (!valid_d(d)) continue;
        const long xr = lroundf((float)x - d);
        if (xr >= 0 && xr < w && d < key_value(nearest[xr]) - tol) row[x] = kInvalid;
    }
}

// Spikes (stereo.cpp:305-323), out of place: in is the disparity after the occlusion filter.
__global__ __launch_bounds__(kTw* kTh) void spike_kernel(const float* in, float* out, int w, int h, float thr) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    const float d = in[(int64_t)y * w + x];
    float o = d;
    if (valid_d(d)) {
        int n;
        const float m = median9(in, w, h, x, y, n);
        if (n >= 3 && fabsf(d - m) > thr) o = kInvalid;
    }
    out[(int64_t)y * w + x] = o;
}

// ------------------------------------------------------------------------------- speckles
// remove_speckles (stereo.cpp:465-494) labels the connected components of the graph whose edges
// join 4-neighbours that are both valid with |d_a - d_b| <= max_diff; a flood fill and a union-find
// find the same components. Labels are pixel indices and every label points at a smaller or equal
// index, so every find ends; unions link the larger root under the smaller by atomic min.
//
// Three passes: (1) each 64x64 tile is labelled in LDS by one block of 1024 threads, every pixel then
// points at its tile-local root and each local root holds its local pixel count; (2) edges that cross
// tile borders unite the local roots in global memory (large tiles keep these unions few and their
// chains short); (3) each local root adds its count to its global root and points straight at it, so
// a pixel's global root is L[L[i]].

template <typename T>
__device__ __forceinline__ T atomic_ld(const T* p, int scope) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, scope);
}

// find_root with path halving: each node passed is pointed at its grandparent by an atomic minimum, so
// it only ever moves to a smaller ancestor and never undoes a concurrent union. (The tile-local pass
// uses it; in the global passes the extra atomic per hop cost more than the shorter paths saved.)
template <int Scope>
__device__ __forceinline__ int32_t find_root_h(int32_t* L, int32_t a) {
    for (;;) {
        const int32_t p = __hip_atomic_load(L + a, __ATOMIC_RELAXED, Scope);
        if (p == a) return a;
        const int32_t gp = __hip_atomic_load(L + p, __ATOMIC_RELAXED, Scope);
        if (gp == p) return p;
        __hip_atomic_fetch_min(L + a, gp, __ATOMIC_RELAXED, Scope);
        a = gp;
    }
}

template <int Scope>
__device__ __forceinline__ int32_t find_root(const int32_t* L, int32_t a) {
    int32_t p = __hip_atomic_load(L + a, __ATOMIC_RELAXED, Scope);
    while (p != a) {
        a = p;
        p = __hip_atomic_load(L + a, __ATOMIC_RELAXED, Scope);
    }
    return a;
}

template <int Scope>
__device__ void unite(int32_t* L, int32_t a, int32_t b) {
    for (;;) {
        a = Scope == __HIP_MEMORY_SCOPE_WORKGROUP ? find_root_h<Scope>(L, a) : find_root<Scope>(L, a);
        b = Scope == __HIP_MEMORY_SCOPE_WORKGROUP ? find_root_h<Scope>(L, b) : find_root<Scope>(L, b);
        if (a == b) return;
        if (a < b) {
            const int32_t old = atomicMin(L + b, a);
            if (old == b) return;
            b = old;
        } else {
            const int32_t old = atomicMin(L + a, b);
            if (old == a) return;
            a = old;
        }
    }
}

__device__ __forceinline__ bool joined(float d, float e, float max_diff) {
    return valid_d(e) && !(fabsf(e - d) > max_diff);
}

constexpr int kCcl = 64, kCclThreads = 1024, kCclRows = kCcl * kCcl / kCclThreads;
static_assert(kCcl == 64, "a wave spans one tile row: the ballots below are one bit a column");

// Tile-local components of a 64x64 tile. Wave v takes tile rows v, v + 16, v + 32, v + 48, a lane a
// column. A ballot of the row's horizontal joins gives every pixel its run (the longest stretch of
// joined neighbours), whose first pixel becomes its label at once: no atomics within a row. Vertical
// joins are then united once per pair of runs: a pixel whose left neighbour joins it in both rows and
// also joins down leaves the union to that neighbour (the same two runs). Labels point at smaller
// indices throughout, so every find ends. Each run's first pixel adds the run's length to its root.
__global__ __launch_bounds__(kCclThreads) void ccl_local_kernel(const float* disp, int32_t* L, uint32_t* local,
                                                                uint32_t* sizes, int w, int h, float max_diff) {
    __shared__ int32_t ll[kCcl * kCcl];
    __shared__ uint32_t cnt[kCcl * kCcl];
    __shared__ uint64_t rowj[kCcl];  // bit c: pixel c of the tile row joins pixel c + 1
    const int bx = (int)blockIdx.x * kCcl, by = (int)blockIdx.y * kCcl;
    const int lane = (int)threadIdx.x & (kCcl - 1), wv = (int)threadIdx.x / kCcl;
    const int x = bx + lane;
    bool v[kCclRows], jd[kCclRows];
    uint64_t runj[kCclRows];
#pragma unroll
    for (int j = 0; j < kCclRows; ++j) {
        const int ly = wv + j * (kCclThreads / kCcl), y = by + ly, t = ly * kCcl + lane;
        const bool in = x < w && y < h;
        const int64_t i = (int64_t)y * w + x;
        const float d = in ? disp[i] : kInvalid;
        v[j] = valid_d(d);
        const float dr = __shfl_down(d, 1, kCcl);  // the right neighbour when it is in the tile
        const bool jr = v[j] && lane + 1 < kCcl && x + 1 < w && joined(d, dr, max_diff);
        jd[j] = v[j] && ly + 1 < kCcl && y + 1 < h && joined(d, disp[i + w], max_diff);
        const uint64_t M = __ballot(jr);
        runj[j] = M;
        // The run's first column: the last s <= lane with s == 0 or pixel s - 1 not joining s.
        const uint64_t starts = ~(M << 1);
        const uint64_t upto = lane == kCcl - 1 ? ~0ull : ((2ull << lane) - 1);
        const int s0 = 63 - __builtin_clzll(starts & upto);
        ll[t] = v[j] ? ly * kCcl + s0 : -1;
        cnt[t] = 0;
        if (lane == 0) rowj[ly] = M;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCclRows; ++j) {
        const int ly = wv + j * (kCclThreads / kCcl), t = ly * kCcl + lane;
        const uint64_t D = __ballot(jd[j]);
        if (jd[j]) {
            const uint64_t below = rowj[ly + 1];  // ly + 1 < kCcl wherever jd holds
            const bool dup = lane > 0 && ((D & runj[j] & below) >> (lane - 1) & 1u);
            if (!dup) unite<__HIP_MEMORY_SCOPE_WORKGROUP>(ll, t, t + kCcl);
        }
    }
    __syncthreads();
    int32_t r[kCclRows];
#pragma unroll
    for (int j = 0; j < kCclRows; ++j) {
        const int ly = wv + j * (kCclThreads / kCcl), t = ly * kCcl + lane;
        r[j] = -1;
        if (v[j]) {
            r[j] = find_root<__HIP_MEMORY_SCOPE_WORKGROUP>(ll, t);
            const uint64_t M = runj[j];
            const bool head = lane == 0 || !((M >> (lane - 1)) & 1u);
            if (head) {
                // The run ends at the first column >= lane that does not join its right neighbour.
                const int e = __builtin_ctzll(~M & (~0ull << lane));
                atomicAdd(&cnt[r[j]], (uint32_t)(e - lane + 1));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kCclRows; ++j) {
        const int ly = wv + j * (kCclThreads / kCcl), y = by + ly, t = ly * kCcl + lane;
        if (x >= w || y >= h) continue;
        const int64_t i = (int64_t)y * w + x;
        L[i] = v[j] ? (by + r[j] / kCcl) * w + bx + (r[j] & (kCcl - 1)) : -1;
        local[i] = v[j] && r[j] == t ? cnt[t] : 0u;
        sizes[i] = 0;
    }
}

// Joins across tile borders, one thread a border pair: the first (tiles_x - 1) * h threads take the
// vertical borders (pixel (64k - 1, y) and its right neighbour), the rest the horizontal ones (pixel
// (x, 64k - 1) and the one below). A pair whose predecessor along the same tile edge (the row above,
// or the column to the left) is joined too, with both sides' links between the two pairs inside one
// tile, leaves the union to that predecessor: the same two components. (Links across another border
// are never relied on, so no two pairs can leave their union to each other.)
__global__ __launch_bounds__(kBlock) void ccl_border_kernel(const float* disp, int32_t* L, int w, int h, float max_diff,
                                                           int64_t n_vertical, int64_t n_total) {
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n_total) return;
    int x, y, xn, yn, xp, yp, xpn, ypn;  // the pair, its predecessor
    bool pred_local;
    if (k < n_vertical) {
        const int b = (int)(k / h);
        y = (int)(k - (int64_t)b * h);
        x = (b + 1) * kCcl - 1;
        xn = x + 1;
        yn = y;
        xp = x, yp = y - 1, xpn = xn, ypn = y - 1;
        pred_local = y % kCcl != 0;
    } else {
        const int64_t q = k - n_vertical;
        const int b = (int)(q / w);
        x = (int)(q - (int64_t)b * w);
        y = (b + 1) * kCcl - 1;
        xn = x;
        yn = y + 1;
        xp = x - 1, yp = y, xpn = x - 1, ypn = yn;
        pred_local = x % kCcl != 0;
    }
    if (xn >= w || yn >= h) return;
    const int32_t i = y * w + x, in_ = yn * w + xn;
    const float d = disp[i];
    if (!valid_d(d) || !joined(d, disp[in_], max_diff)) return;
    if (pred_local) {
        const int32_t ip = yp * w + xp, ipn = ypn * w + xpn;
        const float dp = disp[ip], dpn = disp[ipn];
        if (valid_d(dp) && joined(dp, dpn, max_diff) && joined(dp, d, max_diff) && valid_d(dpn) &&
            joined(dpn, disp[in_], max_diff))
            return;
    }
    unite<__HIP_MEMORY_SCOPE_AGENT>(L, i, in_);
}

__global__ __launch_bounds__(kTw* kTh) void ccl_accum_kernel(int32_t* L, const uint32_t* local, uint32_t* sizes, int w, int h) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    const int32_t i = y * w + x;
    const uint32_t n = local[i];
    if (n == 0) return;  // not a tile-local root
    const int32_t root = find_root<__HIP_MEMORY_SCOPE_AGENT>(L, i);
    atomicAdd(sizes + root, n);
    __hip_atomic_store(L + i, root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The end of compute_disparity (stereo.cpp:527-534): speckle regions removed, confidence 0 where
// the disparity is invalid (conf_src null: 1 elsewhere, the level-0 SGM case).
__global__ __launch_bounds__(kTw* kTh) void finalize_kernel(const float* disp, const int32_t* L, const uint32_t* sizes,
                                                            int max_region, const float* conf_src, float* out_disp,
                                                            float* out_conf, int w, int h) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    const int32_t i = y * w + x;
    float d = disp[i];
    if (L && valid_d(d) && sizes[L[L[i]]] < (uint32_t)max_region) d = kInvalid;  // L[i]: local root
    out_disp[i] = d;
    if (out_conf) out_conf[i] = valid_d(d) ? (conf_src ? conf_src[i] : 1.0f) : 0.0f;
}

// ------------------------------------------------------------------------------- rectify_full
// calib::remap at the maps' resolution (rectify.cpp:133-148), kept for the detail refinement; the
// stereo-resolution pair is then downsample2 of it, which equals the fused rectify_kernel's output.

__global__ __launch_bounds__(kTw* kTh) void rectify_full_kernel(RectifyDesc r, uint8_t* f0, uint8_t* f1, int64_t fpitch,
                                                                int fw, int fh) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= fw || y >= fh) return;
    const int side = (int)blockIdx.z;
    const uint8_t* src = r.raw[side];
    const float* m = r.map[side] + ((int64_t)y * r.map_w + x) * 2;
    const float sx = m[0], sy = m[1];
    uint32_t v = 0;
    if (sx >= 0.0f && sy >= 0.0f && sx < (float)(r.raw_w - 1) && sy < (float)(r.raw_h - 1)) {
        const float flx = __builtin_floorf(sx), fly = __builtin_floorf(sy);
        const float fx = sx - flx, fy = sy - fly;
        const uint8_t* p = src + (int64_t)(int)fly * r.raw_pitch + (int)flx;
        const float a = p[0], b = p[1], c = p[r.raw_pitch], d = p[r.raw_pitch + 1];
        const float s = (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
        v = (uint32_t)__builtin_fminf(s + 0.5f, 255.0f);
    }
    (side ? f1 : f0)[(int64_t)y * fpitch + x] = (uint8_t)v;
}

dim3 grid2d(int w, int h, int z = 1) { return dim3((unsigned)((w + kTw - 1) / kTw), (unsigned)((h + kTh - 1) / kTh), (unsigned)z); }

}  // namespace

hipError_t launch_downsample2(const uint8_t* s0, const uint8_t* s1, int64_t spitch, int sw, int sh, uint8_t* d0, uint8_t* d1,
                              int64_t dpitch, hipStream_t stream) {
    const int ow = sw / 2, oh = sh / 2;
    if (ow <= 0 || oh <= 0) return hipSuccess;
    hipLaunchKernelGGL(downsample2_kernel, grid2d(ow, oh, 2), dim3(kTw, kTh), 0, stream, s0, s1, spitch, d0, d1, dpitch, ow, oh);
    return hipGetLastError();
}

hipError_t launch_median3(const float* in, float* out, int w, int h, hipStream_t stream) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(median3_kernel, grid2d(w, h), dim3(kTw, kTh), 0, stream, in, out, w, h);
    return hipGetLastError();
}

hipError_t launch_refine(const RefineDesc& r, hipStream_t stream) {
    if (r.w <= 0 || r.h <= 0) return hipSuccess;
    const dim3 g = grid2d(r.w, r.h), b(kTw, kTh);
    switch (r.radius) {  // the window in registers up to 11 x 11
        case 1: hipLaunchKernelGGL(refine_kernel<1>, g, b, 0, stream, r); break;
        case 2: hipLaunchKernelGGL(refine_kernel<2>, g, b, 0, stream, r); break;
        case 3: hipLaunchKernelGGL(refine_kernel<3>, g, b, 0, stream, r); break;
        case 4: hipLaunchKernelGGL(refine_kernel<4>, g, b, 0, stream, r); break;
        case 5: hipLaunchKernelGGL(refine_kernel<5>, g, b, 0, stream, r); break;
        default: hipLaunchKernelGGL(refine_kernel<0>, g, b, 0, stream, r); break;
    }
    return hipGetLastError();
}

hipError_t launch_detail(const DetailDesc& d, hipStream_t stream) {
    if (d.w <= 0 || d.h <= 0) return hipSuccess;
    const dim3 g = grid2d(d.w, d.h), b(kTw, kTh);
    switch (d.radius) {  // the window in registers up to 12 x 12
        case 1: hipLaunchKernelGGL(detail_kernel<2>, g, b, 0, stream, d); break;
        case 2: hipLaunchKernelGGL(detail_kernel<4>, g, b, 0, stream, d); break;
        case 3: hipLaunchKernelGGL(detail_kernel<6>, g, b, 0, stream, d); break;
        case 4: hipLaunchKernelGGL(detail_kernel<8>, g, b, 0, stream, d); break;
        case 5: hipLaunchKernelGGL(detail_kernel<10>, g, b, 0, stream, d); break;
        case 6: hipLaunchKernelGGL(detail_kernel<12>, g, b, 0, stream, d); break;
        default: hipLaunchKernelGGL(detail_kernel<0>, g, b, 0, stream, d); break;
    }
    return hipGetLastError();
}

hipError_t launch_occlusion(float* disp, int w, int h, float tol, hipStream_t stream) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(occlusion_kernel, dim3((unsigned)h), dim3(kBlock), (size_t)w * 4, stream, disp, w, tol);
    return hipGetLastError();
}

hipError_t launch_spikes(const float* in, float* out, int w, int h, float thr, hipStream_t stream) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(spike_kernel, grid2d(w, h), dim3(kTw, kTh), 0, stream, in, out, w, h, thr);
    return hipGetLastError();
}

hipError_t launch_ccl(const float* disp, int32_t* labels, uint32_t* local, uint32_t* sizes, int w, int h, float max_diff,
                      hipStream_t stream) {
    if (w <= 0 || h <= 0) return hipSuccess;
    const int tx = (w + kCcl - 1) / kCcl, ty = (h + kCcl - 1) / kCcl;
    hipLaunchKernelGGL(ccl_local_kernel, dim3((unsigned)tx, (unsigned)ty), dim3(kCclThreads), 0, stream, disp, labels, local,
                       sizes, w, h, max_diff);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t n_vertical = (int64_t)(tx - 1) * h, n_total = n_vertical + (int64_t)(ty - 1) * w;
    if (n_total > 0) {
        hipLaunchKernelGGL(ccl_border_kernel, dim3((unsigned)((n_total + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, disp,
                           labels, w, h, max_diff, n_vertical, n_total);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(ccl_accum_kernel, grid2d(w, h), dim3(kTw, kTh), 0, stream, labels, local, sizes, w, h);
    return hipGetLastError();
}

hipError_t launch_finalize(const float* disp, const int32_t* labels, const uint32_t* sizes, int max_region,
                           const float* conf_src, float* out_disp, float* out_conf, int w, int h, hipStream_t stream) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(finalize_kernel, grid2d(w, h), dim3(kTw, kTh), 0, stream, disp, labels, sizes, max_region, conf_src,
                       out_disp, out_conf, w, h);
    return hipGetLastError();
}

hipError_t launch_rectify_full(const RectifyDesc& r, uint8_t* f0, uint8_t* f1, int64_t fpitch, int fw, int fh,
                               hipStream_t stream) {
    if (fw <= 0 || fh <= 0) return hipSuccess;
    hipLaunchKernelGGL(rectify_full_kernel, grid2d(fw, fh, 2), dim3(kTw, kTh), 0, stream, r, f0, f1, fpitch, fw, fh);
    return hipGetLastError();
}

}  // namespace sgm
