/*
 * CPU oracle for the SGM stereo hot path -- TEST INFRASTRUCTURE ONLY (see sgm_oracle.h).
 * Restates libs/depth/src/stereo.cpp (SGM subset) in C; each function cites the lines it follows.
 * Threads only split independent rows or scanlines, as tbb::parallel_for does in the reference
 * (stereo.cpp:18-23, 403-408), so results never depend on the thread count.
 */
#define _GNU_SOURCE
#include "sgm_oracle.h"

#include <limits.h>
#include <sched.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef uint16_t cost_t;
#define MAX_COST ((cost_t)0xFFFF) /* kMaxCost, stereo.cpp:16 */

/* ---------------------------------------------------------------- parallel_for */

typedef void (*range_fn)(void* ctx, int begin, int end);

typedef struct {
    range_fn fn;
    void* ctx;
    int n;
    int chunk;
    int next; /* claimed with __atomic_fetch_add */
} pf_job;

static void* pf_worker(void* arg) {
    pf_job* job = (pf_job*)arg;
    for (;;) {
        const int b = __atomic_fetch_add(&job->next, job->chunk, __ATOMIC_RELAXED);
        if (b >= job->n) break;
        const int e = b + job->chunk < job->n ? b + job->chunk : job->n;
        job->fn(job->ctx, b, e);
    }
    return NULL;
}

/* threads <= 0: the CPUs this process may run on (its affinity mask), not every online CPU. */
static int resolve_threads(int threads) {
    if (threads > 0) return threads;
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) {
        const int n = CPU_COUNT(&set);
        if (n > 0) return n;
    }
    const long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static void parallel_for(int n, int threads, range_fn fn, void* ctx) {
    if (n <= 0) return;
    threads = resolve_threads(threads);
    if (threads > n) threads = n;
    if (threads <= 1) {
        fn(ctx, 0, n);
        return;
    }
    pf_job job = {fn, ctx, n, 0, 0};
    job.chunk = n / (threads * 8);
    if (job.chunk < 1) job.chunk = 1;
    pthread_t* tid = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)(threads - 1));
    int started = 0;
    for (int i = 0; tid && i < threads - 1; ++i)
        if (pthread_create(&tid[i], NULL, pf_worker, &job) == 0) ++started;
    pf_worker(&job);
    for (int i = 0; i < started; ++i) pthread_join(tid[i], NULL);
    free(tid);
}

/* ---------------------------------------------------------------- census (stereo.cpp:328-346) */

typedef struct {
    const uint8_t* img;
    int w, h, rx, ry;
    ptrdiff_t stride;
    uint64_t* out;
} census_ctx;

static void census_rows(void* vctx, int y0, int y1) {
    const census_ctx* c = (const census_ctx*)vctx;
    for (int y = y0; y < y1; ++y) {
        uint64_t* orow = c->out + (size_t)y * (size_t)c->w;
        memset(orow, 0, sizeof(uint64_t) * (size_t)c->w); /* border pixels stay 0 */
        if (y < c->ry || y >= c->h - c->ry) continue;
        for (int x = c->rx; x < c->w - c->rx; ++x) {
            const uint8_t center = c->img[(ptrdiff_t)y * c->stride + x];
            uint64_t bits = 0;
            for (int v = -c->ry; v <= c->ry; ++v) { /* rows outer, stereo.cpp:335 */
                const uint8_t* row = c->img + (ptrdiff_t)(y + v) * c->stride;
                for (int u = -c->rx; u <= c->rx; ++u) { /* columns inner, :337 */
                    if (u == 0 && v == 0) continue;
                    bits = (bits << 1) | (uint64_t)(row[x + u] < center);
                }
            }
            orow[x] = bits;
        }
    }
}

void orc_census(const uint8_t* img, int w, int h, ptrdiff_t stride, int rx, int ry, uint64_t* out, int threads) {
    census_ctx c = {img, w, h, rx, ry, stride, out};
    parallel_for(h, threads, census_rows, &c);
}

/* ---------------------------------------------------------------- cost fill (stereo.cpp:366-380) */

typedef struct {
    const uint64_t *cl, *cr;
    int w, h, min_d, nd;
    cost_t invalid;
    cost_t* cost;
} cost_ctx;

static void cost_rows(void* vctx, int y0, int y1) {
    const cost_ctx* c = (const cost_ctx*)vctx;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < c->w; ++x) {
            cost_t* cv = c->cost + ((size_t)y * (size_t)c->w + (size_t)x) * (size_t)c->nd;
            for (int k = 0; k < c->nd; ++k) cv[k] = c->invalid; /* CostVolume fill, :368 */
            const uint64_t a = c->cl[(size_t)y * (size_t)c->w + (size_t)x];
            for (int k = 0; k < c->nd; ++k) {
                const int xr = x - (c->min_d + k);
                if (xr < 0) break; /* the rest stay invalid, :375 */
                if (xr >= c->w) continue;
                cv[k] = (cost_t)__builtin_popcountll(a ^ c->cr[(size_t)y * (size_t)c->w + (size_t)xr]);
            }
        }
    }
}

void orc_cost_volume(const uint64_t* cl, const uint64_t* cr, int w, int h, int min_d, int nd, uint16_t invalid_cost,
                     uint16_t* cost, int threads) {
    cost_ctx c = {cl, cr, w, h, min_d, nd, invalid_cost, cost};
    parallel_for(h, threads, cost_rows, &c);
}

/* ---------------------------------------------------------------- aggregation (stereo.cpp:44-89, 385-409) */

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* aggregate_path, stereo.cpp:44-89: walks from (x0, y0) in steps (dx, dy) until it leaves the image. */
static void aggregate_path(const cost_t* cost, const uint8_t* guide, ptrdiff_t gstride, int w, int h, int nd,
                           const orc_params* p, cost_t* sum, int x0, int y0, int dx, int dy, cost_t* prev, cost_t* cur) {
    int x = x0, y = y0;
    const cost_t* c = cost + ((size_t)y * (size_t)w + (size_t)x) * (size_t)nd;
    cost_t prev_min = MAX_COST;
    for (int k = 0; k < nd; ++k) {
        prev[k] = c[k];
        if (c[k] < prev_min) prev_min = c[k];
    }
    {
        cost_t* s = sum + ((size_t)y * (size_t)w + (size_t)x) * (size_t)nd;
        for (int k = 0; k < nd; ++k) s[k] = (cost_t)(s[k] + prev[k]);
    }
    int prev_intensity = guide[(ptrdiff_t)y * gstride + x];
    x += dx;
    y += dy;
    while (x >= 0 && y >= 0 && x < w && y < h) {
        const int intensity = guide[(ptrdiff_t)y * gstride + x];
        int p2 = p->p2;
        if (p->adaptive_p2) p2 = imax(p->p1 + 1, p->p2 / (abs(intensity - prev_intensity) / 8 + 1)); /* :64 */
        prev_intensity = intensity;
        c = cost + ((size_t)y * (size_t)w + (size_t)x) * (size_t)nd;
        const int jump = prev_min + p2;
        cost_t cur_min = MAX_COST;
        for (int k = 0; k < nd; ++k) {
            int best = imin(prev[k], jump);
            if (k > 0) best = imin(best, prev[k - 1] + p->p1);
            if (k + 1 < nd) best = imin(best, prev[k + 1] + p->p1);
            if (p->slant_steps) {
                if (k > 1) best = imin(best, prev[k - 2] + p->p1);
                if (k + 2 < nd) best = imin(best, prev[k + 2] + p->p1);
            }
            const cost_t v = (cost_t)(c[k] + best - prev_min);
            cur[k] = v;
            if (v < cur_min) cur_min = v;
        }
        cost_t* s = sum + ((size_t)y * (size_t)w + (size_t)x) * (size_t)nd;
        for (int k = 0; k < nd; ++k) s[k] = (cost_t)(s[k] + cur[k]);
        cost_t* t = prev;
        prev = cur;
        cur = t;
        prev_min = cur_min;
        x += dx;
        y += dy;
    }
}

typedef struct {
    const cost_t* cost;
    const uint8_t* guide;
    ptrdiff_t gstride;
    int w, h, nd, dx, dy;
    const orc_params* p;
    cost_t* sum;
} dir_ctx;

/* Start pixel i of a direction, in the order of stereo.cpp:391-402: the x-entry column for every y,
 * then the y-entry row for every x except the corner already taken. */
static void start_pixel(const dir_ctx* c, int i, int* sx, int* sy) {
    const int ncol = c->dx != 0 ? c->h : 0;
    if (i < ncol) {
        *sx = c->dx > 0 ? 0 : c->w - 1;
        *sy = i;
        return;
    }
    int j = i - ncol;
    const int skip = c->dx != 0 ? (c->dx > 0 ? 0 : c->w - 1) : -1;
    if (skip >= 0 && j >= skip) ++j;
    *sx = j;
    *sy = c->dy > 0 ? 0 : c->h - 1;
}

static int num_starts(int w, int h, int dx, int dy) {
    int n = 0;
    if (dx != 0) n += h;
    if (dy != 0) n += dx != 0 ? w - 1 : w;
    return n;
}

static void dir_lines(void* vctx, int i0, int i1) {
    const dir_ctx* c = (const dir_ctx*)vctx;
    cost_t* prev = (cost_t*)malloc(sizeof(cost_t) * (size_t)c->nd * 2);
    if (!prev) abort();
    cost_t* cur = prev + c->nd;
    for (int i = i0; i < i1; ++i) {
        int sx, sy;
        start_pixel(c, i, &sx, &sy);
        aggregate_path(c->cost, c->guide, c->gstride, c->w, c->h, c->nd, c->p, c->sum, sx, sy, c->dx, c->dy, prev, cur);
    }
    free(prev);
}

void orc_aggregate_direction(const uint16_t* cost, const uint8_t* guide, ptrdiff_t guide_stride, int w, int h, int nd,
                             const orc_params* p, int dx, int dy, uint16_t* sum, int threads) {
    dir_ctx c = {cost, guide, guide_stride, w, h, nd, dx, dy, p, sum};
    parallel_for(num_starts(w, h, dx, dy), threads, dir_lines, &c);
}

void orc_aggregate(const uint16_t* cost, const uint8_t* guide, ptrdiff_t guide_stride, int w, int h, const orc_params* p,
                   uint16_t* sum, int threads) {
    static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}, {1, -1}, {-1, 1}}; /* :387-388 */
    const int n = p->eight_paths ? 8 : 4;
    for (int i = 0; i < n; ++i)
        orc_aggregate_direction(cost, guide, guide_stride, w, h, p->num_disparities, p, dirs[i][0], dirs[i][1], sum, threads);
}

/* ---------------------------------------------------------------- WTA (stereo.cpp:92-96, 411-448) */

float orc_subpixel_offset(float cm, float c0, float cp) {
    const float denom = cm - 2.0f * c0 + cp;
    if (denom <= 0.0f) return 0.0f;
    const float v = 0.5f * (cm - cp) / denom;
    return v < -0.5f ? -0.5f : (0.5f < v ? 0.5f : v); /* std::clamp(v, -0.5f, 0.5f) */
}

typedef struct {
    const cost_t* sum;
    int w, h;
    const orc_params* p;
    float* disp;
    int32_t *right, *dint;
    uint8_t* status;
    cost_t *best_cost, *second_cost;
} wta_ctx;

static void wta_rows(void* vctx, int y0, int y1) {
    const wta_ctx* c = (const wta_ctx*)vctx;
    const int w = c->w, nd = c->p->num_disparities, min_d = c->p->min_disparity;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * (size_t)w + (size_t)x;
            const cost_t* s = c->sum + i * (size_t)nd;
            c->disp[i] = ORC_INVALID_DISPARITY;
            if (c->dint) c->dint[i] = ORC_NO_MATCH;
            if (c->status) c->status[i] = 0;
            if (c->best_cost) c->best_cost[i] = MAX_COST;
            if (c->second_cost) c->second_cost[i] = MAX_COST;
            const int min_k = imax(0, x - min_d - (w - 1)); /* :419 */
            const int max_k = imin(nd, x - min_d + 1);      /* :420 */
            if (max_k <= min_k) continue;
            int best_k = min_k;
            for (int k = min_k + 1; k < max_k; ++k)
                if (s[k] < s[best_k]) best_k = k;
            cost_t second = MAX_COST;
            for (int k = min_k; k < max_k; ++k)
                if (abs(k - best_k) > 1 && s[k] < second) second = s[k];
            if (c->dint) c->dint[i] = min_d + best_k;
            if (c->best_cost) c->best_cost[i] = s[best_k];
            if (c->second_cost) c->second_cost[i] = second;
            uint8_t st = ORC_ST_WINDOW;
            if (second != MAX_COST && (float)s[best_k] >= c->p->uniqueness * (float)second) { /* :428 */
                if (c->status) c->status[i] = st;
                continue;
            }
            st |= ORC_ST_UNIQUE;
            if (c->status) c->status[i] = st;
            float d = (float)(min_d + best_k);
            if (c->p->subpixel && best_k > min_k && best_k + 1 < max_k)
                d += orc_subpixel_offset((float)s[best_k - 1], (float)s[best_k], (float)s[best_k + 1]);
            c->disp[i] = d;
        }
        for (int xr = 0; xr < w; ++xr) { /* right WTA, :434-447 */
            int best_k = -1;
            cost_t best = MAX_COST;
            for (int k = imax(0, -(xr + min_d)); k < nd; ++k) {
                const int xl = xr + min_d + k;
                if (xl >= w) break;
                const cost_t v = c->sum[((size_t)y * (size_t)w + (size_t)xl) * (size_t)nd + (size_t)k];
                if (v < best) {
                    best = v;
                    best_k = k;
                }
            }
            c->right[(size_t)y * (size_t)w + (size_t)xr] = best_k < 0 ? ORC_NO_MATCH : min_d + best_k;
        }
    }
}

void orc_wta(const uint16_t* sum, int w, int h, const orc_params* p, float* disp, int32_t* right_disp,
             int32_t* disp_int, uint8_t* status, uint16_t* best_cost, uint16_t* second_cost, int threads) {
    wta_ctx c = {sum, w, h, p, disp, right_disp, disp_int, status, best_cost, second_cost};
    parallel_for(h, threads, wta_rows, &c);
}

/* ---------------------------------------------------------------- left/right check (stereo.cpp:450-461) */

typedef struct {
    float* disp;
    const int32_t* right;
    int w, lr;
    uint8_t* status;
} lr_ctx;

static void lr_rows(void* vctx, int y0, int y1) {
    const lr_ctx* c = (const lr_ctx*)vctx;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < c->w; ++x) {
            const size_t i = (size_t)y * (size_t)c->w + (size_t)x;
            const float d = c->disp[i];
            if (!(d > -1.0e8f)) continue; /* valid_disparity, stereo.hpp:24 */
            const int xr = (int)lroundf((float)x - d);
            if (xr < 0 || xr >= c->w || c->right[(size_t)y * (size_t)c->w + (size_t)xr] == ORC_NO_MATCH ||
                fabsf((float)c->right[(size_t)y * (size_t)c->w + (size_t)xr] - d) > (float)c->lr) {
                c->disp[i] = ORC_INVALID_DISPARITY;
                continue;
            }
            if (c->status) c->status[i] |= ORC_ST_LR;
        }
    }
}

void orc_lr_check(float* disp, const int32_t* right_disp, int w, int h, int lr_max_diff, uint8_t* status, int threads) {
    lr_ctx c = {disp, right_disp, w, lr_max_diff, status};
    parallel_for(h, threads, lr_rows, &c);
}

/* ---------------------------------------------------------------- median3 (stereo.cpp:129-147) */

typedef struct {
    const float* in;
    float* out;
    int w, h;
} med_ctx;

static void med_rows(void* vctx, int y0, int y1) {
    const med_ctx* c = (const med_ctx*)vctx;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < c->w; ++x) {
            const size_t i = (size_t)y * (size_t)c->w + (size_t)x;
            c->out[i] = ORC_INVALID_DISPARITY;
            if (!(c->in[i] > -1.0e8f)) continue;
            float buf[9];
            int n = 0;
            for (int v = -1; v <= 1; ++v)
                for (int u = -1; u <= 1; ++u) {
                    const int xx = x + u, yy = y + v;
                    if (xx < 0 || yy < 0 || xx >= c->w || yy >= c->h) continue;
                    const float q = c->in[(size_t)yy * (size_t)c->w + (size_t)xx];
                    if (!(q > -1.0e8f)) continue;
                    buf[n++] = q;
                }
            /* std::nth_element(buf, buf + n/2, buf + n) leaves the n/2-th smallest at n/2: sort and take it. */
            for (int a = 1; a < n; ++a) {
                const float t = buf[a];
                int b = a - 1;
                while (b >= 0 && buf[b] > t) {
                    buf[b + 1] = buf[b];
                    --b;
                }
                buf[b + 1] = t;
            }
            c->out[i] = buf[n / 2];
        }
    }
}

void orc_median3(const float* in, int w, int h, float* out, int threads) {
    med_ctx c = {in, out, w, h};
    parallel_for(h, threads, med_rows, &c);
}

/* ---------------------------------------------------------------- rectification (rectify.cpp:133-148, stereo.cpp:348-359) */

typedef struct {
    const uint8_t* src;
    int sw, sh;
    ptrdiff_t sstride;
    const float *mx, *my;
    int w;
    uint8_t* dst;
} remap_ctx;

static void remap_rows(void* vctx, int y0, int y1) {
    const remap_ctx* c = (const remap_ctx*)vctx;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < c->w; ++x) {
            const size_t i = (size_t)y * (size_t)c->w + (size_t)x;
            const float sx = c->mx[i], sy = c->my[i];
            const int x0 = (int)floorf(sx), yy0 = (int)floorf(sy);
            if (x0 < 0 || yy0 < 0 || x0 + 1 >= c->sw || yy0 + 1 >= c->sh) {
                c->dst[i] = 0;
                continue;
            }
            const float fx = sx - (float)x0, fy = sy - (float)yy0;
            const uint8_t* r0 = c->src + (ptrdiff_t)yy0 * c->sstride;
            const uint8_t* r1 = r0 + c->sstride;
            const float a = r0[x0], b = r0[x0 + 1], cc = r1[x0], d = r1[x0 + 1];
            const float v = (a * (1 - fx) + b * fx) * (1 - fy) + (cc * (1 - fx) + d * fx) * fy;
            const float r = v + 0.5f;
            c->dst[i] = (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r)); /* std::clamp then truncation */
        }
}

void orc_remap(const uint8_t* src, int src_w, int src_h, ptrdiff_t src_stride, const float* map_x, const float* map_y,
               int w, int h, uint8_t* dst, int threads) {
    remap_ctx c = {src, src_w, src_h, src_stride, map_x, map_y, w, dst};
    parallel_for(h, threads, remap_rows, &c);
}

typedef struct {
    const uint8_t* src;
    ptrdiff_t stride;
    int ow;
    uint8_t* dst;
} down_ctx;

static void down_rows(void* vctx, int y0, int y1) {
    const down_ctx* c = (const down_ctx*)vctx;
    for (int y = y0; y < y1; ++y) {
        const uint8_t* r0 = c->src + (ptrdiff_t)(2 * y) * c->stride;
        const uint8_t* r1 = r0 + c->stride;
        for (int x = 0; x < c->ow; ++x)
            c->dst[(size_t)y * (size_t)c->ow + (size_t)x] = (uint8_t)((r0[2 * x] + r0[2 * x + 1] + r1[2 * x] + r1[2 * x + 1] + 2) / 4);
    }
}

void orc_downsample2(const uint8_t* src, int w, int h, ptrdiff_t stride, uint8_t* dst, int threads) {
    down_ctx c = {src, stride, w / 2, dst};
    parallel_for(h / 2, threads, down_rows, &c);
}

int orc_rectify_half(const uint8_t* raw, int raw_w, int raw_h, ptrdiff_t raw_stride, const float* map_x, const float* map_y,
                     int map_w, int map_h, uint8_t* dst, int threads) {
    uint8_t* full = (uint8_t*)malloc((size_t)map_w * (size_t)map_h + 1);
    if (!full) return -1;
    orc_remap(raw, raw_w, raw_h, raw_stride, map_x, map_y, map_w, map_h, full, threads);
    orc_downsample2(full, map_w, map_h, map_w, dst, threads);
    free(full);
    return 0;
}

/* ---------------------------------------------------------------- sgm_disparity + median3 (stereo.cpp:361-463) */

int orc_sgm(const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h, const orc_params* p,
            const orc_outputs* out, int threads) {
    const size_t npx = (size_t)w * (size_t)h;
    const size_t nvol = npx * (size_t)p->num_disparities;
    uint64_t* cl = out->census_left ? out->census_left : (uint64_t*)malloc(sizeof(uint64_t) * (npx ? npx : 1));
    uint64_t* cr = out->census_right ? out->census_right : (uint64_t*)malloc(sizeof(uint64_t) * (npx ? npx : 1));
    cost_t* cost = out->cost_volume ? out->cost_volume : (cost_t*)malloc(sizeof(cost_t) * (nvol ? nvol : 1));
    cost_t* sum = out->sum_volume ? out->sum_volume : (cost_t*)malloc(sizeof(cost_t) * (nvol ? nvol : 1));
    int32_t* rdisp = out->right_disparity ? out->right_disparity : (int32_t*)malloc(sizeof(int32_t) * (npx ? npx : 1));
    int rc = 0;
    if (!cl || !cr || !cost || !sum || !rdisp) {
        rc = -1;
        goto done;
    }
    orc_census(left, w, h, stride, p->census_radius_x, p->census_radius_y, cl, threads);
    orc_census(right, w, h, stride, p->census_radius_x, p->census_radius_y, cr, threads);
    const cost_t invalid_cost = (cost_t)((2 * p->census_radius_x + 1) * (2 * p->census_radius_y + 1)); /* :366 */
    orc_cost_volume(cl, cr, w, h, p->min_disparity, p->num_disparities, invalid_cost, cost, threads);
    memset(sum, 0, sizeof(cost_t) * nvol);
    orc_aggregate(cost, left, stride, w, h, p, sum, threads);
    orc_wta(sum, w, h, p, out->disparity, rdisp, out->disparity_int, out->status, out->best_cost, out->second_cost, threads);
    orc_lr_check(out->disparity, rdisp, w, h, p->lr_max_diff, out->status, threads);
    if (out->disparity_median) orc_median3(out->disparity, w, h, out->disparity_median, threads);
done:
    if (cl != out->census_left) free(cl);
    if (cr != out->census_right) free(cr);
    if (cost != out->cost_volume) free(cost);
    if (sum != out->sum_volume) free(sum);
    if (rdisp != out->right_disparity) free(rdisp);
    return rc;
}

/* ---------------------------------------------------------------- refinement (stereo.cpp:98-284) */

static inline int valid_d(float d) { return d > -1.0e8f; } /* valid_disparity, stereo.hpp:24 */

static inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); } /* std::clamp */

float orc_zncc_slanted(const uint8_t* a, const uint8_t* b, ptrdiff_t stride, int bw, int x, int y, float d, float gx,
                       float gy, int r) {
    double sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
    const int n = (2 * r + 1) * (2 * r + 1);
    for (int v = -r; v <= r; ++v) {
        const uint8_t* ra = a + (ptrdiff_t)(y + v) * stride;
        const uint8_t* rb = b + (ptrdiff_t)(y + v) * stride;
        for (int u = -r; u <= r; ++u) {
            const float xb = (float)(x + u) - (d + gx * (float)u + gy * (float)v);
            const int x0 = (int)floorf(xb);
            if (x0 < 0 || x0 + 1 >= bw) return -2.0f;
            const float f = xb - (float)x0;
            const double ib = rb[x0] * (1.0f - f) + rb[x0 + 1] * f;
            const double ia = ra[x + u];
            sa += ia;
            sb += ib;
            saa += ia * ia;
            sbb += ib * ib;
            sab += ia * ib;
        }
    }
    const double va = saa - sa * sa / n;
    const double vb = sbb - sb * sb / n;
    if (va < 1e-6 || vb < 1e-6) return -1.0f;
    return (float)((sab - sa * sb / n) / sqrt(va * vb));
}

typedef struct {
    const uint8_t *left, *right;
    ptrdiff_t stride;
    int w, h;
    const float* coarse; /* median3 of the coarse disparity */
    int cw, ch;
    const orc_stereo_params* sp;
    int subpixel;
    float *out_disp, *out_conf;
} refine_ctx;

/* The slope lambda of refine_level (stereo.cpp:159-170). */
static void coarse_slope(const refine_ctx* c, int cx, int cy, float* gx, float* gy) {
    const float* C = c->coarse;
    const int cw = c->cw, ch = c->ch;
    *gx = *gy = 0.0f;
    const float m = C[(size_t)cy * cw + cx];
    if (cx > 0 && cx + 1 < cw && valid_d(C[(size_t)cy * cw + cx - 1]) && valid_d(C[(size_t)cy * cw + cx + 1])) {
        const float l = C[(size_t)cy * cw + cx - 1], r = C[(size_t)cy * cw + cx + 1];
        const float g = 0.5f * (r - l);
        if (fabsf(r - m) < 2 && fabsf(m - l) < 2) *gx = g;
    }
    if (cy > 0 && cy + 1 < ch && valid_d(C[(size_t)(cy - 1) * cw + cx]) && valid_d(C[(size_t)(cy + 1) * cw + cx])) {
        const float u = C[(size_t)(cy - 1) * cw + cx], dn = C[(size_t)(cy + 1) * cw + cx];
        const float g = 0.5f * (dn - u);
        if (fabsf(dn - m) < 2 && fabsf(m - u) < 2) *gy = g;
    }
    *gx = clampf(*gx, -0.9f, 0.9f);
    *gy = clampf(*gy, -0.9f, 0.9f);
}

static void refine_rows(void* vctx, int y0, int y1) {
    const refine_ctx* c = (const refine_ctx*)vctx;
    const int r = c->sp->refine_zncc_radius;
    const float kStep = 0.5f;
    for (int y = y0; y < y1; ++y) {
        if (y < r || y >= c->h - r) continue;
        const int cy = imin(y / 2, c->ch - 1);
        float scores[128];
        for (int x = r; x < c->w - r; ++x) {
            const int cx = imin(x / 2, c->cw - 1);
            const float d0 = c->coarse[(size_t)cy * c->cw + cx];
            if (!valid_d(d0)) continue;
            float gx, gy;
            coarse_slope(c, cx, cy, &gx, &gy);
            const float center = 2.0f * d0;
            const int nd = imin(127, (int)((float)(2 * c->sp->refine_search_radius) / kStep) + 1);
            const float lo = center - (float)c->sp->refine_search_radius;
            float best = -2.0f;
            int best_k = -1;
            for (int k = 0; k < nd; ++k) {
                const float d = lo + kStep * (float)k;
                scores[k] = orc_zncc_slanted(c->left, c->right, c->stride, c->w, x, y, d, gx, gy, r);
                if (scores[k] > best) {
                    best = scores[k];
                    best_k = k;
                }
            }
            if (best_k < 0 || best < c->sp->refine_min_zncc) continue;
            if (c->sp->refine_reject_edge_winners && (best_k == 0 || best_k == nd - 1)) continue;
            float d = lo + kStep * (float)best_k;
            if (c->subpixel && best_k > 0 && best_k + 1 < nd && scores[best_k - 1] > -2.0f && scores[best_k + 1] > -2.0f)
                d += kStep * orc_subpixel_offset(1.0f - scores[best_k - 1], 1.0f - best, 1.0f - scores[best_k + 1]);
            c->out_disp[(size_t)y * c->w + x] = d;
            c->out_conf[(size_t)y * c->w + x] = best;
        }
    }
}

void orc_refine_level(const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h, const float* coarse_raw,
                      int cw, int ch, const orc_stereo_params* sp, int subpixel, float* out_disp, float* out_conf,
                      int threads) {
    for (size_t i = 0; i < (size_t)w * (size_t)h; ++i) {
        out_disp[i] = ORC_INVALID_DISPARITY;
        out_conf[i] = 0.0f;
    }
    float* coarse = (float*)malloc(sizeof(float) * ((size_t)cw * (size_t)ch + 1));
    if (!coarse) return;
    orc_median3(coarse_raw, cw, ch, coarse, threads);
    refine_ctx c = {left, right, stride, w, h, coarse, cw, ch, sp, subpixel, out_disp, out_conf};
    parallel_for(h, threads, refine_rows, &c);
    free(coarse);
}

float orc_zncc_window(const uint8_t* a, const uint8_t* b, ptrdiff_t stride, int bw, int x0, int y0, int n, float cx,
                      float cy, float D, float gx, float gy) {
    double sa = 0, sb = 0, saa = 0, sbb = 0, sab = 0;
    for (int j = 0; j < n; ++j) {
        const int Y = y0 + j;
        const uint8_t* ra = a + (ptrdiff_t)Y * stride;
        const uint8_t* rb = b + (ptrdiff_t)Y * stride;
        const float dy = gy * ((float)Y - cy);
        for (int i = 0; i < n; ++i) {
            const int X = x0 + i;
            const float xb = (float)X - (D + gx * ((float)X - cx) + dy);
            const int xi = (int)floorf(xb);
            if (xi < 0 || xi + 1 >= bw) return -2.0f;
            const float f = xb - (float)xi;
            const double ib = rb[xi] * (1.0f - f) + rb[xi + 1] * f;
            const double ia = ra[X];
            sa += ia;
            sb += ib;
            saa += ia * ia;
            sbb += ib * ib;
            sab += ia * ib;
        }
    }
    const double m = (double)n * n;
    const double va = saa - sa * sa / m;
    const double vb = sbb - sb * sb / m;
    if (va < 1e-6 || vb < 1e-6) return -1.0f;
    return (float)((sab - sa * sb / m) / sqrt(va * vb));
}

typedef struct {
    const uint8_t *fl, *fr;
    ptrdiff_t fstride;
    int W, H;
    const float* first;
    float* disparity;
    int w, h;
    const orc_stereo_params* sp;
    int subpixel;
} detail_ctx;

/* The grad lambda of refine_detail (stereo.cpp:252-255). */
static float detail_grad(float l, float c, float r) {
    if (!valid_d(l) || !valid_d(r) || fabsf(r - c) >= 1.0f || fabsf(c - l) >= 1.0f) return 0.0f;
    return clampf(0.5f * (r - l), -0.9f, 0.9f);
}

static void detail_rows(void* vctx, int ya, int yb) {
    const detail_ctx* c = (const detail_ctx*)vctx;
    const orc_stereo_params* p = c->sp;
    const int K = imin(16, (int)lroundf(p->detail_search / p->detail_step));
    const int nc = 2 * K + 1, n = 2 * p->detail_radius;
    const float* F = c->first;
    const int w = c->w;
    for (int y = ya; y < yb; ++y) {
        float scores[33];
        const int y0 = 2 * y - p->detail_radius + 1;
        if (y < 1 || y + 1 >= c->h || y0 < 0 || y0 + n > c->H) continue;
        for (int x = 1; x + 1 < w; ++x) {
            const float d = F[(size_t)y * w + x];
            if (!valid_d(d)) continue;
            const int x0 = 2 * x - p->detail_radius + 1;
            if (x0 < 0 || x0 + n > c->W) continue;
            const float gx = detail_grad(F[(size_t)y * w + x - 1], d, F[(size_t)y * w + x + 1]);
            const float gy = detail_grad(F[(size_t)(y - 1) * w + x], d, F[(size_t)(y + 1) * w + x]);
            const float cx = 2.0f * (float)x + 0.5f, cy = 2.0f * (float)y + 0.5f;
            const float D0 = 2.0f * d;
            float best = -2.0f;
            int best_k = -1;
            for (int k = 0; k < nc; ++k) {
                scores[k] = orc_zncc_window(c->fl, c->fr, c->fstride, c->W, x0, y0, n, cx, cy,
                                            D0 + p->detail_step * (float)(k - K), gx, gy);
                if (scores[k] > best) {
                    best = scores[k];
                    best_k = k;
                }
            }
            if (best_k <= 0 || best_k + 1 >= nc || best < p->detail_min_zncc || scores[best_k - 1] <= -2.0f ||
                scores[best_k + 1] <= -2.0f)
                continue;
            float D = D0 + p->detail_step * (float)(best_k - K);
            if (c->subpixel)
                D += p->detail_step * orc_subpixel_offset(1.0f - scores[best_k - 1], 1.0f - best, 1.0f - scores[best_k + 1]);
            c->disparity[(size_t)y * w + x] = 0.5f * D;
        }
    }
}

void orc_refine_detail(const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride, int W, int H,
                       float* disparity, int w, int h, const orc_stereo_params* sp, int subpixel, int threads) {
    if (!sp->detail_enabled || sp->detail_radius < 1 || sp->detail_step <= 0) return;
    float* first = (float*)malloc(sizeof(float) * ((size_t)w * (size_t)h + 1));
    if (!first) return;
    memcpy(first, disparity, sizeof(float) * (size_t)w * (size_t)h);
    detail_ctx c = {full_left, full_right, full_stride, W, H, first, disparity, w, h, sp, subpixel};
    parallel_for(h, threads, detail_rows, &c);
    free(first);
}

/* ---------------------------------------------------------------- filters (stereo.cpp:286-326, 465-494) */

typedef struct {
    float* disp;
    const float* in; /* spikes: the disparity after the occlusion filter */
    int w, h;
    const orc_stereo_params* sp;
} filt_ctx;

static void occlusion_rows(void* vctx, int y0, int y1) {
    const filt_ctx* c = (const filt_ctx*)vctx;
    const int w = c->w;
    float* nearest = (float*)malloc(sizeof(float) * ((size_t)w + 1));
    if (!nearest) return;
    for (int y = y0; y < y1; ++y) {
        float* row = c->disp + (size_t)y * w;
        for (int x = 0; x < w; ++x) nearest[x] = ORC_INVALID_DISPARITY;
        for (int x = 0; x < w; ++x) {
            const float d = row[x];
            if (!valid_d(d)) continue;
            const long xr = lroundf((float)x - d);
            if (xr >= 0 && xr < w) nearest[xr] = nearest[xr] < d ? d : nearest[xr]; /* std::max(nearest, d) */
        }
        for (int x = 0; x < w; ++x) {
            const float d = row[x];
            if (!valid_d(d)) continue;
            const long xr = lroundf((float)x - d);
            if (xr >= 0 && xr < w && d < nearest[xr] - c->sp->filter_occlusion_tolerance) row[x] = ORC_INVALID_DISPARITY;
        }
    }
    free(nearest);
}

static void spike_rows(void* vctx, int y0, int y1) {
    const filt_ctx* c = (const filt_ctx*)vctx;
    const int w = c->w, h = c->h;
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < w; ++x) {
            const float d = c->in[(size_t)y * w + x];
            if (!valid_d(d)) continue;
            float buf[9];
            int n = 0;
            for (int v = -1; v <= 1; ++v)
                for (int u = -1; u <= 1; ++u) {
                    const int xx = x + u, yy = y + v;
                    if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                    const float q = c->in[(size_t)yy * w + xx];
                    if (!valid_d(q)) continue;
                    buf[n++] = q;
                }
            if (n < 3) continue;
            for (int a = 1; a < n; ++a) { /* n/2-th smallest, as std::nth_element leaves it */
                const float t = buf[a];
                int b = a - 1;
                while (b >= 0 && buf[b] > t) {
                    buf[b + 1] = buf[b];
                    --b;
                }
                buf[b + 1] = t;
            }
            if (fabsf(d - buf[n / 2]) > c->sp->filter_spike_threshold) c->disp[(size_t)y * w + x] = ORC_INVALID_DISPARITY;
        }
    }
}

void orc_filter_disparity(float* disparity, int w, int h, const orc_stereo_params* sp, int threads) {
    filt_ctx c = {disparity, NULL, w, h, sp};
    if (sp->filter_occlusion) parallel_for(h, threads, occlusion_rows, &c);
    if (sp->filter_spikes) {
        float* in = (float*)malloc(sizeof(float) * ((size_t)w * (size_t)h + 1));
        if (!in) return;
        memcpy(in, disparity, sizeof(float) * (size_t)w * (size_t)h);
        c.in = in;
        parallel_for(h, threads, spike_rows, &c);
        free(in);
    }
}

int orc_remove_speckles(float* disparity, int w, int h, const orc_stereo_params* sp) {
    const size_t npx = (size_t)w * (size_t)h;
    uint8_t* seen = (uint8_t*)calloc(npx + 1, 1);
    int* stack = (int*)malloc(sizeof(int) * (npx + 1));
    int* region = (int*)malloc(sizeof(int) * (npx + 1));
    if (!seen || !stack || !region) {
        free(seen);
        free(stack);
        free(region);
        return -1;
    }
    for (int start = 0; start < (int)npx; ++start) {
        if (seen[start] || !valid_d(disparity[start])) continue;
        int ns = 0, nr = 0;
        stack[ns++] = start;
        seen[start] = 1;
        while (ns > 0) {
            const int i = stack[--ns];
            region[nr++] = i;
            const int x = i % w, y = i / w;
            const float d = disparity[i];
            const int nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
            for (int k = 0; k < 4; ++k) {
                if (nb[k][0] < 0 || nb[k][1] < 0 || nb[k][0] >= w || nb[k][1] >= h) continue;
                const int j = nb[k][1] * w + nb[k][0];
                const float dj = disparity[j];
                if (seen[j] || !valid_d(dj) || fabsf(dj - d) > sp->speckle_max_diff) continue;
                seen[j] = 1;
                stack[ns++] = j;
            }
        }
        if (nr < sp->speckle_max_region_size)
            for (int k = 0; k < nr; ++k) disparity[region[k]] = ORC_INVALID_DISPARITY;
    }
    free(seen);
    free(stack);
    free(region);
    return 0;
}

/* ---------------------------------------------------------------- compute_disparity (stereo.cpp:496-536) */

int orc_compute_disparity(const uint8_t* left, const uint8_t* right, ptrdiff_t stride, int w, int h,
                          const uint8_t* full_left, const uint8_t* full_right, ptrdiff_t full_stride, int full_w,
                          int full_h, const orc_params* sgm, const orc_stereo_params* sp, float* disparity,
                          float* confidence, int threads) {
    const int L = sp->pyramid_levels;
    int rc = -1;
    /* Level i (1..L) images, compact; level 0 is the caller's. */
    uint8_t* pl[17] = {0};
    uint8_t* pr[17] = {0};
    int lw[17], lh[17];
    float *disp = NULL, *finer = NULL;
    if (L < 0 || L > 16) return -1;
    lw[0] = w;
    lh[0] = h;
    for (int i = 1; i <= L; ++i) {
        lw[i] = lw[i - 1] / 2;
        lh[i] = lh[i - 1] / 2;
        pl[i] = (uint8_t*)malloc((size_t)lw[i] * (size_t)lh[i] + 1);
        pr[i] = (uint8_t*)malloc((size_t)lw[i] * (size_t)lh[i] + 1);
        if (!pl[i] || !pr[i]) goto done;
        const uint8_t* sl = i == 1 ? left : pl[i - 1];
        const uint8_t* sr = i == 1 ? right : pr[i - 1];
        const ptrdiff_t ss = i == 1 ? stride : lw[i - 1];
        orc_downsample2(sl, lw[i - 1], lh[i - 1], ss, pl[i], threads);
        orc_downsample2(sr, lw[i - 1], lh[i - 1], ss, pr[i], threads);
    }
    {
        const int tw = lw[L], th = lh[L];
        disp = (float*)malloc(sizeof(float) * ((size_t)tw * (size_t)th + 1));
        if (!disp) goto done;
        orc_outputs o;
        memset(&o, 0, sizeof(o));
        o.disparity = disp;
        if (orc_sgm(L ? pl[L] : left, L ? pr[L] : right, L ? tw : stride, tw, th, sgm, &o, threads) != 0) goto done;
        const size_t npx = (size_t)w * (size_t)h;
        if (L == 0)
            for (size_t i = 0; i < npx; ++i) confidence[i] = valid_d(disp[i]) ? 1.0f : 0.0f;
        for (int level = L - 1; level >= 0; --level) {
            const size_t n = (size_t)lw[level] * (size_t)lh[level];
            finer = (float*)malloc(sizeof(float) * (n + 1));
            if (!finer) goto done;
            float* cf = level == 0 ? confidence : (float*)malloc(sizeof(float) * (n + 1));
            if (!cf) goto done;
            orc_refine_level(level == 0 ? left : pl[level], level == 0 ? right : pr[level], level == 0 ? stride : lw[level],
                             lw[level], lh[level], disp, lw[level + 1], lh[level + 1], sp, sgm->subpixel, finer, cf,
                             threads);
            if (cf != confidence) free(cf);
            free(disp);
            disp = finer;
            finer = NULL;
        }
        if (sp->detail_enabled && full_left && full_right && full_w == 2 * w && full_h == 2 * h)
            orc_refine_detail(full_left, full_right, full_stride, full_w, full_h, disp, w, h, sp, sgm->subpixel, threads);
        orc_filter_disparity(disp, w, h, sp, threads);
        if (orc_remove_speckles(disp, w, h, sp) != 0) goto done;
        for (size_t i = 0; i < npx; ++i) {
            disparity[i] = disp[i];
            if (!valid_d(disp[i])) confidence[i] = 0.0f;
        }
    }
    rc = 0;
done:
    for (int i = 1; i <= L; ++i) {
        free(pl[i]);
        free(pr[i]);
    }
    free(disp);
    free(finer);
    return rc;
}

/* ---------------------------------------------------------------- points (point_image.cpp:10-55) */

typedef struct {
    const float *disp, *conf;
    int w, h;
    const orc_geometry* g;
    const orc_point_params* p;
    float *points, *normals, *weights;
} pts_ctx;

static void points_rows(void* vctx, int y0, int y1) {
    const pts_ctx* c = (const pts_ctx*)vctx;
    const orc_geometry* g = c->g;
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < c->w; ++x) {
            const size_t i = (size_t)y * c->w + x;
            const float d = c->disp[i];
            if (!valid_d(d) || d + g->cx_offset <= 0.0) continue;
            const double z = g->f * g->baseline / (d + g->cx_offset); /* depth_from_disparity */
            if (z < c->p->min_depth || z > c->p->max_depth) continue;
            c->points[3 * i] = (float)((x - g->cx) * z / g->f);
            c->points[3 * i + 1] = (float)((y - g->cy) * z / g->f);
            c->points[3 * i + 2] = (float)z;
            c->weights[i] = c->conf ? c->conf[i] : 1.0f;
        }
}

/* Eigen's Vector3f operations as the reference spells them: dot = (x x' + y y') + z z',
 * norm = sqrt of the dot with itself, normalized = v / norm, cross component by component. */
static inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static float depth_noise_weight(float z, float reference) { /* point_image.hpp:51-55 */
    if (reference <= 0.0f) return 1.0f;
    const float r = reference / z;
    return r * r > 1.0f ? 1.0f : (r * r < 0.1f ? 0.1f : r * r);
}

static void normals_rows(void* vctx, int y0, int y1) {
    const pts_ctx* c = (const pts_ctx*)vctx;
    const int w = c->w, h = c->h;
    const float* P = c->points;
    const float rc[3] = {(float)c->g->baseline, 0.0f, 0.0f};
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const float* p = P + 3 * i;
            if (p[2] == 0.0f) continue;
            float view = c->p->edge_weight;
            if (x > 0 && y > 0 && x + 1 < w && y + 1 < h) {
                const float* l = P + 3 * (i - 1);
                const float* r = P + 3 * (i + 1);
                const float* u = P + 3 * (i - w);
                const float* dn = P + 3 * (i + w);
                if (l[2] != 0 && r[2] != 0 && u[2] != 0 && dn[2] != 0 && fabsf(l[2] - r[2]) <= c->p->max_depth_jump &&
                    fabsf(u[2] - dn[2]) <= c->p->max_depth_jump) {
                    const float a[3] = {r[0] - l[0], r[1] - l[1], r[2] - l[2]};
                    const float b[3] = {dn[0] - u[0], dn[1] - u[1], dn[2] - u[2]};
                    float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
                    const float len = sqrtf(dot3(n, n));
                    if (len >= 1e-9f) {
                        n[0] /= len;
                        n[1] /= len;
                        n[2] /= len;
                        if (dot3(n, p) > 0) {
                            n[0] = -n[0];
                            n[1] = -n[1];
                            n[2] = -n[2];
                        }
                        c->normals[3 * i] = n[0];
                        c->normals[3 * i + 1] = n[1];
                        c->normals[3 * i + 2] = n[2];
                        const float pn = sqrtf(dot3(p, p));
                        const float ph[3] = {p[0] / pn, p[1] / pn, p[2] / pn};
                        const float q[3] = {p[0] - rc[0], p[1] - rc[1], p[2] - rc[2]};
                        const float qn = sqrtf(dot3(q, q));
                        const float qh[3] = {q[0] / qn, q[1] / qn, q[2] / qn};
                        const float v1 = -dot3(n, ph), v2 = -dot3(n, qh);
                        const float m = v2 < v1 ? v2 : v1; /* std::min(v1, v2) */
                        view = 0.0f < m ? m : 0.0f;        /* std::max(0.0f, m) */
                    }
                }
            }
            c->weights[i] *= view * depth_noise_weight(p[2], c->p->weight_reference_depth);
        }
}

void orc_disparity_to_points(const float* disparity, const float* confidence, int w, int h, const orc_geometry* g,
                             const orc_point_params* p, float* points, float* normals, float* weights, int threads) {
    const size_t n = (size_t)w * (size_t)h;
    memset(points, 0, sizeof(float) * 3 * n);
    memset(normals, 0, sizeof(float) * 3 * n);
    memset(weights, 0, sizeof(float) * n);
    pts_ctx c = {disparity, confidence, w, h, g, p, points, normals, weights};
    parallel_for(h, threads, points_rows, &c);
    parallel_for(h, threads, normals_rows, &c);
}

/* ---------------------------------------------------------------- depth packing (session.cpp:143-185) */

void orc_pack_depth(const float* points, const float* weights, int w, int h, uint8_t* out) {
    const size_t n = (size_t)w * (size_t)h;
    uint16_t* d16 = (uint16_t*)out; /* out is at least 2-byte aligned */
    for (int y = 0; y < h; ++y) {
        uint16_t prev = 0;
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            float z = points[3 * i + 2];
            if (!(z > 0.0f)) z = 0.0f; /* capture_depth: max(0, z) */
            long q = 0;
            if (z > 0) {
                q = lround(z * 50.0); /* kDepthScale = 50.0 (double) */
                q = q < 1 ? 1 : (q > 65535 ? 65535 : q);
            }
            d16[i] = (uint16_t)((uint16_t)q - prev);
            prev = (uint16_t)q;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        const float c = weights ? weights[i] : 1.0f;
        long b = lroundf(c * 255.0f);
        b = b < 0 ? 0 : (b > 255 ? 255 : b);
        out[2 * n + i] = (uint8_t)b;
    }
}

void orc_unpack_depth(const uint8_t* in, int w, int h, float* depth, float* confidence) {
    const size_t n = (size_t)w * (size_t)h;
    const uint16_t* d16 = (const uint16_t*)in;
    for (int y = 0; y < h; ++y) {
        uint16_t acc = 0;
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            acc = (uint16_t)(acc + d16[i]);
            depth[i] = acc ? (float)(acc / 50.0) : 0.0f;
        }
    }
    for (size_t i = 0; i < n; ++i) confidence[i] = in[2 * n + i] / 255.0f;
}
