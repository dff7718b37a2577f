/*
 * Synthetic rectified IR speckle stereo pairs (host code, deterministic across machines).
 *
 * Scene: piecewise-planar patches (Voronoi cells around 12-24 seeds), each a plane
 * z = z0 + gx*(x-sx) + gy*(y-sy) clamped to [z_min, z_max] mm, seen through a rig whose
 * focal length times baseline maps [z_min, z_max] into [min_d+2, min_d+D-2):
 *     fB = (D-5) / (1/z_min - 1/z_max),  min_d = floor(fB / z_max) - 2,  d = fB / z.
 * Texture: Bernoulli(density) dots per pixel, Gaussian blur (sigma px), mapped linearly so the
 * field's maximum is 230 DN and zero is 20 DN. The left image is that texture. The right image
 * forward-warps it row by row, x_r = x - d(x, y), linearly interpolating between neighbouring
 * left pixels of the same patch, with a z-buffer (the larger disparity wins); right pixels that
 * nothing covers are occluded and get fresh speckle from an independent stream. Gaussian noise
 * (noise_sigma DN) is added to both images, which are rounded and clamped to u8.
 *
 * Randomness is SplitMix64 with Box-Muller normals (no library distributions), so a seed gives
 * the same bytes with any C library.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t width;
    int32_t height;
    int32_t num_disparities;
    int32_t min_cells;       /* Voronoi seeds: uniform in [min_cells, max_cells] */
    int32_t max_cells;
    uint64_t seed;
    double z_min;            /* mm */
    double z_max;            /* mm */
    double max_gradient;     /* |dz/dx|, |dz/dy| bound per patch, mm/px */
    double dot_density;
    double blur_sigma;
    double noise_sigma;
} synth_params;

typedef struct {
    uint64_t s;
} splitmix;

static uint64_t sm_next(splitmix* r) {
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double sm_uniform(splitmix* r) { return (double)(sm_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static double sm_normal(splitmix* r) {
    double u1 = sm_uniform(r);
    const double u2 = sm_uniform(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Blurred dot field in [20, 230] DN. */
static int speckle_field(int w, int h, double density, double sigma, splitmix* r, float* out) {
    const size_t n = (size_t)w * (size_t)h;
    float* dots = (float*)calloc(n, sizeof(float));
    float* tmp = (float*)calloc(n, sizeof(float));
    if (!dots || !tmp) {
        free(dots);
        free(tmp);
        return -1;
    }
    for (size_t i = 0; i < n; ++i) dots[i] = sm_uniform(r) < density ? 1.0f : 0.0f;
    int rad = (int)ceil(3.0 * sigma);
    if (rad < 1) rad = 1;
    float k[64];
    if (rad > 31) rad = 31;
    double ks = 0;
    for (int i = -rad; i <= rad; ++i) {
        k[i + rad] = (float)exp(-0.5 * (double)(i * i) / (sigma * sigma));
        ks += k[i + rad];
    }
    for (int i = 0; i <= 2 * rad; ++i) k[i] = (float)(k[i] / ks);
    for (int y = 0; y < h; ++y) /* horizontal pass, clamped borders */
        for (int x = 0; x < w; ++x) {
            float acc = 0;
            for (int i = -rad; i <= rad; ++i) {
                int xx = x + i;
                xx = xx < 0 ? 0 : (xx >= w ? w - 1 : xx);
                acc += k[i + rad] * dots[(size_t)y * (size_t)w + (size_t)xx];
            }
            tmp[(size_t)y * (size_t)w + (size_t)x] = acc;
        }
    float vmax = 0;
    for (int y = 0; y < h; ++y) /* vertical pass */
        for (int x = 0; x < w; ++x) {
            float acc = 0;
            for (int i = -rad; i <= rad; ++i) {
                int yy = y + i;
                yy = yy < 0 ? 0 : (yy >= h ? h - 1 : yy);
                acc += k[i + rad] * tmp[(size_t)yy * (size_t)w + (size_t)x];
            }
            out[(size_t)y * (size_t)w + (size_t)x] = acc;
            if (acc > vmax) vmax = acc;
        }
    const float scale = vmax > 0 ? 210.0f / vmax : 0.0f;
    for (size_t i = 0; i < n; ++i) out[i] = 20.0f + out[i] * scale;
    free(dots);
    free(tmp);
    return 0;
}

static uint8_t to_u8(double v) {
    const double r = floor(v + 0.5);
    return (uint8_t)(r < 0 ? 0 : (r > 255 ? 255 : r));
}

/* Returns min_disparity (>= INT32_MIN/2) on success, INT32_MIN on bad input or allocation failure.
 * gt_disparity (W*H floats, may be NULL) receives the left image's true disparity. */
int32_t synth_speckle_pair(const synth_params* p, uint8_t* left, uint8_t* right, float* gt_disparity) {
    const int w = p->width, h = p->height, nd = p->num_disparities;
    if (w <= 0 || h <= 0 || nd < 6 || p->z_min <= 0 || p->z_max <= p->z_min || p->min_cells < 1 ||
        p->max_cells < p->min_cells)
        return INT32_MIN;
    const double fb = (double)(nd - 5) / (1.0 / p->z_min - 1.0 / p->z_max);
    const int32_t min_d = (int32_t)floor(fb / p->z_max) - 2;
    const size_t n = (size_t)w * (size_t)h;
    splitmix scene = {p->seed * 0x100000001B3ull + 1}, tex_rng = {p->seed ^ 0xA5A5A5A55A5A5A5Aull},
             occ_rng = {p->seed ^ 0x3C3C3C3CC3C3C3C3ull}, noise_rng = {p->seed ^ 0x0F0F0F0FF0F0F0F0ull};

    const int ncell = p->min_cells + (int)(sm_next(&scene) % (uint64_t)(p->max_cells - p->min_cells + 1));
    double* cell = (double*)malloc(sizeof(double) * 5 * (size_t)ncell); /* sx, sy, z0, gx, gy */
    float* disp = (float*)malloc(sizeof(float) * n);
    int* label = (int*)malloc(sizeof(int) * n);
    float* tex = (float*)malloc(sizeof(float) * n);
    float* occ = (float*)malloc(sizeof(float) * n);
    float* rimg = (float*)malloc(sizeof(float) * n);
    float* zbuf = (float*)malloc(sizeof(float) * n);
    int32_t rc = min_d;
    if (!cell || !disp || !label || !tex || !occ || !rimg || !zbuf) {
        rc = INT32_MIN;
        goto done;
    }
    for (int c = 0; c < ncell; ++c) {
        cell[5 * c + 0] = sm_uniform(&scene) * w;
        cell[5 * c + 1] = sm_uniform(&scene) * h;
        cell[5 * c + 2] = p->z_min + sm_uniform(&scene) * (p->z_max - p->z_min);
        cell[5 * c + 3] = (2.0 * sm_uniform(&scene) - 1.0) * p->max_gradient;
        cell[5 * c + 4] = (2.0 * sm_uniform(&scene) - 1.0) * p->max_gradient;
    }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int best = 0;
            double bd = 1e300;
            for (int c = 0; c < ncell; ++c) {
                const double ddx = x - cell[5 * c], ddy = y - cell[5 * c + 1];
                const double d2 = ddx * ddx + ddy * ddy;
                if (d2 < bd) {
                    bd = d2;
                    best = c;
                }
            }
            double z = cell[5 * best + 2] + cell[5 * best + 3] * (x - cell[5 * best]) + cell[5 * best + 4] * (y - cell[5 * best + 1]);
            z = z < p->z_min ? p->z_min : (z > p->z_max ? p->z_max : z);
            label[(size_t)y * (size_t)w + (size_t)x] = best;
            disp[(size_t)y * (size_t)w + (size_t)x] = (float)(fb / z);
        }
    if (speckle_field(w, h, p->dot_density, p->blur_sigma, &tex_rng, tex) != 0 ||
        speckle_field(w, h, p->dot_density, p->blur_sigma, &occ_rng, occ) != 0) {
        rc = INT32_MIN;
        goto done;
    }
    for (size_t i = 0; i < n; ++i) zbuf[i] = -1.0f;
    for (int y = 0; y < h; ++y) {
        const size_t row = (size_t)y * (size_t)w;
        for (int x = 0; x + 1 < w; ++x) {
            if (label[row + (size_t)x] != label[row + (size_t)x + 1]) continue; /* depth edge: leave a gap */
            const double d0 = disp[row + (size_t)x], d1 = disp[row + (size_t)x + 1];
            const double xa = x - d0, xb = x + 1 - d1;
            if (!(xb > xa)) continue;
            int xr = (int)ceil(xa);
            if (xr < 0) xr = 0;
            for (; xr < w && (double)xr < xb; ++xr) {
                const double t = (xr - xa) / (xb - xa);
                const float dd = (float)(d0 + t * (d1 - d0));
                if (dd > zbuf[row + (size_t)xr]) {
                    zbuf[row + (size_t)xr] = dd;
                    rimg[row + (size_t)xr] = (float)(tex[row + (size_t)x] * (1.0 - t) + tex[row + (size_t)x + 1] * t);
                }
            }
        }
        for (int xr = 0; xr < w; ++xr)
            if (zbuf[row + (size_t)xr] < 0) rimg[row + (size_t)xr] = occ[row + (size_t)xr];
    }
    for (size_t i = 0; i < n; ++i) left[i] = to_u8(tex[i] + p->noise_sigma * sm_normal(&noise_rng));
    for (size_t i = 0; i < n; ++i) right[i] = to_u8(rimg[i] + p->noise_sigma * sm_normal(&noise_rng));
    if (gt_disparity) memcpy(gt_disparity, disp, sizeof(float) * n);
done:
    free(cell);
    free(disp);
    free(label);
    free(tex);
    free(occ);
    free(rimg);
    free(zbuf);
    return rc;
}
