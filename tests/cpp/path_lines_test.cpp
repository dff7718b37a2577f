// The path kernel's scanline enumeration (open_source_einstar_amd/csrc/path_lines.h) on the CPU:
// for every direction of many image sizes, the lines are whole scanlines (each starts where its
// line enters the image and ends where it leaves) and together they walk every pixel of the
// direction exactly once. Prints "ok" and exits 0, or names the first failure and exits 1.
#define __host__
#define __device__
#include <cstdio>
#include <vector>

#include "path_lines.h"

namespace {

bool inside(int x, int y, int w, int h) { return x >= 0 && y >= 0 && x < w && y < h; }

bool check(int w, int h) {
    for (int dir = 0; dir < 8; ++dir) {
        const int dx = sgm::path_dx(dir), dy = sgm::path_dy(dir);
        std::vector<int> hits((size_t)w * h, 0);
        const int lines = sgm::path_lines(dir, w, h);
        for (int li = 0; li < lines; ++li) {
            const sgm::PathLine s = sgm::path_line(dir, li, w, h);
            int x = s.x, y = s.y;
            if (s.n < 1 || !inside(x, y, w, h) || inside(x - dx, y - dy, w, h)) {
                std::printf("%dx%d dir %d line %d: (%d,%d) n=%d is not where a line enters\n", w, h, dir, li, x, y, s.n);
                return false;
            }
            for (int i = 0; i < s.n; ++i, x += dx, y += dy) {
                if (!inside(x, y, w, h)) {
                    std::printf("%dx%d dir %d line %d: step %d leaves the image\n", w, h, dir, li, i);
                    return false;
                }
                ++hits[(size_t)y * w + x];
            }
            if (inside(x, y, w, h)) {
                std::printf("%dx%d dir %d line %d: ends inside the image at (%d,%d)\n", w, h, dir, li, x, y);
                return false;
            }
        }
        for (int i = 0; i < w * h; ++i)
            if (hits[(size_t)i] != 1) {
                std::printf("%dx%d dir %d: pixel (%d,%d) walked %d times\n", w, h, dir, i % w, i / w, hits[(size_t)i]);
                return false;
            }
    }
    return true;
}

}  // namespace

int main() {
    for (int w = 1; w <= 48; ++w)
        for (int h = 1; h <= 48; ++h)
            if (!check(w, h)) return 1;
    const int big[][2] = {{1280, 1024}, {1024, 1280}, {2048, 1536}, {640, 512}, {4096, 7}, {3, 2000}, {320, 240}};
    for (const auto& s : big)
        if (!check(s[0], s[1])) return 1;
    std::printf("ok\n");
    return 0;
}
