// Scanlines of the path kernel: which pixels each line index of a direction walks.
//
// Directions follow stereo.cpp:387-388 (dx, dy): 0 (+1, 0), 1 (-1, 0), 2 (0, +1), 3 (0, -1),
// 4 (+1, +1), 5 (-1, -1), 6 (+1, -1), 7 (-1, +1). A direction has one line per pixel where a line
// enters the image, in the order of stereo.cpp:391-402: the pixels of the start column (when
// dx != 0), then those of the start row not already counted, each walked until the line leaves.
//
// Host and device code share these functions (a host-only build defines __host__ and __device__
// away).
#pragma once

namespace sgm {

struct PathLine {
    int x, y, n;  // start pixel and pixel count
};

__host__ __device__ inline int path_dx(int dir) {
    return (dir == 0 || dir == 4 || dir == 6) ? 1 : ((dir == 1 || dir == 5 || dir == 7) ? -1 : 0);
}
__host__ __device__ inline int path_dy(int dir) {
    return (dir == 2 || dir == 4 || dir == 7) ? 1 : ((dir == 3 || dir == 5 || dir == 6) ? -1 : 0);
}

// Lines of direction dir on a w x h image.
__host__ __device__ inline int path_lines(int dir, int w, int h) {
    const int dx = path_dx(dir), dy = path_dy(dir);
    return (dx != 0 ? h : 0) + (dy != 0 ? (dx != 0 ? w - 1 : w) : 0);
}

// Line li (0 <= li < path_lines(dir, w, h)).
__host__ __device__ inline PathLine path_line(int dir, int li, int w, int h) {
    const int dx = path_dx(dir), dy = path_dy(dir);
    PathLine s;
    const int ncol = dx != 0 ? h : 0;
    if (li < ncol) {
        s.x = dx > 0 ? 0 : w - 1;
        s.y = li;
    } else {
        int j = li - ncol;
        if (dx != 0 && j >= (dx > 0 ? 0 : w - 1)) ++j;  // the corner is a column start
        s.x = j;
        s.y = dy > 0 ? 0 : h - 1;
    }
    int n = 0x7FFFFFFF;
    if (dx > 0) n = w - s.x;
    if (dx < 0) n = s.x + 1;
    if (dy > 0) n = n < h - s.y ? n : h - s.y;
    if (dy < 0) n = n < s.y + 1 ? n : s.y + 1;
    s.n = n;
    return s;
}

}  // namespace sgm
