// gfx950 kernels after the disparity: the organised point image with normals and weights
// (depth::disparity_to_points, libs/depth/src/point_image.cpp:10-55) and the packing of a depth
// frame into the session file's uncompressed layout (libs/session/src/session.cpp:143-160,
// libs/depth_metal/src/depth_packer.cpp:12-26).
//
// Points are formed in double exactly as the CPU forms them; normals and weights keep the CPU's
// float expressions in its order (Eigen's cross, dot, norm and normalized spelled out), so with
// -ffp-contract=off and correctly rounded division and square root every value equals the CPU's.
// The packing quantises z * 50 in double, as the session writer does, not in float.
#include "sgm_internal.h"

namespace sgm {
namespace {

constexpr int kTw = 16, kTh = 16;

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return ax * bx + ay * by + az * bz;
}

__global__ __launch_bounds__(kTw* kTh) void points_kernel(PointsDesc P) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= P.w || y >= P.h) return;
    const int64_t i = (int64_t)y * P.w + x;
    const float d = P.disp[i];
    f32x4 p = {0.0f, 0.0f, 0.0f, 0.0f};
    float wt = 0.0f;
    if (d > -1.0e8f && d + P.cx_offset > 0.0) {
        const double z = P.f * P.baseline / (d + P.cx_offset);
        if (!(z < P.min_depth || z > P.max_depth)) {
            p = f32x4{(float)((x - P.cx) * z / P.f), (float)((y - P.cy) * z / P.f), (float)z, 1.0f};
            wt = P.conf ? P.conf[i] : 1.0f;
        }
    }
    reinterpret_cast<f32x4*>(P.points)[i] = p;
    P.weights[i] = wt;
}

__global__ __launch_bounds__(kTw* kTh) void normals_kernel(PointsDesc P) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= P.w || y >= P.h) return;
    const int w = P.w;
    const int64_t i = (int64_t)y * w + x;
    const f32x4* pts = reinterpret_cast<const f32x4*>(P.points);
    const f32x4 p = pts[i];
    f32x4 n = {0.0f, 0.0f, 0.0f, 0.0f};
    if (p.z != 0.0f) {
        float view = P.edge_weight;
        if (x > 0 && y > 0 && x + 1 < w && y + 1 < P.h) {
            const f32x4 l = pts[i - 1], r = pts[i + 1], u = pts[i - w], dn = pts[i + w];
            if (l.z != 0 && r.z != 0 && u.z != 0 && dn.z != 0 && fabsf(l.z - r.z) <= P.max_depth_jump &&
                fabsf(u.z - dn.z) <= P.max_depth_jump) {
                const float ax = r.x - l.x, ay = r.y - l.y, az = r.z - l.z;
                const float bx = dn.x - u.x, by = dn.y - u.y, bz = dn.z - u.z;
                float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
                const float len = __builtin_sqrtf(dot3(nx, ny, nz, nx, ny, nz));
                if (len >= 1e-9f) {
                    nx /= len;
                    ny /= len;
                    nz /= len;
                    if (dot3(nx, ny, nz, p.x, p.y, p.z) > 0) {
                        nx = -nx;
                        ny = -ny;
                        nz = -nz;
                    }
                    n = f32x4{nx, ny, nz, 1.0f};
                    const float pn = __builtin_sqrtf(dot3(p.x, p.y, p.z, p.x, p.y, p.z));
                    const float qx = p.x - P.rc_x, qy = p.y - 0.0f, qz = p.z - 0.0f;
                    const float qn = __builtin_sqrtf(dot3(qx, qy, qz, qx, qy, qz));
                    const float v1 = -dot3(nx, ny, nz, p.x / pn, p.y / pn, p.z / pn);
                    const float v2 = -dot3(nx, ny, nz, qx / qn, qy / qn, qz / qn);
                    const float m = v2 < v1 ? v2 : v1;  // std::min(v1, v2)
                    view = 0.0f < m ? m : 0.0f;        // std::max(0.0f, m)
                }
            }
        }
        // depth_noise_weight, point_image.hpp:51-55
        float dnw = 1.0f;
        if (P.weight_reference_depth > 0.0f) {
            const float rr = P.weight_reference_depth / p.z;
            dnw = rr * rr > 1.0f ? 1.0f : (rr * rr < 0.1f ? 0.1f : rr * rr);
        }
        P.weights[i] *= view * dnw;
    }
    reinterpret_cast<f32x4*>(P.normals)[i] = n;
}

// q = lround(max(0, z) * 50) in [1, 65535], 0 where z <= 0 (session.cpp:150-153, :304).
__device__ __forceinline__ uint32_t quantize(float z) {
    if (!(z > 0.0f)) return 0;
    const double q = __builtin_round((double)z * 50.0);  // half away from zero, as lround
    return q < 1.0 ? 1u : (q > 65535.0 ? 65535u : (uint32_t)q);
}

__global__ __launch_bounds__(kTw* kTh) void pack_kernel(const float* points, const float* weights, uint8_t* out, int w, int h) {
    const int x = (int)blockIdx.x * kTw + (int)threadIdx.x, y = (int)blockIdx.y * kTh + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    const int64_t i = (int64_t)y * w + x;
    const int64_t n = (int64_t)w * h;
    const f32x4* pts = reinterpret_cast<const f32x4*>(points);
    const uint32_t q = quantize(pts[i].z);
    const uint32_t left = x > 0 ? quantize(pts[i - 1].z) : 0u;
    reinterpret_cast<uint16_t*>(out)[i] = (uint16_t)((q - left) & 0xFFFFu);
    const float c = weights ? weights[i] : 1.0f;
    const float r = __builtin_roundf(c * 255.0f);  // lround of the float product
    out[2 * n + i] = (uint8_t)(r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r));
}

dim3 grid2d(int w, int h) { return dim3((unsigned)((w + kTw - 1) / kTw), (unsigned)((h + kTh - 1) / kTh)); }

}  // namespace

hipError_t launch_points(const PointsDesc& p, hipStream_t stream) {
    if (p.w <= 0 || p.h <= 0) return hipSuccess;
    hipLaunchKernelGGL(points_kernel, grid2d(p.w, p.h), dim3(kTw, kTh), 0, stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(normals_kernel, grid2d(p.w, p.h), dim3(kTw, kTh), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_pack(const float* points, const float* weights, uint8_t* out, int w, int h, hipStream_t stream) {
    if (w <= 0 || h <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, grid2d(w, h), dim3(kTw, kTh), 0, stream, points, weights, out, w, h);
    return hipGetLastError();
}

}  // namespace sgm
