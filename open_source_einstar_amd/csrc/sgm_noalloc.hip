// sgm_compute_device_noalloc (include/sgm_hip.h): the SGM pipeline of sgm_compute_device without a
// context and without any device allocation. Its scratch -- census words, step penalties, the byte
// cost volume, the per-direction path costs and, when the caller passes none, the disparity map --
// lives in one file-scope device array of kScratchBytes, which every GPU that runs this code holds
// once, from the loading of this code object to the end of the process.
//
// Host code never passes the array's name to a kernel: the name is a host-side stand-in, not a device
// address. Its device address is looked up once per device (hipGetSymbolAddress, with the device
// current) and that address is what the kernels receive.
#include <mutex>
#include <string>

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

constexpr size_t kScratchBytes = SGM_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once looked up

size_t up256(size_t v) { return (v + 255) & ~size_t(255); }

// Byte offsets of one call's buffers inside the scratch array.
struct Layout {
    size_t census, cvol, lvol, disp, total;
    bool use_cvol;
};

Layout layout_for(int w, int h, int min_d, int nd, int ndirs, bool need_disp, bool wide) {
    const size_t npx = (size_t)w * h;
    const size_t dp = (size_t)dpl_for(nd) * kGroup;
    Layout L{};
    size_t o = 0;
    L.census = o;
    o = up256(o + 2 * census_words_ext(w, h, min_d, nd) * 8 + npx * edge_bytes_per_px(wide));  // both images, then penalties
    L.lvol = o;
    o = up256(o + (size_t)ndirs * npx * dp * (wide ? 2 : 1));
    L.disp = o;
    if (need_disp) o = up256(o + npx * 4);
    // The byte cost volume (read once per direction instead of recomputing the costs) where it still fits.
    L.cvol = o;
    L.use_cvol = o + npx * dp <= kScratchBytes && npx * dp <= kMaxViewBytes;
    L.total = L.use_cvol ? up256(o + npx * dp) : o;
    return L;
}

}  // namespace

extern "C" {

size_t sgm_noalloc_scratch_bytes(void) { return kScratchBytes; }

int sgm_compute_device_noalloc(const sgm_params* params, const uint8_t* d_left, const uint8_t* d_right, ptrdiff_t pitch,
                               int width, int height, const sgm_outputs* d_out, void* stream) {
    std::string err;
    if (int rc = validate_params_ext(params, err); rc != SGM_OK) {
        set_thread_error("sgm_compute_device_noalloc: " + err);
        return rc;
    }
    if (width < 1 || height < 1 || width > kMaxWidth || !d_left || !d_right || pitch < width) {
        set_thread_error("sgm_compute_device_noalloc: width must be in [1, " + std::to_string(kMaxWidth) +
                         "], height >= 1, images non-null and pitch >= width");
        return SGM_ERR_INVALID_ARGUMENT;
    }
    if (d_out && (d_out->census_left || d_out->census_right)) {
        set_thread_error("sgm_compute_device_noalloc: census outputs are only available from sgm_compute");
        return SGM_ERR_UNSUPPORTED;
    }
    if (int rc = check_row_lds_ext(width, err); rc != SGM_OK) {
        set_thread_error("sgm_compute_device_noalloc: " + err);
        return rc;
    }
    const int ndirs = params->eight_paths ? 8 : 4;
    const bool need_disp = !(d_out && d_out->disparity);
    const SgmConsts consts = make_consts_ext(*params);
    const Layout L = layout_for(width, height, params->min_disparity, params->num_disparities, ndirs, need_disp,
                                consts.wide_costs != 0);
    if (L.total > kScratchBytes) {
        set_thread_error("sgm_compute_device_noalloc: " + std::to_string(width) + "x" + std::to_string(height) + " with " +
                         std::to_string(params->num_disparities) + " disparities and " + std::to_string(ndirs) +
                         " paths needs " + std::to_string(L.total) + " bytes of scratch, more than the " +
                         std::to_string(kScratchBytes) + " this entry point has (use sgm_compute_device)");
        return SGM_ERR_INVALID_ARGUMENT;
    }
    // The array's device address on the current device, looked up on the first call there.
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_thread_error(std::string("sgm_compute_device_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return SGM_ERR_IO;
    }
    uint8_t* base = nullptr;
    {
        std::lock_guard lk(g_mu);
        if (!g_scratch[dev]) {
            void* ptr = nullptr;
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                set_thread_error(std::string("sgm_compute_device_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
                return SGM_ERR_IO;
            }
            g_scratch[dev] = ptr;
        }
        base = static_cast<uint8_t*>(g_scratch[dev]);
    }
    PairDesc d{};
    d.left = d_left;
    d.right = d_right;
    d.pitch = pitch;
    set_geometry_ext(d, width, height, params->min_disparity, params->num_disparities, ndirs, consts.wide_costs ? 2 : 1);
    const size_t cw = census_words_ext(width, height, params->min_disparity, params->num_disparities);
    d.census_l = reinterpret_cast<uint64_t*>(base + L.census);
    d.census_r = d.census_l + cw;
    d.edge = reinterpret_cast<uint32_t*>(d.census_r + cw);
    d.lvol = base + L.lvol;
    d.cvol = L.use_cvol ? base + L.cvol : nullptr;
    d.disp = need_disp ? reinterpret_cast<float*>(base + L.disp) : d_out->disparity;
    if (d_out) {
        d.median = d_out->disparity_median;
        d.conf = d_out->confidence;
        d.dint = d_out->disparity_int;
        d.status = d_out->status;
        d.rdisp = d_out->right_disparity;
        d.best = d_out->best_cost;
        d.second = d_out->second_cost;
        d.sum_volume = d_out->sum_volume;
        d.cost_volume = d_out->cost_volume;
    }
    // Every scratch byte a kernel reads was written earlier in the same call (census pads are read
    // only where the costs they give are replaced by the invalid cost), so nothing is cleared.
    e = enqueue_pair_ext(d, consts, d.dp != d.nd, static_cast<hipStream_t>(stream), nullptr, 0);
    if (e != hipSuccess) {
        set_thread_error(std::string("sgm_compute_device_noalloc: launch: ") + hipGetErrorString(e));
        return SGM_ERR_IO;
    }
    return SGM_OK;
}

}  // extern "C"
