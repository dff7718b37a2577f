// Path-kernel instantiations: 16 lanes per scanline, 12..32 disparities per lane (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_paths_g16_hi(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    switch (dpl) {
        case 12: return launch_paths_variant<12, 16>(pl, t, c, stream);
        case 16: return launch_paths_variant<16, 16>(pl, t, c, stream);
        case 24: return launch_paths_variant<24, 16>(pl, t, c, stream);
        case 32: return launch_paths_variant<32, 16>(pl, t, c, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sgm
