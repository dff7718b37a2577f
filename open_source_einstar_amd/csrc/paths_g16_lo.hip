// Path-kernel instantiations: 16 lanes per scanline, 2, 4 or 8 disparities per lane (sgm_kernels.h).
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_paths_g16_lo(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    switch (dpl) {
        case 2: return launch_paths_variant<2, 16>(pl, t, c, stream);
        case 4: return launch_paths_variant<4, 16>(pl, t, c, stream);
        case 8: return launch_paths_variant<8, 16>(pl, t, c, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sgm
