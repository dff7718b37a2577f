// Path-kernel instantiations: 64 lanes per scanline (the whole wave), each lane holding a quarter of
// the disparities of the 16-lane layout `dpl` (sgm_kernels.h). Only layouts that leave 4 or 8 per
// lane: with 2 or 6 a lane would move its path costs in 2-byte pieces.
#include "sgm_kernels.h"

namespace sgm {

hipError_t launch_paths_g64(int dpl, const PathLaunch& pl, const LaunchTable& t, const SgmConsts& c, hipStream_t stream) {
    switch (dpl) {
        case 16: return launch_paths_variant<4, 64>(pl, t, c, stream);
        case 32: return launch_paths_variant<8, 64>(pl, t, c, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace sgm
