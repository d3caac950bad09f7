// Open-addressing kernels for the 4-byte key / 2-byte payload slot layout.
#include "oa_launch.h"

namespace mxco {
hipError_t launch_oa_k4v2(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    return launch_oa_layout<4, 2>(g, op, a, grid, s, inflight);
}
}  // namespace mxco
