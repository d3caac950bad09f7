// Open-addressing kernels for the 2-byte key set slot layout.
#include "oa_launch.h"

namespace mxco {
hipError_t launch_oa_k2v0(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    return launch_oa_layout<2, 0>(g, op, a, grid, s, inflight);
}
}  // namespace mxco
