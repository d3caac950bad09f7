// Open-addressing kernels for the 8-byte key / 0-byte payload slot layout.
#include "oa_launch.h"

namespace mxco {
hipError_t launch_oa_k8(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    return launch_oa_layout<8, 0>(g, op, a, grid, s, inflight);
}
}  // namespace mxco
