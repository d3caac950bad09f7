// Open-addressing kernels for the 8-byte key / 1-byte payload slot layout.
#include "oa_launch.h"

namespace mxco {
hipError_t launch_oa_k8v1(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    return launch_oa_layout<8, 1>(g, op, a, grid, s, inflight);
}
}  // namespace mxco
