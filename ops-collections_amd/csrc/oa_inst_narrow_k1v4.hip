// Open-addressing kernels for the 1-byte key / 4-byte payload slot layout.
#include "oa_launch.h"

namespace mxco {
hipError_t launch_oa_k1v4(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    return launch_oa_layout<1, 4>(g, op, a, grid, s, inflight);
}
}  // namespace mxco
