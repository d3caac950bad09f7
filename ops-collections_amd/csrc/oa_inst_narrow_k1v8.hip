// Open-addressing kernels for the 1-byte key / 8-byte payload slot layout.
#include "oa_launch.h"

namespace mxco {
hipError_t launch_oa_k1v8(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    return launch_oa_layout<1, 8>(g, op, a, grid, s, inflight);
}
}  // namespace mxco
