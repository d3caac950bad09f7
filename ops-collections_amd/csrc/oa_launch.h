// Per-layout launchers of the open-addressing kernels.  Each slot layout is
// instantiated in its own translation unit (oa_inst_*.hip) so the kernels
// compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "mxco/detail/oa_kernels.h"

namespace mxco {

template <int KS, int VS, int G>
hipError_t launch_oa_g(int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    const dim3 block(256);
    if constexpr (!kWideLayout<KS, VS>) inflight = 1;  // 1- and 2-byte fields: the plain kernel only
    if constexpr (kWideLayout<KS, VS>) {
    if (inflight == 4) {
        // wave-staged outputs: the bulk Find / Contains paths
        switch (op) {
        case OP_FIND: oa_wave_staged_kernel<KS, VS, G, OP_FIND><<<grid, block, 0, s>>>(a); return hipGetLastError();
        case OP_CONTAINS:
            oa_wave_staged_kernel<KS, VS, G, OP_CONTAINS><<<grid, block, 0, s>>>(a);
            return hipGetLastError();
        default: break;
        }
    }
    if (inflight == 3) {
        // LDS-staged chunks: the streaming Find / Contains / Insert / Count paths
        switch (op) {
        case OP_FIND: oa_staged_kernel<KS, VS, G, OP_FIND><<<grid, block, 0, s>>>(a); return hipGetLastError();
        case OP_CONTAINS: oa_staged_kernel<KS, VS, G, OP_CONTAINS><<<grid, block, 0, s>>>(a); return hipGetLastError();
        case OP_INSERT: oa_staged_kernel<KS, VS, G, OP_INSERT><<<grid, block, 0, s>>>(a); return hipGetLastError();
        case OP_COUNT: oa_staged_kernel<KS, VS, G, OP_COUNT><<<grid, block, 0, s>>>(a); return hipGetLastError();
        default: break;
        }
    }
    if (inflight == 2) {
        // two keys in flight per group: the bulk Find / Contains / Insert paths
        switch (op) {
        case OP_FIND: oa_kernel<KS, VS, G, OP_FIND, 2><<<grid, block, 0, s>>>(a); return hipGetLastError();
        case OP_CONTAINS: oa_kernel<KS, VS, G, OP_CONTAINS, 2><<<grid, block, 0, s>>>(a); return hipGetLastError();
        case OP_INSERT: oa_kernel<KS, VS, G, OP_INSERT, 2><<<grid, block, 0, s>>>(a); return hipGetLastError();
        default: break;
        }
    }
    }
    switch (op) {
    case OP_FIND: oa_kernel<KS, VS, G, OP_FIND><<<grid, block, 0, s>>>(a); break;
    case OP_CONTAINS: oa_kernel<KS, VS, G, OP_CONTAINS><<<grid, block, 0, s>>>(a); break;
    case OP_COUNT: oa_kernel<KS, VS, G, OP_COUNT><<<grid, block, 0, s>>>(a); break;
    case OP_INSERT: oa_kernel<KS, VS, G, OP_INSERT><<<grid, block, 0, s>>>(a); break;
    case OP_INSERT_AND_FIND: oa_kernel<KS, VS, G, OP_INSERT_AND_FIND><<<grid, block, 0, s>>>(a); break;
    case OP_ERASE: oa_kernel<KS, VS, G, OP_ERASE><<<grid, block, 0, s>>>(a); break;
    case OP_PROBE_STATS: oa_kernel<KS, VS, G, OP_PROBE_STATS><<<grid, block, 0, s>>>(a); break;
    case OP_INSERT_STATS: oa_kernel<KS, VS, G, OP_INSERT_STATS><<<grid, block, 0, s>>>(a); break;
    case OP_INSERT_OR_ASSIGN:
        if constexpr (VS == 0) return hipErrorInvalidValue;
        else oa_kernel<KS, VS, G, OP_INSERT_OR_ASSIGN><<<grid, block, 0, s>>>(a);
        break;
    case OP_ASSIGN:
        if constexpr (VS == 0) return hipErrorInvalidValue;
        else oa_kernel<KS, VS, G, OP_ASSIGN><<<grid, block, 0, s>>>(a);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int KS, int VS>
hipError_t launch_oa_layout(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight)
{
    switch (g) {
    case 1: return launch_oa_g<KS, VS, 1>(op, a, grid, s, inflight);
    case 2: return launch_oa_g<KS, VS, 2>(op, a, grid, s, inflight);
    case 4: return launch_oa_g<KS, VS, 4>(op, a, grid, s, inflight);
    case 8: return launch_oa_g<KS, VS, 8>(op, a, grid, s, inflight);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_oa_k4v4(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
hipError_t launch_oa_k8v8(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
hipError_t launch_oa_k4v8(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
hipError_t launch_oa_k8v4(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
hipError_t launch_oa_k4(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
hipError_t launch_oa_k8(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
// layouts with a 1- or 2-byte key or payload (oa_inst_narrow_*.hip)
#define MXCO_NARROW_LAYOUTS(X) \
    X(1, 0) X(1, 1) X(1, 2) X(1, 4) X(1, 8) X(2, 0) X(2, 1) X(2, 2) X(2, 4) X(2, 8) X(4, 1) X(4, 2) X(8, 1) X(8, 2)
#define MXCO_DECLARE_NARROW(K, V) \
    hipError_t launch_oa_k##K##v##V(int g, int op, const OaArgs& a, dim3 grid, hipStream_t s, int inflight);
MXCO_NARROW_LAYOUTS(MXCO_DECLARE_NARROW)
#undef MXCO_DECLARE_NARROW

inline hipError_t launch_oa(int ks, int vs, int g, int op, const OaArgs& a, dim3 grid, hipStream_t s,
                            int inflight = 1)
{
    if (ks == 4 && vs == 4) return launch_oa_k4v4(g, op, a, grid, s, inflight);
    if (ks == 8 && vs == 8) return launch_oa_k8v8(g, op, a, grid, s, inflight);
    if (ks == 4 && vs == 8) return launch_oa_k4v8(g, op, a, grid, s, inflight);
    if (ks == 8 && vs == 4) return launch_oa_k8v4(g, op, a, grid, s, inflight);
    if (ks == 4 && vs == 0) return launch_oa_k4(g, op, a, grid, s, inflight);
    if (ks == 8 && vs == 0) return launch_oa_k8(g, op, a, grid, s, inflight);
#define MXCO_DISPATCH_NARROW(K, V) \
    if (ks == K && vs == V) return launch_oa_k##K##v##V(g, op, a, grid, s, inflight);
    MXCO_NARROW_LAYOUTS(MXCO_DISPATCH_NARROW)
#undef MXCO_DISPATCH_NARROW
    return hipErrorInvalidValue;
}

}  // namespace mxco
