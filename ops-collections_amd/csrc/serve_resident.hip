// Instantiations and launch of the resident serving kernel (serve_resident.h)
// for the StaticMap layouts the serving engine takes: 4- and 8-byte keys and
// payloads, 1, 2, 4 or 8 lanes per key.
#include "serve_resident.h"

namespace mxco {

namespace {

template <int KS, int VS, int G>
hipError_t launch_g(const OaArgs& a, ServeMailbox* mb, ServeCtl* ctl,
                    uint64_t seq_base, uint64_t idle, uint64_t life, uint64_t hard, unsigned grid, hipStream_t s)
{
    serve_resident_kernel<KS, VS, G><<<grid, kServeThreads, 0, s>>>(a, mb, ctl, seq_base, idle, life, hard);
    return hipGetLastError();
}

template <int KS, int VS, int G>
int occupancy_g()
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void*)serve_resident_kernel<KS, VS, G>,
                                                     kServeThreads, 0) !=
        hipSuccess)
        return 0;
    return n;
}

template <int KS, int VS>
hipError_t launch_l(int g, const OaArgs& a, ServeMailbox* mb,
                    ServeCtl* ctl, uint64_t seq_base, uint64_t idle, uint64_t life, uint64_t hard, unsigned grid,
                    hipStream_t s)
{
    switch (g) {
    case 1: return launch_g<KS, VS, 1>(a, mb, ctl, seq_base, idle, life, hard, grid, s);
    case 2: return launch_g<KS, VS, 2>(a, mb, ctl, seq_base, idle, life, hard, grid, s);
    case 4: return launch_g<KS, VS, 4>(a, mb, ctl, seq_base, idle, life, hard, grid, s);
    case 8: return launch_g<KS, VS, 8>(a, mb, ctl, seq_base, idle, life, hard, grid, s);
    default: return hipErrorInvalidValue;
    }
}

template <int KS, int VS>
int occupancy_l(int g)
{
    switch (g) {
    case 1: return occupancy_g<KS, VS, 1>();
    case 2: return occupancy_g<KS, VS, 2>();
    case 4: return occupancy_g<KS, VS, 4>();
    case 8: return occupancy_g<KS, VS, 8>();
    default: return 0;
    }
}

}  // namespace

hipError_t launch_serve_resident(int ks, int vs, int g, const OaArgs& a, ServeMailbox* mb,
                                 ServeCtl* ctl, uint64_t seq_base, uint64_t idle, uint64_t life,
                                 uint64_t hard, unsigned grid, hipStream_t s)
{
    if (ks == 8 && vs == 8) return launch_l<8, 8>(g, a, mb, ctl, seq_base, idle, life, hard, grid, s);
    if (ks == 4 && vs == 4) return launch_l<4, 4>(g, a, mb, ctl, seq_base, idle, life, hard, grid, s);
    if (ks == 4 && vs == 8) return launch_l<4, 8>(g, a, mb, ctl, seq_base, idle, life, hard, grid, s);
    if (ks == 8 && vs == 4) return launch_l<8, 4>(g, a, mb, ctl, seq_base, idle, life, hard, grid, s);
    return hipErrorInvalidValue;
}

int serve_resident_occupancy(int ks, int vs, int g)
{
    if (ks == 8 && vs == 8) return occupancy_l<8, 8>(g);
    if (ks == 4 && vs == 4) return occupancy_l<4, 4>(g);
    if (ks == 4 && vs == 8) return occupancy_l<4, 8>(g);
    if (ks == 8 && vs == 4) return occupancy_l<8, 4>(g);
    return 0;
}

}  // namespace mxco
