// tracegen.hip -- seeded synthetic trace generation on the GPU (one thread
// per ZMW, per-ZMW counter-based streams; identical to the host generator
// because both compile include/hwpa_tracegen.h).

#include "internal.h"
#include "../../include/hwpa_tracegen.h"

namespace hwpa {

__global__ __launch_bounds__(64) void TracegenKernel(hwpa_tracegen_params p, uint32_t lane_offset, uint32_t frames,
                                                     uint32_t first_frame, hwpa_tracegen_state* state, int32_t init,
                                                     int16_t* out)
{
    const uint32_t lane = blockIdx.x;
    const int i = threadIdx.x;
    hwpa_tracegen_state& st = state[lane];
    const uint64_t z = (uint64_t)(lane + lane_offset) * LS + (uint64_t)i;
    if (init) hw_tracegen_init_zmw(&p, z, &st, i);
    int16_t* o = out + (size_t)lane * frames * LS + i;
    for (uint32_t t = 0; t < frames; ++t) o[(size_t)t * LS] = hw_tracegen_frame(&p, z, first_frame + t, &st, i);
}

int32_t TracegenRun(const hwpa_tracegen_params& p, uint32_t lanes, uint32_t lane_offset, uint32_t frames,
                    uint32_t first_frame, hwpa_tracegen_state* state, int32_t init, int16_t* out, hipStream_t s)
{
    if (lanes == 0 || frames == 0) return HWPA_OK;
    hipLaunchKernelGGL(TracegenKernel, dim3(lanes), dim3(LS), 0, s, p, lane_offset, frames, first_frame, state, init,
                       out);
    return CheckLaunch("TracegenKernel");
}

}  // namespace hwpa
