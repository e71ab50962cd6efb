// Film post-processing for gfx950: the picture a save writes.
//
// Pipeline::final_picture (src/base/mgr/pipeline.cpp:313-330) runs the
// tone-mapping pass of Postprocessor::compile_tone_mapping
// (src/base/mgr/postprocessor.cpp:13-21) over the collector's output buffer
// (which already holds tone_map(1 - exp(-v * exposure))): the tone mapper is
// applied a second time, then linear_to_srgb unless the file is EXR or HDR,
// and w is forced to 1.  One thread per pixel, float4 loads and stores: the
// pass reads 16 B and writes 16 B per pixel and is HBM-bound.
#include <hip/hip_runtime.h>

#include "dev_pt.h"
#include "launch.h"

namespace vs {

using namespace dev;

__global__ void __launch_bounds__(256) final_picture_kernel(const float4 *__restrict__ in, float4 *__restrict__ out,
                                                             uint32_t n, uint32_t tone_mapper, uint32_t gamma) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = in[i];
    float4 o;
    o.x = tone_map(v.x, tone_mapper);
    o.y = tone_map(v.y, tone_mapper);
    o.z = tone_map(v.z, tone_mapper);
    if (gamma) {
        o.x = linear_to_srgb(o.x);
        o.y = linear_to_srgb(o.y);
        o.z = linear_to_srgb(o.z);
    }
    o.w = 1.f;
    out[i] = o;
}

hipError_t launch_final_picture(const float4 *in, float4 *out, uint32_t n, uint32_t tone_mapper, bool gamma,
                                hipStream_t stream) {
    if (n == 0) return hipSuccess;
    final_picture_kernel<<<(n + 255u) / 256u, 256, 0, stream>>>(in, out, n, tone_mapper, gamma ? 1u : 0u);
    return hipGetLastError();
}

}// namespace vs
