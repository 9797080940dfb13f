// tgk_fast.hip -- the fast tracking path: persistent lanes that refill from a
// device seed queue, seed-major dense output, and its compaction.
#include <hip/hip_runtime.h>

#include "tgk_internal.h"

namespace tgk {

namespace {

// Dense seed-major rows -> ragged points. One wavefront per streamline copies
// its lengths[g] * 3 floats, coalesced on both sides.
__global__ __launch_bounds__(256) void compact_seed_major_kernel(const float *__restrict__ dense, int64_t n_seeds,
                                                                 uint32_t n_points,
                                                                 const uint32_t *__restrict__ lengths,
                                                                 const int64_t *__restrict__ offsets,
                                                                 float *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t g = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < n_seeds; g += waves) {
        const int64_t count = (int64_t)lengths[g] * 3;
        const float *src = dense + g * (int64_t)n_points * 3;
        float *dst = out + offsets[g] * 3;
        for (int64_t i = lane; i < count; i += 64)
            dst[i] = src[i];
    }
}

}  // namespace

hipError_t launch_compact_seed_major(const float *dense, int64_t n_seeds, uint32_t n_points,
                                     const uint32_t *lengths, const int64_t *offsets, float *out,
                                     hipStream_t stream)
{
    if (n_seeds == 0)
        return hipSuccess;
    const int64_t blocks = (n_seeds + 3) / 4;
    compact_seed_major_kernel<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, stream>>>(
        dense, n_seeds, n_points, lengths, offsets, out);
    return hipGetLastError();
}

}  // namespace tgk
