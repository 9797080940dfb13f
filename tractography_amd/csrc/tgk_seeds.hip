// tgk_seeds.hip -- seeding from FOD images (seeds.cl:12-118), VALU method.
//
// Every seed evaluates its voxel's model on the same 1000 Fibonacci
// directions, twice (once for the total weight, once for the cumulative
// search). The reference recomputes the direction and its 45 SH basis values
// for every evaluation; here they are computed once per launch into a table
// (basis_table_kernel, with the plain path's bit-exact basis code) and each
// evaluation is the in-order contraction of the seed's coefficients with a
// table row. The table index is uniform across a wavefront, so rows are read
// through the scalar cache. Results are bit-identical to the reference's
// arithmetic with the tgk_math.h elementary functions.
//
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "tgk_internal.h"
#include "tgk_plain_model.h"

namespace tgk {

namespace {

constexpr int N_DIRECTIONS = 1000;
constexpr int TABLE_STRIDE = 48;

// seeds.cl:12-25
__device__ __forceinline__ V3 fibonacci_direction(uint32_t n, uint32_t n_points)
{
    const float phi = (tgk_sqrtf(5.0f) + 1.0f) * 0.5f;
    const float golden_angle = 2.0f * 3.14159265359f / phi;
    const float theta = (float)n * golden_angle;
    const float z = 1.0f - (2.0f * (float)n / (float)(n_points - 1));
    const float radius = tgk_sqrtf(1.0f - z * z);
    float st, ct;
    tgk_sincosf(theta, &st, &ct);
    return v3(radius * ct, radius * st, z);
}

// Row i: the 45 SH basis values of direction i (spharm.cl:216-413 values,
// through cart2sph as models/symmetric-real-spherical-harmonics.cl:287-295),
// then the direction itself in the padding (45, 46, 47).
__global__ void basis_table_kernel(float *__restrict__ table)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N_DIRECTIONS)
        return;
    const V3 d = fibonacci_direction((uint32_t)i, N_DIRECTIONS);
    float az, col;
    plain::cart2sph(d, az, col);
    // Unit coefficient vectors pick the basis values out of the contraction
    // one at a time, with no rounding (c = 1 times y, plus zeros).
    float c[45];
    for (int k = 0; k < 45; k++) {
        for (int j = 0; j < 45; j++)
            c[j] = j == k ? 1.0f : 0.0f;
        float v, th, ph;
        plain::sh45_contract(c, az, col, v, th, ph);
        table[i * TABLE_STRIDE + k] = v;
    }
    table[i * TABLE_STRIDE + 45] = d.x;
    table[i * TABLE_STRIDE + 46] = d.y;
    table[i * TABLE_STRIDE + 47] = d.z;
}

// Model value at direction i (models/*.cl evaluate_interpolated_model):
// SH: softplus of the in-order contraction; DTI: o . (D o).
template <int MODEL>
__device__ __forceinline__ float model_value(const float *__restrict__ row, const float c[45], int K)
{
    if constexpr (MODEL == MODEL_DTI) {
        const float ox = row[45], oy = row[46], oz = row[47];
        const float t0 = (c[0] * ox + c[3] * oy) + c[4] * oz;
        const float t1 = (c[3] * ox + c[1] * oy) + c[5] * oz;
        const float t2 = (c[4] * ox + c[5] * oy) + c[2] * oz;
        return (ox * t0 + oy * t1) + oz * t2;
    } else {
        float v = 0.0f;
        for (int k = 0; k < K; k++)
            v += c[k] * row[k];
        return plain::softplus(v);
    }
}

// seeds.cl:68-118
template <int MODEL>
__global__ __launch_bounds__(256) void seeds_from_fod_kernel(const float *__restrict__ fod, int K,
                                                             const float4 *__restrict__ voxels, uint32_t nnz,
                                                             const float *__restrict__ affine,
                                                             const float *__restrict__ table, uint2 *__restrict__ states,
                                                             int64_t n, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    uint2 state = states[g];
    const uint32_t index = randi(state, nnz);
    float4 voxel = voxels[index];
    const float jx = randu(state) - 0.5f;
    const float jy = randu(state) - 0.5f;
    const float jz = randu(state) - 0.5f;
    voxel.x += jx;
    voxel.y += jy;
    voxel.z += jz;
    float c[45];
    for (int k = 0; k < 45; k++)
        c[k] = k < K ? fod[(int64_t)index * K + k] : 0.0f;

    // seeds.cl:28-62
    float sum = 0.0f;
    for (int i = 0; i < N_DIRECTIONS; i++)
        sum += fmaxf(model_value<MODEL>(table + i * TABLE_STRIDE, c, K), 0.0f);
    const float threshold = randu(state) * sum;
    float cs = 0.0f;
    int selected = N_DIRECTIONS;
    for (int i = 0; i < N_DIRECTIONS; i++) {
        cs += fmaxf(model_value<MODEL>(table + i * TABLE_STRIDE, c, K), 0.0f);
        if (selected == N_DIRECTIONS && cs > threshold)
            selected = i;
    }
    float ox = 0.0f, oy = 0.0f, oz = 1.0f;
    if (selected < N_DIRECTIONS) {
        ox = table[selected * TABLE_STRIDE + 45];
        oy = table[selected * TABLE_STRIDE + 46];
        oz = table[selected * TABLE_STRIDE + 47];
    }
    float *row = out + 8 * g;
    for (int i = 0; i < 4; i++)
        row[i] = ((affine[4 * i] * voxel.x + affine[4 * i + 1] * voxel.y) + affine[4 * i + 2] * voxel.z) +
                 affine[4 * i + 3] * voxel.w;
    row[4] = ox;
    row[5] = oy;
    row[6] = oz;
    row[7] = 0.0f;
    states[g] = state;
}

}  // namespace

size_t seeds_from_fod_workspace_bytes(int model, int K, int64_t n, int32_t method)
{
    (void)model;
    (void)K;
    (void)n;
    (void)method;
    return (size_t)N_DIRECTIONS * TABLE_STRIDE * sizeof(float);
}

hipError_t launch_seeds_from_fod(int model, int K, const float *fod, const float *voxels, uint32_t nnz,
                                 const float *affine, uint32_t *states, int64_t n, int32_t method, float *out,
                                 void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    (void)method;
    if (n == 0)
        return hipSuccess;
    if (workspace_bytes < seeds_from_fod_workspace_bytes(model, K, n, method))
        return hipErrorInvalidValue;
    float *table = static_cast<float *>(workspace);
    basis_table_kernel<<<(N_DIRECTIONS + 255) / 256, 256, 0, stream>>>(table);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (model == MODEL_DTI)
        seeds_from_fod_kernel<MODEL_DTI><<<blocks, 256, 0, stream>>>(
            fod, K, reinterpret_cast<const float4 *>(voxels), nnz, affine, table, reinterpret_cast<uint2 *>(states),
            n, out);
    else
        seeds_from_fod_kernel<MODEL_SH><<<blocks, 256, 0, stream>>>(
            fod, K, reinterpret_cast<const float4 *>(voxels), nnz, affine, table, reinterpret_cast<uint2 *>(states),
            n, out);
    return hipGetLastError();
}

}  // namespace tgk
