// tgk_seeds.hip -- seeding from FOD images (seeds.cl:12-118).
//
// VALU method (bit-identical to the reference's arithmetic):
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
// MFMA method: the model values of 64 seeds (one wavefront) on 32 directions
// are one 32x32 tile of the product (directions x K) . (K x seeds), computed
// with v_mfma_f32_32x32x16_bf16 on split operands (x = hi + lo in bfloat16;
// hi.hi + hi.lo + lo.hi, FP32 accumulation), i.e. about 16 significant bits
// per operand. The basis table is built once per launch on the host in FP64
// at the reference's FP32 Fibonacci directions. After each tile the lane
// pairs (l, l ^ 32) swap halves so that every lane holds the 32 values of its
// own seed in direction order; sums, thresholds and the cumulative search
// then follow the reference's sequence exactly (seeds.cl:28-62), so the only
// difference to the VALU method is the rounding of the model values.
//
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "tgk_internal.h"
#include "tgk_plain_model.h"
#include "tgk_sh_basis.h"

namespace tgk {

namespace {

constexpr int N_DIRECTIONS = 1000;
constexpr int TABLE_STRIDE = 48;
constexpr int DIRECTION_BLOCK = 4;  // directions evaluated together by the VALU method

// seeds.cl:12-25 (host and device: the MFMA table is built on the host at
// the same FP32 directions).
__host__ __device__ inline void fibonacci_xyz(uint32_t n, uint32_t n_points, float &x, float &y, float &z)
{
    const float phi = (tgk_sqrtf(5.0f) + 1.0f) * 0.5f;
    const float golden_angle = 2.0f * 3.14159265359f / phi;
    const float theta = (float)n * golden_angle;
    z = 1.0f - (2.0f * (float)n / (float)(n_points - 1));
    const float radius = tgk_sqrtf(1.0f - z * z);
    float st, ct;
    tgk_sincosf(theta, &st, &ct);
    x = radius * ct;
    y = radius * st;
}

__device__ __forceinline__ V3 fibonacci_direction(uint32_t n, uint32_t n_points)
{
    float x, y, z;
    fibonacci_xyz(n, n_points, x, y, z);
    return v3(x, y, z);
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
// SH: softplus of the in-order contraction of the first K terms; DTI:
// o . (D o). The contraction is unrolled so that c stays in registers (a
// loop to a run-time K indexed it in scratch memory); KT = 45 needs no
// test, otherwise each term is taken while k < K (a uniform branch).
template <int MODEL, int KT>
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
#pragma unroll
        for (int k = 0; k < 45; k++)
            if (KT == 45 || k < K)
                v += c[k] * row[k];
        return plain::softplus(v);
    }
}

// seeds.cl:68-118
template <int MODEL, int KT>
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
    const uint32_t index = min(randi(state, nnz), nnz - 1);
    float4 voxel = voxels[index];
    const float jx = randu(state) - 0.5f;
    const float jy = randu(state) - 0.5f;
    const float jz = randu(state) - 0.5f;
    voxel.x += jx;
    voxel.y += jy;
    voxel.z += jz;
    float c[45];
#pragma unroll
    for (int k = 0; k < 45; k++)
        c[k] = k < K ? fod[(int64_t)index * K + k] : 0.0f;

    // seeds.cl:28-62. Directions are evaluated DIRECTION_BLOCK at a time: their
    // contractions are independent chains the wavefront can interleave (one
    // at a time, each waited on its own dependent adds and table loads),
    // and their weights are then added to the sums in direction order, so
    // every rounding is the reference's.
    static_assert(N_DIRECTIONS % DIRECTION_BLOCK == 0, "directions come in whole blocks");
    float sum = 0.0f;
    for (int i = 0; i < N_DIRECTIONS; i += DIRECTION_BLOCK) {
        float w[DIRECTION_BLOCK];
#pragma unroll
        for (int d = 0; d < DIRECTION_BLOCK; d++)
            w[d] = fmaxf(model_value<MODEL, KT>(table + (i + d) * TABLE_STRIDE, c, K), 0.0f);
#pragma unroll
        for (int d = 0; d < DIRECTION_BLOCK; d++)
            sum += w[d];
    }
    const float threshold = randu(state) * sum;
    float cs = 0.0f;
    int selected = N_DIRECTIONS;
    for (int i = 0; i < N_DIRECTIONS; i += DIRECTION_BLOCK) {
        float w[DIRECTION_BLOCK];
#pragma unroll
        for (int d = 0; d < DIRECTION_BLOCK; d++)
            w[d] = fmaxf(model_value<MODEL, KT>(table + (i + d) * TABLE_STRIDE, c, K), 0.0f);
#pragma unroll
        for (int d = 0; d < DIRECTION_BLOCK; d++) {
            cs += w[d];
            if (selected == N_DIRECTIONS && cs > threshold)
                selected = i + d;
        }
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


// ------------------------------------------------------------ MFMA method ---

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MFMA_TILES = 32;  // 32 tiles of 32 directions cover the 1000
constexpr int MFMA_K = 16;      // bfloat16 k per v_mfma_f32_32x32x16_bf16
constexpr size_t MFMA_DIRECTIONS_BYTES = 16384;  // 1000 float4, rounded up

__host__ __device__ constexpr int mfma_ksteps(int K) { return (K + MFMA_K - 1) / MFMA_K; }
size_t mfma_table_bytes(int K) { return (size_t)MFMA_TILES * mfma_ksteps(K) * 2 * 64 * sizeof(bf16x8); }

// Nonnegative sampling weight of a model value (seeds.cl:36, 49): softplus
// (k = 100, core.cl:111-120) for SH from the hardware exp2/log2, with
// log(1 + e) ~ e (1 - e/2) where 1 + e would round away e; DTI as is.
template <int MODEL>
__device__ __forceinline__ float sampling_weight(float x)
{
    if constexpr (MODEL == MODEL_DTI) {
        return fmaxf(x, 0.0f);
    } else {
        if (x * 100.0f >= 30.0f)
            return x;
        const float e = __builtin_amdgcn_exp2f(x * 144.26950408889634f);
        const float lp = e < 1e-3f ? e * (1.0f - 0.5f * e) : __builtin_amdgcn_logf(1.0f + e) * 0.6931471805599453f;
        return fmaxf(lp * 0.01f, 0.0f);
    }
}

// x = hi + lo to about 16 significant bits, both round-to-nearest bfloat16.
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// Weights of the 32 directions of tile t for this lane's seed, in direction
// order. Lane l supplies row/column l % 32 and the k half l / 32 of both
// operands; the 32x32 result puts D[i][j] in lane j + 32 ((i / 4) % 2),
// register 4 (i / 8) + i % 4. Set 0 holds the wave's seeds 0..31, set 1 its
// seeds 32..63; swapping halves between lanes l and l ^ 32 gives lane l < 32
// all rows of set 0 column l and lane l >= 32 all rows of set 1 column l - 32,
// i.e. every lane the weights of seed first + l.
template <int MODEL, int KS>
__device__ __forceinline__ void tile_weights(const bf16x8 *__restrict__ table, int t, int lane, const bf16x8 *b0hi,
                                             const bf16x8 *b0lo, const bf16x8 *b1hi, const bf16x8 *b1lo, float v[32])
{
    f32x16 acc0 = {}, acc1 = {};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const bf16x8 ahi = table[((t * KS + ks) * 2) * 64 + lane];
        const bf16x8 alo = table[((t * KS + ks) * 2 + 1) * 64 + lane];
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b0hi[ks], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b1hi[ks], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b0lo[ks], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, b1lo[ks], acc1, 0, 0, 0);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, b0hi[ks], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, b1hi[ks], acc1, 0, 0, 0);
    }
    const int h = lane >> 5;
    float keep[16], recv[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = 8 * (r >> 2) + 4 * h + (r & 3);
        const bool valid = 32 * t + row < N_DIRECTIONS;
        const float w0 = valid ? sampling_weight<MODEL>(acc0[r]) : 0.0f;
        const float w1 = valid ? sampling_weight<MODEL>(acc1[r]) : 0.0f;
        keep[r] = h ? w1 : w0;
        recv[r] = __shfl_xor(h ? w0 : w1, 32);
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[8 * b + q] = h ? recv[4 * b + q] : keep[4 * b + q];
            v[8 * b + 4 + q] = h ? keep[4 * b + q] : recv[4 * b + q];
        }
    }
}

// One wavefront per 64 consecutive seeds; the launch covers n rounded up to
// whole wavefronts and lanes past n work on a dummy state without writing.
template <int MODEL, int KS>
__global__ __launch_bounds__(256) void seeds_mfma_kernel(const float *__restrict__ fod, int K,
                                                         const float4 *__restrict__ voxels, uint32_t nnz,
                                                         const float *__restrict__ affine,
                                                         const bf16x8 *__restrict__ table,
                                                         const float4 *__restrict__ directions,
                                                         uint2 *__restrict__ states, int64_t n, float *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane;
    if (first >= n)
        return;
    const int64_t g = first + lane;
    const bool active = g < n;
    uint2 state = active ? states[g] : make_uint2(0x2545F491u, 0x9E3779B9u);
    const uint32_t index = min(randi(state, nnz), nnz - 1);
    float4 voxel = voxels[index];
    const float jx = randu(state) - 0.5f;
    const float jy = randu(state) - 0.5f;
    const float jz = randu(state) - 0.5f;
    voxel.x += jx;
    voxel.y += jy;
    voxel.z += jz;

    const int h = lane >> 5, c = lane & 31;
    const uint32_t i0 = (uint32_t)__shfl((int)index, c);
    const uint32_t i1 = (uint32_t)__shfl((int)index, 32 + c);
    bf16x8 b0hi[KS], b0lo[KS], b1hi[KS], b1lo[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = MFMA_K * ks + 8 * h + j;
            const float x0 = k < K ? fod[(int64_t)i0 * K + k] : 0.0f;
            const float x1 = k < K ? fod[(int64_t)i1 * K + k] : 0.0f;
            __bf16 hi, lo;
            split_bf16(x0, hi, lo);
            b0hi[ks][j] = hi;
            b0lo[ks][j] = lo;
            split_bf16(x1, hi, lo);
            b1hi[ks][j] = hi;
            b1lo[ks][j] = lo;
        }
    }

    // seeds.cl:28-62 with the tile weights.
    float v[32];
    float sum = 0.0f;
    for (int t = 0; t < MFMA_TILES; t++) {
        tile_weights<MODEL, KS>(table, t, lane, b0hi, b0lo, b1hi, b1lo, v);
#pragma unroll
        for (int r = 0; r < 32; r++)
            sum += v[r];
    }
    const float threshold = randu(state) * sum;
    float cs = 0.0f;
    int selected = N_DIRECTIONS;
    for (int t = 0; t < MFMA_TILES; t++) {
        tile_weights<MODEL, KS>(table, t, lane, b0hi, b0lo, b1hi, b1lo, v);
#pragma unroll
        for (int r = 0; r < 32; r++) {
            cs += v[r];
            if (selected == N_DIRECTIONS && cs > threshold)
                selected = 32 * t + r;
        }
        if (!__any(selected == N_DIRECTIONS))
            break;
    }
    if (!active)
        return;
    const float4 o = selected < N_DIRECTIONS ? directions[selected] : make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    float *row = out + 8 * g;
    for (int i = 0; i < 4; i++)
        row[i] = ((affine[4 * i] * voxel.x + affine[4 * i + 1] * voxel.y) + affine[4 * i + 2] * voxel.z) +
                 affine[4 * i + 3] * voxel.w;
    row[4] = o.x;
    row[5] = o.y;
    row[6] = o.z;
    row[7] = 0.0f;
    states[g] = state;
}

// Round to nearest even bfloat16 (finite x).
uint16_t bf16_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, sizeof u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_value(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, sizeof f);
    return f;
}

// The split basis in the operand layout of tile_weights (row i of tile t is
// direction 32 t + i; directions past 1000 and k past K are zero) and the
// FP32 directions as float4.
void build_mfma_table(int model, int K, std::vector<uint16_t> &table, std::vector<float> &dirs)
{
    const int KS = mfma_ksteps(K);
    int lmax = 0;
    while (shb::n_coefficients(lmax) < K)
        lmax += 2;
    std::vector<double> basis((size_t)N_DIRECTIONS * K);
    std::vector<double> row((size_t)shb::n_coefficients(lmax));
    dirs.assign(MFMA_DIRECTIONS_BYTES / sizeof(float), 0.0f);
    for (int d = 0; d < N_DIRECTIONS; d++) {
        float x, y, z;
        fibonacci_xyz((uint32_t)d, N_DIRECTIONS, x, y, z);
        dirs[4 * d] = x;
        dirs[4 * d + 1] = y;
        dirs[4 * d + 2] = z;
        double *b = basis.data() + (size_t)d * K;
        if (model == MODEL_DTI) {
            const double X = x, Y = y, Z = z;
            const double q[6] = {X * X, Y * Y, Z * Z, 2.0 * X * Y, 2.0 * X * Z, 2.0 * Y * Z};
            for (int k = 0; k < K; k++)
                b[k] = q[k];
        } else {
            shb::real_sh_basis(lmax, x, y, z, row.data());
            for (int k = 0; k < K; k++)
                b[k] = row[k];
        }
    }
    table.assign(mfma_table_bytes(K) / sizeof(uint16_t), 0);
    for (int t = 0; t < MFMA_TILES; t++)
        for (int ks = 0; ks < KS; ks++)
            for (int lane = 0; lane < 64; lane++) {
                const int d = 32 * t + (lane & 31);
                for (int j = 0; j < 8; j++) {
                    const int k = MFMA_K * ks + 8 * (lane >> 5) + j;
                    const double value = (d < N_DIRECTIONS && k < K) ? basis[(size_t)d * K + k] : 0.0;
                    const uint16_t hi = bf16_bits((float)value);
                    const uint16_t lo = bf16_bits((float)(value - (double)bf16_value(hi)));
                    const size_t base = ((size_t)(t * KS + ks) * 2 * 64 + lane) * 8 + j;
                    table[base] = hi;
                    table[base + 64 * 8] = lo;
                }
            }
}

template <int MODEL, int KS>
void launch_mfma(const float *fod, int K, const float *voxels, uint32_t nnz, const float *affine, const void *table,
                 const void *dirs, uint32_t *states, int64_t n, float *out, hipStream_t stream)
{
    const unsigned blocks = (unsigned)((n + 255) / 256);
    seeds_mfma_kernel<MODEL, KS><<<blocks, 256, 0, stream>>>(
        fod, K, reinterpret_cast<const float4 *>(voxels), nnz, affine, static_cast<const bf16x8 *>(table),
        static_cast<const float4 *>(dirs), reinterpret_cast<uint2 *>(states), n, out);
}

}  // namespace

size_t seeds_from_fod_workspace_bytes(int model, int K, int64_t n, int32_t method)
{
    (void)model;
    (void)n;
    if (method == SEEDS_MFMA)
        return mfma_table_bytes(K) + MFMA_DIRECTIONS_BYTES;
    return (size_t)N_DIRECTIONS * TABLE_STRIDE * sizeof(float);
}

hipError_t prepare_seeds_from_fod(int model, int K, int32_t method, void *table, size_t table_bytes,
                                  hipStream_t stream)
{
    if (table_bytes < seeds_from_fod_workspace_bytes(model, K, 0, method))
        return hipErrorInvalidValue;
    if (method == SEEDS_MFMA) {
        if (K < 1 || mfma_ksteps(K) > 6 || (model == MODEL_DTI && K != 6))
            return hipErrorInvalidValue;
        std::vector<uint16_t> host_table;
        std::vector<float> dirs;
        build_mfma_table(model, K, host_table, dirs);
        char *ws = static_cast<char *>(table);
        hipError_t e = hipMemcpyAsync(ws, host_table.data(), host_table.size() * sizeof(uint16_t),
                                      hipMemcpyHostToDevice, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(ws + mfma_table_bytes(K), dirs.data(), dirs.size() * sizeof(float),
                               hipMemcpyHostToDevice, stream);
        // The host tables must outlive the copies.
        const hipError_t s = hipStreamSynchronize(stream);
        return e != hipSuccess ? e : s;
    }
    if (K < 1 || K > 45 || (model == MODEL_DTI && K != 6))
        return hipErrorInvalidValue;
    basis_table_kernel<<<(N_DIRECTIONS + 255) / 256, 256, 0, stream>>>(static_cast<float *>(table));
    return hipGetLastError();
}

hipError_t launch_seeds_from_fod_prepared(int model, int K, const float *fod, const float *voxels, uint32_t nnz,
                                          const float *affine, uint32_t *states, int64_t n, int32_t method,
                                          float *out, const void *table, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    if (nnz == 0)
        return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (method == SEEDS_MFMA) {
        const void *ddirs = static_cast<const char *>(table) + mfma_table_bytes(K);
        if (model == MODEL_DTI)
            launch_mfma<MODEL_DTI, 1>(fod, K, voxels, nnz, affine, table, ddirs, states, n, out, stream);
        else if (mfma_ksteps(K) <= 1)
            launch_mfma<MODEL_SH, 1>(fod, K, voxels, nnz, affine, table, ddirs, states, n, out, stream);
        else if (mfma_ksteps(K) <= 3)
            launch_mfma<MODEL_SH, 3>(fod, K, voxels, nnz, affine, table, ddirs, states, n, out, stream);
        else
            launch_mfma<MODEL_SH, 6>(fod, K, voxels, nnz, affine, table, ddirs, states, n, out, stream);
        return hipGetLastError();
    }
    const float *t = static_cast<const float *>(table);
    if (model == MODEL_DTI)
        seeds_from_fod_kernel<MODEL_DTI, 0><<<blocks, 256, 0, stream>>>(
            fod, K, reinterpret_cast<const float4 *>(voxels), nnz, affine, t, reinterpret_cast<uint2 *>(states), n,
            out);
    else if (K == 45)
        seeds_from_fod_kernel<MODEL_SH, 45><<<blocks, 256, 0, stream>>>(
            fod, K, reinterpret_cast<const float4 *>(voxels), nnz, affine, t, reinterpret_cast<uint2 *>(states), n,
            out);
    else
        seeds_from_fod_kernel<MODEL_SH, 0><<<blocks, 256, 0, stream>>>(
            fod, K, reinterpret_cast<const float4 *>(voxels), nnz, affine, t, reinterpret_cast<uint2 *>(states), n,
            out);
    return hipGetLastError();
}

hipError_t launch_seeds_from_fod(int model, int K, const float *fod, const float *voxels, uint32_t nnz,
                                 const float *affine, uint32_t *states, int64_t n, int32_t method, float *out,
                                 void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipError_t e = prepare_seeds_from_fod(model, K, method, workspace, workspace_bytes, stream);
    if (e != hipSuccess)
        return e;
    return launch_seeds_from_fod_prepared(model, K, fod, voxels, nnz, affine, states, n, method, out, workspace,
                                          stream);
}

}  // namespace tgk
