// tgk_scan.hip -- exclusive prefix sum of per-streamline saved-point counts
// (u32) into i64 offsets, offsets[n] = total: where each streamline's points
// start in the ragged output (tracking.py:178 builds the same split on the
// host). Three launches of plain kernels, with no library call and no memset,
// so the sequence can be recorded into a HIP graph.
#include <hip/hip_runtime.h>

#include "tgk_internal.h"

namespace tgk {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // items per block

// Exclusive scan of one value per thread across the block; returns the
// thread's exclusive prefix and stores the block total in *total.
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t *total)
{
    __shared__ int64_t wave_totals[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t inclusive = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int64_t up = __shfl_up(inclusive, d);
        if (lane >= d)
            inclusive += up;
    }
    if (lane == 63)
        wave_totals[wave] = inclusive;
    __syncthreads();
    int64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; w++) {
        if (w < wave)
            before += wave_totals[w];
        all += wave_totals[w];
    }
    __syncthreads();
    *total = all;
    return before + inclusive - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void tile_sums_kernel(const uint32_t *__restrict__ lengths, int64_t n,
                                                                 int64_t *__restrict__ sums)
{
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t v = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (base + i < n)
            v += lengths[base + i];
    int64_t total;
    (void)block_exclusive_scan(v, &total);
    if (threadIdx.x == 0)
        sums[blockIdx.x] = total;
}

// One block turns the tile sums into exclusive tile offsets in place and
// writes the grand total to sums[tiles] and offsets[n].
__global__ __launch_bounds__(SCAN_THREADS) void tile_offsets_kernel(int64_t *__restrict__ sums, int64_t tiles,
                                                                    int64_t *__restrict__ offsets, int64_t n)
{
    int64_t carry = 0;
    for (int64_t start = 0; start < tiles; start += SCAN_THREADS) {
        const int64_t i = start + threadIdx.x;
        const int64_t v = i < tiles ? sums[i] : 0;
        int64_t total;
        const int64_t prefix = block_exclusive_scan(v, &total);
        if (i < tiles)
            sums[i] = carry + prefix;
        carry += total;
    }
    if (threadIdx.x == 0) {
        sums[tiles] = carry;
        offsets[n] = carry;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void tile_scan_kernel(const uint32_t *__restrict__ lengths, int64_t n,
                                                                 const int64_t *__restrict__ tile_offsets,
                                                                 int64_t *__restrict__ offsets)
{
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t items[SCAN_ITEMS];
    int64_t v = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        items[i] = base + i < n ? lengths[base + i] : 0u;
        v += items[i];
    }
    int64_t total;
    int64_t running = tile_offsets[blockIdx.x] + block_exclusive_scan(v, &total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (base + i < n)
            offsets[base + i] = running;
        running += items[i];
    }
}

}  // namespace

size_t offsets_workspace_bytes(int64_t n)
{
    const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)(tiles + 1) * sizeof(int64_t);
}

hipError_t launch_offsets(const uint32_t *lengths, int64_t n, int64_t *offsets, void *workspace,
                          size_t workspace_bytes, hipStream_t stream)
{
    if (workspace_bytes < offsets_workspace_bytes(n))
        return hipErrorInvalidValue;
    const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t *sums = static_cast<int64_t *>(workspace);
    if (tiles > 0)
        tile_sums_kernel<<<(unsigned)tiles, SCAN_THREADS, 0, stream>>>(lengths, n, sums);
    tile_offsets_kernel<<<1, SCAN_THREADS, 0, stream>>>(sums, tiles, offsets, n);
    if (tiles > 0)
        tile_scan_kernel<<<(unsigned)tiles, SCAN_THREADS, 0, stream>>>(lengths, n, sums, offsets);
    return hipGetLastError();
}

}  // namespace tgk
