// tgk_capi.hip -- the extern "C" boundary of libtgk.so (declared in
// include/tgk.h): runtime and buffer ownership, argument validation, chunked
// host-array tracking and the one-op drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <atomic>
#include <mutex>
#include <new>
#include <thread>
#include <string>
#include <deque>
#include <vector>

#include "../../include/tgk.h"
#include "tgk_internal.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buffer[512];
    va_list args;
    va_start(args, fmt);
    vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    g_last_error = buffer;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    const int code = (e == hipErrorOutOfMemory) ? TGK_ERR_NO_MEMORY : TGK_ERR_DEVICE;
    return fail(code, "%s: %s", what, hipGetErrorString(e));
}

#define TGK_HIP(expr, what)                   \
    do {                                      \
        const hipError_t e_ = (expr);         \
        if (e_ != hipSuccess)                 \
            return hip_fail(e_, what);        \
    } while (0)

// Owning device allocation.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t allocate(size_t n)
    {
        if (n <= bytes && ptr)
            return hipSuccess;
        release();
        if (n == 0)
            return hipSuccess;
        hipError_t e = hipMalloc(&ptr, n);
        if (e == hipSuccess)
            bytes = n;
        else
            ptr = nullptr;
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace

// Byte offsets in the runtime's 256-byte queue buffer (offset 0: the seed
// counter of the connectome and TOD kernels): the fast kernel's last-launch
// read totals (two u64: coefficient rows, mask values) and its two
// saved-streamline counts (u32).
constexpr size_t FAST_STATS_OFFSET = 64;
constexpr size_t FAST_CONT_COUNT_OFFSET = 136;

// tgk_track's device buffers for one chunk of seeds, kept by the runtime.
struct TrackBuffers {
    DeviceBuffer seeds, rng, lengths, steps, hits, dense, offsets, points;
};

struct tgk_runtime {
    int32_t device = 0;
    hipStream_t stream = nullptr;
    DeviceBuffer workspace;
    DeviceBuffer queue;        // counters of the persistent kernels (layout: FAST_*_OFFSET)
    DeviceBuffer continuations[2];  // streamlines saved between generations of the fast kernel
    DeviceBuffer fast_counters;     // claim counter and ended-streamline shards of the fast kernel
    TrackBuffers track_buffers;     // tgk_track's working buffers
    DeviceBuffer seeds_table;  // prepared FOD-seeding table for seeds_key
    int32_t seeds_key[3] = {-1, -1, -1};  // model, K, method
    bool capturing = false;    // the stream is being recorded into a graph
    int32_t fast_grid = 0;     // blocks of the fast kernel; <= 0: sized by occupancy
    int32_t fast_shape = 0;    // launch shape of the order-8 SH Transport / Diffusion kernels
    uint32_t fast_min_repack = 0;  // 0: the model's default (fast_launch); see tgk_runtime_set_fast_repack
};

struct tgk_graph {
    tgk_runtime *rt = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct tgk_field {
    tgk_runtime *rt = nullptr;
    DeviceBuffer data;
    uint32_t dims[4] = {0, 0, 0, 0};
    uint32_t stride = 0;
    float inverse_affine[16];
    int32_t fast_precision = TGK_PRECISION_FP32;
    DeviceBuffer fast;              // transformed coefficients for the fast path
    int32_t fast_cached = -1;       // model the transform in `fast` was made for
    uint32_t fast_stride = 0;
};

struct tgk_domain {
    tgk_runtime *rt = nullptr;
    int32_t kind = tgk::DOMAIN_MASK;
    DeviceBuffer mask;
    uint32_t mask_dims[3] = {0, 0, 0};
    float mask_inverse_affine[16];
    // boundary surface (seeding) and BVH (mesh domains)
    DeviceBuffer vertices, triangles, cumulative;
    uint32_t n_vertices = 0, n_triangles = 0;
    DeviceBuffer bvh_min, bvh_max, bvh_nodes, bvh_order, near, cell_start, cell_entries;
    uint32_t n_nodes = 0;
    uint32_t near_dims[3] = {0, 0, 0};
    float near_inverse_affine[16];
    float reach = 0.0f;
};

// Page-locked host memory (hipHostMalloc): the device writes it by DMA at
// the link's full rate, and the host reads it as ordinary cached memory.
struct PinnedBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer &) = delete;
    PinnedBuffer &operator=(const PinnedBuffer &) = delete;
    ~PinnedBuffer()
    {
        if (ptr)
            (void)hipHostFree(ptr);
    }
    hipError_t allocate(size_t n)
    {
        hipError_t e = hipHostMalloc(&ptr, n, hipHostMallocDefault);
        if (e == hipSuccess)
            bytes = n;
        else
            ptr = nullptr;
        return e;
    }
};

// Copies spans of host memory with up to 8 threads, in slices of at least
// 64 MiB (smaller copies stay on the calling thread).
template <class Span> static void parallel_copy(const Span *spans, size_t n)
{
    constexpr size_t SLICE = (size_t)64 << 20;
    struct Slice {
        char *dst;
        const char *src;
        size_t bytes;
    };
    std::vector<Slice> slices;
    for (size_t i = 0; i < n; i++)
        for (size_t off = 0; off < spans[i].bytes; off += SLICE)
            slices.push_back({spans[i].dst + off, spans[i].src + off, std::min(SLICE, spans[i].bytes - off)});
    const size_t workers = std::min<size_t>(8, slices.size());
    if (workers <= 1) {
        for (const Slice &s : slices)
            memcpy(s.dst, s.src, s.bytes);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i = next.fetch_add(1); i < slices.size(); i = next.fetch_add(1))
            memcpy(slices[i].dst, slices[i].src, slices[i].bytes);
    };
    std::vector<std::thread> threads;
    for (size_t t = 1; t < workers; t++)
        threads.emplace_back(work);
    work();
    for (std::thread &t : threads)
        t.join();
}

struct tgk_tracks {
    int device = 0;
    std::vector<uint32_t> lengths, steps, hits;
    // The ragged points of each chunk: copied by DMA into exactly sized
    // page-locked host memory as soon as the chunk is compacted (a single
    // chunk's buffer is then the result itself, see tgk_tracks_points);
    // left in device memory instead when no page-locked memory is to be had,
    // and in pageable host memory when neither has room.
    struct Piece {
        std::unique_ptr<PinnedBuffer> pinned;
        std::unique_ptr<DeviceBuffer> device;
        std::vector<float> host;
        int64_t n_points = 0;
    };
    std::vector<Piece> pieces;
    int64_t total_points = 0;
};

namespace {

// ---------------------------------------- allocation-free compaction ---
//
// tgk_compact_device_noalloc keeps the offsets scan's per-tile sums (one
// int64 per 2048 seeds, plus the total) in this file-scope device array
// instead of the runtime's growable workspace. Each device holds its own
// copy at its own address until the process ends. In host code `scratch`
// names a host-side stand-in, not device memory: kernels and copies get the
// device address, looked up once per device when the first runtime on that
// device is created (scratch_setup) and kept in g_scratch_address.
constexpr int64_t NOALLOC_SCRATCH_WORDS = 8192;  // 64 KiB: up to 16,775,168 seeds per call
__device__ __align__(256) int64_t scratch[NOALLOC_SCRATCH_WORDS];

constexpr int MAX_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_address[MAX_SCRATCH_DEVICES] = {};

// With `device` current: the device address of `scratch` on it, looked up
// on the first call for that device.
int scratch_setup(int32_t device)
{
    if (device < 0 || device >= MAX_SCRATCH_DEVICES)
        return fail(TGK_ERR_UNSUPPORTED, "device %d: the allocation-free scratch serves devices 0..%d", device,
                    MAX_SCRATCH_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_address[device])
        return TGK_OK;
    void *ptr = nullptr;
    TGK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr)
        return fail(TGK_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_address[device] = ptr;
    return TGK_OK;
}

void *scratch_address(int32_t device)
{
    if (device < 0 || device >= MAX_SCRATCH_DEVICES)
        return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_address[device];
}

void copy_rows3(const float m[16], float out[12])
{
    for (int i = 0; i < 12; i++)
        out[i] = m[i];
}

tgk::FieldView field_view(const tgk_field *f)
{
    tgk::FieldView v;
    v.data = f->data.as<const float>();
    v.dims[0] = f->dims[0];
    v.dims[1] = f->dims[1];
    v.dims[2] = f->dims[2];
    v.n_parameters = f->dims[3];
    v.stride = f->stride;
    copy_rows3(f->inverse_affine, v.inv);
    return v;
}

tgk::MaskView mask_view(const tgk_domain *d)
{
    tgk::MaskView v;
    v.mask = d->mask.as<const uint8_t>();
    for (int i = 0; i < 3; i++)
        v.dims[i] = d->mask_dims[i];
    copy_rows3(d->mask_inverse_affine, v.inv);
    return v;
}

tgk::MeshView mesh_view(const tgk_domain *d)
{
    tgk::MeshView v;
    v.vertices = d->vertices.as<const float4>();
    v.triangles = d->triangles.as<const uint4>();
    v.bvh_min = d->bvh_min.as<const float4>();
    v.bvh_max = d->bvh_max.as<const float4>();
    v.bvh_nodes = d->bvh_nodes.as<const int4>();
    v.bvh_order = d->bvh_order.as<const uint32_t>();
    v.n_nodes = d->n_nodes;
    v.near = d->near.as<const uint8_t>();
    for (int i = 0; i < 3; i++)
        v.near_dims[i] = d->near_dims[i];
    copy_rows3(d->near_inverse_affine, v.near_inv);
    v.reach = d->near.ptr ? d->reach : -1.0f;
    v.cell_start = d->cell_start.as<const uint32_t>();
    v.cell_entries = d->cell_entries.as<const float4>();
    return v;
}

// max_sh: the largest SH coefficient count the caller's kernels take (45 for
// the reference's l <= 8 basis; 91 where the order-12 polynomial form is
// available).
int check_algorithm(const tgk_algorithm *a, const tgk_field *f, int max_sh = 45)
{
    if (!a)
        return fail(TGK_ERR_INVALID, "algorithm is NULL");
    if (a->model != TGK_MODEL_SH && a->model != TGK_MODEL_DTI)
        return fail(TGK_ERR_INVALID, "unknown model %d", a->model);
    if (a->algorithm < TGK_ALGO_TRANSPORT || a->algorithm > TGK_ALGO_DETERMINISTIC)
        return fail(TGK_ERR_INVALID, "unknown algorithm %d", a->algorithm);
    if (a->algorithm == TGK_ALGO_PROBABILISTIC) {
        if (a->n_samples < 1)
            return fail(TGK_ERR_INVALID, "n_samples must be positive");
        if (!std::isfinite(a->max_angle) || !(a->max_angle >= 0.0f) || !(a->max_angle <= 3.14159265358979f))
            return fail(TGK_ERR_INVALID, "max_angle must be finite and between 0 and pi");
    }
    if (a->algorithm == TGK_ALGO_DETERMINISTIC && a->max_iterations < 0)
        return fail(TGK_ERR_INVALID, "max_iterations must be non-negative");
    if (a->model == TGK_MODEL_DTI && a->n_parameters != 6)
        return fail(TGK_ERR_INVALID, "DTI needs 6 parameters, got %d", a->n_parameters);
    if (a->model == TGK_MODEL_SH && (a->n_parameters < 1 || a->n_parameters > max_sh))
        return fail(TGK_ERR_UNSUPPORTED, "SH with %d coefficients: 1..%d supported%s", a->n_parameters, max_sh,
                    max_sh == 45 ? " (up to 91 on the fast path)" : "");
    if (f && (uint32_t)a->n_parameters != f->dims[3])
        return fail(TGK_ERR_INVALID, "incompatible local model and field: %d != %u", a->n_parameters, f->dims[3]);
    return TGK_OK;
}

tgk::AlgoParams algo_params(const tgk_algorithm *a)
{
    tgk::AlgoParams p;
    p.model = a->model;
    p.algorithm = a->algorithm;
    p.n_parameters = a->n_parameters;
    p.step_size = a->step_size;
    p.gamma = a->gamma;
    p.noise_variance = a->noise_variance;
    p.n_samples = a->n_samples;
    p.max_angle = a->max_angle;
    p.max_iterations = a->max_iterations;
    p.tolerance = a->tolerance;
    p.max_step_size = a->max_step_size;
    return p;
}

template <class T> int upload(DeviceBuffer &buf, const T *host, size_t count, hipStream_t stream, const char *what)
{
    TGK_HIP(buf.allocate(count * sizeof(T)), what);
    if (count)
        TGK_HIP(hipMemcpyAsync(buf.ptr, host, count * sizeof(T), hipMemcpyHostToDevice, stream), what);
    return TGK_OK;
}

int select_device(tgk_runtime *rt)
{
    TGK_HIP(hipSetDevice(rt->device), "hipSetDevice");
    return TGK_OK;
}

// Grows a runtime-owned buffer to `bytes`; while the stream is being
// recorded into a graph no allocation may happen, so growth is an error
// (the caller should have prepared the buffer before recording).
int reserve(tgk_runtime *rt, DeviceBuffer &buf, size_t bytes, const char *what)
{
    if (bytes == 0 || (buf.ptr && bytes <= buf.bytes))
        return TGK_OK;
    if (rt->capturing)
        return fail(TGK_ERR_INVALID, "%s needs %zu bytes while a graph is being recorded; prepare it first", what,
                    bytes);
    TGK_HIP(buf.allocate(bytes), what);
    return TGK_OK;
}

int ensure_workspace(tgk_runtime *rt, size_t bytes) { return reserve(rt, rt->workspace, bytes, "workspace"); }

}  // namespace



extern "C" {

const char *tgk_last_error(void) { return g_last_error.c_str(); }

int tgk_version(void) { return 1; }

int tgk_runtime_create(int32_t device, tgk_runtime **out)
{
    if (!out)
        return fail(TGK_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(TGK_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(TGK_ERR_INVALID, "device %d out of range (have %d)", device, count);
    std::unique_ptr<tgk_runtime> rt(new (std::nothrow) tgk_runtime());
    if (!rt)
        return fail(TGK_ERR_NO_MEMORY, "runtime allocation");
    rt->device = device;
    TGK_HIP(hipSetDevice(device), "hipSetDevice");
    TGK_HIP(hipStreamCreateWithFlags(&rt->stream, hipStreamNonBlocking), "hipStreamCreate");
    // The seed counter of the persistent kernels, allocated up front so that
    // no launch allocates (e.g. while a graph is being recorded).
    TGK_HIP(rt->queue.allocate(256), "queue");
    TGK_HIP(rt->fast_counters.allocate(tgk::FAST_COUNTER_BYTES), "fast counters");
    // The device address of the allocation-free entry point's scratch array,
    // once per device (a device past the table's size only loses that entry
    // point, which then fails with TGK_ERR_UNSUPPORTED).
    if (device < MAX_SCRATCH_DEVICES)
        if (int r = scratch_setup(device))
            return r;
    *out = rt.release();
    return TGK_OK;
}

int tgk_runtime_destroy(tgk_runtime *rt)
{
    if (!rt)
        return TGK_OK;
    (void)hipSetDevice(rt->device);
    if (rt->stream) {
        (void)hipStreamSynchronize(rt->stream);
        (void)hipStreamDestroy(rt->stream);
    }
    delete rt;
    return TGK_OK;
}

int tgk_runtime_fast_reads(tgk_runtime *rt, uint64_t *rows, uint64_t *mask_values)
{
    if (!rt || !rows || !mask_values)
        return fail(TGK_ERR_INVALID, "NULL argument");
    if (int r = select_device(rt))
        return r;
    uint64_t totals[2] = {0, 0};
    TGK_HIP(hipMemcpyAsync(totals, rt->queue.as<char>() + FAST_STATS_OFFSET, sizeof totals, hipMemcpyDeviceToHost,
                           rt->stream),
            "read totals");
    TGK_HIP(hipStreamSynchronize(rt->stream), "hipStreamSynchronize");
    *rows = totals[0];
    *mask_values = totals[1];
    return TGK_OK;
}

int tgk_runtime_synchronize(tgk_runtime *rt)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    TGK_HIP(hipStreamSynchronize(rt->stream), "hipStreamSynchronize");
    return TGK_OK;
}

int tgk_runtime_info(tgk_runtime *rt, char *name, int32_t name_len, int32_t *compute_units, int64_t *total_bytes)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    hipDeviceProp_t prop;
    TGK_HIP(hipGetDeviceProperties(&prop, rt->device), "hipGetDeviceProperties");
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units)
        *compute_units = prop.multiProcessorCount;
    if (total_bytes)
        *total_bytes = (int64_t)prop.totalGlobalMem;
    return TGK_OK;
}

int tgk_runtime_set_fast_grid(tgk_runtime *rt, int32_t blocks)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    rt->fast_grid = blocks > 0 ? blocks : 0;
    return TGK_OK;
}

int tgk_runtime_set_fast_shape(tgk_runtime *rt, int32_t shape)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (shape < 0 || shape >= tgk::FAST_SHAPE_COUNT)
        return fail(TGK_ERR_INVALID, "fast launch shape %d: shapes are 0..%d", shape, tgk::FAST_SHAPE_COUNT - 1);
    rt->fast_shape = shape;
    return TGK_OK;
}

int tgk_fast_shape_info(int32_t shape, int32_t *lanes_per_block, int32_t *lds_coefficients)
{
    static const int32_t table[tgk::FAST_SHAPE_COUNT][2] = {{256, 40}, {128, 40}, {512, 40}, {256, 32}, {256, 44}};
    if (shape < 0 || shape >= tgk::FAST_SHAPE_COUNT)
        return fail(TGK_ERR_INVALID, "fast launch shape %d: shapes are 0..%d", shape, tgk::FAST_SHAPE_COUNT - 1);
    if (lanes_per_block)
        *lanes_per_block = table[shape][0];
    if (lds_coefficients)
        *lds_coefficients = table[shape][1];
    return TGK_OK;
}

int tgk_runtime_set_fast_repack(tgk_runtime *rt, int32_t min_streamlines)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (min_streamlines < 1)
        return fail(TGK_ERR_INVALID, "min_streamlines must be positive, got %d", min_streamlines);
    rt->fast_min_repack = (uint32_t)min_streamlines;
    return TGK_OK;
}

int tgk_runtime_stream(tgk_runtime *rt, void **stream)
{
    if (!rt || !stream)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *stream = (void *)rt->stream;
    return TGK_OK;
}

int tgk_field_create(tgk_runtime *rt, const float *data, const uint32_t dims[4], const float inverse_affine[16],
                     tgk_field **out)
{
    if (!rt || !data || !dims || !inverse_affine || !out)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0 || dims[3] == 0)
        return fail(TGK_ERR_INVALID, "empty field (%u, %u, %u, %u)", dims[0], dims[1], dims[2], dims[3]);
    if (int r = select_device(rt))
        return r;
    std::unique_ptr<tgk_field> f(new (std::nothrow) tgk_field());
    if (!f)
        return fail(TGK_ERR_NO_MEMORY, "field allocation");
    f->rt = rt;
    memcpy(f->dims, dims, sizeof f->dims);
    memcpy(f->inverse_affine, inverse_affine, sizeof f->inverse_affine);
    const uint32_t K = dims[3];
    f->stride = (K + 3u) & ~3u;
    const size_t n_voxels = (size_t)dims[0] * dims[1] * dims[2];
    TGK_HIP(f->data.allocate(n_voxels * f->stride * sizeof(float)), "field allocation");
    // K floats per voxel into 16-byte aligned rows, pad zeroed: contiguous
    // slabs of whole rows go to a staging buffer in one copy each and a
    // kernel pads them in HBM (a pitched copy from pageable memory moves one
    // 180-byte row at a time: 33 s for a 145x174x145x45 volume).
    const size_t row_bytes = (size_t)K * sizeof(float);
    const size_t slab_rows = std::max<size_t>(1, std::min<size_t>(n_voxels, ((size_t)64 << 20) / row_bytes));
    DeviceBuffer staging;
    TGK_HIP(staging.allocate(slab_rows * row_bytes), "field staging");
    for (size_t first = 0; first < n_voxels; first += slab_rows) {
        const size_t rows = std::min(slab_rows, n_voxels - first);
        TGK_HIP(hipMemcpyAsync(staging.ptr, data + first * K, rows * row_bytes, hipMemcpyHostToDevice, rt->stream),
                "field upload");
        TGK_HIP(tgk::launch_pad_rows(staging.as<const float>(), (int64_t)rows, K,
                                     f->data.as<float>() + first * f->stride, f->stride, rt->stream),
                "field padding");
    }
    TGK_HIP(hipStreamSynchronize(rt->stream), "field upload");
    *out = f.release();
    return TGK_OK;
}

int tgk_field_destroy(tgk_field *field)
{
    if (!field)
        return TGK_OK;
    (void)hipSetDevice(field->rt->device);
    delete field;
    return TGK_OK;
}

int tgk_field_set_fast_precision(tgk_field *field, int32_t precision)
{
    if (!field)
        return fail(TGK_ERR_INVALID, "field is NULL");
    if (precision < TGK_PRECISION_FP32 || precision > TGK_PRECISION_BF16)
        return fail(TGK_ERR_INVALID, "unknown precision %d", precision);
    field->fast_precision = precision;
    return TGK_OK;
}

int tgk_mask_domain_create(tgk_runtime *rt, const uint8_t *mask, const uint32_t dims[3],
                           const float inverse_affine[16], tgk_domain **out)
{
    if (!rt || !mask || !dims || !inverse_affine || !out)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0)
        return fail(TGK_ERR_INVALID, "empty mask");
    if (int r = select_device(rt))
        return r;
    std::unique_ptr<tgk_domain> d(new (std::nothrow) tgk_domain());
    if (!d)
        return fail(TGK_ERR_NO_MEMORY, "domain allocation");
    d->rt = rt;
    d->kind = tgk::DOMAIN_MASK;
    memcpy(d->mask_dims, dims, sizeof d->mask_dims);
    memcpy(d->mask_inverse_affine, inverse_affine, sizeof d->mask_inverse_affine);
    if (int r = upload(d->mask, mask, (size_t)dims[0] * dims[1] * dims[2], rt->stream, "mask upload"))
        return r;
    TGK_HIP(hipStreamSynchronize(rt->stream), "mask upload");
    *out = d.release();
    return TGK_OK;
}

int tgk_domain_set_boundary(tgk_domain *d, const float *vertices, uint32_t n_vertices, const uint32_t *triangles,
                            const float *cumulative, uint32_t n_triangles)
{
    if (!d || !vertices || !triangles || !cumulative)
        return fail(TGK_ERR_INVALID, "NULL argument");
    if (n_triangles == 0 || n_vertices == 0)
        return fail(TGK_ERR_INVALID, "empty boundary");
    if (int r = select_device(d->rt))
        return r;
    hipStream_t s = d->rt->stream;
    if (int r = upload(d->vertices, vertices, (size_t)n_vertices * 4, s, "boundary upload"))
        return r;
    if (int r = upload(d->triangles, triangles, (size_t)n_triangles * 4, s, "boundary upload"))
        return r;
    if (int r = upload(d->cumulative, cumulative, (size_t)n_triangles, s, "boundary upload"))
        return r;
    d->n_vertices = n_vertices;
    d->n_triangles = n_triangles;
    TGK_HIP(hipStreamSynchronize(s), "boundary upload");
    return TGK_OK;
}

int tgk_mesh_domain_create(tgk_runtime *rt, const float *vertices, uint32_t n_vertices, const uint32_t *triangles,
                           const float *cumulative, uint32_t n_triangles, const float *bvh_min, const float *bvh_max,
                           const int32_t *bvh_nodes, uint32_t n_nodes, const uint32_t *bvh_order, uint32_t n_order,
                           const uint8_t *near, const uint32_t near_dims[3], const float near_inverse_affine[16],
                           float reach, tgk_domain **out)
{
    if (!rt || !out || !bvh_min || !bvh_max || !bvh_nodes || !bvh_order)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (n_nodes == 0 || n_order == 0)
        return fail(TGK_ERR_INVALID, "empty BVH");
    for (uint32_t i = 0; i < n_nodes; i++) {
        const int32_t *nd = bvh_nodes + 4 * (size_t)i;
        if (nd[3] > 0) {
            if (nd[2] < 0 || (uint64_t)nd[2] + (uint64_t)nd[3] > n_order)
                return fail(TGK_ERR_INVALID, "BVH leaf %u out of range", i);
        } else if (nd[0] < 0 || nd[1] < 0 || (uint32_t)nd[0] >= n_nodes || (uint32_t)nd[1] >= n_nodes) {
            return fail(TGK_ERR_INVALID, "BVH node %u has invalid children", i);
        }
    }
    for (uint32_t i = 0; i < n_order; i++)
        if (bvh_order[i] >= n_triangles)
            return fail(TGK_ERR_INVALID, "BVH triangle slot out of range");
    for (uint32_t i = 0; i < n_triangles; i++)
        for (int k = 0; k < 3; k++)
            if (triangles[4 * (size_t)i + k] >= n_vertices)
                return fail(TGK_ERR_INVALID, "triangle %u references vertex out of range", i);
    if (int r = select_device(rt))
        return r;
    std::unique_ptr<tgk_domain> d(new (std::nothrow) tgk_domain());
    if (!d)
        return fail(TGK_ERR_NO_MEMORY, "domain allocation");
    d->rt = rt;
    d->kind = tgk::DOMAIN_MESH;
    hipStream_t s = rt->stream;
    if (int r = upload(d->vertices, vertices, (size_t)n_vertices * 4, s, "mesh upload"))
        return r;
    if (int r = upload(d->triangles, triangles, (size_t)n_triangles * 4, s, "mesh upload"))
        return r;
    if (cumulative) {
        if (int r = upload(d->cumulative, cumulative, (size_t)n_triangles, s, "mesh upload"))
            return r;
    }
    if (int r = upload(d->bvh_min, bvh_min, (size_t)n_nodes * 4, s, "BVH upload"))
        return r;
    if (int r = upload(d->bvh_max, bvh_max, (size_t)n_nodes * 4, s, "BVH upload"))
        return r;
    if (int r = upload(d->bvh_nodes, bvh_nodes, (size_t)n_nodes * 4, s, "BVH upload"))
        return r;
    if (int r = upload(d->bvh_order, bvh_order, (size_t)n_order, s, "BVH upload"))
        return r;
    d->n_vertices = n_vertices;
    d->n_triangles = n_triangles;
    d->n_nodes = n_nodes;
    if (near && near_dims && near_inverse_affine) {
        memcpy(d->near_dims, near_dims, sizeof d->near_dims);
        memcpy(d->near_inverse_affine, near_inverse_affine, sizeof d->near_inverse_affine);
        if (int r = upload(d->near, near, (size_t)near_dims[0] * near_dims[1] * near_dims[2], s, "grid upload"))
            return r;
        d->reach = reach;
    }
    TGK_HIP(hipStreamSynchronize(s), "mesh upload");
    *out = d.release();
    return TGK_OK;
}

int tgk_domain_set_segment_grid(tgk_domain *domain, const uint32_t *cell_start, uint64_t n_cells,
                                const float *entries, uint64_t n_entries)
{
    if (!domain || !cell_start || (n_entries > 0 && !entries))
        return fail(TGK_ERR_INVALID, "NULL argument");
    if (domain->kind != tgk::DOMAIN_MESH || !domain->near.ptr)
        return fail(TGK_ERR_INVALID, "candidate lists need a mesh domain with a culling grid");
    const uint64_t cells = (uint64_t)domain->near_dims[0] * domain->near_dims[1] * domain->near_dims[2];
    if (n_cells != cells)
        return fail(TGK_ERR_INVALID, "candidate lists for %llu cells, the culling grid has %llu",
                    (unsigned long long)n_cells, (unsigned long long)cells);
    // Tracking lanes keep a 32-bit cell id with two values reserved.
    if (n_cells >= 0xfffffffeull)
        return fail(TGK_ERR_INVALID, "candidate lists need a culling grid of fewer than 2^32 - 2 cells");
    if (cell_start[0] != 0 || cell_start[n_cells] != n_entries || n_entries > 0xffffffffull)
        return fail(TGK_ERR_INVALID, "candidate offsets do not span the entries");
    for (uint64_t i = 0; i < n_cells; i++) {
        if (cell_start[i + 1] < cell_start[i])
            return fail(TGK_ERR_INVALID, "candidate offsets decrease at cell %llu", (unsigned long long)i);
        // A wavefront sums its 64 lanes' list lengths in 32 bits.
        if (cell_start[i + 1] - cell_start[i] >= (1u << 26))
            return fail(TGK_ERR_INVALID, "cell %llu lists %u candidates; a cell takes fewer than 2^26",
                        (unsigned long long)i, cell_start[i + 1] - cell_start[i]);
    }
    for (uint64_t i = 0; i < n_entries; i++) {
        uint32_t slot;
        memcpy(&slot, entries + 12 * i + 3, sizeof slot);
        if (slot >= domain->n_triangles)
            return fail(TGK_ERR_INVALID, "candidate %llu names triangle %u of %u", (unsigned long long)i, slot,
                        domain->n_triangles);
    }
    if (int r = select_device(domain->rt))
        return r;
    hipStream_t s = domain->rt->stream;
    if (int r = upload(domain->cell_start, cell_start, (size_t)n_cells + 1, s, "candidate upload"))
        return r;
    if (int r = upload(domain->cell_entries, entries, (size_t)n_entries * 12, s, "candidate upload"))
        return r;
    TGK_HIP(hipStreamSynchronize(s), "candidate upload");
    return TGK_OK;
}

int tgk_domain_destroy(tgk_domain *domain)
{
    if (!domain)
        return TGK_OK;
    (void)hipSetDevice(domain->rt->device);
    delete domain;
    return TGK_OK;
}

}  // extern "C"

namespace {

// Seeds per call: seed indices are 32-bit (the reference's get_global_id,
// rng_init's stream id) and the persistent kernels claim them from a 32-bit
// counter that may run past n by one claim per wavefront.
constexpr int64_t MAX_SEEDS = INT32_MAX;

// A field or domain holds device memory of the runtime it was made on and
// was uploaded on that runtime's stream (as was the fast path's field
// transform): a call on another runtime could read it before its stream has
// written it, so calls take only handles of their own runtime.
int check_owner(const tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field,
                const tgk_domain *boundary = nullptr)
{
    if ((domain && domain->rt != rt) || (field && field->rt != rt) || (boundary && boundary->rt != rt))
        return fail(TGK_ERR_INVALID, "field and domain handles must belong to the runtime of the call");
    return TGK_OK;
}

int check_track_args(const tgk_domain *domain, const tgk_field *field, const tgk_algorithm *algorithm, int32_t path,
                     int64_t n_seeds, int64_t batch_size, uint32_t n_points, float save_at)
{
    if (!domain || !field)
        return fail(TGK_ERR_INVALID, "domain and field are required");
    if (path != TGK_PATH_PLAIN && path != TGK_PATH_FAST)
        return fail(TGK_ERR_INVALID, "unknown path %d", path);
    if (int r = check_algorithm(algorithm, field, path == TGK_PATH_FAST ? 91 : 45))
        return r;
    if (n_seeds < 0)
        return fail(TGK_ERR_INVALID, "negative seed count");
    if (n_seeds > MAX_SEEDS)
        return fail(TGK_ERR_INVALID, "%lld seeds: at most %lld per call", (long long)n_seeds, (long long)MAX_SEEDS);
    if (batch_size <= 0)
        return fail(TGK_ERR_INVALID, "batch_size must be positive");
    if (n_points == 0)
        return fail(TGK_ERR_INVALID, "n_points must be positive");
    if (!(save_at > 0.0f) || !std::isfinite(save_at))
        return fail(TGK_ERR_INVALID, "save_at must be finite and positive");
    if (!(algorithm->step_size > 0.0f) || !std::isfinite(algorithm->step_size))
        return fail(TGK_ERR_INVALID, "step_size must be finite and positive");
    if (domain->kind == tgk::DOMAIN_MESH && domain->n_nodes == 0)
        return fail(TGK_ERR_INVALID, "mesh domain without BVH");
    return TGK_OK;
}

tgk::TrackLaunch make_launch(const tgk_domain *domain, const tgk_field *field, const tgk_algorithm *algorithm,
                             const float *d_seeds, int64_t first_seed, int64_t n_seeds, const uint32_t *d_rng_seeds,
                             int64_t batch_size, uint32_t n_points, float save_at, uint32_t *d_lengths,
                             uint32_t *d_steps, uint32_t *d_hits, float *d_dense)
{
    tgk::TrackLaunch p;
    memset(&p, 0, sizeof p);
    p.domain_kind = domain->kind;
    if (domain->kind == tgk::DOMAIN_MASK)
        p.mask = mask_view(domain);
    else
        p.mesh = mesh_view(domain);
    p.field = field_view(field);
    p.algo = algo_params(algorithm);
    p.seeds = d_seeds;
    p.first_seed = first_seed;
    p.n_seeds = n_seeds;
    p.rng_seeds = d_rng_seeds;
    p.batch_size = batch_size;
    p.n_points = n_points;
    p.save_at = save_at;
    const double per_save = std::ceil((double)save_at / (double)algorithm->step_size);
    const double cap = 2.0 * (double)n_points * per_save + 16.0;
    p.max_steps = cap >= 4294967295.0 ? 4294967295u : (uint32_t)cap;
    p.lengths = d_lengths;
    p.steps = d_steps;
    p.hits = d_hits;
    p.points = d_dense;
    return p;
}

// The field as the fast path reads it: transformed once per model (the
// SH -> polynomial map of order 8 or 12, or padded DTI rows) and cached on
// the field.
int ensure_fast_field(tgk_runtime *rt, tgk_field *field, const tgk_algorithm *algorithm, int32_t *sh_order)
{
    if (field->fast_precision != TGK_PRECISION_FP32)
        return fail(TGK_ERR_UNSUPPORTED,
                    "fast-path coefficient storage %d is not served (only FP32): 16-bit coefficients move tracks "
                    "past the 1e-3 voxel budget",
                    field->fast_precision);
    const int32_t model = algorithm->model;
    *sh_order = algorithm->n_parameters > 45 ? 12 : 8;
    if (field->fast_cached != model) {
        if (rt->capturing)
            return fail(TGK_ERR_INVALID, "the field's fast-path transform must be prepared before recording a graph");
        const uint32_t out_stride = model == TGK_MODEL_SH ? (*sh_order == 12 ? 92u : 48u) : 8u;
        const int64_t n_voxels = (int64_t)field->dims[0] * field->dims[1] * field->dims[2];
        TGK_HIP(field->fast.allocate((size_t)n_voxels * out_stride * sizeof(float)), "fast field allocation");
        TGK_HIP(tgk::launch_field_transform(field->data.as<const float>(), n_voxels, field->stride, model,
                                            (int32_t)field->dims[3], *sh_order, field->fast.ptr, out_stride,
                                            rt->stream),
                "field transform");
        field->fast_cached = model;
        field->fast_stride = out_stride;
    }
    return TGK_OK;
}

int fast_launch(tgk_runtime *rt, tgk_field *field, const tgk_algorithm *algorithm, const tgk::TrackLaunch &p,
                tgk::FastLaunch *f)
{
    int32_t sh_order = 8;
    if (int r = ensure_fast_field(rt, field, algorithm, &sh_order))
        return r;
    if (int r = reserve(rt, rt->queue, 256, "queue"))
        return r;
    memset(f, 0, sizeof *f);
    f->track = p;
    char *words = rt->queue.as<char>();
    f->claims = rt->fast_counters.as<uint32_t>();
    f->ended = rt->fast_counters.as<uint32_t>() + 32;
    f->stats = reinterpret_cast<unsigned long long *>(words + FAST_STATS_OFFSET);
    f->precision = field->fast_precision;
    f->coeffs = field->fast.ptr;
    f->coeff_stride = field->fast_stride;
    f->sh_order = sh_order;
    // Unless set: SH launches repack their tail from 16384 streamlines on
    // (65536 on mesh domains: 38.0 -> 37.3 ms per 2M-seed shell launch);
    // DTI launches run in one generation. The DTI step is short, so packing
    // its tail costs more than the idle lanes it removes (145x174x145 DTI
    // Transport, 1M seeds: 13.9 ms per launch repacked, 12.8 ms in one).
    f->min_repack = rt->fast_min_repack != 0                ? rt->fast_min_repack
                    : algorithm->model == TGK_MODEL_DTI          ? 0xffffffffu
                    : p.domain_kind == tgk::DOMAIN_MESH          ? tgk::FAST_MIN_REPACK_MESH
                                                                 : tgk::FAST_MIN_REPACK;
    f->shape = rt->fast_shape;
    // The tracking kernel's cell ids are 32-bit.
    const uint64_t field_voxels = (uint64_t)p.field.dims[0] * p.field.dims[1] * p.field.dims[2];
    const uint64_t mask_voxels = (uint64_t)p.mask.dims[0] * p.mask.dims[1] * p.mask.dims[2];
    if (field_voxels >= 0xfffffffeull || (p.domain_kind == tgk::DOMAIN_MASK && mask_voxels >= 0xfffffffeull))
        return fail(TGK_ERR_UNSUPPORTED, "the fast path takes grids of fewer than 2^32 - 2 voxels");
    f->shared_affine = p.domain_kind == tgk::DOMAIN_MASK && memcmp(p.mask.inv, p.field.inv, sizeof p.mask.inv) == 0;
    // Room for one generation's saved streamlines: one per lane of the grid.
    int64_t lanes = 0;
    TGK_HIP(tgk::prepare_track_fast(*f, rt->fast_grid, &lanes), "fast kernel occupancy");
    const size_t bytes = (size_t)lanes * tgk::FAST_CONTINUATION_VEC4 * sizeof(float4);
    for (int i = 0; i < 2; i++) {
        if (int r = reserve(rt, rt->continuations[i], bytes, "saved streamlines"))
            return r;
        f->cont[i] = rt->continuations[i].as<float4>();
        f->cont_count[i] = reinterpret_cast<uint32_t *>(words + FAST_CONT_COUNT_OFFSET + 4 * i);
    }
    return TGK_OK;
}

int run_track(tgk_runtime *rt, const tgk_domain *domain, tgk_field *field, const tgk_algorithm *algorithm,
              int32_t path, const float *d_seeds, int64_t first_seed, int64_t n_seeds, const uint32_t *d_rng_seeds,
              int64_t batch_size, uint32_t n_points, float save_at, uint32_t *d_lengths, uint32_t *d_steps,
              uint32_t *d_hits, float *d_dense);

int run_compact(tgk_runtime *rt, int32_t path, const float *d_dense, int64_t n_seeds, uint32_t n_points,
                const uint32_t *d_lengths, int64_t *d_offsets, float *d_points)
{
    const size_t scan_bytes = tgk::offsets_workspace_bytes(n_seeds);
    if (int r = ensure_workspace(rt, scan_bytes))
        return r;
    TGK_HIP(tgk::launch_offsets(d_lengths, n_seeds, d_offsets, rt->workspace.ptr, scan_bytes, rt->stream),
            "offsets scan");
    // Both paths write seed-major dense rows.
    TGK_HIP(tgk::launch_compact_seed_major(d_dense, n_seeds, n_points, d_lengths, d_offsets, d_points, rt->stream),
            "compaction");
    return TGK_OK;
}

}  // namespace

extern "C" {

int tgk_track_device(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field,
                     const tgk_algorithm *algorithm, int32_t path, const float *d_seeds, int64_t first_seed,
                     int64_t n_seeds, const uint32_t *d_rng_seeds, int64_t batch_size, uint32_t n_points,
                     float save_at, uint32_t *d_lengths, uint32_t *d_steps, uint32_t *d_hits, float *d_dense)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (int r = check_track_args(domain, field, algorithm, path, n_seeds, batch_size, n_points, save_at))
        return r;
    if (int r = check_owner(rt, domain, field))
        return r;
    if (n_seeds > 0 && (!d_seeds || !d_rng_seeds || !d_lengths || !d_steps || !d_hits || !d_dense))
        return fail(TGK_ERR_INVALID, "NULL device buffer");
    if (first_seed < 0)
        return fail(TGK_ERR_INVALID, "negative first_seed");
    if (int r = select_device(rt))
        return r;
    return run_track(rt, domain, const_cast<tgk_field *>(field), algorithm, path, d_seeds, first_seed, n_seeds,
                     d_rng_seeds, batch_size, n_points, save_at, d_lengths, d_steps, d_hits, d_dense);
}

int tgk_compact_device(tgk_runtime *rt, int32_t path, const float *d_dense, int64_t n_seeds, uint32_t n_points,
                       const uint32_t *d_lengths, int64_t *d_offsets, float *d_points)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (path != TGK_PATH_PLAIN && path != TGK_PATH_FAST)
        return fail(TGK_ERR_INVALID, "unknown path %d", path);
    if (n_seeds < 0 || (n_seeds > 0 && (!d_dense || !d_lengths || !d_offsets || !d_points)))
        return fail(TGK_ERR_INVALID, "bad compaction arguments");
    if (n_seeds == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    return run_compact(rt, path, d_dense, n_seeds, n_points, d_lengths, d_offsets, d_points);
}

int tgk_compact_device_noalloc(tgk_runtime *rt, int32_t path, const float *d_dense, int64_t n_seeds,
                               uint32_t n_points, const uint32_t *d_lengths, int64_t *d_offsets, float *d_points)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (path != TGK_PATH_PLAIN && path != TGK_PATH_FAST)
        return fail(TGK_ERR_INVALID, "unknown path %d", path);
    if (n_seeds < 0 || (n_seeds > 0 && (!d_dense || !d_lengths || !d_offsets || !d_points)))
        return fail(TGK_ERR_INVALID, "bad compaction arguments");
    const size_t scan_bytes = tgk::offsets_workspace_bytes(n_seeds);
    if (scan_bytes > sizeof(scratch))
        return fail(TGK_ERR_INVALID,
                    "%lld seeds exceed the allocation-free compaction's maximum of %lld per call (use "
                    "tgk_compact_device)",
                    (long long)n_seeds, (long long)TGK_COMPACT_NOALLOC_MAX_SEEDS);
    if (n_seeds == 0)
        return TGK_OK;
    void *sums = scratch_address(rt->device);
    if (!sums)
        return fail(TGK_ERR_UNSUPPORTED, "no allocation-free scratch on device %d", rt->device);
    if (int r = select_device(rt))
        return r;
    // The scan writes every per-tile sum before reading it, so the scratch
    // needs no clearing between calls.
    TGK_HIP(tgk::launch_offsets(d_lengths, n_seeds, d_offsets, sums, scan_bytes, rt->stream), "offsets scan");
    TGK_HIP(tgk::launch_compact_seed_major(d_dense, n_seeds, n_points, d_lengths, d_offsets, d_points, rt->stream),
            "compaction");
    return TGK_OK;
}

int tgk_track(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field, const tgk_algorithm *algorithm,
              int32_t path, const float *seeds, int64_t n_seeds, const uint32_t *rng_seeds, int64_t batch_size,
              uint32_t n_points, float save_at, tgk_tracks **out)
{
    if (!rt || !out)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (int r = check_track_args(domain, field, algorithm, path, n_seeds, batch_size, n_points, save_at))
        return r;
    if (int r = check_owner(rt, domain, field))
        return r;
    if (n_seeds > 0 && (!seeds || !rng_seeds))
        return fail(TGK_ERR_INVALID, "NULL seeds");
    if (int r = select_device(rt))
        return r;
    std::unique_ptr<tgk_tracks> tracks(new (std::nothrow) tgk_tracks());
    if (!tracks)
        return fail(TGK_ERR_NO_MEMORY, "tracks allocation");
    tracks->lengths.resize((size_t)n_seeds);
    tracks->steps.resize((size_t)n_seeds);
    tracks->hits.resize((size_t)n_seeds);
    if (n_seeds == 0) {
        *out = tracks.release();
        return TGK_OK;
    }
    // Chunk size: the dense per-seed rows of a chunk and its compacted copy
    // take chunk * per_seed bytes each; a chunk may use a quarter of the free
    // device memory, between 2 and 48 GiB each (3.5M seeds of 1200 points),
    // so that most calls are one chunk, whose page-locked points become the
    // result without a host copy.
    const size_t per_seed = (size_t)n_points * 3 * sizeof(float);
    size_t budget = (size_t)2 << 30;
    size_t free_bytes = 0, total_bytes = 0;
    if (hipMemGetInfo(&free_bytes, &total_bytes) == hipSuccess)
        budget = std::min<size_t>((size_t)48 << 30, std::max<size_t>(budget, free_bytes / 8));
    if (const char *env = getenv("TGK_CHUNK_BYTES"))
        budget = std::max<size_t>((size_t)strtoull(env, nullptr, 10), per_seed);
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n_seeds, (int64_t)(budget / per_seed)));
    const int64_t n_batches = (n_seeds + batch_size - 1) / batch_size;
    tracks->device = rt->device;

    // Working buffers are the runtime's, kept between calls (grown, never
    // shrunk until the runtime is destroyed): a 1M-seed call needs about
    // 29 GB of them, whose allocation would otherwise recur per call.
    TrackBuffers &w = rt->track_buffers;
    if (int r = reserve(rt, w.seeds, (size_t)chunk * 8 * sizeof(float), "seed buffer"))
        return r;
    if (int r = reserve(rt, w.rng, (size_t)n_batches * sizeof(uint32_t), "rng seeds"))
        return r;
    TGK_HIP(hipMemcpyAsync(w.rng.ptr, rng_seeds, (size_t)n_batches * sizeof(uint32_t), hipMemcpyHostToDevice,
                           rt->stream),
            "rng seeds");
    if (int r = reserve(rt, w.lengths, (size_t)chunk * 4, "lengths buffer"))
        return r;
    if (int r = reserve(rt, w.steps, (size_t)chunk * 4, "steps buffer"))
        return r;
    if (int r = reserve(rt, w.hits, (size_t)chunk * 4, "hits buffer"))
        return r;
    if (int r = reserve(rt, w.dense, (size_t)chunk * per_seed, "point buffer"))
        return r;
    if (int r = reserve(rt, w.offsets, (size_t)(chunk + 1) * 8, "offsets buffer"))
        return r;
    if (int r = reserve(rt, w.points, (size_t)chunk * per_seed, "point buffer"))
        return r;
    DeviceBuffer &d_seeds = w.seeds, &d_rng = w.rng, &d_lengths = w.lengths, &d_steps = w.steps, &d_hits = w.hits,
                 &d_dense = w.dense, &d_offsets = w.offsets, &d_points = w.points;

    for (int64_t start = 0; start < n_seeds; start += chunk) {
        const int64_t count = std::min(chunk, n_seeds - start);
        TGK_HIP(hipMemcpyAsync(d_seeds.ptr, seeds + 8 * start, (size_t)count * 8 * sizeof(float),
                               hipMemcpyHostToDevice, rt->stream),
                "seed upload");
        if (int r = run_track(rt, domain, const_cast<tgk_field *>(field), algorithm, path, d_seeds.as<float>(),
                              start, count, d_rng.as<uint32_t>(), batch_size, n_points, save_at,
                              d_lengths.as<uint32_t>(), d_steps.as<uint32_t>(), d_hits.as<uint32_t>(),
                              d_dense.as<float>()))
            return r;
        if (int r = run_compact(rt, path, d_dense.as<float>(), count, n_points, d_lengths.as<uint32_t>(),
                                d_offsets.as<int64_t>(), d_points.as<float>()))
            return r;
        int64_t total = 0;
        TGK_HIP(hipMemcpyAsync(&total, d_offsets.as<int64_t>() + count, sizeof total, hipMemcpyDeviceToHost,
                               rt->stream),
                "offsets copy");
        TGK_HIP(hipMemcpyAsync(tracks->lengths.data() + start, d_lengths.ptr, (size_t)count * 4,
                               hipMemcpyDeviceToHost, rt->stream),
                "lengths copy");
        TGK_HIP(hipMemcpyAsync(tracks->steps.data() + start, d_steps.ptr, (size_t)count * 4, hipMemcpyDeviceToHost,
                               rt->stream),
                "steps copy");
        TGK_HIP(hipMemcpyAsync(tracks->hits.data() + start, d_hits.ptr, (size_t)count * 4, hipMemcpyDeviceToHost,
                               rt->stream),
                "hits copy");
        TGK_HIP(hipStreamSynchronize(rt->stream), "tracking");
        tgk_tracks::Piece piece;
        piece.n_points = total;
        const size_t piece_bytes = (size_t)total * 3 * sizeof(float);
        piece.pinned.reset(new (std::nothrow) PinnedBuffer());
        if (!piece.pinned)
            return fail(TGK_ERR_NO_MEMORY, "tracks allocation");
        if (piece_bytes && piece.pinned->allocate(piece_bytes) == hipSuccess) {
            // In stream order: the next chunk's compaction, which reuses
            // d_points, comes after this copy.
            TGK_HIP(hipMemcpyAsync(piece.pinned->ptr, d_points.ptr, piece_bytes, hipMemcpyDeviceToHost, rt->stream),
                    "points copy");
            tracks->pieces.push_back(std::move(piece));
            tracks->total_points += total;
            continue;
        }
        (void)hipGetLastError();  // clear a failed page-locked allocation's error
        piece.pinned.reset();
        piece.device.reset(new (std::nothrow) DeviceBuffer());
        if (!piece.device)
            return fail(TGK_ERR_NO_MEMORY, "tracks allocation");
        const hipError_t e = piece.device->allocate(piece_bytes);
        if (e == hipSuccess) {
            if (piece_bytes)
                TGK_HIP(hipMemcpyAsync(piece.device->ptr, d_points.ptr, piece_bytes, hipMemcpyDeviceToDevice,
                                       rt->stream),
                        "points copy");
        } else if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();  // clear the failed allocation's error
            piece.device.reset();
            try {
                piece.host.resize((size_t)total * 3);
            } catch (const std::bad_alloc &) {
                return fail(TGK_ERR_NO_MEMORY, "points: no room in device or host memory");
            }
            TGK_HIP(hipMemcpy(piece.host.data(), d_points.ptr, piece_bytes, hipMemcpyDeviceToHost), "points copy");
        } else {
            return hip_fail(e, "points allocation");
        }
        tracks->pieces.push_back(std::move(piece));
        tracks->total_points += total;
    }
    // The working buffers are freed on return: finish the piece copies first.
    TGK_HIP(hipStreamSynchronize(rt->stream), "points copy");
    *out = tracks.release();
    return TGK_OK;
}

int tgk_tracks_size(const tgk_tracks *tracks, int64_t *n_seeds, int64_t *total_points)
{
    if (!tracks)
        return fail(TGK_ERR_INVALID, "tracks is NULL");
    if (n_seeds)
        *n_seeds = (int64_t)tracks->lengths.size();
    if (total_points)
        *total_points = tracks->total_points;
    return TGK_OK;
}

int tgk_tracks_copy(const tgk_tracks *tracks, uint32_t *lengths, uint32_t *steps, uint32_t *hits, float *points)
{
    if (!tracks)
        return fail(TGK_ERR_INVALID, "tracks is NULL");
    const size_t n = tracks->lengths.size();
    if (lengths && n)
        memcpy(lengths, tracks->lengths.data(), n * 4);
    if (steps && n)
        memcpy(steps, tracks->steps.data(), n * 4);
    if (hits && n)
        memcpy(hits, tracks->hits.data(), n * 4);
    if (points && tracks->total_points > 0) {
        TGK_HIP(hipSetDevice(tracks->device), "hipSetDevice");
        // Host-resident pieces are copied by several threads: a single
        // thread moves ~3 GB/s into freshly mapped pages (each first touch
        // a page fault), several fault and copy in parallel.
        struct Span {
            char *dst;
            const char *src;
            size_t bytes;
        };
        std::vector<Span> spans;
        float *dst = points;
        for (const tgk_tracks::Piece &piece : tracks->pieces) {
            const size_t bytes = (size_t)piece.n_points * 3 * sizeof(float);
            if (bytes && piece.device)
                TGK_HIP(hipMemcpy(dst, piece.device->ptr, bytes, hipMemcpyDeviceToHost), "points copy");
            else if (bytes)
                spans.push_back({reinterpret_cast<char *>(dst),
                                 static_cast<const char *>(piece.pinned ? piece.pinned->ptr
                                                                        : static_cast<const void *>(piece.host.data())),
                                 bytes});
            dst += (size_t)piece.n_points * 3;
        }
        parallel_copy(spans.data(), spans.size());
    }
    return TGK_OK;
}

int tgk_tracks_points(const tgk_tracks *tracks, const float **points)
{
    if (!tracks || !points)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *points = nullptr;
    if (tracks->pieces.size() == 1 && tracks->pieces[0].pinned)
        *points = static_cast<const float *>(tracks->pieces[0].pinned->ptr);
    return TGK_OK;
}

int tgk_tracks_destroy(tgk_tracks *tracks)
{
    if (tracks && !tracks->pieces.empty())
        (void)hipSetDevice(tracks->device);
    delete tracks;
    return TGK_OK;
}

}  // extern "C"

namespace {

int run_track(tgk_runtime *rt, const tgk_domain *domain, tgk_field *field, const tgk_algorithm *algorithm,
              int32_t path, const float *d_seeds, int64_t first_seed, int64_t n_seeds, const uint32_t *d_rng_seeds,
              int64_t batch_size, uint32_t n_points, float save_at, uint32_t *d_lengths, uint32_t *d_steps,
              uint32_t *d_hits, float *d_dense)
{
    if (n_seeds == 0)
        return TGK_OK;
    tgk::TrackLaunch p = make_launch(domain, field, algorithm, d_seeds, first_seed, n_seeds, d_rng_seeds, batch_size,
                                     n_points, save_at, d_lengths, d_steps, d_hits, d_dense);
    // The plain kernel's seed counter: the queue buffer's first word, which
    // the connectome and TOD kernels also use on this stream, one launch at a
    // time.
    p.queue = rt->queue.as<uint32_t>();
    if (path == TGK_PATH_PLAIN) {
        TGK_HIP(tgk::launch_track_plain(p, rt->stream), "plain tracking kernel");
        return TGK_OK;
    }
    tgk::FastLaunch f;
    if (int r = fast_launch(rt, field, algorithm, p, &f))
        return r;
    TGK_HIP(tgk::launch_track_fast(f, rt->fast_grid, rt->stream), "fast tracking kernel");
    return TGK_OK;
}

// Scratch device copies for the one-op drivers.
struct Staging {
    std::deque<DeviceBuffer> buffers;
    template <class T> int in(const T *host, size_t count, hipStream_t s, T **dev)
    {
        buffers.emplace_back();
        if (int r = upload(buffers.back(), host, count, s, "input upload"))
            return r;
        *dev = buffers.back().as<T>();
        return TGK_OK;
    }
    template <class T> int out(size_t count, T **dev)
    {
        buffers.emplace_back();
        TGK_HIP(buffers.back().allocate(std::max<size_t>(count, 1) * sizeof(T)), "output allocation");
        *dev = buffers.back().as<T>();
        return TGK_OK;
    }
};

template <class T> int fetch(T *host, const T *dev, size_t count, hipStream_t s)
{
    if (count)
        TGK_HIP(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, s), "output copy");
    TGK_HIP(hipStreamSynchronize(s), "one-op driver");
    return TGK_OK;
}

}  // namespace

extern "C" {

int tgk_track_timed(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field, const tgk_algorithm *algorithm,
                    int32_t path, const float *seeds, int64_t n_seeds, uint32_t n_points, float save_at,
                    int32_t repeats, float *milliseconds, uint64_t *steps)
{
    if (!rt || !milliseconds || !steps)
        return fail(TGK_ERR_INVALID, "NULL argument");
    if (int r = check_track_args(domain, field, algorithm, path, n_seeds, 1, n_points, save_at))
        return r;
    if (int r = check_owner(rt, domain, field))
        return r;
    if (n_seeds <= 0 || !seeds || repeats < 1)
        return fail(TGK_ERR_INVALID, "timing needs seeds and at least one repeat");
    if (rt->capturing)
        return fail(TGK_ERR_INVALID, "timing allocates; not while a graph is being recorded");
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *d_seeds;
    uint32_t *d_rng, *d_lengths, *d_steps, *d_hits;
    float *d_dense;
    const uint32_t rng_seed = 0;
    if (int r = st.in(seeds, (size_t)n_seeds * 8, rt->stream, const_cast<float **>(&d_seeds)))
        return r;
    if (int r = st.in(&rng_seed, 1, rt->stream, &d_rng))
        return r;
    if (int r = st.out((size_t)n_seeds, &d_lengths))
        return r;
    if (int r = st.out((size_t)n_seeds, &d_steps))
        return r;
    if (int r = st.out((size_t)n_seeds, &d_hits))
        return r;
    if (int r = st.out((size_t)n_seeds * n_points * 3, &d_dense))
        return r;
    hipEvent_t start = nullptr, stop = nullptr;
    TGK_HIP(hipEventCreate(&start), "hipEventCreate");
    int status = TGK_OK;
    float best = -1.0f;
    if (hipEventCreate(&stop) != hipSuccess) {
        (void)hipEventDestroy(start);
        return fail(TGK_ERR_DEVICE, "hipEventCreate");
    }
    // One untimed launch (field transform, first-use setup), then the timed ones.
    for (int i = 0; i <= repeats && status == TGK_OK; i++) {
        if (hipEventRecord(start, rt->stream) != hipSuccess) {
            status = fail(TGK_ERR_DEVICE, "hipEventRecord");
            break;
        }
        status = run_track(rt, domain, const_cast<tgk_field *>(field), algorithm, path, d_seeds, 0, n_seeds, d_rng,
                           n_seeds, n_points, save_at, d_lengths, d_steps, d_hits, d_dense);
        if (status != TGK_OK)
            break;
        float ms = 0.0f;
        if (hipEventRecord(stop, rt->stream) != hipSuccess || hipEventSynchronize(stop) != hipSuccess ||
            hipEventElapsedTime(&ms, start, stop) != hipSuccess) {
            status = fail(TGK_ERR_DEVICE, "timing events");
            break;
        }
        if (i > 0 && (best < 0.0f || ms < best))
            best = ms;
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    if (status != TGK_OK)
        return status;
    std::vector<uint32_t> counts((size_t)n_seeds);
    if (int r = fetch(counts.data(), d_steps, counts.size(), rt->stream))
        return r;
    uint64_t total = 0;
    for (uint32_t c : counts)
        total += c;
    *milliseconds = best;
    *steps = total;
    return TGK_OK;
}

int tgk_evaluate(tgk_runtime *rt, int32_t model, int32_t n_parameters, const float *params, const float *orientations,
                 int64_t n, float *out)
{
    if (!rt || n < 0 || (n > 0 && (!params || !orientations || !out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    tgk_algorithm a = {model, TGK_ALGO_TRANSPORT, n_parameters, 1.0f, 1.0f, 0.0f};
    if (int r = check_algorithm(&a, nullptr))
        return r;
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *dp, *dor;
    float *dout;
    if (int r = st.in(params, (size_t)n * n_parameters, rt->stream, const_cast<float **>(&dp)))
        return r;
    if (int r = st.in(orientations, (size_t)n * 3, rt->stream, const_cast<float **>(&dor)))
        return r;
    if (int r = st.out((size_t)n * 3, &dout))
        return r;
    TGK_HIP(tgk::launch_evaluate(model, n_parameters, dp, dor, n, dout, rt->stream), "evaluate kernel");
    return fetch(out, dout, (size_t)n * 3, rt->stream);
}

int tgk_update_orientation(tgk_runtime *rt, const tgk_algorithm *algorithm, uint32_t rng_seed, const float *params,
                           const float *orientations, int64_t n, float *out)
{
    if (!rt || n < 0 || (n > 0 && (!params || !orientations || !out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_algorithm(algorithm, nullptr))
        return r;
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *dp, *dor;
    float *dout;
    if (int r = st.in(params, (size_t)n * algorithm->n_parameters, rt->stream, const_cast<float **>(&dp)))
        return r;
    if (int r = st.in(orientations, (size_t)n * 3, rt->stream, const_cast<float **>(&dor)))
        return r;
    if (int r = st.out((size_t)n * 3, &dout))
        return r;
    TGK_HIP(tgk::launch_update(algo_params(algorithm), rng_seed, dp, dor, n, dout, rt->stream), "update kernel");
    return fetch(out, dout, (size_t)n * 3, rt->stream);
}

int tgk_field_evaluate(tgk_runtime *rt, const tgk_field *field, const float *points, int64_t n, float *out)
{
    if (!rt || !field || n < 0 || (n > 0 && (!points || !out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_owner(rt, nullptr, field))
        return r;
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *dpts;
    float *dout;
    if (int r = st.in(points, (size_t)n * 3, rt->stream, const_cast<float **>(&dpts)))
        return r;
    if (int r = st.out((size_t)n * field->dims[3], &dout))
        return r;
    TGK_HIP(tgk::launch_field_evaluate(field_view(field), dpts, n, dout, rt->stream), "field kernel");
    return fetch(out, dout, (size_t)n * field->dims[3], rt->stream);
}

int tgk_domain_contains(tgk_runtime *rt, const tgk_domain *domain, const float *points, int64_t n, uint8_t *out)
{
    if (!rt || !domain || n < 0 || (n > 0 && (!points || !out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_owner(rt, domain, nullptr))
        return r;
    if (domain->kind != tgk::DOMAIN_MASK)
        return fail(TGK_ERR_UNSUPPORTED, "point membership is defined for mask domains");
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *dpts;
    uint8_t *dout;
    if (int r = st.in(points, (size_t)n * 3, rt->stream, const_cast<float **>(&dpts)))
        return r;
    if (int r = st.out((size_t)n, &dout))
        return r;
    TGK_HIP(tgk::launch_contains(mask_view(domain), dpts, n, dout, rt->stream), "contains kernel");
    return fetch(out, dout, (size_t)n, rt->stream);
}

}  // extern "C"

namespace {

int segment_hit(tgk_runtime *rt, const tgk_domain *domain, int method, const float *starts, const float *ends,
                int64_t n, uint8_t *found, uint32_t *hit, float *t)
{
    if (!rt || !domain || n < 0 || (n > 0 && (!starts || !ends || !found || !hit || !t)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_owner(rt, domain, nullptr))
        return r;
    if (domain->n_nodes == 0)
        return fail(TGK_ERR_INVALID, "domain has no BVH");
    if (method == 1 && !domain->cell_start.ptr)
        return fail(TGK_ERR_INVALID, "domain has no candidate lists");
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *ds, *de;
    uint8_t *dfound;
    uint32_t *dhit;
    float *dt;
    if (int r = st.in(starts, (size_t)n * 3, rt->stream, const_cast<float **>(&ds)))
        return r;
    if (int r = st.in(ends, (size_t)n * 3, rt->stream, const_cast<float **>(&de)))
        return r;
    if (int r = st.out((size_t)n, &dfound))
        return r;
    if (int r = st.out((size_t)n, &dhit))
        return r;
    if (int r = st.out((size_t)n, &dt))
        return r;
    TGK_HIP(tgk::launch_bvh_hit(mesh_view(domain), method, ds, de, n, dfound, dhit, dt, rt->stream),
            "segment hit kernel");
    if (int r = fetch(found, dfound, (size_t)n, rt->stream))
        return r;
    if (int r = fetch(hit, dhit, (size_t)n, rt->stream))
        return r;
    return fetch(t, dt, (size_t)n, rt->stream);
}

}  // namespace

extern "C" {

int tgk_bvh_segment_hit(tgk_runtime *rt, const tgk_domain *domain, const float *starts, const float *ends, int64_t n,
                        uint8_t *found, uint32_t *hit, float *t)
{
    return segment_hit(rt, domain, 0, starts, ends, n, found, hit, t);
}

int tgk_grid_segment_hit(tgk_runtime *rt, const tgk_domain *domain, const float *starts, const float *ends,
                         int64_t n, uint8_t *found, uint32_t *hit, float *t)
{
    return segment_hit(rt, domain, 1, starts, ends, n, found, hit, t);
}

int tgk_math(tgk_runtime *rt, int32_t fn, const float *x, const float *y, int64_t n, float *out)
{
    if (!rt || fn < 0 || fn > 7 || n < 0 || (n > 0 && (!x || !y || !out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    const float *dx, *dy;
    float *dout;
    if (int r = st.in(x, (size_t)n, rt->stream, const_cast<float **>(&dx)))
        return r;
    if (int r = st.in(y, (size_t)n, rt->stream, const_cast<float **>(&dy)))
        return r;
    if (int r = st.out((size_t)n, &dout))
        return r;
    TGK_HIP(tgk::launch_math(fn, dx, dy, n, dout, rt->stream), "math kernel");
    return fetch(out, dout, (size_t)n, rt->stream);
}

int tgk_connectome(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
                   const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds,
                   uint32_t n_steps, float minimum_length, int32_t *rows, int32_t *cols, float *weights)
{
    if (!rt || !domain || !boundary || !field || n_seeds < 0 || (n_seeds > 0 && (!rows || !cols || !weights)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (path != TGK_PATH_PLAIN && path != TGK_PATH_FAST)
        return fail(TGK_ERR_INVALID, "unknown path %d", path);
    if (n_seeds > MAX_SEEDS)
        return fail(TGK_ERR_INVALID, "%lld seeds: at most %lld per call", (long long)n_seeds, (long long)MAX_SEEDS);
    if (int r = check_algorithm(algorithm, field, path == TGK_PATH_FAST ? 91 : 45))
        return r;
    if (int r = check_owner(rt, domain, field, boundary))
        return r;
    if (domain->kind != tgk::DOMAIN_MASK)
        return fail(TGK_ERR_INVALID, "the connectome's domain must be a mask domain");
    if (boundary->kind != tgk::DOMAIN_MESH || boundary->n_nodes == 0 || !boundary->cumulative.ptr ||
        boundary->n_triangles == 0)
        return fail(TGK_ERR_INVALID, "the boundary needs triangles, cumulative areas and a BVH");
    if (!(algorithm->step_size > 0.0f) || !std::isfinite(algorithm->step_size))
        return fail(TGK_ERR_INVALID, "step_size must be finite and positive");
    if (!(minimum_length >= 0.0f) || !std::isfinite(minimum_length))
        return fail(TGK_ERR_INVALID, "minimum_length must be finite and non-negative");
    if (n_seeds == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    int32_t *drows, *dcols;
    float *dweights;
    if (int r = st.out((size_t)n_seeds * 9, &drows))
        return r;
    if (int r = st.out((size_t)n_seeds * 9, &dcols))
        return r;
    if (int r = st.out((size_t)n_seeds * 9, &dweights))
        return r;
    tgk::ConnectomeLaunch c;
    memset(&c, 0, sizeof c);
    c.mask = mask_view(domain);
    c.field = field_view(field);
    c.boundary = mesh_view(boundary);
    c.cumulative = boundary->cumulative.as<const float>();
    c.n_triangles = boundary->n_triangles;
    c.algo = algo_params(algorithm);
    c.rng_seed = rng_seed;
    c.n_seeds = n_seeds;
    c.n_steps = n_steps;
    c.min_length = minimum_length;
    c.rows = drows;
    c.cols = dcols;
    c.weights = dweights;
    c.queue = rt->queue.as<uint32_t>();
    if (path == TGK_PATH_PLAIN) {
        TGK_HIP(tgk::launch_connectome(c, rt->stream), "connectome kernel");
    } else {
        tgk::FastConnectomeLaunch f;
        int32_t sh_order = 8;
        if (int r = ensure_fast_field(rt, const_cast<tgk_field *>(field), algorithm, &sh_order))
            return r;
        f.connectome = c;
        f.coeffs = field->fast.ptr;
        f.coeff_stride = field->fast_stride;
        f.sh_order = sh_order;
        TGK_HIP(tgk::launch_connectome_fast(f, rt->stream), "fast connectome kernel");
    }
    if (int r = fetch(rows, drows, (size_t)n_seeds * 9, rt->stream))
        return r;
    if (int r = fetch(cols, dcols, (size_t)n_seeds * 9, rt->stream))
        return r;
    return fetch(weights, dweights, (size_t)n_seeds * 9, rt->stream);
}

}  // extern "C"

namespace {

// Checks shared by the TOD entry points; fills the launch on success.
int tod_launch(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
               const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds, uint32_t n_steps,
               const uint32_t output_dims[3], const float output_inverse_affine[16], tgk::TodLaunch *t,
               int64_t *n_values)
{
    if (!rt || !domain || !boundary || !field || n_seeds < 0 || !output_dims || !output_inverse_affine)
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (path != TGK_PATH_PLAIN && path != TGK_PATH_FAST)
        return fail(TGK_ERR_INVALID, "unknown path %d", path);
    if (n_seeds > MAX_SEEDS)
        return fail(TGK_ERR_INVALID, "%lld seeds: at most %lld per call", (long long)n_seeds, (long long)MAX_SEEDS);
    if (int r = check_algorithm(algorithm, field, path == TGK_PATH_FAST ? 91 : 45))
        return r;
    if (int r = check_owner(rt, domain, field, boundary))
        return r;
    if (domain->kind != tgk::DOMAIN_MASK)
        return fail(TGK_ERR_INVALID, "the TOD's domain must be a mask domain");
    if (boundary->n_triangles == 0 || !boundary->cumulative.ptr)
        return fail(TGK_ERR_INVALID, "the boundary needs triangles and cumulative areas");
    if (!(algorithm->step_size > 0.0f) || !std::isfinite(algorithm->step_size))
        return fail(TGK_ERR_INVALID, "step_size must be finite and positive");
    for (int i = 0; i < 12; i++)
        if (!std::isfinite(output_inverse_affine[i]))
            return fail(TGK_ERR_INVALID, "output inverse affine must be finite");
    // Cells and bytes of the output grid, refusing any product that would
    // not fit (each factor is below 2^32, so the first product cannot wrap).
    const uint64_t limit = (uint64_t)INT64_MAX / (45 * sizeof(float));
    const uint64_t plane = (uint64_t)output_dims[0] * output_dims[1];
    if (output_dims[2] != 0 && plane > limit / output_dims[2])
        return fail(TGK_ERR_INVALID, "output grid too large");
    const uint64_t cells = plane * output_dims[2];
    memset(t, 0, sizeof *t);
    t->mask = mask_view(domain);
    t->field = field_view(field);
    t->vertices = boundary->vertices.as<const float4>();
    t->triangles = boundary->triangles.as<const uint4>();
    t->cumulative = boundary->cumulative.as<const float>();
    t->n_triangles = boundary->n_triangles;
    t->algo = algo_params(algorithm);
    t->rng_seed = rng_seed;
    t->n_seeds = n_seeds;
    t->n_steps = n_steps;
    for (int i = 0; i < 3; i++)
        t->grid.dims[i] = output_dims[i];
    copy_rows3(output_inverse_affine, t->grid.inv);
    t->queue = rt->queue.as<uint32_t>();
    *n_values = (int64_t)cells * 45;
    return TGK_OK;
}

// Zero, deposit, normalise: kernels only, on the runtime stream.
int enqueue_tod(tgk_runtime *rt, tgk_field *field, const tgk_algorithm *algorithm, int32_t path, tgk::TodLaunch &t,
                int64_t n_values, int32_t normalize, float *d_coefficients)
{
    t.coefficients = d_coefficients;
    tgk::FastTodLaunch f;
    if (path == TGK_PATH_FAST) {
        memset(&f, 0, sizeof f);
        if (int r = ensure_fast_field(rt, field, algorithm, &f.sh_order))
            return r;
        f.coeffs = field->fast.ptr;
        f.coeff_stride = field->fast_stride;
    }
    TGK_HIP(tgk::launch_zero(d_coefficients, n_values, rt->stream), "TOD clear");
    if (path == TGK_PATH_FAST) {
        f.tod = t;
        TGK_HIP(tgk::launch_tod_fast(f, rt->stream), "fast TOD kernel");
    } else {
        TGK_HIP(tgk::launch_tod(t, rt->stream), "TOD kernel");
    }
    if (normalize)
        TGK_HIP(tgk::launch_normalize_odf(d_coefficients, n_values / 45, rt->stream), "TOD normalisation");
    return TGK_OK;
}

}  // namespace

extern "C" {

int tgk_tod(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
            const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds, uint32_t n_steps,
            const uint32_t output_dims[3], const float output_inverse_affine[16], int32_t normalize,
            float *coefficients)
{
    tgk::TodLaunch t;
    int64_t n_values = 0;
    if (int r = tod_launch(rt, domain, boundary, field, algorithm, path, rng_seed, n_seeds, n_steps, output_dims,
                           output_inverse_affine, &t, &n_values))
        return r;
    if (n_values > 0 && !coefficients)
        return fail(TGK_ERR_INVALID, "coefficients is NULL");
    if (n_values == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    float *dout;
    if (int r = st.out((size_t)n_values, &dout))
        return r;
    if (int r = enqueue_tod(rt, const_cast<tgk_field *>(field), algorithm, path, t, n_values, normalize, dout))
        return r;
    return fetch(coefficients, dout, (size_t)n_values, rt->stream);
}

int tgk_tod_device(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
                   const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds,
                   uint32_t n_steps, const uint32_t output_dims[3], const float output_inverse_affine[16],
                   int32_t normalize, float *d_coefficients)
{
    tgk::TodLaunch t;
    int64_t n_values = 0;
    if (int r = tod_launch(rt, domain, boundary, field, algorithm, path, rng_seed, n_seeds, n_steps, output_dims,
                           output_inverse_affine, &t, &n_values))
        return r;
    if (n_values > 0 && !d_coefficients)
        return fail(TGK_ERR_INVALID, "coefficients is NULL");
    if (n_values == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    return enqueue_tod(rt, const_cast<tgk_field *>(field), algorithm, path, t, n_values, normalize, d_coefficients);
}

int tgk_normalize_odf(tgk_runtime *rt, int64_t n_cells, float *coefficients)
{
    if (!rt || n_cells < 0 || (n_cells > 0 && !coefficients))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (n_cells == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    float *dev;
    if (int r = st.in(coefficients, (size_t)n_cells * 45, rt->stream, &dev))
        return r;
    TGK_HIP(tgk::launch_normalize_odf(dev, n_cells, rt->stream), "ODF normalisation");
    return fetch(coefficients, dev, (size_t)n_cells * 45, rt->stream);
}

int tgk_domain_seed(tgk_runtime *rt, const tgk_domain *domain, uint32_t rng_seed, int64_t n, float *out)
{
    if (!rt || !domain || n < 0 || (n > 0 && !out))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_owner(rt, domain, nullptr))
        return r;
    if (domain->n_triangles == 0 || !domain->cumulative.ptr)
        return fail(TGK_ERR_INVALID, "domain has no boundary surface");
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    Staging st;
    float *dout;
    if (int r = st.out((size_t)n * 8, &dout))
        return r;
    TGK_HIP(tgk::launch_domain_seed(domain->vertices.as<const float4>(), domain->triangles.as<const uint4>(),
                                    domain->cumulative.as<const float>(), domain->n_triangles, rng_seed, n, dout,
                                    rt->stream),
            "boundary seeding kernel");
    return fetch(out, dout, (size_t)n * 8, rt->stream);
}

}  // extern "C"

namespace {

int check_seeds_args(int32_t model, int32_t n_parameters, int32_t method)
{
    if (method != TGK_SEEDS_VALU && method != TGK_SEEDS_MFMA)
        return fail(TGK_ERR_INVALID, "unknown seeding method %d", method);
    if (method == TGK_SEEDS_MFMA && model == TGK_MODEL_SH && n_parameters > 45) {
        // The MFMA method takes SH to order 12 (91 coefficients) and beyond,
        // up to 96 coefficients.
        if (n_parameters > 96)
            return fail(TGK_ERR_UNSUPPORTED, "SH with %d coefficients: 1..96 supported", n_parameters);
        return TGK_OK;
    }
    tgk_algorithm a = {model, TGK_ALGO_TRANSPORT, n_parameters, 1.0f, 1.0f, 0.0f};
    return check_algorithm(&a, nullptr);
}

// The runtime's seeding table for (model, K, method), built on first use.
int prepare_seeds(tgk_runtime *rt, int32_t model, int32_t n_parameters, int32_t method)
{
    if (rt->seeds_key[0] == model && rt->seeds_key[1] == n_parameters && rt->seeds_key[2] == method)
        return TGK_OK;
    if (rt->capturing)
        return fail(TGK_ERR_INVALID, "the seeding table must be prepared before recording a graph");
    const size_t bytes = tgk::seeds_from_fod_workspace_bytes(model, n_parameters, 0, method);
    if (int r = reserve(rt, rt->seeds_table, bytes, "seeding table"))
        return r;
    rt->seeds_key[0] = -1;
    TGK_HIP(tgk::prepare_seeds_from_fod(model, n_parameters, method, rt->seeds_table.ptr, bytes, rt->stream),
            "seeding table");
    rt->seeds_key[0] = model;
    rt->seeds_key[1] = n_parameters;
    rt->seeds_key[2] = method;
    return TGK_OK;
}

}  // namespace

extern "C" {

int tgk_seeds_from_fod(tgk_runtime *rt, int32_t model, int32_t n_parameters, const float *fod, const float *voxels,
                       uint32_t nnz, const float affine[16], uint32_t *states, int64_t n, int32_t method, float *out)
{
    if (!rt || n < 0 || (n > 0 && (!fod || !voxels || !affine || !states || !out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_seeds_args(model, n_parameters, method))
        return r;
    if (n > 0 && nnz == 0)
        return fail(TGK_ERR_INVALID, "the FOD has no nonzero voxel");
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    if (int r = prepare_seeds(rt, model, n_parameters, method))
        return r;
    Staging st;
    float *dfod, *dvox, *daff, *dout;
    uint32_t *dstates;
    if (int r = st.in(fod, (size_t)nnz * n_parameters, rt->stream, &dfod))
        return r;
    if (int r = st.in(voxels, (size_t)nnz * 4, rt->stream, &dvox))
        return r;
    if (int r = st.in(affine, 16, rt->stream, &daff))
        return r;
    if (int r = st.in(states, (size_t)n * 2, rt->stream, &dstates))
        return r;
    if (int r = st.out((size_t)n * 8, &dout))
        return r;
    TGK_HIP(tgk::launch_seeds_from_fod_prepared(model, n_parameters, dfod, dvox, nnz, daff, dstates, n, method, dout,
                                                rt->seeds_table.ptr, rt->stream),
            "FOD seeding kernel");
    TGK_HIP(hipMemcpyAsync(states, dstates, (size_t)n * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, rt->stream),
            "state copy");
    return fetch(out, dout, (size_t)n * 8, rt->stream);
}

int tgk_seeds_prepare(tgk_runtime *rt, int32_t model, int32_t n_parameters, int32_t method)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (int r = check_seeds_args(model, n_parameters, method))
        return r;
    if (int r = select_device(rt))
        return r;
    if (int r = prepare_seeds(rt, model, n_parameters, method))
        return r;
    TGK_HIP(hipStreamSynchronize(rt->stream), "seeding table");
    return TGK_OK;
}

int tgk_seeds_from_fod_device(tgk_runtime *rt, int32_t model, int32_t n_parameters, const float *d_fod,
                              const float *d_voxels, uint32_t nnz, const float *d_affine, uint32_t *d_states,
                              int64_t n, int32_t method, float *d_out)
{
    if (!rt || n < 0 || (n > 0 && (!d_fod || !d_voxels || !d_affine || !d_states || !d_out)))
        return fail(TGK_ERR_INVALID, "bad arguments");
    if (int r = check_seeds_args(model, n_parameters, method))
        return r;
    if (n > 0 && nnz == 0)
        return fail(TGK_ERR_INVALID, "the FOD has no nonzero voxel");
    if (n == 0)
        return TGK_OK;
    if (int r = select_device(rt))
        return r;
    if (int r = prepare_seeds(rt, model, n_parameters, method))
        return r;
    TGK_HIP(tgk::launch_seeds_from_fod_prepared(model, n_parameters, d_fod, d_voxels, nnz, d_affine, d_states, n,
                                                method, d_out, rt->seeds_table.ptr, rt->stream),
            "FOD seeding kernel");
    return TGK_OK;
}

int tgk_track_prepare(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field,
                      const tgk_algorithm *algorithm, int32_t path, int64_t max_seeds)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (rt->capturing)
        return fail(TGK_ERR_INVALID, "prepare before recording a graph");
    if (int r = check_track_args(domain, field, algorithm, path, max_seeds, 1, 1, 1.0f))
        return r;
    if (int r = check_owner(rt, domain, field))
        return r;
    if (int r = select_device(rt))
        return r;
    if (int r = ensure_workspace(rt, tgk::offsets_workspace_bytes(max_seeds)))
        return r;
    if (path == TGK_PATH_FAST) {
        const tgk::TrackLaunch p = make_launch(domain, field, algorithm, nullptr, 0, 0, nullptr, 1, 1, 1.0f, nullptr,
                                               nullptr, nullptr, nullptr);
        tgk::FastLaunch f;
        if (int r = fast_launch(rt, const_cast<tgk_field *>(field), algorithm, p, &f))
            return r;
    }
    TGK_HIP(hipStreamSynchronize(rt->stream), "track preparation");
    return TGK_OK;
}

int tgk_graph_begin(tgk_runtime *rt)
{
    if (!rt)
        return fail(TGK_ERR_INVALID, "runtime is NULL");
    if (rt->capturing)
        return fail(TGK_ERR_INVALID, "a graph is already being recorded on this runtime");
    if (int r = select_device(rt))
        return r;
    TGK_HIP(hipStreamBeginCapture(rt->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    rt->capturing = true;
    return TGK_OK;
}

int tgk_graph_end(tgk_runtime *rt, tgk_graph **out)
{
    if (!rt || !out)
        return fail(TGK_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (!rt->capturing)
        return fail(TGK_ERR_INVALID, "no graph is being recorded on this runtime");
    rt->capturing = false;
    hipGraph_t graph = nullptr;
    TGK_HIP(hipStreamEndCapture(rt->stream, &graph), "hipStreamEndCapture");
    std::unique_ptr<tgk_graph> g(new (std::nothrow) tgk_graph());
    if (!g) {
        (void)hipGraphDestroy(graph);
        return fail(TGK_ERR_NO_MEMORY, "graph allocation");
    }
    g->rt = rt;
    g->graph = graph;
    const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(e, "hipGraphInstantiate");
    }
    *out = g.release();
    return TGK_OK;
}

int tgk_graph_launch(tgk_graph *graph)
{
    if (!graph)
        return fail(TGK_ERR_INVALID, "graph is NULL");
    if (int r = select_device(graph->rt))
        return r;
    TGK_HIP(hipGraphLaunch(graph->exec, graph->rt->stream), "hipGraphLaunch");
    return TGK_OK;
}

int tgk_graph_destroy(tgk_graph *graph)
{
    if (!graph)
        return TGK_OK;
    if (graph->exec)
        (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph)
        (void)hipGraphDestroy(graph->graph);
    delete graph;
    return TGK_OK;
}

}  // extern "C"

namespace tgk {

// The error path of host code in other translation units (same thread-local
// message as the C-ABI layer's).
int report_error(int code, const char *fmt, ...)
{
    char buffer[512];
    va_list args;
    va_start(args, fmt);
    vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    g_last_error = buffer;
    return code;
}

}  // namespace tgk
