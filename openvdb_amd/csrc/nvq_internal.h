// Host-side entry points shared between the kernel translation units of libnvq.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nvq_device.h"
#include "../../include/nvq.h"

namespace nvq {

// ------------------------------------------------------------------ stream-ordered scratch
// Every stream-ordered allocation of the library comes from one memory pool per device that
// keeps up to kPoolKeepBytes of freed memory mapped for the next call, instead of returning it to
// the driver at each synchronisation; the process's default pool is left as the caller set it.
// Freed with hipFreeAsync.
constexpr uint64_t kPoolKeepBytes = uint64_t(32) << 30;
hipError_t poolMalloc(void** p, size_t bytes, hipStream_t s);

// Small device-to-host reads (counts, flags, a bbox) go through a pinned host buffer of
// kStageBytes per host thread, allocated on its first use and kept until the process ends.  A
// copy into pageable host memory is not queued: the runtime waits for it in the call, and a
// hipFreeAsync issued right after such a copy stalled the host for about 230 ms per call once
// some tens of multi-GiB point-query calls had run.
constexpr size_t kStageBytes = 256;
// Queues a copy of `bytes` <= kStageBytes (whole 32-bit words, 4-byte aligned) at device address
// src into the calling thread's pinned buffer and returns it (valid once the stream has drained),
// or nullptr.
void* stageCopy(const void* src, size_t bytes, hipStream_t s);
// stageCopy, wait for the stream, then copy into dst.
hipError_t readBack(void* dst, const void* src, size_t bytes, hipStream_t s);

// ------------------------------------------------------------------ ray-tile scheduling
// Rays of a row-major image are processed in 8x8 pixel tiles (one tile per 64-wide wavefront).
// Hardware dispatch is round-robin over the 8 XCDs (workgroup b runs on XCD b % 8), each with its
// own L2.  Workgroups are remapped so that the image is cut into bands of whole tile rows, 16
// bands per XCD (at least one row each), and band k runs on XCD k % 8: the tiles of a band share
// an L2, and every XCD gets a share of each part of the image.  With one contiguous band per
// XCD, the XCDs holding the rows through the middle of a volume worked on long after the others
// had run out of tiles.  Linear batches keep one contiguous range per XCD.
constexpr int kBlock = 256;

struct Sched {
    uint64_t n;
    uint32_t width, height, tilesX;
    uint64_t tiles;
    uint32_t blocks;  // logical blocks
    uint32_t perXcd;  // workgroups launched per XCD
    uint32_t group;   // consecutive logical blocks (a band) given to one XCD
};

inline Sched makeSched(uint64_t n, uint32_t width)
{
    Sched s{};
    s.n = n;
    s.width = width;
    if (width) {
        s.height = uint32_t(n / width);
        s.tilesX = (width + 7) / 8;
        s.tiles = uint64_t(s.tilesX) * ((s.height + 7) / 8);
        s.blocks = uint32_t((s.tiles + (kBlock / 64) - 1) / (kBlock / 64));
        const uint32_t row = (s.tilesX + (kBlock / 64) - 1) / (kBlock / 64); // workgroups per tile row
        const uint32_t rows = s.blocks / (8u * 16u) / row;
        s.group = (rows > 1 ? rows : 1) * row;
    } else {
        s.blocks = uint32_t((n + kBlock - 1) / kBlock);
        s.group = (s.blocks + 7) / 8;
        s.tiles = (n + 63) / 64;
    }
    if (s.group == 0) s.group = 1;
    s.perXcd = (s.blocks + 8 * s.group - 1) / (8 * s.group) * s.group;
    return s;
}

__device__ __forceinline__ bool schedule(const Sched& s, uint64_t& idx)
{
    const uint32_t b = blockIdx.x, x = b & 7u, q = b >> 3;
    const uint32_t lb = ((q / s.group) * 8u + x) * s.group + q % s.group;
    if (lb >= s.blocks) return false;
    if (s.width == 0) {
        idx = uint64_t(lb) * kBlock + threadIdx.x;
        return idx < s.n;
    }
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t tile = uint64_t(lb) * (kBlock / 64) + wave;
    if (tile >= s.tiles) return false;
    const uint32_t tx = uint32_t(tile % s.tilesX), ty = uint32_t(tile / s.tilesX);
    const uint32_t px = tx * 8 + (lane & 7u), py = ty * 8 + (lane >> 3);
    if (px >= s.width || py >= s.height) return false;
    idx = uint64_t(py) * s.width + px;
    return true;
}

inline uint32_t launchBlocks(const Sched& s) { return s.perXcd * 8u; }

// Ray `lane` of tile `tile`: 8x8 pixels of a row-major image, or 64 consecutive rays of a linear
// batch.
__device__ __forceinline__ bool tileRay(const Sched& s, uint64_t tile, uint32_t lane, uint64_t& idx)
{
    if (s.width == 0) {
        idx = tile * 64u + lane;
        return idx < s.n;
    }
    const uint32_t tx = uint32_t(tile % s.tilesX), ty = uint32_t(tile / s.tilesX);
    const uint32_t px = tx * 8 + (lane & 7u), py = ty * 8 + (lane >> 3);
    if (px >= s.width || py >= s.height) return false;
    idx = uint64_t(py) * s.width + px;
    return true;
}

// Longest-first tile schedule of the fog march and its adjoint.  A march costs about its ray's
// length inside the grid's bbox, and tiles of a camera image differ several-fold in cost (rays
// through the middle of a volume against rays that graze it).  Dispatched in image order, the
// costly tiles of the last rows started late and ran alone at the end of the launch.  Taking the
// tiles in descending order of their longest clipped ray instead, wavefront k of the launch runs
// tile order[k]: the launch ends with its cheapest tiles, and consecutive workgroups (one per
// XCD in turn) get tiles of equal cost.
// A launch of at most kOrderMinTiles tiles has every tile resident at once: it runs in image
// order (order == nullptr) without the two ordering kernels.
constexpr uint64_t kOrderMinTiles = 4096;
__device__ __forceinline__ bool scheduleOrdered(const Sched& s, const uint32_t* __restrict__ order, uint64_t& idx)
{
    const uint64_t k = uint64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    if (k >= s.tiles) return false;
    return tileRay(s, order ? order[k] : k, threadIdx.x & 63u, idx);
}
inline uint32_t orderedBlocks(const Sched& s) { return uint32_t((s.tiles + kBlock / 64 - 1) / (kBlock / 64)); }

// Cost keys of tiles: the top 12 bits of a positive float cost (exponent and 4 mantissa bits:
// 6 % classes), inverted so that key 0 is the costliest.  Tiles of cost 0 (every ray misses the
// bbox), often most of an image, are spread over the last 64 keys by tile index: the counting
// sort's atomics on one key serialise.  No positive cost reaches those keys (it would need to be
// below 2^-123).
constexpr uint32_t kOrderKeys = 4096, kZeroKeys = 64;
__device__ __forceinline__ uint32_t costKey(float cost, uint64_t tile)
{
    if (!(cost > 0.0f)) return kOrderKeys - kZeroKeys + uint32_t(tile % kZeroKeys);
    const uint32_t k = kOrderKeys - 1u - (__float_as_uint(cost) >> 19);
    return k < kOrderKeys - kZeroKeys ? k : kOrderKeys - kZeroKeys - 1u;
}

// Index-space length of a world ray inside the grid's bbox (the march's own map and clip); 0 when
// it misses.
__device__ __forceinline__ float clippedLength(const GridView& g, const float* e, const float* d)
{
    const float we[3] = {e[0], e[1], e[2]}, wd[3] = {d[0], d[1], d[2]};
    Ray R = worldToIndexRay(g, we, wd);
    return rayClip(R, g.bbox) && double(R.t1) <= 1e20 && R.t1 > R.t0 ? R.t1 - R.t0 : 0.0f;
}

// lanes 0, 7, 56 and 63: the corners of an 8x8 ray tile
__device__ __forceinline__ bool cornerLane(uint32_t lane) { return ((lane & 7u) == 0u || (lane & 7u) == 7u) && (lane < 8u || lane >= 56u); }

__device__ __forceinline__ float waveMax(float x)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x = fmaxf(x, __shfl_xor(x, m));
    return x;
}

constexpr int kOrderThreads = 1024, kOrderBatch = 16;
// The counting sort runs over up to kOrderParts contiguous parts of the tile range (at least
// kOrderPartMin tiles each), one workgroup each: part histograms in LDS, one workgroup turning
// them into per-part offsets, then each part placing its tiles with LDS atomics.  For 262K tiles
// it takes 30 us; one workgroup for the whole range took 161 us, its atomics on one key serialised.
constexpr uint32_t kOrderParts = 128, kOrderPartMin = 8192;
inline uint32_t orderParts(uint64_t tiles)
{
    const uint64_t p = (tiles + kOrderPartMin - 1) / kOrderPartMin;
    return uint32_t(p < 1 ? 1 : (p > kOrderParts ? kOrderParts : p));
}
// bytes of per-part counts the sort needs besides the keys and the order
inline size_t orderScratchBytes(uint64_t tiles) { return size_t(orderParts(tiles)) * kOrderKeys * 4; }

namespace {
// Part w = blockIdx.x covers tiles [w * per, min((w + 1) * per, tiles)); its key counts go to
// count[w * kOrderKeys + k].
__global__ __launch_bounds__(kOrderThreads) void k_order_hist(const uint32_t* __restrict__ key, uint64_t tiles,
                                                              uint64_t per, uint32_t* __restrict__ count)
{
    __shared__ uint32_t c[kOrderKeys];
    for (int k = threadIdx.x; k < int(kOrderKeys); k += kOrderThreads) c[k] = 0;
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * per, hi = lo + per < tiles ? lo + per : tiles;
    // kOrderBatch keys per thread are loaded before their atomics, so each thread keeps that many
    // loads in flight
    for (uint64_t t0 = lo; t0 < hi; t0 += uint64_t(kOrderThreads) * kOrderBatch) {
        uint32_t k[kOrderBatch];
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j) {
            const uint64_t t = t0 + uint64_t(j) * kOrderThreads + threadIdx.x;
            k[j] = t < hi ? key[t] : kOrderKeys;
        }
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j)
            if (k[j] < kOrderKeys) atomicAdd(&c[k[j]], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < int(kOrderKeys); k += kOrderThreads) count[uint64_t(blockIdx.x) * kOrderKeys + k] = c[k];
}

// One workgroup: the counts of `parts` parts become each part's first slot per key, in place (key
// order first, then part order within a key).  Thread x owns keys 4x..4x+3 and reads them as one
// 16-byte word per part, 16 parts in flight at a time.
__global__ __launch_bounds__(kOrderThreads) void k_order_offsets(uint32_t* __restrict__ count, uint32_t parts)
{
    static_assert(kOrderKeys == 4 * kOrderThreads, "four keys per thread");
    constexpr uint32_t kBatch = 16;
    __shared__ uint32_t part[kOrderThreads];
    uint4* rows = reinterpret_cast<uint4*>(count); // row w: kOrderThreads words of four keys
    uint32_t tot[4] = {0u, 0u, 0u, 0u};
    for (uint32_t w0 = 0; w0 < parts; w0 += kBatch) {
        uint4 c[kBatch];
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b)
            c[b] = w0 + b < parts ? rows[uint64_t(w0 + b) * kOrderThreads + threadIdx.x] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
            tot[0] += c[b].x;
            tot[1] += c[b].y;
            tot[2] += c[b].z;
            tot[3] += c[b].w;
        }
    }
    const uint32_t sum = tot[0] + tot[1] + tot[2] + tot[3];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kOrderThreads; off <<= 1) { // inclusive scan of the per-thread sums
        const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run[4];
    run[0] = part[threadIdx.x] - sum;
    run[1] = run[0] + tot[0];
    run[2] = run[1] + tot[1];
    run[3] = run[2] + tot[2];
    for (uint32_t w0 = 0; w0 < parts; w0 += kBatch) {
        uint4 c[kBatch];
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b)
            c[b] = w0 + b < parts ? rows[uint64_t(w0 + b) * kOrderThreads + threadIdx.x] : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
            if (w0 + b >= parts) continue;
            rows[uint64_t(w0 + b) * kOrderThreads + threadIdx.x] = make_uint4(run[0], run[1], run[2], run[3]);
            run[0] += c[b].x;
            run[1] += c[b].y;
            run[2] += c[b].z;
            run[3] += c[b].w;
        }
    }
}

// Part w places its tiles from its first slots, in the order its atomics hand them out.
__global__ __launch_bounds__(kOrderThreads) void k_order_place(const uint32_t* __restrict__ key, uint64_t tiles,
                                                               uint64_t per, const uint32_t* __restrict__ first,
                                                               uint32_t* __restrict__ order)
{
    __shared__ uint32_t slot[kOrderKeys];
    for (int k = threadIdx.x; k < int(kOrderKeys); k += kOrderThreads) slot[k] = first[uint64_t(blockIdx.x) * kOrderKeys + k];
    __syncthreads();
    const uint64_t lo = uint64_t(blockIdx.x) * per, hi = lo + per < tiles ? lo + per : tiles;
    for (uint64_t t0 = lo; t0 < hi; t0 += uint64_t(kOrderThreads) * kOrderBatch) {
        uint32_t k[kOrderBatch];
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j) {
            const uint64_t t = t0 + uint64_t(j) * kOrderThreads + threadIdx.x;
            k[j] = t < hi ? key[t] : kOrderKeys;
        }
#pragma unroll
        for (int j = 0; j < kOrderBatch; ++j)
            if (k[j] < kOrderKeys) order[atomicAdd(&slot[k[j]], 1u)] = uint32_t(t0 + uint64_t(j) * kOrderThreads + threadIdx.x);
    }
}
} // namespace

// Queues the counting sort of `tiles` keys (< kOrderKeys each) into `order`, with `count` of
// orderScratchBytes(tiles) bytes.
inline void launchTileOrder(const uint32_t* key, uint64_t tiles, uint32_t* order, uint32_t* count, hipStream_t s)
{
    const uint32_t parts = orderParts(tiles);
    const uint64_t per = (tiles + parts - 1) / parts;
    hipLaunchKernelGGL(k_order_hist, dim3(parts), dim3(kOrderThreads), 0, s, key, tiles, per, count);
    hipLaunchKernelGGL(k_order_offsets, dim3(1), dim3(kOrderThreads), 0, s, count, parts);
    hipLaunchKernelGGL(k_order_place, dim3(parts), dim3(kOrderThreads), 0, s, key, tiles, per, count, order);
}

// Per-call scratch of the query entry points is leased from one grow-only block per (device,
// stream) that stays allocated between calls: on this runtime hipFreeAsync waits for the stream's
// queued work, so a stream-ordered allocation and free per call made every call synchronous (and,
// for the multi-GiB scratch of binned point queries, stalled the host for about 230 ms per call
// once some tens of calls had run).  Calls on one stream execute in stream order, so they can
// share the block; the lease's lock serialises host threads that issue calls on the same stream.
// When a call needs more, the old block is freed in stream order and a larger one taken; the last
// block of each stream is kept until the process ends.
struct StreamWorkspace;
class WorkspaceLease {
public:
    WorkspaceLease() = default;
    WorkspaceLease(const WorkspaceLease&) = delete;
    WorkspaceLease& operator=(const WorkspaceLease&) = delete;
    ~WorkspaceLease();
    // The stream's block, at least `bytes` long, held until the lease ends; nullptr on failure.
    // One lease per call: a second acquire on the same lease fails.
    char* acquire(size_t bytes, hipStream_t s);

private:
    StreamWorkspace* w_ = nullptr;
};

// The longest-first order of a ray launch (nullptr up to kOrderMinTiles tiles), in the stream's
// workspace, with `extra` more bytes, 256-byte aligned, for the caller.
struct TileOrder {
    WorkspaceLease lease;
    const uint32_t* order = nullptr;
    char* extra = nullptr;
};
int tileOrder(const GridView& g, const float* eye, const float* dir, const Sched& s, TileOrder& out, hipStream_t st,
              size_t extraBytes = 0);


// Large-batch point sampling through spatial bins and LDS bricks (nvq_binned.hip).
bool sampleBinnedSupported(const GridView& g, uint64_t n);
int sampleBinned(const GridView& g, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
                 int mode, hipStream_t s);
// Adjoint of the point queries into an int64 fixed-point accumulator (binned for large batches,
// lane per query otherwise; bit-identical).  0, -2 allocation, -3 type, -4 HIP error,
// -6 a contribution saturated or was not finite.  Waits for s.
int sampleBackward(const GridView& g, const float* xyz, uint64_t n, int world, const float* gnn, const float* gval,
                   const float* ggrad, int mode, int64_t* accum, int shift, hipStream_t s);
// NVQ_SAMPLE_DIRECT=1 in the environment forces the lane-per-query kernels (read once).
bool forceDirectSampling();

// Persistent multi-grid ray queries (nvq_multi.hip)
constexpr uint32_t kMaxGrids = 32;
constexpr uint32_t kOpZeroCrossing = 0, kOpFogMarch = 1;
struct MultiSlot {
    const GridView* view; // device copy of the grid's view
    uint64_t rayBase;     // first ray of this grid in the concatenated arrays
    uint64_t rayEnd;      // one past its last ray
    uint32_t tileBase;    // first 8x8 tile of this grid in the global tile order
    uint32_t width, height, tilesX, op;
};
struct MultiArgs {
    MultiSlot slot[kMaxGrids];
    uint32_t k;
    uint32_t totalTiles;
};
int multiQuery(const MultiArgs& A, const float* eye, const float* dir, float dt, float sigma, uint8_t* hit,
               int32_t* ijk, float* a, float* b, float* normal, hipStream_t s);
// The same with the queue counters in the library's per-device array: no allocation.
int multiQueryNoAlloc(const MultiArgs& A, const float* eye, const float* dir, float dt, float sigma, uint8_t* hit,
                      int32_t* ijk, float* a, float* b, float* normal, hipStream_t s);
// Argument checks and tile layout of a multi-grid query (nvq_kernels.hip); NVQ_OK or a status.
int makeMultiArgs(const nvq_grid* const* host_grids, uint32_t k, const uint64_t* host_ray_offsets,
                  const uint32_t* host_image_width, const uint32_t* host_ops, const float* eye, const float* dir,
                  float dt, MultiArgs& A);

// ------------------------------------------------------------------ persistent work queue
// Tiles [0, total) are handed out in chunks of kQueueChunk through 8 counters, one per XCD
// (workgroup b runs on XCD b % 8), each 256 bytes from the next: the i-th grab on counter r
// yields chunk 8*i + r.  All XCDs therefore advance through the tile order together (the
// wavefronts in flight share one region of the grids) while the atomics are spread over 8
// addresses instead of serialising on one.  A wavefront starts on its XCD's counter and moves to
// the next once that counter's chunks pass the end; it returns after all 8, so every wavefront
// exits.
constexpr uint32_t kQueueBands = 8, kQueueStride = 64, kQueueChunk = 1; // stride in uint32 (256 B)
constexpr size_t kQueueBytes = size_t(kQueueBands) * kQueueStride * sizeof(uint32_t);

template<class F>
__device__ __forceinline__ void forEachQueuedTile(uint32_t* __restrict__ counters, uint32_t total, F&& f)
{
    const uint32_t lane = threadIdx.x & 63u, home = blockIdx.x & (kQueueBands - 1);
    const uint32_t chunks = (total + kQueueChunk - 1) / kQueueChunk;
    for (uint32_t k = 0; k < kQueueBands; ++k) {
        const uint32_t r = (home + k) & (kQueueBands - 1);
        for (;;) {
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(counters + r * kQueueStride, 1u);
            b = __shfl(b, 0, 64);
            const uint64_t q = uint64_t(b) * kQueueBands + r;
            if (q >= chunks) break; // counter r has handed out all its chunks
            const uint32_t t0 = uint32_t(q) * kQueueChunk, t1 = t0 + kQueueChunk < total ? t0 + kQueueChunk : total;
            for (uint32_t t = t0; t < t1; ++t) f(t);
        }
    }
}

namespace {
__global__ void k_queue_reset(uint32_t* counters) { counters[threadIdx.x * kQueueStride] = 0u; }
} // namespace

// Iso-surface renderer camera (nvq_iso_camera of nvq.h, field for field).
struct IsoCamera {
    int32_t width, height;
    float iso, voxelSize, wbboxDimZ, tanHalfFov;
    float origin[3];
};
int renderIso(const GridView& g, const IsoCamera& c, int persistent, float* image, uint8_t* hit, float* t,
              hipStream_t s);

// Adjoint of the fog march into an int64 fixed-point accumulator (nvq_adjoint.hip).
int fogMarchBackward(const GridView& g, const float* eye, const float* dir, uint64_t n, uint32_t width, float dt,
                     float sigma, const float* L, const float* T, const float* gL, const float* gT, int64_t* accum,
                     int shift, hipStream_t s);
int fixedToFloat(const int64_t* accum, uint64_t count, int shift, float* out, hipStream_t s);

// Handle for a grid buffer of `bytes` at device address d, its headers read from the host copy b
// (nvq_kernels.hip).  `owner` says how nvq_grid_free releases d.
enum GridOwner { kOwnerCaller = 0, kOwnerMalloc = 1, kOwnerPool = 2 };
int makeGridHandle(const uint8_t* b, uint64_t bytes, void* d, int owner, hipStream_t s, nvq_grid** out);

// Where a grid's node arrays lie in its device buffer (offsets relative to the tree, i.e. to
// base + 672), for the kernels that rewrite nodes in place.
struct GridLayout {
    uint8_t* base;
    int64_t nodeOff[4]; // leaf, lower, upper, root
    uint32_t count[3];  // leaf, lower, upper
    uint32_t tableSize, type;
};
GridLayout gridLayout(const nvq_grid* g);
int gridRefreshBBox(nvq_grid* g, hipStream_t s);

} // namespace nvq
