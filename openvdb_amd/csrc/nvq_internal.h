// Host-side entry points shared between the kernel translation units of libnvq.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nvq_device.h"
#include "../../include/nvq.h"

namespace nvq {

// ------------------------------------------------------------------ ray-tile scheduling
// Rays of a row-major image are processed in 8x8 pixel tiles (one tile per 64-wide wavefront);
// workgroups are remapped so each XCD (hardware dispatch is round-robin over the 8 XCDs) walks a
// contiguous band of tiles and keeps its own L2 warm with the nodes that band touches.
constexpr int kBlock = 256;

struct Sched {
    uint64_t n;
    uint32_t width, height, tilesX;
    uint64_t tiles;
    uint32_t blocks;  // logical blocks
    uint32_t perXcd;  // logical blocks per XCD
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
    } else {
        s.blocks = uint32_t((n + kBlock - 1) / kBlock);
    }
    s.perXcd = (s.blocks + 7) / 8;
    return s;
}

__device__ __forceinline__ bool schedule(const Sched& s, uint64_t& idx)
{
    const uint32_t b = blockIdx.x;
    const uint32_t lb = (b & 7u) * s.perXcd + (b >> 3);
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

// Keep freed stream-ordered allocations in the device's default pool instead of returning them
// to the driver at every synchronisation, so repeated large calls do not re-map memory.
inline void keepPoolMemory()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
    uint64_t thr = ~uint64_t(0);
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
}

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
