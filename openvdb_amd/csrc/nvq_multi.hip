// Persistent multi-grid ray query: one launch answers the rays of up to kMaxGrids grids of
// mixed leaf codecs and classes.  Wavefronts pull 8x8 ray tiles from a device work queue
// (forEachQueuedTile: chunks interleaved over one counter per XCD), costliest tiles first (a
// per-tile estimate, counting-sorted before the launch; the no-allocation variant keeps grid
// and image order), until every grid's tiles are taken (the persistent-warp pattern of the reference's iso-surface renderer,
// examples/ex_raytrace_iso_surface/common.h:79-110, 166-189), so small and large grids share
// the machine without per-grid launches or tail effects.  Level-set
// grids run zeroCrossing, fog volumes run the emission-absorption march.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "nvq_rays.h"

#include <hip/hip_runtime.h>

#include <mutex>

// Work-queue counters of nvq_multi_query_noalloc (kQueueBands counters, kQueueStride words apart:
// 2 KiB).  Each device holds its own copy until the process ends; host code reaches it only
// through the device address hipGetSymbolAddress returns for that device.
static __device__ __align__(256) uint32_t scratch[nvq::kQueueBands * nvq::kQueueStride];

namespace nvq {
namespace {

template<uint32_t GT>
__device__ __forceinline__ void runRay(const GridView& g, uint32_t op, const float* eye, const float* dir, uint64_t i,
                                       float dt, float sigma, uint8_t* hit, int32_t* ijk, float* a, float* b,
                                       float* normal)
{
    if (op == kOpZeroCrossing) zeroCrossingRay<GT>(g, eye, dir, i, 0.0f, hit, ijk, a, b, normal);
    else fogMarchRay<GT>(g, eye, dir, i, dt, sigma, a, b);
}

__global__ __launch_bounds__(kBlock) void k_multi_query(MultiArgs A, const float* __restrict__ eye,
                                                        const float* __restrict__ dir, float dt, float sigma,
                                                        uint32_t* __restrict__ counters,
                                                        const uint32_t* __restrict__ order, uint8_t* __restrict__ hit,
                                                        int32_t* __restrict__ ijk, float* __restrict__ a,
                                                        float* __restrict__ b, float* __restrict__ normal)
{
    const uint32_t lane = threadIdx.x & 63u;
    forEachQueuedTile(counters, A.totalTiles, [&](uint32_t q) {
        const uint32_t t = order ? order[q] : q;
        uint32_t j = 0;
        while (j + 1 < A.k && A.slot[j + 1].tileBase <= t) ++j;
        const MultiSlot& S = A.slot[j];
        const uint32_t local = t - S.tileBase;
        const uint32_t tx = local % S.tilesX, ty = local / S.tilesX;
        const uint32_t px = tx * 8 + (lane & 7u), py = ty * 8 + (lane >> 3);
        const uint64_t i = S.rayBase + uint64_t(py) * S.width + px;
        // lanes outside the image or past a ragged last row idle; the wave stays together
        if (px < S.width && py < S.height && i < S.rayEnd) {
            const GridView g = *S.view;
            switch (g.type) {
            case GT_FLOAT: runRay<GT_FLOAT>(g, S.op, eye, dir, i, dt, sigma, hit, ijk, a, b, normal); break;
            case GT_FP4: runRay<GT_FP4>(g, S.op, eye, dir, i, dt, sigma, hit, ijk, a, b, normal); break;
            case GT_FP8: runRay<GT_FP8>(g, S.op, eye, dir, i, dt, sigma, hit, ijk, a, b, normal); break;
            case GT_FP16: runRay<GT_FP16>(g, S.op, eye, dir, i, dt, sigma, hit, ijk, a, b, normal); break;
            case GT_FPN: runRay<GT_FPN>(g, S.op, eye, dir, i, dt, sigma, hit, ijk, a, b, normal); break;
            default: runRay<GT_ONINDEX>(g, S.op, eye, dir, i, dt, sigma, hit, ijk, a, b, normal); break;
            }
        }
    });
}

// Cost of a tile of the multi-grid launch for the longest-first queue order: the longest of its
// four corner rays inside its grid's bbox, a sixteenth of that for zeroCrossing.  A march steps every half voxel
// with a trilinear sample per step; zeroCrossing moves by nodes and leaves to the surface and
// stops at the first crossing, at far fewer probes per unit length.
__global__ __launch_bounds__(kBlock) void k_multi_tile_cost(MultiArgs A, const float* __restrict__ eye,
                                                            const float* __restrict__ dir, uint32_t* __restrict__ key)
{
    // the tile, its slot and the slot's grid view are the same for the whole wavefront: scalar
    const uint32_t t = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63u;
    if (t >= A.totalTiles) return; // whole wavefronts leave together
    uint32_t j = 0;
    while (j + 1 < A.k && A.slot[j + 1].tileBase <= t) ++j;
    const MultiSlot& S = A.slot[j];
    const uint32_t local = t - S.tileBase;
    const uint32_t tx = local % S.tilesX, ty = local / S.tilesX;
    const uint32_t px = tx * 8 + (lane & 7u), py = ty * 8 + (lane >> 3);
    const uint64_t i = S.rayBase + uint64_t(py) * S.width + px;
    float c = 0.0f;
    if (cornerLane(lane) && px < S.width && py < S.height && i < S.rayEnd) {
        c = clippedLength(*S.view, eye + 3 * i, dir + 3 * i);
        if (S.op == kOpZeroCrossing) c *= 0.0625f;
    }
    c = waveMax(c);
    if (lane == 0) key[t] = costKey(c, t);
}

constexpr int kMaxDevices = 64;
std::mutex gScratchLock;
uint32_t* gScratch[kMaxDevices]; // device address of `scratch` on each device, set on first use

// Device address of `scratch` on device `dev`, which must be current: looked up once per device.
int scratchOn(int dev, uint32_t** out)
{
    if (dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lock(gScratchLock);
    if (!gScratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
        gScratch[dev] = static_cast<uint32_t*>(p);
    }
    *out = gScratch[dev];
    return 0;
}

uint32_t queryBlocks(const MultiArgs& A, int cus)
{
    const uint64_t waves = A.totalTiles; // one wave per tile is the most that can be busy
    uint32_t blocks = uint32_t(cus) * 8u;
    if (uint64_t(blocks) * (kBlock / 64) > waves) blocks = uint32_t((waves + (kBlock / 64) - 1) / (kBlock / 64));
    return blocks == 0 ? 1 : blocks;
}

} // namespace

int multiQueryNoAlloc(const MultiArgs& A, const float* eye, const float* dir, float dt, float sigma, uint8_t* hit,
                      int32_t* ijk, float* a, float* b, float* normal, hipStream_t s)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    uint32_t* counter = nullptr;
    if (scratchOn(dev, &counter) != 0) return -2;
    // the counters keep the previous call's values: clear them with a kernel (no memset node)
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(kQueueBands), 0, s, counter);
    hipLaunchKernelGGL(k_multi_query, dim3(queryBlocks(A, cus)), dim3(kBlock), 0, s, A, eye, dir, dt, sigma, counter,
                       nullptr, hit, ijk, a, b, normal);
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

int multiQuery(const MultiArgs& A, const float* eye, const float* dir, float dt, float sigma, uint8_t* hit,
               int32_t* ijk, float* a, float* b, float* normal, hipStream_t s)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    // queue counters, then the tile cost keys and the longest-first order (4 bytes per tile each;
    // none up to kOrderMinTiles tiles, which are all resident at once)
    const bool ordered = A.totalTiles > kOrderMinTiles;
    const size_t keyBytes = ordered ? (size_t(A.totalTiles) * 4 + 255) & ~size_t(255) : 0;
    const size_t countBytes = ordered ? orderScratchBytes(A.totalTiles) : 0;
    WorkspaceLease lease;
    char* base = lease.acquire(kQueueBytes + 2 * keyBytes + countBytes, s);
    if (!base) return -2;
    uint32_t* counter = reinterpret_cast<uint32_t*>(base);
    uint32_t* key = reinterpret_cast<uint32_t*>(base + kQueueBytes);
    uint32_t* order = ordered ? reinterpret_cast<uint32_t*>(base + kQueueBytes + keyBytes) : nullptr;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(kQueueBands), 0, s, counter);
    if (ordered) {
        ProfScope prof("tile_order", s);
        hipLaunchKernelGGL(k_multi_tile_cost, dim3((A.totalTiles + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0,
                           s, A, eye, dir, key);
        launchTileOrder(key, uint64_t(A.totalTiles), order,
                        reinterpret_cast<uint32_t*>(base + kQueueBytes + 2 * keyBytes), s);
    }
    {
        ProfScope prof("k_multi_query", s);
        hipLaunchKernelGGL(k_multi_query, dim3(queryBlocks(A, cus)), dim3(kBlock), 0, s, A, eye, dir, dt, sigma,
                           counter, order, hit, ijk, a, b, normal);
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace nvq
