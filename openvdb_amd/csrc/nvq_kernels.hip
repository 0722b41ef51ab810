// gfx950 kernels and the C ABI of libnvq.so (declared in include/nvq.h).
//
// One lane per query / ray.  Every kernel templates on the leaf codec so the
// decode is branch-free in the inner loops.  Ray kernels schedule 8x8 pixel
// tiles per 64-wide wavefront and remap workgroups so that each of the eight
// XCDs (hardware dispatch is round-robin over XCDs) walks a contiguous band of
// tiles and keeps its own L2 warm with the nodes that band touches.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "nvq_sample.h"
#include "nvq_rays.h"
#include "../../include/nvq.h"

#include <cstdlib>

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <new>

using namespace nvq;

struct nvq_grid {
    void* dbuf;
    GridView* dview; // device copy of `view` (persistent multi-grid queries)
    uint64_t bytes;
    int device;
    int owner; // GridOwner: who frees dbuf (the caller for nvq_grid_upload_into)
    GridView view;
    nvq_grid_info info;
    const float* ownChannel; // index grids: the buffer's own float channel (blind data #0) or null
    uint64_t valueCount;     // index grids: GridData::mData1
    int64_t nodeOff[4];      // TreeData::mNodeOffset: leaf, lower, upper, root (from the tree)
};

nvq::GridLayout nvq::gridLayout(const nvq_grid* g)
{
    GridLayout L;
    L.base = static_cast<uint8_t*>(g->dbuf);
    for (int i = 0; i < 4; ++i) L.nodeOff[i] = g->nodeOff[i];
    for (int i = 0; i < 3; ++i) L.count[i] = g->info.node_count[i];
    L.tableSize = g->info.root_tiles;
    L.type = g->info.grid_type;
    return L;
}

// The device buffer's root bbox may have changed (statistics recomputed on the device): re-read
// it into the handle's view and info, on stream s, and wait for s.  The caller's in-place update
// already requires that no query on other streams reads the grid meanwhile.
int nvq::gridRefreshBBox(nvq_grid* g, hipStream_t s)
{
    int32_t box[6];
    if (readBack(box, static_cast<const uint8_t*>(g->dbuf) + kGridSize + g->nodeOff[3] + kRootBBox, sizeof(box), s) !=
        hipSuccess)
        return NVQ_ERR_HIP;
    for (int i = 0; i < 6; ++i) g->view.bbox[i] = g->info.index_bbox[i] = box[i];
    if (hipMemcpyAsync(g->dview, &g->view, sizeof(GridView), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return NVQ_ERR_HIP;
    return NVQ_OK;
}

// A grid can be queried when the kernels decode its type and its values are reachable: value grids
// always, index grids once they have a float channel.  Point grids (built on the device) are not.
static bool queryable(const nvq_grid* g)
{
    switch (g->view.type) {
    case GT_FLOAT:
    case GT_FP4:
    case GT_FP8:
    case GT_FP16:
    case GT_FPN: return true;
    case GT_ONINDEX: return g->view.channel != nullptr;
    default: return false;
    }
}

namespace {


inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ getValue
template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_get_value(GridView g, const int32_t* __restrict__ ijk, uint64_t n,
                                                      float* __restrict__ val, uint8_t* __restrict__ active,
                                                      uint32_t* __restrict__ dim)
{
    Accessor<GT> acc(g);
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        const Probe p = acc.probe(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
        if (val) val[i] = p.value;
        if (active) active[i] = p.active ? 1 : 0;
        if (dim) dim[i] = p.dim;
    }
}

// ------------------------------------------------------------------ point sampling
template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_sample(GridView g, const float* __restrict__ xyz, uint64_t n, int world,
                                                   float* __restrict__ nn, float* __restrict__ val,
                                                   float* __restrict__ grad, int mode, SampleConsts k)
{
    Accessor<GT> acc(g);
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        float p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        if (world) worldToIndexF64(g, p);
        sampleOne(acc, p, i, nn, val, grad, mode, k);
    }
}

// ------------------------------------------------------------------ GradStencil
__device__ __forceinline__ float sqf(float x) { return x * x; }

template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_grad_stencil(GridView g, const int32_t* __restrict__ ijk, uint64_t n,
                                                         float* __restrict__ grad, float* __restrict__ normsq,
                                                         float* __restrict__ lap, uint8_t* __restrict__ zc,
                                                         float inv2dx, float invdx2)
{
    Accessor<GT> acc(g);
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        const int32_t a = ijk[3 * i], b = ijk[3 * i + 1], c = ijk[3 * i + 2];
        float v[7];
        v[0] = acc.getValue(a, b, c);
        v[1] = acc.getValue(a - 1, b, c);
        v[2] = acc.getValue(a + 1, b, c);
        v[3] = acc.getValue(a, b - 1, c);
        v[4] = acc.getValue(a, b + 1, c);
        v[5] = acc.getValue(a, b, c - 1);
        v[6] = acc.getValue(a, b, c + 1);
        if (grad) {
            grad[3 * i] = inv2dx * (v[2] - v[1]);
            grad[3 * i + 1] = inv2dx * (v[4] - v[3]);
            grad[3 * i + 2] = inv2dx * (v[6] - v[5]);
        }
        if (normsq) {
            // GodunovsNormSqrd, math/Stencils.h:62-80
            const float xm = v[0] - v[1], xp = v[2] - v[0], ym = v[0] - v[3], yp = v[4] - v[0], zm = v[0] - v[5],
                        zp = v[6] - v[0];
            float d;
            if (v[0] > 0.0f) {
                d = fmaxf(sqf(fmaxf(xm, 0.0f)), sqf(fminf(xp, 0.0f)));
                d += fmaxf(sqf(fmaxf(ym, 0.0f)), sqf(fminf(yp, 0.0f)));
                d += fmaxf(sqf(fmaxf(zm, 0.0f)), sqf(fminf(zp, 0.0f)));
            } else {
                d = fmaxf(sqf(fminf(xm, 0.0f)), sqf(fmaxf(xp, 0.0f)));
                d += fmaxf(sqf(fminf(ym, 0.0f)), sqf(fmaxf(yp, 0.0f)));
                d += fmaxf(sqf(fminf(zm, 0.0f)), sqf(fmaxf(zp, 0.0f)));
            }
            normsq[i] = invdx2 * d;
        }
        if (lap) lap[i] = invdx2 * (v[1] + v[2] + v[3] + v[4] + v[5] + v[6] - 6.0f * v[0]);
        if (zc)
            zc[i] = (v[0] > 0.0f ? (v[1] < 0 || v[2] < 0 || v[3] < 0 || v[4] < 0 || v[5] < 0 || v[6] < 0)
                                 : (v[1] > 0 || v[2] > 0 || v[3] > 0 || v[4] > 0 || v[5] > 0 || v[6] > 0))
                        ? 1
                        : 0;
    }
}

// ------------------------------------------------------------------ zeroCrossing / fog march
template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_zero_crossing(GridView g, const float* __restrict__ eye,
                                                          const float* __restrict__ dir, Sched s, float iso,
                                                          uint8_t* __restrict__ hit, int32_t* __restrict__ ijk,
                                                          float* __restrict__ tout, float* __restrict__ vout,
                                                          float* __restrict__ normal)
{
    uint64_t i;
    if (!schedule(s, i)) return;
    zeroCrossingRay<GT>(g, eye, dir, i, iso, hit, ijk, tout, vout, normal);
}

// 75 VGPRs uncapped (6 wavefronts per SIMD).  Capped at 8 wavefronts (64 VGPRs, a 28-byte spill):
// 1024^2 rays through the 512^3 fog 163 -> 168 M rays/s, 4.19 M rays through the 256^3 fog
// 12.7 -> 11.2 ms; at 7 wavefronts 166 M rays/s and 11.6 ms.
template<uint32_t GT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_fog_march(GridView g, const float* __restrict__ eye,
                                                      const float* __restrict__ dir, Sched s,
                                                      const uint32_t* __restrict__ order, float dt, float sigma,
                                                      float* __restrict__ Lout, float* __restrict__ Tout)
{
    uint64_t i;
    if (!scheduleOrdered(s, order, i)) return;
    fogMarchRay<GT>(g, eye, dir, i, dt, sigma, Lout, Tout);
}

// ------------------------------------------------------------------ longest-first tile order
// Cost of a tile: the longest of its four corner rays inside the grid's bbox.
__global__ __launch_bounds__(kBlock) void k_tile_cost(GridView g, const float* __restrict__ eye,
                                                      const float* __restrict__ dir, Sched s,
                                                      uint32_t* __restrict__ key)
{
    const uint64_t tile = uint64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    if (tile >= s.tiles) return; // whole wavefronts leave together
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t i;
    float len = 0.0f;
    if (cornerLane(lane) && tileRay(s, tile, lane, i)) len = clippedLength(g, eye + 3 * i, dir + 3 * i);
    len = waveMax(len);
    if (lane == 0) key[tile] = costKey(len, tile);
}

// ------------------------------------------------------------------ host helpers
template<template<uint32_t> class K, typename... Args>
int launchTyped(uint32_t type, Args&&... args)
{
    switch (type) {
    case GT_FLOAT: K<GT_FLOAT>::run(args...); break;
    case GT_FP4: K<GT_FP4>::run(args...); break;
    case GT_FP8: K<GT_FP8>::run(args...); break;
    case GT_FP16: K<GT_FP16>::run(args...); break;
    case GT_FPN: K<GT_FPN>::run(args...); break;
    case GT_ONINDEX: K<GT_ONINDEX>::run(args...); break;
    default: return NVQ_ERR_UNSUPPORTED_TYPE;
    }
    return hipGetLastError() == hipSuccess ? NVQ_OK : NVQ_ERR_HIP;
}

uint32_t gridStrideBlocks(uint64_t n)
{
    const uint64_t b = (n + kBlock - 1) / kBlock;
    const uint64_t cap = 1u << 20;
    return uint32_t(b < 1 ? 1 : (b > cap ? cap : b));
}

template<uint32_t GT> struct RunGetValue {
    static void run(const GridView& g, const int32_t* ijk, uint64_t n, float* val, uint8_t* act, uint32_t* dim, hipStream_t s)
    {
        hipLaunchKernelGGL(k_get_value<GT>, dim3(gridStrideBlocks(n)), dim3(kBlock), 0, s, g, ijk, n, val, act, dim);
    }
};
template<uint32_t GT> struct RunSample {
    static void run(const GridView& g, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
                    int mode, hipStream_t s)
    {
        ProfScope prof("k_sample", s);
        hipLaunchKernelGGL(k_sample<GT>, dim3(gridStrideBlocks(n)), dim3(kBlock), 0, s, g, xyz, n, world, nn, val, grad,
                           mode, sampleConsts(g));
    }
};
template<uint32_t GT> struct RunGradStencil {
    static void run(const GridView& g, const int32_t* ijk, uint64_t n, float* grad, float* ns, float* lap, uint8_t* zc,
                    hipStream_t s)
    {
        const float inv2dx = float(0.5 / g.voxelSize[0]);
        const float invdx2 = float(4.0 * double(inv2dx) * double(inv2dx));
        hipLaunchKernelGGL(k_grad_stencil<GT>, dim3(gridStrideBlocks(n)), dim3(kBlock), 0, s, g, ijk, n, grad, ns, lap, zc,
                           inv2dx, invdx2);
    }
};
template<uint32_t GT> struct RunZeroCrossing {
    static void run(const GridView& g, const float* eye, const float* dir, const Sched& sc, float iso, uint8_t* hit,
                    int32_t* ijk, float* t, float* v, float* nrm, hipStream_t s)
    {
        ProfScope prof("k_zero_crossing", s);
        hipLaunchKernelGGL(k_zero_crossing<GT>, dim3(launchBlocks(sc)), dim3(kBlock), 0, s, g, eye, dir, sc, iso, hit,
                           ijk, t, v, nrm);
    }
};
template<uint32_t GT> struct RunFogMarch {
    static void run(const GridView& g, const float* eye, const float* dir, const Sched& sc, const uint32_t* order,
                    float dt, float sigma, float* L, float* T, hipStream_t s)
    {
        ProfScope prof("k_fog_march", s);
        hipLaunchKernelGGL(k_fog_march<GT>, dim3(orderedBlocks(sc)), dim3(kBlock), 0, s, g, eye, dir, sc, order, dt,
                           sigma, L, T);
    }
};

template<typename T> T rd(const uint8_t* p)
{
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}

uint32_t leafSizeOf(uint32_t type)
{
    switch (type) {
    case GT_FLOAT: return 2144u;
    case GT_FP4: return 352u;
    case GT_FP8: return 608u;
    case GT_FP16: return 1120u;
    case GT_ONINDEX: return 96u;
    default: return 0u;
    }
}

// Host-side validation of a NanoVDB buffer before upload (GridData::isValid, NanoVDB.h:2002-2013,
// plus bounds checks of every node offset the kernels follow).
// Byte offset of an index grid's float channel: blind data #0 when its type is Float and it holds
// at least GridData::mData1 (the value count) floats inside the buffer, as ChannelAccessor(grid, 0)
// takes it (NanoVDB.h:5737-5790, GridBlindMetaData :1591-1600); -1 when there is none.
int64_t indexChannelOffset(const uint8_t* b, uint64_t bytes)
{
    const int64_t metaOff = rd<int64_t>(b + 640);
    const uint32_t metaCount = rd<uint32_t>(b + 648);
    if (metaCount == 0 || metaOff < 0 || uint64_t(metaOff) + 288u > bytes) return -1;
    const uint8_t* m = b + metaOff;
    if (rd<uint32_t>(m + 28) != GT_FLOAT || rd<uint32_t>(m + 16) != 4u) return -1;
    const int64_t data = metaOff + rd<int64_t>(m);
    const uint64_t count = rd<uint64_t>(m + 8), need = rd<uint64_t>(b + 656);
    if (data < 0 || count < need || uint64_t(data) + 4u * count > bytes || (data & 3)) return -1;
    return data;
}

int validate(const uint8_t* b, uint64_t bytes)
{
    if (!b || bytes < kGridSize + 64 + kIdxRootSize) return b ? NVQ_ERR_TRUNCATED : NVQ_ERR_INVALID_ARG;
    if (reinterpret_cast<uintptr_t>(b) % 32u) return NVQ_ERR_MISALIGNED;
    const uint64_t magic = rd<uint64_t>(b) & 0x00FFFFFFFFFFFFFFull;
    if (magic != (0x314244566f6e614eull & 0x00FFFFFFFFFFFFFFull)) return NVQ_ERR_BAD_MAGIC;
    if ((rd<uint32_t>(b + 16) >> 21) != 32u) return NVQ_ERR_BAD_VERSION;
    const uint32_t type = rd<uint32_t>(b + kGridType);
    if (type != GT_FLOAT && type != GT_FP4 && type != GT_FP8 && type != GT_FP16 && type != GT_FPN && type != GT_ONINDEX)
        return NVQ_ERR_UNSUPPORTED_TYPE;
    const uint64_t gridSize = rd<uint64_t>(b + 32);
    if (gridSize > bytes) return NVQ_ERR_TRUNCATED;
    const uint8_t* tree = b + kGridSize;
    const int64_t offLeaf = rd<int64_t>(tree), offLower = rd<int64_t>(tree + 8), offUpper = rd<int64_t>(tree + 16),
                  offRoot = rd<int64_t>(tree + 24);
    const uint32_t nLeaf = rd<uint32_t>(tree + 32), nLower = rd<uint32_t>(tree + 36), nUpper = rd<uint32_t>(tree + 40);
    const uint64_t base = kGridSize;
    auto inside = [&](int64_t off, uint64_t len) {
        return off >= 0 && base + uint64_t(off) + len <= bytes;
    };
    const uint32_t rootSize = rootHeaderSize(type);
    if (!inside(offRoot, rootSize)) return NVQ_ERR_TRUNCATED;
    const uint32_t tableSize = rd<uint32_t>(tree + offRoot + kRootTableSize);
    if (!inside(offRoot, rootSize + uint64_t(kTileSize) * tableSize)) return NVQ_ERR_TRUNCATED;
    if (nUpper && !inside(offUpper, uint64_t(kUpSize) * nUpper)) return NVQ_ERR_TRUNCATED;
    if (nLower && !inside(offLower, uint64_t(kLoSize) * nLower)) return NVQ_ERR_TRUNCATED;
    const uint32_t ls = leafSizeOf(type);
    if (nLeaf && ls && !inside(offLeaf, uint64_t(ls) * nLeaf)) return NVQ_ERR_TRUNCATED;
    if (nLeaf && !ls && !inside(offLeaf, 96u)) return NVQ_ERR_TRUNCATED;
    return NVQ_OK;
}

} // namespace

hipError_t nvq::poolMalloc(void** p, size_t bytes, hipStream_t s)
{
    constexpr int kDevices = 64;
    static std::mutex mu;
    static hipMemPool_t pools[kDevices];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kDevices) return hipMallocAsync(p, bytes, s);
    hipMemPool_t pool = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!pools[dev]) {
            hipMemPoolProps props = {};
            props.allocType = hipMemAllocationTypePinned;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            hipMemPool_t made = nullptr;
            e = hipMemPoolCreate(&made, &props);
            if (e != hipSuccess) return e;
            uint64_t keep = kPoolKeepBytes;
            e = hipMemPoolSetAttribute(made, hipMemPoolAttrReleaseThreshold, &keep);
            if (e != hipSuccess) {
                (void)hipMemPoolDestroy(made);
                return e;
            }
            pools[dev] = made;
        }
        pool = pools[dev];
    }
    return hipMallocFromPoolAsync(p, bytes, pool, s);
}

namespace {
// copies `words` 32-bit words from device memory into mapped pinned host memory
__global__ void k_stage_words(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint32_t words)
{
    if (threadIdx.x < words) dst[threadIdx.x] = src[threadIdx.x];
}
} // namespace

void* nvq::stageCopy(const void* src, size_t bytes, hipStream_t s)
{
    thread_local void* stage = nullptr;
    if (bytes > kStageBytes || bytes % 4 || reinterpret_cast<uintptr_t>(src) % 4) return nullptr;
    if (!stage && hipHostMalloc(&stage, kStageBytes, hipHostMallocDefault) != hipSuccess) {
        stage = nullptr;
        return nullptr;
    }
    // a kernel store rather than a copy command: no copy is queued ahead of the caller's next
    // stream-ordered free
    hipLaunchKernelGGL(k_stage_words, dim3(1), dim3(kStageBytes / 4), 0, s, static_cast<uint32_t*>(stage),
                       static_cast<const uint32_t*>(src), uint32_t(bytes / 4));
    return hipGetLastError() == hipSuccess ? stage : nullptr;
}

hipError_t nvq::readBack(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    const void* st = stageCopy(src, bytes, s);
    if (!st) return hipErrorUnknown;
    const hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess) std::memcpy(dst, st, bytes);
    return e;
}

struct nvq::StreamWorkspace {
    std::mutex use;
    char* ptr = nullptr;
    size_t bytes = 0;
};

namespace {
std::mutex gWorkspaceLock;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<nvq::StreamWorkspace>> gWorkspaces;
} // namespace

nvq::WorkspaceLease::~WorkspaceLease()
{
    if (w_) w_->use.unlock();
}

char* nvq::WorkspaceLease::acquire(size_t bytes, hipStream_t s)
{
    int dev = 0;
    if (w_ || hipGetDevice(&dev) != hipSuccess) return nullptr;
    StreamWorkspace* w;
    {
        std::lock_guard<std::mutex> lock(gWorkspaceLock);
        std::unique_ptr<StreamWorkspace>& e = gWorkspaces[std::make_pair(dev, s)];
        if (!e) e = std::make_unique<StreamWorkspace>();
        w = e.get();
    }
    w->use.lock();
    w_ = w;
    if (w->bytes < bytes || !w->ptr) {
        if (w->ptr) (void)hipFreeAsync(w->ptr, s); // after the earlier calls' kernels on s
        w->ptr = nullptr;
        w->bytes = 0;
        const size_t want = std::max<size_t>(bytes + bytes / 8, size_t(1) << 20);
        void* p = nullptr;
        if (poolMalloc(&p, want, s) != hipSuccess) return nullptr;
        w->ptr = static_cast<char*>(p);
        w->bytes = want;
    }
    return w->ptr;
}

int nvq::tileOrder(const GridView& g, const float* eye, const float* dir, const Sched& s, TileOrder& out, hipStream_t st,
                   size_t extraBytes)
{
    if (s.tiles == 0 || s.tiles > (uint64_t(1) << 32)) return -1;
    const bool ordered = s.tiles > kOrderMinTiles;
    const size_t keyBytes = ordered ? (size_t(s.tiles) * 4 + 255) & ~size_t(255) : 0;
    const size_t orderBytes = keyBytes;
    const size_t countBytes = ordered ? orderScratchBytes(s.tiles) : 0;
    char* base = nullptr;
    if (keyBytes + orderBytes + countBytes + extraBytes) {
        base = out.lease.acquire(keyBytes + orderBytes + countBytes + extraBytes, st);
        if (!base) return -1;
    }
    out.extra = extraBytes ? base + keyBytes + orderBytes + countBytes : nullptr;
    if (!ordered) {
        out.order = nullptr;
        return 0;
    }
    uint32_t* key = reinterpret_cast<uint32_t*>(base);
    uint32_t* order = reinterpret_cast<uint32_t*>(base + keyBytes);
    {
        ProfScope prof("tile_order", st);
        hipLaunchKernelGGL(k_tile_cost, dim3(orderedBlocks(s)), dim3(kBlock), 0, st, g, eye, dir, s, key);
        launchTileOrder(key, s.tiles, order, reinterpret_cast<uint32_t*>(base + keyBytes + orderBytes), st);
    }
    out.order = order;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ======================================================================= C ABI
bool nvq::forceDirectSampling()
{
    static const bool force = std::getenv("NVQ_SAMPLE_DIRECT") && std::getenv("NVQ_SAMPLE_DIRECT")[0] == '1';
    return force;
}

extern "C" {

const char* nvq_status_string(int status)
{
    switch (status) {
    case NVQ_OK: return "ok";
    case NVQ_ERR_INVALID_ARG: return "invalid argument";
    case NVQ_ERR_BAD_MAGIC: return "not a NanoVDB grid buffer (bad magic number)";
    case NVQ_ERR_BAD_VERSION: return "unsupported NanoVDB major version (expected 32)";
    case NVQ_ERR_UNSUPPORTED_TYPE: return "unsupported grid type for this operation";
    case NVQ_ERR_MISALIGNED: return "grid buffer is not 32-byte aligned";
    case NVQ_ERR_TRUNCATED: return "grid buffer is truncated or has out-of-range node offsets";
    case NVQ_ERR_HIP: return "HIP runtime error";
    case NVQ_ERR_NO_DEVICE: return "no usable GPU device";
    case NVQ_ERR_OVERFLOW: return "a value exceeded the fixed-point range of the gradient accumulator (saturated)";
    default: return "unknown status";
    }
}

} // extern "C"

// Build the handle of a grid whose `bytes`-long buffer already is (or is being copied) at device
// address d.  `b` is a host copy of at least its grid, tree and root headers, or of the whole
// buffer when an index grid's blind channel must be located.  On failure nothing is freed.
int nvq::makeGridHandle(const uint8_t* b, uint64_t bytes, void* d, int owner, hipStream_t stream, nvq_grid** out)
{
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return NVQ_ERR_NO_DEVICE;
    nvq_grid* g = new (std::nothrow) nvq_grid;
    if (!g) return NVQ_ERR_HIP;
    g->owner = owner;
    g->dbuf = d;
    g->dview = nullptr;
    g->bytes = bytes;
    g->device = device;
    const uint8_t* tree = b + kGridSize;
    const int64_t offLeaf = rd<int64_t>(tree), offRoot = rd<int64_t>(tree + 24);
    for (int i = 0; i < 4; ++i) g->nodeOff[i] = rd<int64_t>(tree + 8 * i);
    const uint8_t* root = tree + offRoot;
    const uint8_t* dbase = static_cast<const uint8_t*>(d);
    GridView& v = g->view;
    std::memset(&v, 0, sizeof(v));
    v.root = dbase + kGridSize + offRoot;
    v.leaf0 = dbase + kGridSize + offLeaf;
    v.tableSize = rd<uint32_t>(root + kRootTableSize);
    v.type = rd<uint32_t>(b + kGridType);
    v.background = rd<float>(root + kRootBackground);
    for (int i = 0; i < 6; ++i) v.bbox[i] = rd<int32_t>(root + kRootBBox + 4 * i);
    for (int i = 0; i < 9; ++i) {
        v.invMatF[i] = rd<float>(b + kGridMap + kMapInvMatF + 4 * i);
        v.invMatD[i] = rd<double>(b + kGridMap + kMapInvMatD + 8 * i);
    }
    for (int i = 0; i < 3; ++i) {
        v.vecF[i] = rd<float>(b + kGridMap + kMapVecF + 4 * i);
        v.vecD[i] = rd<double>(b + kGridMap + kMapVecD + 8 * i);
        v.voxelSize[i] = rd<double>(b + kGridVoxelSize + 8 * i);
    }
    v.leafCount = rd<uint32_t>(tree + 32);
    v.leafSize = leafSizeOf(v.type);
    v.channel = nullptr;
    if (v.type == GT_ONINDEX) {
        const int64_t ch = indexChannelOffset(b, bytes);
        if (ch >= 0) v.channel = reinterpret_cast<const float*>(dbase + ch);
    }
    g->ownChannel = v.channel;
    g->valueCount = v.type == GT_ONINDEX ? rd<uint64_t>(b + 656) : 0;
    nvq_grid_info& in = g->info;
    std::memset(&in, 0, sizeof(in));
    in.grid_type = v.type;
    in.grid_class = rd<uint32_t>(b + kGridClass);
    in.bytes = bytes;
    for (int i = 0; i < 3; ++i) in.node_count[i] = rd<uint32_t>(tree + 32 + 4 * i);
    in.root_tiles = v.tableSize;
    for (int i = 0; i < 6; ++i) in.index_bbox[i] = v.bbox[i];
    for (int i = 0; i < 3; ++i) in.voxel_size[i] = v.voxelSize[i];
    in.background = v.background;
    in.device_buffer = d;
    if (poolMalloc(reinterpret_cast<void**>(&g->dview), sizeof(GridView), stream) != hipSuccess ||
        hipMemcpyAsync(g->dview, &g->view, sizeof(GridView), hipMemcpyHostToDevice, stream) != hipSuccess) {
        if (g->dview) (void)hipFreeAsync(g->dview, stream);
        delete g;
        return NVQ_ERR_HIP;
    }
    *out = g;
    return NVQ_OK;
}

extern "C" {

static int uploadImpl(const void* host_buf, uint64_t bytes, void* device_dst, void* stream, nvq_grid** out)
{
    if (!out) return NVQ_ERR_INVALID_ARG;
    *out = nullptr;
    const uint8_t* b = static_cast<const uint8_t*>(host_buf);
    const int st = validate(b, bytes);
    if (st != NVQ_OK) return st;
    void* d = device_dst;
    if (!d && hipMalloc(&d, bytes) != hipSuccess) return NVQ_ERR_HIP;
    int rc = NVQ_ERR_HIP;
    if (hipMemcpyAsync(d, host_buf, bytes, hipMemcpyHostToDevice, S(stream)) == hipSuccess)
        rc = makeGridHandle(b, bytes, d, device_dst ? kOwnerCaller : kOwnerMalloc, S(stream), out);
    if (rc != NVQ_OK && !device_dst) (void)hipFree(d);
    return rc;
}

int nvq_grid_upload(const void* host_buf, uint64_t bytes, void* stream, nvq_grid** out)
{
    return uploadImpl(host_buf, bytes, nullptr, stream, out);
}

int nvq_grid_upload_into(const void* host_buf, uint64_t bytes, void* device_dst, void* stream, nvq_grid** out)
{
    if (!device_dst || reinterpret_cast<uintptr_t>(device_dst) % 32u) return NVQ_ERR_INVALID_ARG;
    return uploadImpl(host_buf, bytes, device_dst, stream, out);
}

int nvq_grid_free(nvq_grid* grid)
{
    if (!grid) return NVQ_ERR_INVALID_ARG;
    const hipError_t e1 = hipDeviceSynchronize();
    hipError_t e2 = hipSuccess;
    if (grid->owner == kOwnerMalloc) e2 = hipFree(grid->dbuf);
    if (grid->owner == kOwnerPool) e2 = hipFreeAsync(grid->dbuf, nullptr); // the device is idle
    if (grid->dview) (void)hipFreeAsync(grid->dview, nullptr);
    delete grid;
    return (e1 == hipSuccess && e2 == hipSuccess) ? NVQ_OK : NVQ_ERR_HIP;
}

int nvq_grid_set_channel(nvq_grid* grid, const float* channel, uint64_t count)
{
    if (!grid || grid->view.type != GT_ONINDEX) return NVQ_ERR_INVALID_ARG;
    if (channel && count < grid->valueCount) return NVQ_ERR_INVALID_ARG;
    // queued work may still read the current view: wait for the device, then swap
    if (hipDeviceSynchronize() != hipSuccess) return NVQ_ERR_HIP;
    grid->view.channel = channel ? channel : grid->ownChannel;
    if (hipMemcpy(grid->dview, &grid->view, sizeof(GridView), hipMemcpyHostToDevice) != hipSuccess) return NVQ_ERR_HIP;
    return NVQ_OK;
}

int nvq_grid_set_ray_map(nvq_grid* grid, int mode)
{
    if (!grid || (mode != NVQ_RAY_MAP_F32 && mode != NVQ_RAY_MAP_F64)) return NVQ_ERR_INVALID_ARG;
    if (grid->view.rayMapF64 == uint32_t(mode)) return NVQ_OK;
    // queued work may still read the current view: wait for the device, then swap
    if (hipDeviceSynchronize() != hipSuccess) return NVQ_ERR_HIP;
    grid->view.rayMapF64 = uint32_t(mode);
    if (hipMemcpy(grid->dview, &grid->view, sizeof(GridView), hipMemcpyHostToDevice) != hipSuccess) return NVQ_ERR_HIP;
    return NVQ_OK;
}

int nvq_grid_get_info(const nvq_grid* grid, nvq_grid_info* info)
{
    if (!grid || !info) return NVQ_ERR_INVALID_ARG;
    *info = grid->info;
    return NVQ_OK;
}

int nvq_grid_get_view(const nvq_grid* grid, void* host_view, uint64_t view_bytes)
{
    if (!grid || !host_view || view_bytes != sizeof(GridView)) return NVQ_ERR_INVALID_ARG;
    std::memcpy(host_view, &grid->view, sizeof(GridView));
    return NVQ_OK;
}

int nvq_get_value(const nvq_grid* grid, const int32_t* ijk, uint64_t n, float* val, uint8_t* active, uint32_t* dim,
                  void* stream)
{
    if (!grid || !queryable(grid) || (n && !ijk)) return NVQ_ERR_INVALID_ARG;
    if (n == 0) return NVQ_OK;
    return launchTyped<RunGetValue>(grid->view.type, grid->view, ijk, n, val, active, dim, S(stream));
}

int nvq_sample(const nvq_grid* grid, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
               int grad_mode, void* stream)
{
    if (!grid || !queryable(grid) || (n && !xyz) || (grad && (grad_mode < 1 || grad_mode > 3))) return NVQ_ERR_INVALID_ARG;
    if (n == 0) return NVQ_OK;
    // Large batches go through spatial binning + LDS bricks (bit-identical results);
    // NVQ_SAMPLE_DIRECT=1 forces the lane-per-query kernel.
    if (!forceDirectSampling() && sampleBinnedSupported(grid->view, n)) {
        const int st = sampleBinned(grid->view, xyz, n, world, nn, val, grad, grad_mode, S(stream));
        if (st == 0) return NVQ_OK;
        if (st == -2) return NVQ_ERR_HIP; // scratch allocation failed
        return st == -3 ? NVQ_ERR_UNSUPPORTED_TYPE : NVQ_ERR_HIP;
    }
    return launchTyped<RunSample>(grid->view.type, grid->view, xyz, n, world, nn, val, grad, grad_mode, S(stream));
}

int nvq_grad_stencil(const nvq_grid* grid, const int32_t* ijk, uint64_t n, float* grad, float* norm_sq, float* lap,
                     uint8_t* zero_crossing, void* stream)
{
    if (!grid || !queryable(grid) || (n && !ijk)) return NVQ_ERR_INVALID_ARG;
    if (grid->view.type != GT_FLOAT) return NVQ_ERR_UNSUPPORTED_TYPE;
    if (n == 0) return NVQ_OK;
    return launchTyped<RunGradStencil>(grid->view.type, grid->view, ijk, n, grad, norm_sq, lap, zero_crossing, S(stream));
}

int nvq_zero_crossing(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                      uint8_t* hit, int32_t* ijk, float* t, float* v, float* normal, void* stream)
{
    if (!grid || !queryable(grid) || (n && (!eye || !dir))) return NVQ_ERR_INVALID_ARG;
    if (image_width && n % image_width) return NVQ_ERR_INVALID_ARG;
    if (n == 0) return NVQ_OK;
    if (n > (uint64_t(1) << 38)) return NVQ_ERR_INVALID_ARG;
    const Sched sc = makeSched(n, image_width);
    return launchTyped<RunZeroCrossing>(grid->view.type, grid->view, eye, dir, sc, 0.0f, hit, ijk, t, v, normal,
                                        S(stream));
}

int nvq_iso_crossing(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                     float iso, uint8_t* hit, int32_t* ijk, float* t, float* v, float* normal, void* stream)
{
    if (!grid || !queryable(grid) || (n && (!eye || !dir))) return NVQ_ERR_INVALID_ARG;
    if (image_width && n % image_width) return NVQ_ERR_INVALID_ARG;
    if (n == 0) return NVQ_OK;
    if (n > (uint64_t(1) << 38)) return NVQ_ERR_INVALID_ARG;
    const Sched sc = makeSched(n, image_width);
    return launchTyped<RunZeroCrossing>(grid->view.type, grid->view, eye, dir, sc, iso, hit, ijk, t, v, normal,
                                        S(stream));
}

int nvq_render_iso_surface(const nvq_grid* grid, const nvq_iso_camera* camera, int persistent, float* image,
                           uint8_t* hit, float* t, void* stream)
{
    if (!grid || !queryable(grid) || !camera || !image || camera->width <= 0 || camera->height <= 0) return NVQ_ERR_INVALID_ARG;
    if (uint64_t(camera->width) * uint64_t(camera->height) >= (uint64_t(1) << 31)) return NVQ_ERR_INVALID_ARG;
    static_assert(sizeof(nvq_iso_camera) == sizeof(IsoCamera), "camera layouts differ");
    IsoCamera c;
    std::memcpy(&c, camera, sizeof(c));
    return renderIso(grid->view, c, persistent, image, hit, t, S(stream)) == 0 ? NVQ_OK : NVQ_ERR_HIP;
}

int nvq_fog_march(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width, float dt,
                  float sigma_t, float* radiance, float* transmittance, void* stream)
{
    if (!grid || !queryable(grid) || (n && (!eye || !dir)) || !(dt > 0.0f)) return NVQ_ERR_INVALID_ARG;
    if (image_width && n % image_width) return NVQ_ERR_INVALID_ARG;
    if (n == 0) return NVQ_OK;
    if (n > (uint64_t(1) << 38)) return NVQ_ERR_INVALID_ARG;
    const Sched sc = makeSched(n, image_width);
    TileOrder to;
    if (tileOrder(grid->view, eye, dir, sc, to, S(stream)) != 0) return NVQ_ERR_HIP;
    return launchTyped<RunFogMarch>(grid->view.type, grid->view, eye, dir, sc, to.order, dt, sigma_t, radiance,
                                    transmittance, S(stream));
}

int nvq_fog_march_bwd(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                      float dt, float sigma_t, const float* radiance, const float* transmittance,
                      const float* dL_dradiance, const float* dL_dtransmittance, int64_t* grad_accum, int fixed_shift,
                      void* stream)
{
    if (!grid || !grad_accum || (n && (!eye || !dir || !radiance)) || !(dt > 0.0f) || fixed_shift < 0 ||
        fixed_shift > 62 || (dL_dtransmittance && !transmittance))
        return NVQ_ERR_INVALID_ARG;
    if (grid->view.type != GT_FLOAT && grid->view.type != GT_ONINDEX) return NVQ_ERR_UNSUPPORTED_TYPE;
    if (!queryable(grid)) return NVQ_ERR_INVALID_ARG;
    // keys are int32: leaf*512+n for Float grids, the channel index for index grids
    if (grid->view.type == GT_FLOAT && grid->info.node_count[0] > (1u << 22)) return NVQ_ERR_INVALID_ARG;
    if (grid->view.type == GT_ONINDEX && grid->valueCount >= (uint64_t(1) << 31)) return NVQ_ERR_INVALID_ARG;
    if (image_width && n % image_width) return NVQ_ERR_INVALID_ARG;
    if (n == 0 || (!dL_dradiance && !dL_dtransmittance)) return NVQ_OK;
    if (n > (uint64_t(1) << 38)) return NVQ_ERR_INVALID_ARG;
    const int st = fogMarchBackward(grid->view, eye, dir, n, image_width, dt, sigma_t, radiance, transmittance,
                                    dL_dradiance, dL_dtransmittance, grad_accum, fixed_shift, S(stream));
    return st == 0 ? NVQ_OK : (st == -2 ? NVQ_ERR_OVERFLOW : NVQ_ERR_HIP);
}

int nvq_sample_bwd(const nvq_grid* grid, const float* xyz, uint64_t n, int world, const float* dL_dnn,
                   const float* dL_dval, const float* dL_dgrad, int grad_mode, int64_t* grad_accum, int fixed_shift,
                   void* stream)
{
    if (!grid || !grad_accum || (n && !xyz) || fixed_shift < 0 || fixed_shift > 62 ||
        (dL_dgrad && (grad_mode < 1 || grad_mode > 3)))
        return NVQ_ERR_INVALID_ARG;
    if (grid->view.type != GT_FLOAT && grid->view.type != GT_ONINDEX) return NVQ_ERR_UNSUPPORTED_TYPE;
    if (!queryable(grid)) return NVQ_ERR_INVALID_ARG;
    if (n == 0 || (!dL_dnn && !dL_dval && !dL_dgrad)) return NVQ_OK;
    const int st = sampleBackward(grid->view, xyz, n, world, dL_dnn, dL_dval, dL_dgrad, grad_mode, grad_accum,
                                  fixed_shift, S(stream));
    switch (st) {
    case 0: return NVQ_OK;
    case -3: return NVQ_ERR_UNSUPPORTED_TYPE;
    case -6: return NVQ_ERR_OVERFLOW;
    default: return NVQ_ERR_HIP;
    }
}

int nvq_grad_fixed_to_float(const int64_t* accum, uint64_t count, int fixed_shift, float* out, void* stream)
{
    if ((count && (!accum || !out)) || fixed_shift < 0 || fixed_shift > 62) return NVQ_ERR_INVALID_ARG;
    if (count == 0) return NVQ_OK;
    return fixedToFloat(accum, count, fixed_shift, out, S(stream)) == 0 ? NVQ_OK : NVQ_ERR_HIP;
}

} // extern "C"

// Validate the arguments of a multi-grid query and lay out its grids' tiles (shared by
// nvq_multi_query and nvq_multi_query_noalloc).  Reads host memory only.
int nvq::makeMultiArgs(const nvq_grid* const* host_grids, uint32_t k, const uint64_t* host_ray_offsets,
                       const uint32_t* host_image_width, const uint32_t* host_ops, const float* eye, const float* dir,
                       float dt, MultiArgs& A)
{
    if (!host_grids || !host_ray_offsets || k == 0 || k > kMaxGrids || !(dt > 0.0f)) return NVQ_ERR_INVALID_ARG;
    A = MultiArgs{};
    A.k = k;
    uint64_t tiles = 0;
    for (uint32_t j = 0; j < k; ++j) {
        const nvq_grid* g = host_grids[j];
        const uint64_t n = host_ray_offsets[j + 1] - host_ray_offsets[j];
        if (!g || !queryable(g) || host_ray_offsets[j + 1] < host_ray_offsets[j]) return NVQ_ERR_INVALID_ARG;
        const uint32_t given = host_image_width ? host_image_width[j] : 0u;
        if (given && n % given) return NVQ_ERR_INVALID_ARG;
        const uint32_t w = given ? given : 64u; // linear batches: rows of 64 rays
        MultiSlot& S = A.slot[j];
        S.view = g->dview;
        S.rayBase = host_ray_offsets[j];
        S.rayEnd = host_ray_offsets[j + 1];
        S.width = w;
        S.height = uint32_t((n + w - 1) / w);
        S.tilesX = (w + 7) / 8;
        S.tileBase = uint32_t(tiles);
        if (host_ops) {
            if (host_ops[j] != NVQ_OP_ZERO_CROSSING && host_ops[j] != NVQ_OP_FOG_MARCH) return NVQ_ERR_INVALID_ARG;
            S.op = host_ops[j] == NVQ_OP_FOG_MARCH ? kOpFogMarch : kOpZeroCrossing;
        } else if (g->info.grid_class == 1u || g->info.grid_class == 2u) {
            S.op = g->info.grid_class == 2u ? kOpFogMarch : kOpZeroCrossing;
        } else {
            return NVQ_ERR_INVALID_ARG; // e.g. index grids: the class does not say which query
        }
        if (n && (!eye || !dir)) return NVQ_ERR_INVALID_ARG;
        tiles += uint64_t(S.tilesX) * ((S.height + 7) / 8);
    }
    if (tiles >= (uint64_t(1) << 31)) return NVQ_ERR_INVALID_ARG;
    A.totalTiles = uint32_t(tiles);
    return NVQ_OK;
}

extern "C" {

int nvq_multi_query(const nvq_grid* const* host_grids, uint32_t k, const uint64_t* host_ray_offsets,
                    const uint32_t* host_image_width, const uint32_t* host_ops, const float* eye, const float* dir,
                    float dt, float sigma_t, uint8_t* hit, int32_t* ijk, float* out_a, float* out_b, float* normal,
                    void* stream)
{
    MultiArgs A;
    const int st = makeMultiArgs(host_grids, k, host_ray_offsets, host_image_width, host_ops, eye, dir, dt, A);
    if (st != NVQ_OK || A.totalTiles == 0) return st;
    return multiQuery(A, eye, dir, dt, sigma_t, hit, ijk, out_a, out_b, normal, S(stream)) == 0 ? NVQ_OK : NVQ_ERR_HIP;
}

int nvq_multi_query_noalloc(const nvq_grid* const* host_grids, uint32_t k, const uint64_t* host_ray_offsets,
                            const uint32_t* host_image_width, const uint32_t* host_ops, const float* eye,
                            const float* dir, float dt, float sigma_t, uint8_t* hit, int32_t* ijk, float* out_a,
                            float* out_b, float* normal, void* stream)
{
    MultiArgs A;
    const int st = makeMultiArgs(host_grids, k, host_ray_offsets, host_image_width, host_ops, eye, dir, dt, A);
    if (st != NVQ_OK || A.totalTiles == 0) return st;
    return multiQueryNoAlloc(A, eye, dir, dt, sigma_t, hit, ijk, out_a, out_b, normal, S(stream)) == 0 ? NVQ_OK
                                                                                                        : NVQ_ERR_HIP;
}

} // extern "C"
