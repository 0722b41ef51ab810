// Iso-surface renderer: one ray per pixel from a pinhole camera, isoCrossing through the tree,
// composited over a checkerboard (the RenderOp of examples/ex_raytrace_iso_surface/common.h:18-157).
// Two launch forms: one logical thread per pixel in 8x8 tiles per wavefront, or a persistent
// grid whose wavefronts pull 8x8 pixel tiles from one device counter until none is left (the
// reference pulls 32-pixel runs per warp, renderIsoSurfacePersistentKernel, common.h:166-189).
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "nvq_rays.h"

#include <hip/hip_runtime.h>

namespace nvq {
namespace {

// RenderOp::getIndexRay (common.h:130-145): x = i % width, y = i / width, u = (x + 0.5) / width,
// P = ((2u - 1) tan(fov/2) aspect, (2v - 1) tan(fov/2), -1), normalised as Vec3::normalize does
// (scaled by 1 / sqrtf(x*x + y*y + z*z), math/Math.h:1423-1427, 1472-1473).  tan(fov/2) is
// evaluated on the host, as the reference's host path does.
__device__ __forceinline__ void pixelRay(const IsoCamera& c, uint32_t x, uint32_t y, float d[3])
{
    const float u = (float(x) + 0.5f) / float(c.width);
    const float v = (float(y) + 0.5f) / float(c.height);
    const float aspect = float(c.width) / float(c.height);
    const float px = (2.f * u - 1.f) * c.tanHalfFov * aspect;
    const float py = (2.f * v - 1.f) * c.tanHalfFov;
    const float pz = -1.f;
    const float inv = 1.0f / sqrtf(px * px + py * py + pz * pz);
    d[0] = px * inv;
    d[1] = py * inv;
    d[2] = pz * inv;
}

template<uint32_t GT>
__device__ __forceinline__ void renderPixel(Accessor<GT>& acc, const GridView& g, const IsoCamera& c, uint32_t x,
                                            uint32_t y, float* __restrict__ image, uint8_t* __restrict__ hit,
                                            float* __restrict__ tout)
{
    float d[3];
    pixelRay(c, x, y, d);
    Ray ray = worldToIndexRay(g, c.origin, d);
    int32_t ijk[3];
    float v = 0.0f, t = 0.0f;
    const bool found = isoCrossingCore<GT>(acc, g, ray, c.iso, ijk, v, t);
    // RenderOp::composite (common.h:147-155): depth (t * dx) / (2 * bboxDimZ) over a checkerboard
    const float value = found ? (t * c.voxelSize) / (c.wbboxDimZ * 2) : 0.0f;
    const float alpha = found ? 1.0f : 0.0f;
    const uint32_t mask = 1u << 7;
    const float bg = ((x & mask) ^ (y & mask)) ? 1.0f : 0.5f;
    const uint64_t i = uint64_t(y) * uint32_t(c.width) + x;
    image[i] = alpha * value + (1.0f - alpha) * bg;
    if (hit) hit[i] = found ? 1 : 0;
    if (tout) tout[i] = found ? t : 0.0f;
}

template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_iso_render(GridView g, IsoCamera c, Sched s, float* __restrict__ image,
                                                       uint8_t* __restrict__ hit, float* __restrict__ tout)
{
    uint64_t i;
    if (!schedule(s, i)) return;
    Accessor<GT> acc(g);
    renderPixel<GT>(acc, g, c, uint32_t(i % uint32_t(c.width)), uint32_t(i / uint32_t(c.width)), image, hit, tout);
}

template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_iso_render_persistent(GridView g, IsoCamera c, uint32_t tilesX,
                                                                  uint32_t tiles, uint32_t* __restrict__ counters,
                                                                  float* __restrict__ image, uint8_t* __restrict__ hit,
                                                                  float* __restrict__ tout)
{
    const uint32_t lane = threadIdx.x & 63u;
    forEachQueuedTile(counters, tiles, [&](uint32_t t) {
        const uint32_t x = (t % tilesX) * 8 + (lane & 7u), y = (t / tilesX) * 8 + (lane >> 3);
        if (x < uint32_t(c.width) && y < uint32_t(c.height)) {
            Accessor<GT> acc(g);
            renderPixel<GT>(acc, g, c, x, y, image, hit, tout);
        }
    });
}

template<uint32_t GT>
int launchIso(const GridView& g, const IsoCamera& c, int persistent, float* image, uint8_t* hit, float* t,
              hipStream_t s)
{
    const uint64_t n = uint64_t(c.width) * uint32_t(c.height);
    if (!persistent) {
        const Sched sc = makeSched(n, uint32_t(c.width));
        ProfScope prof("k_iso_render", s);
        hipLaunchKernelGGL(k_iso_render<GT>, dim3(launchBlocks(sc)), dim3(kBlock), 0, s, g, c, sc, image, hit, t);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return -1;
    const uint32_t tilesX = (uint32_t(c.width) + 7) / 8, tiles = tilesX * ((uint32_t(c.height) + 7) / 8);
    WorkspaceLease lease;
    uint32_t* counter = reinterpret_cast<uint32_t*>(lease.acquire(kQueueBytes, s));
    if (!counter) return -2;
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(kQueueBands), 0, s, counter);
    uint32_t blocks = uint32_t(cus) * 8u;
    const uint32_t wavesNeeded = (tiles + (kBlock / 64) - 1) / (kBlock / 64);
    if (blocks > wavesNeeded) blocks = wavesNeeded;
    if (blocks == 0) blocks = 1;
    {
        ProfScope prof("k_iso_render_persistent", s);
        hipLaunchKernelGGL(k_iso_render_persistent<GT>, dim3(blocks), dim3(kBlock), 0, s, g, c, tilesX, tiles, counter,
                           image, hit, t);
    }
    return hipGetLastError() == hipSuccess ? 0 : -4;
}

} // namespace

int renderIso(const GridView& g, const IsoCamera& c, int persistent, float* image, uint8_t* hit, float* t,
              hipStream_t s)
{
    switch (g.type) {
    case GT_FLOAT: return launchIso<GT_FLOAT>(g, c, persistent, image, hit, t, s);
    case GT_FP4: return launchIso<GT_FP4>(g, c, persistent, image, hit, t, s);
    case GT_FP8: return launchIso<GT_FP8>(g, c, persistent, image, hit, t, s);
    case GT_FP16: return launchIso<GT_FP16>(g, c, persistent, image, hit, t, s);
    case GT_FPN: return launchIso<GT_FPN>(g, c, persistent, image, hit, t, s);
    case GT_ONINDEX: return launchIso<GT_ONINDEX>(g, c, persistent, image, hit, t, s);
    default: return -5;
    }
}

} // namespace nvq
