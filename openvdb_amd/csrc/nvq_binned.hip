// Large-batch point sampling: spatial binning + LDS-staged voxel bricks.
//
// Random query points touch a dozen distinct cache lines each over a grid that
// does not fit in L2, so a lane-per-query kernel is bound by line traffic from
// the last-level cache.  For large batches the queries are first sorted into
// bins of 16^3 voxels (over the grid's active bounding box grown by 2 voxels;
// points outside are clamped into the border bins):
//   k_bin_keys     : bin key per point and the point itself {index-space xyz (after the FP64
//                    world->index map when asked), original index}
//   radix sort     : (key, point) pairs, rocPRIM/hipCUB DeviceRadixSort; the points travel
//                    with their keys, so the per-bin pass reads them coalesced instead of
//                    gathering 12 bytes per 128-byte cache line
//   k_bin_bounds   : [begin, end) of every non-empty bin in the sorted order
//   k_bin_sample   : one workgroup per bin stages the 20^3 brick of voxel values
//                    around the bin in LDS (ReadAccessor::getValue semantics, so
//                    results are bit-identical) and evaluates the bin's queries
//                    from LDS; each query's results go to its original index as
//                    one aligned 32-byte record {nn, val, grad xyz}, a whole memory
//                    sector, instead of three partial-sector writes to the outputs.
//   k_sample_list  : queries whose stencil leaves the brick (border bins) run
//                    through the lane accessor, into the same records.
//   k_unpack       : records -> the requested outputs, coalesced.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "nvq_sample.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>

namespace nvq {
namespace {

constexpr int kThreads = 256;
constexpr int kBinThreads = 1024;       // one bin per workgroup: 16 waves share the 32 KB brick
constexpr int kBinLog2 = 4;
constexpr int kBin = 1 << kBinLog2;     // bin edge in voxels
constexpr int kBrick = kBin + 4;        // brick covers [bin-1, bin+kBin+2]
constexpr int kBrickVox = kBrick * kBrick * kBrick;
constexpr uint32_t kMaxBins = 1u << 22;

struct BinLayout {
    int32_t lo[3];   // index-space origin of bin (0,0,0)
    int32_t nb[3];   // bins per axis
    uint32_t nbins;
    uint32_t bits;   // key bits for the radix sort
};

struct alignas(16) Pt {
    float x, y, z; // index space
    uint32_t id;   // original position in the batch
};

struct alignas(32) Rec {
    float4 a; // nn, val, grad x, grad y
    float4 b; // grad z, unused
};

__device__ __forceinline__ void storeRec(Rec* __restrict__ rec, uint32_t id, float nn, float val, const float g[3])
{
    Rec r;
    r.a = make_float4(nn, val, g[0], g[1]);
    r.b = make_float4(g[2], 0.0f, 0.0f, 0.0f);
    rec[id] = r;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t m, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = v;
}

__device__ __forceinline__ void loadPoint(const GridView& g, const float* xyz, uint64_t i, int world, float p[3])
{
    p[0] = xyz[3 * i];
    p[1] = xyz[3 * i + 1];
    p[2] = xyz[3 * i + 2];
    if (world) worldToIndexF64(g, p);
}

__device__ __forceinline__ uint32_t binOf(const BinLayout& L, const float p[3])
{
    uint32_t b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // NaN and huge coordinates are clamped into the border bins: the key is only a locality hint
        const float rel = floorf(p[a]) - float(L.lo[a]);
        b[a] = (rel > 0.0f) ? (rel < float(L.nb[a] << kBinLog2) ? uint32_t(int32_t(rel)) >> kBinLog2 : uint32_t(L.nb[a] - 1))
                            : 0u;
    }
    return (b[0] * uint32_t(L.nb[1]) + b[1]) * uint32_t(L.nb[2]) + b[2];
}

__global__ __launch_bounds__(kThreads) void k_bin_keys(GridView g, BinLayout L, const float* __restrict__ xyz,
                                                       uint32_t n, int world, uint32_t* __restrict__ keys,
                                                       Pt* __restrict__ pts)
{
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        float p[3];
        loadPoint(g, xyz, i, world, p);
        keys[i] = binOf(L, p);
        Pt t;
        t.x = p[0];
        t.y = p[1];
        t.z = p[2];
        t.id = i;
        pts[i] = t;
    }
}

__global__ __launch_bounds__(kThreads) void k_bin_bounds(const uint32_t* __restrict__ keys, uint32_t n,
                                                         uint32_t* __restrict__ begin, uint32_t* __restrict__ end)
{
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const uint32_t k = keys[i];
        if (i == 0 || keys[i - 1] != k) begin[k] = i;
        if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
    }
}

template<uint32_t GT>
__global__ __launch_bounds__(kBinThreads) void k_bin_sample(GridView g, BinLayout L, const Pt* __restrict__ sorted,
                                                         const uint32_t* __restrict__ begin,
                                                         const uint32_t* __restrict__ end, Rec* __restrict__ rec,
                                                         bool wantNN, bool wantVal, int mode,
                                                         SampleConsts k, uint32_t* __restrict__ spill,
                                                         uint32_t* __restrict__ spillCount)
{
    __shared__ float brick[kBrickVox];
    __shared__ const uint8_t* slotLeaf[64]; // leaf node of each 8^3 slot the brick overlaps, or null
    __shared__ float slotValue[64];         // the slot's tile or background value when it has no leaf
    const uint32_t b = blockIdx.x;
    const uint32_t q0 = begin[b], q1 = end[b];
    if (q0 == q1) return;
    const uint32_t bz = b % uint32_t(L.nb[2]), by = (b / uint32_t(L.nb[2])) % uint32_t(L.nb[1]),
                   bx = b / (uint32_t(L.nb[2]) * uint32_t(L.nb[1]));
    const int32_t o[3] = {L.lo[0] + (int32_t(bx) << kBinLog2) - 1, L.lo[1] + (int32_t(by) << kBinLog2) - 1,
                          L.lo[2] + (int32_t(bz) << kBinLog2) - 1};
    // The brick overlaps at most 4x4x4 aligned 8^3 slots; inside one slot every voxel is read from
    // one leaf or, without a leaf, equals one tile or background value.  One tree walk per slot
    // (64 lanes in parallel), then every brick voxel is an independent load (ReadAccessor::getValue
    // semantics, NanoVDB.h:5370-5377).
    const int32_t s0[3] = {o[0] >> 3, o[1] >> 3, o[2] >> 3};
    const int32_t ns[3] = {((o[0] + kBrick - 1) >> 3) - s0[0] + 1, ((o[1] + kBrick - 1) >> 3) - s0[1] + 1,
                           ((o[2] + kBrick - 1) >> 3) - s0[2] + 1};
    if (threadIdx.x < uint32_t(ns[0] * ns[1] * ns[2])) {
        const int32_t t = int32_t(threadIdx.x);
        const int32_t sx = t / (ns[1] * ns[2]), sy = (t / ns[2]) % ns[1], sz = t % ns[2];
        Accessor<GT> acc(g);
        const Probe pr = acc.probe((s0[0] + sx) << 3, (s0[1] + sy) << 3, (s0[2] + sz) << 3);
        slotLeaf[(sx * 4 + sy) * 4 + sz] = pr.leaf;
        slotValue[(sx * 4 + sy) * 4 + sz] = pr.value;
    }
    __syncthreads();
    for (int v = threadIdx.x; v < kBrickVox; v += kBinThreads) {
        const int32_t x = v / (kBrick * kBrick), y = (v / kBrick) % kBrick, z = v % kBrick;
        const int32_t gx = o[0] + x, gy = o[1] + y, gz = o[2] + z;
        const int slot = (((gx >> 3) - s0[0]) * 4 + ((gy >> 3) - s0[1])) * 4 + ((gz >> 3) - s0[2]);
        const uint8_t* leaf = slotLeaf[slot];
        brick[v] = leaf ? Codec<GT>::get(leaf, Accessor<GT>::leafOffset(gx, gy, gz), g.channel) : slotValue[slot];
    }
    __syncthreads();
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += kBinThreads) {
        const Pt t = sorted[q];
        const uint32_t id = t.id;
        const float p[3] = {t.x, t.y, t.z};
        const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
        const int32_t ci = int32_t(fx) - o[0], cj = int32_t(fy) - o[1], ck = int32_t(fz) - o[2];
        // every FD6 sample cell lies in [cell-1, cell+2] (p +- 1 rounds at most one voxel further up),
        // so the whole stencil is inside the brick iff the cell is in [1, kBrick-4]
        if (uint32_t(ci - 1) > uint32_t(kBrick - 5) || uint32_t(cj - 1) > uint32_t(kBrick - 5) ||
            uint32_t(ck - 1) > uint32_t(kBrick - 5)) {
            spill[atomicAdd(spillCount, 1u)] = id;
            continue;
        }
        auto cell = [&](float qx, float qy, float qz) -> float {
            const float gx = floorf(qx), gy = floorf(qy), gz = floorf(qz);
            const int32_t base = ((int32_t(gx) - o[0]) * kBrick + (int32_t(gy) - o[1])) * kBrick + (int32_t(gz) - o[2]);
            Cell c;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int e = 0; e < 2; ++e) c.v[a][bb][e] = brick[base + (a * kBrick + bb) * kBrick + e];
            return trilerp(c, qx - gx, qy - gy, qz - gz);
        };
        const int32_t c0 = (ci * kBrick + cj) * kBrick + ck;
        float nnv = 0.0f, vv = 0.0f;
        if (wantNN) {
            const int32_t ri = int32_t(rintf(p[0])) - o[0], rj = int32_t(rintf(p[1])) - o[1], rk = int32_t(rintf(p[2])) - o[2];
            nnv = brick[(ri * kBrick + rj) * kBrick + rk];
        }
        if (wantVal) vv = cell(p[0], p[1], p[2]);
        float g3[3] = {0.0f, 0.0f, 0.0f};
        if (mode == 1) {
            g3[0] = (cell(p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f) - cell(p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f)) * k.fdx;
            g3[1] = (cell(p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f) - cell(p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f)) * k.fdy;
            g3[2] = (cell(p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f) - cell(p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f)) * k.fdz;
        } else if (mode == 2) {
            g3[0] = k.gsInv2dx * (brick[c0 + kBrick * kBrick] - brick[c0 - kBrick * kBrick]);
            g3[1] = k.gsInv2dx * (brick[c0 + kBrick] - brick[c0 - kBrick]);
            g3[2] = k.gsInv2dx * (brick[c0 + 1] - brick[c0 - 1]);
        } else if (mode == 3) {
            Cell C;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                    for (int e = 0; e < 2; ++e) C.v[a][bb][e] = brick[c0 + (a * kBrick + bb) * kBrick + e];
            trilinearGradient(C, p[0] - fx, p[1] - fy, p[2] - fz, g3);
        }
        storeRec(rec, id, nnv, vv, g3);
    }
}

template<uint32_t GT>
__global__ __launch_bounds__(kThreads) void k_sample_list(GridView g, const float* __restrict__ xyz, int world,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ count, Rec* __restrict__ rec,
                                                          bool wantNN, bool wantVal, int mode, SampleConsts k)
{
    Accessor<GT> acc(g);
    const uint32_t m = *count;
    for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < m; j += gridDim.x * kThreads) {
        const uint32_t id = list[j];
        float p[3];
        loadPoint(g, xyz, id, world, p);
        const SampleOut r = sampleEval(acc, p, wantNN, wantVal, mode, k);
        storeRec(rec, id, r.nn, r.val, r.g);
    }
}

__global__ __launch_bounds__(kThreads) void k_unpack(const Rec* __restrict__ rec, uint32_t n, float* __restrict__ nn,
                                                     float* __restrict__ val, float* __restrict__ grad)
{
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const float4 a = rec[i].a;
        if (nn) nn[i] = a.x;
        if (val) val[i] = a.y;
        if (grad) {
            const float gz = rec[i].b.x;
            grad[3 * uint64_t(i)] = a.z;
            grad[3 * uint64_t(i) + 1] = a.w;
            grad[3 * uint64_t(i) + 2] = gz;
        }
    }
}

template<uint32_t GT>
void launchSample(const GridView& g, const BinLayout& L, const float* xyz, int world, const Pt* sorted,
                  const uint32_t* begin, const uint32_t* end, Rec* rec, bool wantNN, bool wantVal, int mode,
                  uint32_t* spill, uint32_t* spillCount, hipStream_t s)
{
    const SampleConsts k = sampleConsts(g);
    {
        ProfScope prof("k_bin_sample", s);
        hipLaunchKernelGGL(k_bin_sample<GT>, dim3(L.nbins), dim3(kBinThreads), 0, s, g, L, sorted, begin, end, rec,
                           wantNN, wantVal, mode, k, spill, spillCount);
    }
    hipLaunchKernelGGL(k_sample_list<GT>, dim3(1024), dim3(kThreads), 0, s, g, xyz, world, spill, spillCount, rec,
                       wantNN, wantVal, mode, k);
}

// ------------------------------------------------------------------ adjoint
__device__ __forceinline__ SampleGrad loadGrad(const float* gnn, const float* gval, const float* ggrad, uint64_t i)
{
    SampleGrad d;
    d.gn = gnn ? gnn[i] : 0.0f;
    d.gv = gval ? gval[i] : 0.0f;
    d.gg[0] = ggrad ? ggrad[3 * i] : 0.0f;
    d.gg[1] = ggrad ? ggrad[3 * i + 1] : 0.0f;
    d.gg[2] = ggrad ? ggrad[3 * i + 2] : 0.0f;
    return d;
}

__device__ __forceinline__ bool gradZero(const SampleGrad& d)
{
    return d.gn == 0.0f && d.gv == 0.0f && d.gg[0] == 0.0f && d.gg[1] == 0.0f && d.gg[2] == 0.0f;
}

// Lane per query (small batches): the query's contributions go straight to the accumulator.
template<uint32_t GT>
__global__ __launch_bounds__(kThreads) void k_sample_bwd(GridView g, const float* __restrict__ xyz, uint64_t n,
                                                         int world, const float* __restrict__ gnn,
                                                         const float* __restrict__ gval,
                                                         const float* __restrict__ ggrad, int mode, SampleConsts k,
                                                         float scale, unsigned long long* __restrict__ accum,
                                                         uint32_t* __restrict__ overflow)
{
    Accessor<GT> acc(g);
    bool ok = true;
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
        const SampleGrad d = loadGrad(gnn, gval, ggrad, i);
        if (gradZero(d)) continue;
        float p[3];
        loadPoint(g, xyz, i, world, p);
        sampleAdjointOne(acc, g, p, d, mode, k, scale, accum, ok);
    }
    if (!ok) atomicOr(overflow, 1u);
}

// Queries the binned adjoint handed back (stencil outside the brick, or FD6 cells off the block).
template<uint32_t GT>
__global__ __launch_bounds__(kThreads) void k_sample_bwd_list(GridView g, const float* __restrict__ xyz, int world,
                                                              const uint32_t* __restrict__ list,
                                                              const uint32_t* __restrict__ count,
                                                              const float* __restrict__ gnn,
                                                              const float* __restrict__ gval,
                                                              const float* __restrict__ ggrad, int mode,
                                                              SampleConsts k, float scale,
                                                              unsigned long long* __restrict__ accum,
                                                              uint32_t* __restrict__ overflow)
{
    Accessor<GT> acc(g);
    bool ok = true;
    const uint32_t m = *count;
    for (uint32_t j = blockIdx.x * kThreads + threadIdx.x; j < m; j += gridDim.x * kThreads) {
        const uint32_t id = list[j];
        float p[3];
        loadPoint(g, xyz, id, world, p);
        sampleAdjointOne(acc, g, p, loadGrad(gnn, gval, ggrad, id), mode, k, scale, accum, ok);
    }
    if (!ok) atomicOr(overflow, 1u);
}

// One workgroup per bin: every query's fixed-point contributions are added into an int64 brick in
// LDS (the same 20^3 voxels the forward stages), then each non-zero brick voxel that holds a leaf
// value (Float) or a channel index (index grids) is added once into the accumulator.  Integer
// sums, so the result equals the lane kernel's bit for bit.
template<uint32_t GT>
__global__ __launch_bounds__(kBinThreads) void k_bin_sample_bwd(GridView g, BinLayout L, const Pt* __restrict__ sorted,
                                                             const uint32_t* __restrict__ begin,
                                                             const uint32_t* __restrict__ end,
                                                             const float* __restrict__ gnn,
                                                             const float* __restrict__ gval,
                                                             const float* __restrict__ ggrad, int mode,
                                                             SampleConsts k, float scale,
                                                             unsigned long long* __restrict__ accum,
                                                             uint32_t* __restrict__ spill,
                                                             uint32_t* __restrict__ spillCount,
                                                             uint32_t* __restrict__ overflow)
{
    __shared__ unsigned long long brick[kBrickVox]; // 64,000 B
    __shared__ const uint8_t* slotLeaf[64];
    __shared__ uint64_t slotIndex[64]; // index grids: the slot's tile or background index without a leaf
    __shared__ uint32_t bad;
    const uint32_t b = blockIdx.x;
    const uint32_t q0 = begin[b], q1 = end[b];
    if (q0 == q1) return;
    const uint32_t bz = b % uint32_t(L.nb[2]), by = (b / uint32_t(L.nb[2])) % uint32_t(L.nb[1]),
                   bx = b / (uint32_t(L.nb[2]) * uint32_t(L.nb[1]));
    const int32_t o[3] = {L.lo[0] + (int32_t(bx) << kBinLog2) - 1, L.lo[1] + (int32_t(by) << kBinLog2) - 1,
                          L.lo[2] + (int32_t(bz) << kBinLog2) - 1};
    const int32_t s0[3] = {o[0] >> 3, o[1] >> 3, o[2] >> 3};
    const int32_t ns[3] = {((o[0] + kBrick - 1) >> 3) - s0[0] + 1, ((o[1] + kBrick - 1) >> 3) - s0[1] + 1,
                           ((o[2] + kBrick - 1) >> 3) - s0[2] + 1};
    if (threadIdx.x < uint32_t(ns[0] * ns[1] * ns[2])) {
        const int32_t t = int32_t(threadIdx.x);
        const int32_t sx = t / (ns[1] * ns[2]), sy = (t / ns[2]) % ns[1], sz = t % ns[2];
        Accessor<GT> acc(g);
        const Probe pr = acc.probe((s0[0] + sx) << 3, (s0[1] + sy) << 3, (s0[2] + sz) << 3);
        slotLeaf[(sx * 4 + sy) * 4 + sz] = pr.leaf;
        slotIndex[(sx * 4 + sy) * 4 + sz] = GT == GT_ONINDEX ? pr.index : 0;
    }
    if (threadIdx.x == 0) bad = 0;
    for (int v = threadIdx.x; v < kBrickVox; v += kBinThreads) brick[v] = 0ull;
    __syncthreads();
    bool ok = true;
    for (uint32_t q = q0 + threadIdx.x; q < q1; q += kBinThreads) {
        const Pt t = sorted[q];
        const uint32_t id = t.id;
        const SampleGrad d = loadGrad(gnn, gval, ggrad, id);
        if (gradZero(d)) continue;
        const float p[3] = {t.x, t.y, t.z};
        const int32_t ci = int32_t(floorf(p[0])) - o[0], cj = int32_t(floorf(p[1])) - o[1],
                      ck = int32_t(floorf(p[2])) - o[2];
        // the block [cell-1, cell+2] must lie in the brick and the FD6 cells where the block expects them
        if (uint32_t(ci - 1) > uint32_t(kBrick - 5) || uint32_t(cj - 1) > uint32_t(kBrick - 5) ||
            uint32_t(ck - 1) > uint32_t(kBrick - 5) || !adjointInBlock(p, mode)) {
            spill[atomicAdd(spillCount, 1u)] = id;
            continue;
        }
        auto emit = [&](int32_t x, int32_t y, int32_t z, float sum) {
            const long long v = toFixedSat(sum * scale, ok);
            if (v != 0)
                atomicAdd(&brick[((x - o[0]) * kBrick + (y - o[1])) * kBrick + (z - o[2])], static_cast<unsigned long long>(v));
        };
        sampleAdjointBlock(p, d, mode, k, emit);
    }
    if (!ok) bad = 1;
    __syncthreads();
    for (int v = threadIdx.x; v < kBrickVox; v += kBinThreads) {
        const unsigned long long a = brick[v];
        if (a == 0ull) continue;
        const int32_t x = v / (kBrick * kBrick), y = (v / kBrick) % kBrick, z = v % kBrick;
        const int32_t gx = o[0] + x, gy = o[1] + y, gz = o[2] + z;
        const int slot = (((gx >> 3) - s0[0]) * 4 + ((gy >> 3) - s0[1])) * 4 + ((gz >> 3) - s0[2]);
        const uint8_t* leaf = slotLeaf[slot];
        const uint32_t n = Accessor<GT>::leafOffset(gx, gy, gz);
        int64_t key;
        if constexpr (GT == GT_ONINDEX) key = int64_t(leaf ? Codec<GT_ONINDEX>::index(leaf, n) : slotIndex[slot]);
        else key = leaf ? int64_t((uint64_t(leaf - g.leaf0) >> 5) / 67u) * 512 + n : -1;
        if (key >= 0) atomicAdd(accum + key, a);
    }
    if (threadIdx.x == 0 && bad) atomicOr(overflow, 1u);
}

bool makeLayout(const GridView& g, BinLayout& L)
{
    if (g.bbox[0] > g.bbox[3] || g.bbox[1] > g.bbox[4] || g.bbox[2] > g.bbox[5]) return false; // empty grid
    uint64_t total = 1;
    for (int a = 0; a < 3; ++a) {
        const int64_t lo = int64_t(g.bbox[a]) - 2, hi = int64_t(g.bbox[3 + a]) + 2;
        const int64_t nb = ((hi - lo) >> kBinLog2) + 1;
        if (nb > (int64_t(1) << 20)) return false;
        L.lo[a] = int32_t(lo);
        L.nb[a] = int32_t(nb);
        total *= uint64_t(nb);
    }
    if (total > kMaxBins) return false;
    L.nbins = uint32_t(total);
    L.bits = 1;
    while ((uint64_t(1) << L.bits) < total) ++L.bits;
    return true;
}

// Scratch of one binned call: keys[2n] | pts[2n] | begin[m] | end[m] | spillCount | flag | spill[n]
// | rec[n] (forward only) | sort temporaries, one stream-ordered allocation.
struct Bins {
    BinLayout L;
    uint32_t n = 0;
    char* base = nullptr;
    uint32_t *keys = nullptr, *keys2 = nullptr, *begin = nullptr, *end = nullptr, *spillCount = nullptr,
             *flag = nullptr, *spill = nullptr;
    Pt *idx = nullptr, *idx2 = nullptr;
    Rec* rec = nullptr;
};

// Allocate, key, sort and bound the bins of n points.  0, or -1 (no layout), -2 (allocation),
// -4 (sort: nothing was launched that reads the sorted arrays; the scratch is already freed).
int binPoints(const GridView& g, const float* xyz, uint32_t n, int world, bool withRecords, Bins& B, hipStream_t s)
{
    if (!makeLayout(g, B.L)) return -1;
    const uint32_t m = B.L.nbins;
    B.n = n;
    size_t sortBytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, static_cast<uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), static_cast<Pt*>(nullptr),
                                           static_cast<Pt*>(nullptr), n, 0, int(B.L.bits), s) != hipSuccess)
        return -4;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t oKeys = 0, oKeys2 = up(oKeys + 4ull * n), oIdx = up(oKeys2 + 4ull * n),
                 oIdx2 = up(oIdx + sizeof(Pt) * size_t(n)), oBegin = up(oIdx2 + sizeof(Pt) * size_t(n)),
                 oEnd = up(oBegin + 4ull * m), oCount = up(oEnd + 4ull * m), oSpill = up(oCount + 8),
                 oRec = up(oSpill + 4ull * n), oTemp = up(oRec + (withRecords ? sizeof(Rec) * size_t(n) : 0)),
                 total = oTemp + up(sortBytes);
    keepPoolMemory();
    if (hipMallocAsync(reinterpret_cast<void**>(&B.base), total, s) != hipSuccess) return -2;
    char* base = B.base;
    B.keys = reinterpret_cast<uint32_t*>(base + oKeys);
    B.keys2 = reinterpret_cast<uint32_t*>(base + oKeys2);
    B.idx = reinterpret_cast<Pt*>(base + oIdx);
    B.idx2 = reinterpret_cast<Pt*>(base + oIdx2);
    B.begin = reinterpret_cast<uint32_t*>(base + oBegin);
    B.end = reinterpret_cast<uint32_t*>(base + oEnd);
    B.spillCount = reinterpret_cast<uint32_t*>(base + oCount);
    B.flag = B.spillCount + 1;
    B.spill = reinterpret_cast<uint32_t*>(base + oSpill);
    B.rec = withRecords ? reinterpret_cast<Rec*>(base + oRec) : nullptr;
    void* temp = base + oTemp;
    const uint32_t gridN = std::min<uint32_t>((n + kThreads - 1) / kThreads, 1u << 16);
    const uint32_t zeroWords = uint32_t((oCount + 8) / 4 - oBegin / 4); // begin[], end[], spillCount and flag
    hipLaunchKernelGGL(k_fill_u32, dim3((zeroWords + kThreads - 1) / kThreads), dim3(kThreads), 0, s, B.begin,
                       zeroWords, 0u);
    hipLaunchKernelGGL(k_bin_keys, dim3(gridN), dim3(kThreads), 0, s, g, B.L, xyz, n, world, B.keys, B.idx);
    if (hipcub::DeviceRadixSort::SortPairs(temp, sortBytes, B.keys, B.keys2, B.idx, B.idx2, n, 0, int(B.L.bits), s) !=
        hipSuccess) {
        // the sorted arrays were never written: launch nothing that reads them
        (void)hipGetLastError();
        return hipFreeAsync(B.base, s) == hipSuccess ? -4 : -5;
    }
    hipLaunchKernelGGL(k_bin_bounds, dim3(gridN), dim3(kThreads), 0, s, B.keys2, n, B.begin, B.end);
    return 0;
}

template<uint32_t GT>
void launchSampleBwd(const GridView& g, const Bins& B, const float* xyz, int world, const float* gnn,
                     const float* gval, const float* ggrad, int mode, float scale, unsigned long long* accum,
                     hipStream_t s)
{
    const SampleConsts k = sampleConsts(g);
    {
        ProfScope prof("k_bin_sample_bwd", s);
        hipLaunchKernelGGL(k_bin_sample_bwd<GT>, dim3(B.L.nbins), dim3(kBinThreads), 0, s, g, B.L, B.idx2, B.begin,
                           B.end, gnn, gval, ggrad, mode, k, scale, accum, B.spill, B.spillCount, B.flag);
    }
    hipLaunchKernelGGL(k_sample_bwd_list<GT>, dim3(1024), dim3(kThreads), 0, s, g, xyz, world, B.spill, B.spillCount,
                       gnn, gval, ggrad, mode, k, scale, accum, B.flag);
}

} // namespace

bool sampleBinnedSupported(const GridView& g, uint64_t n)
{
    BinLayout L;
    return n >= (uint64_t(1) << 20) && n < (uint64_t(1) << 31) && makeLayout(g, L);
}

int sampleBinned(const GridView& g, const float* xyz, uint64_t n64, int world, float* nn, float* val, float* grad,
                 int mode, hipStream_t s)
{
    if (n64 >= (uint64_t(1) << 31)) return -1;
    Bins B;
    const int bs = binPoints(g, xyz, uint32_t(n64), world, true, B, s);
    if (bs) return bs;
    const uint32_t n = B.n;
    int st = 0;
    const bool wantNN = nn != nullptr, wantVal = val != nullptr;
    const int gm = grad ? mode : 0;
    switch (g.type) {
    case GT_FLOAT: launchSample<GT_FLOAT>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FP4: launchSample<GT_FP4>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FP8: launchSample<GT_FP8>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FP16: launchSample<GT_FP16>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FPN: launchSample<GT_FPN>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_ONINDEX: launchSample<GT_ONINDEX>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    default: st = -3;
    }
    if (st == 0)
        hipLaunchKernelGGL(k_unpack, dim3(std::min<uint32_t>((n + kThreads - 1) / kThreads, 1u << 14)), dim3(kThreads), 0,
                           s, B.rec, n, nn, val, grad);
    if (hipGetLastError() != hipSuccess) st = -4;
    if (hipFreeAsync(B.base, s) != hipSuccess) st = -5;
    return st;
}

int sampleBackward(const GridView& g, const float* xyz, uint64_t n64, int world, const float* gnn, const float* gval,
                   const float* ggrad, int mode, int64_t* accum, int shift, hipStream_t s)
{
    if (g.type != GT_FLOAT && g.type != GT_ONINDEX) return -3;
    const float scale = ldexpf(1.0f, shift);
    const int gm = ggrad ? mode : 0;
    auto* acc = reinterpret_cast<unsigned long long*>(accum);
    uint32_t over = 0;
    if (sampleBinnedSupported(g, n64) && !forceDirectSampling()) {
        Bins B;
        const int bs = binPoints(g, xyz, uint32_t(n64), world, false, B, s);
        if (bs) return bs;
        if (g.type == GT_FLOAT) launchSampleBwd<GT_FLOAT>(g, B, xyz, world, gnn, gval, ggrad, gm, scale, acc, s);
        else launchSampleBwd<GT_ONINDEX>(g, B, xyz, world, gnn, gval, ggrad, gm, scale, acc, s);
        bool ok = hipGetLastError() == hipSuccess;
        ok = hipMemcpyAsync(&over, B.flag, sizeof(over), hipMemcpyDeviceToHost, s) == hipSuccess && ok;
        ok = hipFreeAsync(B.base, s) == hipSuccess && ok;
        ok = hipStreamSynchronize(s) == hipSuccess && ok;
        if (!ok) return -4;
        return over ? -6 : 0;
    }
    uint32_t* flag = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&flag), sizeof(uint32_t), s) != hipSuccess) return -2;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kThreads), 0, s, flag, 1u, 0u);
    const SampleConsts k = sampleConsts(g);
    const uint32_t blocks = uint32_t(std::min<uint64_t>((n64 + kThreads - 1) / kThreads, 1u << 20));
    {
        ProfScope prof("k_sample_bwd", s);
        if (g.type == GT_FLOAT)
            hipLaunchKernelGGL(k_sample_bwd<GT_FLOAT>, dim3(blocks), dim3(kThreads), 0, s, g, xyz, n64, world, gnn, gval,
                               ggrad, gm, k, scale, acc, flag);
        else
            hipLaunchKernelGGL(k_sample_bwd<GT_ONINDEX>, dim3(blocks), dim3(kThreads), 0, s, g, xyz, n64, world, gnn,
                               gval, ggrad, gm, k, scale, acc, flag);
    }
    bool ok = hipGetLastError() == hipSuccess;
    ok = hipMemcpyAsync(&over, flag, sizeof(over), hipMemcpyDeviceToHost, s) == hipSuccess && ok;
    ok = hipFreeAsync(flag, s) == hipSuccess && ok;
    ok = hipStreamSynchronize(s) == hipSuccess && ok;
    if (!ok) return -4;
    return over ? -6 : 0;
}

} // namespace nvq
