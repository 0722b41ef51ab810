// Device-side grid construction from integer voxels: voxelsToGrid<float | ValueOnIndex>
// (tools/cuda/PointsToGrid.cuh:119-125, :1417-1424).  The buffer written is the reference's,
// byte for byte (tile order, node order, masks, child offsets, bboxes, index offsets).
//
// The reference sorts 63-bit tile keys, run-length encodes them and then sorts the 36-bit
// in-tile keys segment by segment (countNodes, :619-826).  Here the voxels' upper-node tiles are
// numbered densely over the bbox of the input (x-major, which is the order of the reference's
// tile keys), so one radix sort of (tile << 36 | in-tile key) over 36 + log2(tiles) bits gives
// the final order.  The run heads at the four levels (voxel, leaf, lower, upper) are counted per
// chunk of 4096 sorted positions and the chunk totals scanned; each chunk then numbers its own
// positions with a workgroup scan while it writes the nodes that start there.  Each node is
// written by one thread (headers, child tables) or one 64-wide wavefront (leaf masks, bboxes): no
// per-voxel atomics, and no host round trip between the levels.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "../../include/nvq.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

using namespace nvq;

namespace {

constexpr int kB = 256;
constexpr uint32_t kWaves = kB / 64;
constexpr uint32_t kGridTypeIndex = 19; // GridType::Index (ValueIndex build type)
constexpr uint64_t kTreeOff = 672, kRootOff = 736;
constexpr uint64_t kUpperBytes = 270400, kLowerBytes = 33856;

// Inclusive counts of run heads at the voxel, leaf, lower and upper level.
struct Cnt {
    uint32_t v, l, m, u;
};
struct CntSum {
    __host__ __device__ Cnt operator()(const Cnt& a, const Cnt& b) const
    {
        return Cnt{a.v + b.v, a.l + b.l, a.m + b.m, a.u + b.u};
    }
};
// Position p of the sorted keys starts a new voxel / leaf / lower / upper when the key differs
// from its predecessor in the bits above 0 / 9 / 21 / 36 (32-bit keys hold one upper node).
template<typename K>
__host__ __device__ __forceinline__ Cnt headFlags(K k, K q)
{
    const uint64_t a = k, b = q;
    return Cnt{uint32_t(a != b), uint32_t((a >> 9) != (b >> 9)), uint32_t((a >> 21) != (b >> 21)),
               uint32_t((a >> 36) != (b >> 36))};
}

struct Layout {
    uint64_t offUpper, offLower, offLeaf, total;
    uint32_t rootHdr, leafSize, isIndex; // isIndex: ValueOnIndex or ValueIndex
    uint32_t allIndex;                   // ValueIndex: every leaf value has an index
    int32_t tlo[3];   // smallest tile coordinate (x >> 12) per axis
    uint32_t dim[3];  // tiles per axis in the bbox of the input
    int32_t llo[3];   // smallest lower-node coordinate (x >> 7) per axis
    uint32_t ldim[3]; // lower nodes per axis in the bbox of the input
    uint32_t isPoint; // Point grid: leaves hold point counts, the points follow as blind data
    uint64_t offMeta; // blind metadata (point grids) or the end of the leaves
};

// Point grids from float world points (pointsToGrid, tools/cuda/PointsToGrid.cuh:57-64).
struct PointArgs {
    const void* xyz;    // float or double (f64) world points, [n][3]
    uint32_t f64;
    uint32_t type;      // PointType 1..9 (NanoVDB.h:395-405)
    uint32_t size;      // bytes per encoded point
    uint32_t semantic, dataClass, dataType, gridClass;
    const char* name;   // blind-data name
    float invF;         // Map::mInvMatF diagonal
    double invD;        // Map::mInvMatD diagonal
};

// Upper-node tile coordinates t (x >> 12) and upper-node child offset n2 of a sort key.  64-bit
// keys: dense tile number << 36 | n2 << 21 | ...; 32-bit keys (the input lies in one tile):
// dense lower-node number over the input's bbox << 21 | ..., x-major, which is n2's order.
__device__ __forceinline__ void nodeOf(const Layout& L, uint64_t key, int32_t t[3], uint32_t& n2)
{
    const uint64_t tile = key >> 36;
    t[2] = int32_t(tile % L.dim[2]) + L.tlo[2];
    const uint64_t r = tile / L.dim[2];
    t[1] = int32_t(r % L.dim[1]) + L.tlo[1];
    t[0] = int32_t(r / L.dim[1]) + L.tlo[0];
    n2 = uint32_t(key >> 21) & 32767u;
}
__device__ __forceinline__ void nodeOf(const Layout& L, uint32_t key, int32_t t[3], uint32_t& n2)
{
    const uint32_t dl = key >> 21;
    const int32_t z7 = int32_t(dl % L.ldim[2]) + L.llo[2];
    const uint32_t r = dl / L.ldim[2];
    const int32_t y7 = int32_t(r % L.ldim[1]) + L.llo[1], x7 = int32_t(r / L.ldim[1]) + L.llo[0];
    t[0] = x7 >> 5;
    t[1] = y7 >> 5;
    t[2] = z7 >> 5;
    n2 = (uint32_t(x7 & 31) << 10) | (uint32_t(y7 & 31) << 5) | uint32_t(z7 & 31);
}

constexpr uint32_t kBoundBlocks = 1024;

__device__ __forceinline__ void waveMinMax(int32_t lo[3], int32_t hi[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], off, 64));
            hi[a] = max(hi[a], __shfl_xor(hi[a], off, 64));
        }
}

// Exact bbox of the input voxels (= RootData::mBBox of the result, processBBox :1386-1414): one
// partial bbox per workgroup, then one workgroup reduces the partials (no atomics).
__global__ __launch_bounds__(kB) void k_v2g_bounds(const int32_t* __restrict__ ijk, uint64_t n, int32_t* __restrict__ part)
{
    __shared__ int32_t sb[kB / 64][6];
    int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    const uint64_t stride = uint64_t(gridDim.x) * kB;
    for (uint64_t i = uint64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int32_t v = ijk[3 * i + a];
            lo[a] = min(lo[a], v);
            hi[a] = max(hi[a], v);
        }
    }
    waveMinMax(lo, hi);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0)
        for (int a = 0; a < 3; ++a) {
            sb[w][a] = lo[a];
            sb[w][3 + a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        int32_t v = sb[0][threadIdx.x];
        for (uint32_t j = 1; j < kB / 64; ++j) v = threadIdx.x < 3 ? min(v, sb[j][threadIdx.x]) : max(v, sb[j][threadIdx.x]);
        part[6 * blockIdx.x + threadIdx.x] = v;
    }
}

__global__ __launch_bounds__(kB) void k_v2g_bounds_final(const int32_t* __restrict__ part, uint32_t blocks,
                                                         int32_t* __restrict__ bnd)
{
    __shared__ int32_t sb[kB / 64][6];
    int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (uint32_t b = threadIdx.x; b < blocks; b += kB)
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], part[6 * b + a]);
            hi[a] = max(hi[a], part[6 * b + 3 + a]);
        }
    waveMinMax(lo, hi);
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0)
        for (int a = 0; a < 3; ++a) {
            sb[w][a] = lo[a];
            sb[w][3 + a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        int32_t v = sb[0][threadIdx.x];
        for (uint32_t j = 1; j < kB / 64; ++j) v = threadIdx.x < 3 ? min(v, sb[j][threadIdx.x]) : max(v, sb[j][threadIdx.x]);
        bnd[threadIdx.x] = v;
    }
}

// Sort key: dense tile number << 36 | upper offset << 21 | lower offset << 9 | leaf offset
// (InternalNode::CoordToOffset NanoVDB.h:3560-3566, LeafNode::CoordToOffset :4516-4519), or for
// 32-bit keys dense lower-node number << 21 | lower offset << 9 | leaf offset.
template<typename K>
__global__ __launch_bounds__(kB) void k_v2g_keys(const int32_t* __restrict__ ijk, uint32_t n, Layout L,
                                                 K* __restrict__ keys, uint32_t* __restrict__ ids)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const int32_t x = ijk[3ull * i], y = ijk[3ull * i + 1], z = ijk[3ull * i + 2];
    const uint64_t lo = (uint64_t((x & 127) >> 3) << 8) | (uint64_t((y & 127) >> 3) << 4) | uint64_t((z & 127) >> 3);
    const uint64_t lf = (uint64_t(x & 7) << 6) | (uint64_t(y & 7) << 3) | uint64_t(z & 7);
    if constexpr (sizeof(K) == 8) {
        const uint64_t tile = (uint64_t(uint32_t((x >> 12) - L.tlo[0])) * L.dim[1] + uint32_t((y >> 12) - L.tlo[1])) *
                                  L.dim[2] +
                              uint32_t((z >> 12) - L.tlo[2]);
        const uint64_t up = (uint64_t((x & 4095) >> 7) << 10) | (uint64_t((y & 4095) >> 7) << 5) | uint64_t((z & 4095) >> 7);
        keys[i] = (tile << 36) | (up << 21) | (lo << 9) | lf;
    } else {
        const uint32_t dl = (uint32_t((x >> 7) - L.llo[0]) * L.ldim[1] + uint32_t((y >> 7) - L.llo[1])) * L.ldim[2] +
                            uint32_t((z >> 7) - L.llo[2]);
        keys[i] = K((uint64_t(dl) << 21) | (lo << 9) | lf);
    }
    ids[i] = i;
}

constexpr uint32_t kItems = 16, kChunk = kB * kItems; // sorted positions per workgroup

// Run-head totals of each chunk of kChunk sorted positions.
template<typename K>
__global__ __launch_bounds__(kB) void k_v2g_count(const K* __restrict__ keys, uint32_t n, Cnt* __restrict__ chunkSums)
{
    using Reduce = hipcub::BlockReduce<Cnt, kB>;
    __shared__ typename Reduce::TempStorage tmp;
    const uint32_t base = blockIdx.x * kChunk;
    Cnt c{0u, 0u, 0u, 0u};
#pragma unroll 4
    for (uint32_t i = 0; i < kItems; ++i) {
        const uint32_t p = base + i * kB + threadIdx.x;
        if (p < n) c = CntSum()(c, p ? headFlags(keys[p], keys[p - 1]) : Cnt{1u, 1u, 1u, 1u});
    }
    const Cnt total = Reduce(tmp).Reduce(c, CntSum());
    if (threadIdx.x == 0) chunkSums[blockIdx.x] = total;
}

// One workgroup per chunk: each thread takes kItems consecutive sorted positions, numbers them
// (chunk prefix + block scan + running count) and, for each node that starts there, writes its
// root tile / parent child-table entry / child-mask bit (BuildUpperNodesFunctor :1034-1061,
// BuildLowerNodesFunctor :1085-1104, ProcessLeafMetaDataFunctor :1126-1154) and records where its
// children start.  Every voxel's value index goes to its input position.
template<typename K>
__global__ __launch_bounds__(kB) void k_v2g_heads(const K* __restrict__ keys, const uint32_t* __restrict__ ids,
                                                  const Cnt* __restrict__ chunkIncl, uint32_t n, Layout L,
                                                  uint8_t* __restrict__ buf, uint32_t* __restrict__ leafFirst,
                                                  uint32_t* __restrict__ leafRank, uint32_t* __restrict__ lowerFirst,
                                                  uint32_t* __restrict__ upperFirst, uint64_t* __restrict__ index)
{
    using Scan = hipcub::BlockScan<Cnt, kB>;
    __shared__ typename Scan::TempStorage tmp;
    const uint32_t p0 = blockIdx.x * kChunk + threadIdx.x * kItems;
    K k[kItems];
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) k[j] = p0 + j < n ? keys[p0 + j] : K(0);
    const K before = p0 > 0 && p0 < n ? keys[p0 - 1] : K(0);
    auto flags = [&](uint32_t j) {
        const uint32_t p = p0 + j;
        if (p >= n) return Cnt{0u, 0u, 0u, 0u};
        if (p == 0) return Cnt{1u, 1u, 1u, 1u};
        return headFlags(k[j], j ? k[j - 1] : before);
    };
    Cnt mine{0u, 0u, 0u, 0u};
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) mine = CntSum()(mine, flags(j));
    const Cnt chunkPrefix = blockIdx.x ? chunkIncl[blockIdx.x - 1] : Cnt{0u, 0u, 0u, 0u};
    Cnt c;
    Scan(tmp).ExclusiveScan(mine, c, chunkPrefix, CntSum());
#pragma unroll
    for (uint32_t j = 0; j < kItems; ++j) {
        const uint32_t p = p0 + j;
        if (p >= n) break;
        const Cnt h = flags(j);
        c = CntSum()(c, h);
        const K key = k[j];
        const uint32_t u = c.u - 1u, m = c.m - 1u, l = c.l - 1u;
        int32_t t[3];
        uint32_t n2;
        nodeOf(L, key, t, n2);
        const uint32_t n1 = uint32_t(key >> 9) & 4095u;
        const uint64_t oU = L.offUpper + kUpperBytes * u, oL = L.offLower + kLowerBytes * m,
                       oF = L.offLeaf + uint64_t(L.leafSize) * l;
        if (h.u) {
            uint8_t* tile = buf + kRootOff + L.rootHdr + 32ull * u;
            // RootData::CoordToKey (NanoVDB.h:2633-2640) of the tile origin, Tile::setChild
            *reinterpret_cast<uint64_t*>(tile) = uint64_t(uint32_t(t[2] * 4096) >> 12) |
                                                 (uint64_t(uint32_t(t[1] * 4096) >> 12) << 21) |
                                                 (uint64_t(uint32_t(t[0] * 4096) >> 12) << 42);
            *reinterpret_cast<int64_t*>(tile + 8) = int64_t(oU - kRootOff);
            *reinterpret_cast<uint64_t*>(buf + oU + 24) = 2u; // GridFlags::HasBBox
            upperFirst[u] = m;
        }
        if (h.m) {
            atomicOr(reinterpret_cast<unsigned long long*>(buf + oU + 4128 + 8u * (n2 >> 6)), 1ull << (n2 & 63u));
            *reinterpret_cast<int64_t*>(buf + oU + 8256 + 8ull * n2) = int64_t(oL - oU);
            *reinterpret_cast<uint64_t*>(buf + oL + 24) = 2u;
            lowerFirst[m] = l;
        }
        if (h.l) {
            atomicOr(reinterpret_cast<unsigned long long*>(buf + oL + 544 + 8u * (n1 >> 6)), 1ull << (n1 & 63u));
            *reinterpret_cast<int64_t*>(buf + oL + 1088 + 8ull * n1) = int64_t(oF - oL);
            leafFirst[l] = p;
            leafRank[l] = c.v; // 1 + rank of the leaf's first voxel: LeafData<ValueOnIndex>::mOffset
        }
        // the voxel's value index (index 0 is the background): 1 + its rank among the active voxels
        // (ValueOnIndex), or 1 + 512 * leaf + its offset in the leaf (ValueIndex)
        if (index) index[ids[p]] = L.allIndex ? 1ull + 512ull * l + (uint32_t(key) & 511u) : uint64_t(c.v);
        if (p == n - 1) {
            leafFirst[c.l] = n;
            lowerFirst[c.m] = c.l;
            upperFirst[c.u] = c.m;
        }
    }
}

// Map(dx)::applyInverseMapF of a float point (matMult with mInvMatF, math/Math.h:1704-1709)
__device__ __forceinline__ void invMapF(float inv, const float p[3], float q[3])
{
    q[0] = __builtin_fmaf(p[0], inv, __builtin_fmaf(p[1], 0.0f, p[2] * 0.0f));
    q[1] = __builtin_fmaf(p[0], 0.0f, __builtin_fmaf(p[1], inv, p[2] * 0.0f));
    q[2] = __builtin_fmaf(p[0], 0.0f, __builtin_fmaf(p[1], 0.0f, p[2] * inv));
}
// Map(dx)::applyInverseMap of a float point: FP64 matMult (math/Math.h:1716-1722), returned as float
__device__ __forceinline__ void invMapD(double inv, const float p[3], float q[3])
{
    const double x = p[0], y = p[1], z = p[2];
    q[0] = float(__builtin_fma(x, inv, __builtin_fma(y, 0.0, z * 0.0)));
    q[1] = float(__builtin_fma(x, 0.0, __builtin_fma(y, inv, z * 0.0)));
    q[2] = float(__builtin_fma(x, 0.0, __builtin_fma(y, 0.0, z * inv)));
}

// Map(dx)::applyInverseMap of a double point: FP64 matMult, kept in double
__device__ __forceinline__ void invMapDD(double inv, const double p[3], double q[3])
{
    q[0] = __builtin_fma(p[0], inv, __builtin_fma(p[1], 0.0, p[2] * 0.0));
    q[1] = __builtin_fma(p[0], 0.0, __builtin_fma(p[1], inv, p[2] * 0.0));
    q[2] = __builtin_fma(p[0], 0.0, __builtin_fma(p[1], 0.0, p[2] * inv));
}

// Voxel of each point (TileKeyFunctor :553-580): Round(applyInverseMapF(p)) for float points,
// Round(applyInverseMap(p)) in FP64 for double points (Vec3<T>::round, math/Math.h:1515-1524).
template<typename T>
__global__ __launch_bounds__(kB) void k_p2g_voxels(const T* __restrict__ xyz, uint32_t n, float invF, double invD,
                                                   int32_t* __restrict__ ijk)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const T p[3] = {xyz[3ull * i], xyz[3ull * i + 1], xyz[3ull * i + 2]};
    if constexpr (std::is_same<T, double>::value) {
        double q[3];
        invMapDD(invD, p, q);
#pragma unroll
        for (int a = 0; a < 3; ++a) ijk[3ull * i + a] = int32_t(floor(q[a] + 0.5));
    } else {
        float q[3];
        invMapF(invF, p, q);
#pragma unroll
        for (int a = 0; a < 3; ++a) ijk[3ull * i + a] = int32_t(floorf(q[a] + 0.5f));
    }
}

// Index-space voxel coordinates given as Vec3f / Vec3d: Vec3<T>::round (math/Math.h:1515-1524),
// as TileKeyFunctor does for non-Point build types (tools/cuda/PointsToGrid.cuh:576-577).
template<typename T>
__global__ __launch_bounds__(kB) void k_round_voxels(const T* __restrict__ xyz, uint32_t n, int32_t* __restrict__ ijk)
{
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T x = xyz[3ull * i + a];
        ijk[3ull * i + a] = std::is_same<T, double>::value ? int32_t(floor(double(x) + 0.5)) : int32_t(floorf(float(x) + 0.5f));
    }
}

// One wavefront per Point leaf: mOffset (points before the leaf), mPointCount, and per voxel the
// number of the leaf's points up to and including it (SetLeafActiveVoxelStateAndValuesFunctor
// :1156-1174); an inactive voxel repeats the previous active voxel's value, else 0
// (SetLeafInactiveVoxelValuesFunctor :1176-1191).
template<typename K>
__global__ __launch_bounds__(kB) void k_p2g_leaf_values(const K* __restrict__ keys, const uint32_t* __restrict__ leafFirst,
                                                        uint32_t nLeaf, Layout L, uint8_t* __restrict__ buf,
                                                        uint32_t* __restrict__ maxCount)
{
    __shared__ uint16_t vals[kWaves][512];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, l = blockIdx.x * kWaves + wv;
    const bool valid = l < nLeaf;
    const uint32_t a = valid ? leafFirst[l] : 0u, b = valid ? leafFirst[l + 1] : 0u;
    for (uint32_t j = lane; j < 512u; j += 64u) vals[wv][j] = 0;
    __syncthreads();
    for (uint32_t q = a + lane; q < b; q += 64u) {
        const K k = keys[q];
        if (q + 1u == b || keys[q + 1u] != k) vals[wv][uint32_t(k) & 511u] = uint16_t(q + 1u - a);
    }
    __syncthreads();
    if (!valid) return;
    uint8_t* F = buf + L.offLeaf + uint64_t(L.leafSize) * l;
    const uint64_t* mask = reinterpret_cast<const uint64_t*>(F + 16);
    uint16_t out[8];
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) {
        const uint32_t v = lane * 8u + e;
        // Mask::findPrev<true>(v), NanoVDB.h:1399-1409
        int32_t wi = int32_t(v >> 6);
        uint64_t bits = mask[wi] & (~0ull >> (63u - (v & 63u)));
        while (!bits && --wi >= 0) bits = mask[wi];
        out[e] = bits ? vals[wv][uint32_t(wi) * 64u + 63u - uint32_t(__builtin_clzll(bits))] : uint16_t(0);
    }
    uint4 o;
    o.x = uint32_t(out[0]) | (uint32_t(out[1]) << 16);
    o.y = uint32_t(out[2]) | (uint32_t(out[3]) << 16);
    o.z = uint32_t(out[4]) | (uint32_t(out[5]) << 16);
    o.w = uint32_t(out[6]) | (uint32_t(out[7]) << 16);
    *reinterpret_cast<uint4*>(F + 96 + 16u * lane) = o;
    if (lane == 0) {
        *reinterpret_cast<uint64_t*>(F + 80) = a;
        *reinterpret_cast<uint64_t*>(F + 88) = b - a;
        // only an overfull leaf touches the flag: one address hit by every leaf serialises them
        if (b - a > 65535u) atomicMax(maxCount, b - a);
    }
}

// The points in sorted order as blind data (encodePoints :1262-1317, worldToVoxel :197-236).
// Double points keep FP64 through Grid64 and the voxel offsets; Grid32 maps the float-cast
// coordinates (applyInverseMapF casts each one, math/Math.h:1704-1709); Default stores the input.
template<typename T>
__global__ __launch_bounds__(kB) void k_p2g_blind(const uint32_t* __restrict__ ids, uint32_t n, PointArgs P,
                                                  uint8_t* __restrict__ out)
{
    constexpr bool kF64 = std::is_same<T, double>::value;
    const uint32_t q = blockIdx.x * kB + threadIdx.x;
    if (q >= n) return;
    const uint32_t id = ids[q];
    const T* xyz = static_cast<const T*>(P.xyz);
    const T w[3] = {xyz[3ull * id], xyz[3ull * id + 1], xyz[3ull * id + 2]};
    const float wf[3] = {float(w[0]), float(w[1]), float(w[2])};
    uint8_t* o = out + uint64_t(P.size) * q;
    // index-space position: FP64 for double points, FP64 rounded to float for float points
    auto indexPos = [&](double g[3]) {
        if constexpr (kF64) {
            invMapDD(P.invD, w, g);
        } else {
            float gf[3];
            invMapD(P.invD, wf, gf);
            for (int a = 0; a < 3; ++a) g[a] = double(gf[a]);
        }
    };
    double g[3];
    float gf[3];
    switch (P.type) {
    case 1: *reinterpret_cast<uint32_t*>(o) = id; break;
    case 2:
        for (int a = 0; a < 3; ++a) reinterpret_cast<double*>(o)[a] = double(w[a]);
        break;
    case 4:
        indexPos(g);
        for (int a = 0; a < 3; ++a) reinterpret_cast<double*>(o)[a] = g[a];
        break;
    case 5:
        invMapF(P.invF, wf, gf);
        for (int a = 0; a < 3; ++a) reinterpret_cast<float*>(o)[a] = gf[a];
        break;
    case 6:
    case 7:
    case 8:
        indexPos(g);
        for (int a = 0; a < 3; ++a) {
            const double f = g[a] - double(int32_t(floor(g[a] + 0.5)));
            if (P.type == 6) reinterpret_cast<float*>(o)[a] = float(f);
            else if (P.type == 7) reinterpret_cast<uint16_t*>(o)[a] = uint16_t(65535.0 * (f + 0.5));
            else o[a] = uint8_t(255.0 * (f + 0.5));
        }
        break;
    case 9:
        for (int a = 0; a < 3; ++a) reinterpret_cast<T*>(o)[a] = w[a];
        break;
    default: // 3: World32
        for (int a = 0; a < 3; ++a) reinterpret_cast<float*>(o)[a] = wf[a];
    }
}

// Grid, tree and root headers, passed by value so the launch owns a copy of them.
constexpr uint32_t kHeaderMax = 832; // 672 + 64 + the 96-byte index-grid root header
struct Header {
    uint32_t bytes;
    uint32_t w[kHeaderMax / 4];
};
__global__ void k_v2g_header(uint8_t* __restrict__ buf, Header h)
{
    for (uint32_t i = threadIdx.x; i < h.bytes / 4u; i += blockDim.x) reinterpret_cast<uint32_t*>(buf)[i] = h.w[i];
}

template<typename T>
__device__ __forceinline__ T laneSelect8(const T (&w)[8], uint32_t j)
{
    T r = w[0];
#pragma unroll
    for (uint32_t i = 1; i < 8; ++i) r = j == i ? w[i] : r;
    return r;
}

// One wavefront per leaf: value mask from its voxels, LeafNode::updateBBox (NanoVDB.h:4596-4624),
// flags, and the index offset + prefix sums (leafPrefixSumKernel :459-479) or the float values
// (1 active, 0 inactive; SetLeafActiveVoxelStateAndValuesFunctor :1156-1174).
template<typename K>
__global__ __launch_bounds__(kB) void k_v2g_leaves(const K* __restrict__ keys, const uint32_t* __restrict__ leafRank,
                                                   const uint32_t* __restrict__ leafFirst, uint32_t nLeaf, Layout L,
                                                   uint8_t* __restrict__ buf)
{
    const uint32_t lane = threadIdx.x & 63u, l = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (l >= nLeaf) return;
    const uint32_t a = leafFirst[l], b = leafFirst[l + 1];
    uint64_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t q = a + lane; q < b; q += 64) {
        const uint32_t n0 = uint32_t(keys[q]) & 511u;
        const uint64_t bit = 1ull << (n0 & 63u);
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) w[j] |= (n0 >> 6) == j ? bit : 0ull;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) w[j] |= __shfl_xor(w[j], off, 64);

    uint8_t* F = buf + L.offLeaf + uint64_t(L.leafSize) * l;
    if (lane == 0) {
        const K k0 = keys[a];
        int32_t t[3];
        uint32_t n2;
        nodeOf(L, k0, t, n2);
        const uint32_t n1 = uint32_t(k0 >> 9) & 4095u;
        const int32_t origin[3] = {t[0] * 4096 + int32_t((n2 >> 10) & 31u) * 128 + int32_t((n1 >> 8) & 15u) * 8,
                                   t[1] * 4096 + int32_t((n2 >> 5) & 31u) * 128 + int32_t((n1 >> 4) & 15u) * 8,
                                   t[2] * 4096 + int32_t(n2 & 31u) * 128 + int32_t(n1 & 15u) * 8};
        uint64_t word = w[0];
        uint32_t xmin = w[0] ? 0u : 8u, xmax = xmin;
#pragma unroll
        for (uint32_t j = 1; j < 8; ++j)
            if (w[j]) {
                word |= w[j];
                if (xmin == 8u) xmin = j;
                xmax = j;
            }
        const uint32_t ymin = uint32_t(__builtin_ctzll(word)) >> 3, ymax = (63u - uint32_t(__builtin_clzll(word))) >> 3;
        const uint32_t w32 = uint32_t(word) | uint32_t(word >> 32);
        const uint32_t w16 = (w32 & 0xffffu) | (w32 >> 16);
        const uint32_t byte = (w16 & 0xffu) | (w16 >> 8);
        const uint32_t zmin = uint32_t(__builtin_ctz(byte)), zmax = 31u - uint32_t(__builtin_clz(byte));
        int4 h;
        h.x = origin[0] + int32_t(xmin);
        h.y = origin[1] + int32_t(ymin);
        h.z = origin[2] + int32_t(zmin);
        h.w = int32_t((xmax - xmin) | ((ymax - ymin) << 8) | ((zmax - zmin) << 16) | (2u << 24)); // dif[3], flags HasBBox
        *reinterpret_cast<int4*>(F) = h;
        if (!L.isIndex && !L.isPoint) // float leaves: min/max/avg/std, zero as the builder leaves them
            *reinterpret_cast<uint4*>(F + 80) = make_uint4(0u, 0u, 0u, 0u);
        if (L.isIndex) {
            uint64_t sum = uint64_t(__popcll(w[0])), prefix = sum;
#pragma unroll
            for (int j = 1; j < 7; ++j) {
                sum += uint64_t(__popcll(w[j]));
                prefix |= sum << (9 * j);
            }
            // 1 + values before the leaf: active voxels (ValueOnIndex), or all 512 per leaf with a zero
            // prefix sum (ValueIndex, tools/cuda/PointsToGrid.cuh:1147-1148)
            *reinterpret_cast<uint64_t*>(F + 80) = L.allIndex ? 1ull + 512ull * l : uint64_t(leafRank[l]);
            *reinterpret_cast<uint64_t*>(F + 88) = L.allIndex ? 0ull : prefix;
        }
    }
    if (lane < 8) *reinterpret_cast<uint64_t*>(F + 16 + 8 * lane) = laneSelect8(w, lane);
    if (!L.isIndex && !L.isPoint) { // float leaves only: 2048 B of values
        // values 8*lane .. 8*lane+7 live in byte (lane & 7) of mask word lane >> 3
        const uint32_t bits = uint32_t(laneSelect8(w, lane >> 3) >> (8u * (lane & 7u))) & 255u;
        float4 v0, v1;
        v0.x = (bits & 1u) ? 1.0f : 0.0f;
        v0.y = (bits & 2u) ? 1.0f : 0.0f;
        v0.z = (bits & 4u) ? 1.0f : 0.0f;
        v0.w = (bits & 8u) ? 1.0f : 0.0f;
        v1.x = (bits & 16u) ? 1.0f : 0.0f;
        v1.y = (bits & 32u) ? 1.0f : 0.0f;
        v1.z = (bits & 64u) ? 1.0f : 0.0f;
        v1.w = (bits & 128u) ? 1.0f : 0.0f;
        float4* dst = reinterpret_cast<float4*>(F + 96 + 32u * lane);
        dst[0] = v0;
        dst[1] = v1;
    }
}

// Sparse inputs (fewer than 64 sorted keys per leaf on average): eight lanes per leaf, lane j
// owning voxels 64j .. 64j+63, i.e. mask word j and that stretch of the value array.  The leaf's
// keys are sorted by voxel offset, so each lane binary-searches its stretch and walks it once.
// Writes the same bytes as k_v2g_leaves and, for Point leaves, k_p2g_leaf_values: a voxel's
// point value is the number of the leaf's points at or before it, which is the count at the
// previous active voxel for an inactive one, or 0 before the first.
template<typename K>
__global__ __launch_bounds__(kB) void k_v2g_leaves_sparse(const K* __restrict__ keys, const uint32_t* __restrict__ leafRank,
                                                          const uint32_t* __restrict__ leafFirst, uint32_t nLeaf, Layout L,
                                                          uint8_t* __restrict__ buf, uint32_t* __restrict__ maxCount)
{
    const uint32_t lane = threadIdx.x & 63u, sub = lane & 7u;
    const uint32_t l = (blockIdx.x * kB + threadIdx.x) >> 3;
    if (l >= nLeaf) return; // whole groups of eight leave together
    const uint32_t a = leafFirst[l], b = leafFirst[l + 1];
    const uint32_t lo = 64u * sub;
    // first key of the stretch: lower bound of lo among the leaf's voxel offsets
    uint32_t q0 = a, hiq = b;
    while (q0 < hiq) {
        const uint32_t mid = (q0 + hiq) >> 1;
        if ((uint32_t(keys[mid]) & 511u) < lo) q0 = mid + 1u;
        else hiq = mid;
    }
    uint64_t word = 0;
    uint32_t p = q0, n0 = p < b ? uint32_t(keys[p]) & 511u : 512u;
    uint8_t* F = buf + L.offLeaf + uint64_t(L.leafSize) * l;
    if (L.isPoint) {
        // values lo .. lo+63: points of the leaf with voxel offset <= v
        uint32_t packed[32];
#pragma unroll
        for (uint32_t e = 0; e < 64u; ++e) {
            while (n0 <= lo + e) {
                word |= 1ull << e;
                n0 = ++p < b ? uint32_t(keys[p]) & 511u : 512u;
            }
            const uint32_t c = (p - a) & 0xffffu;
            if (e & 1u) packed[e >> 1] |= c << 16;
            else packed[e >> 1] = c;
        }
        uint4* dst = reinterpret_cast<uint4*>(F + 96 + 128u * sub);
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) dst[k] = make_uint4(packed[4 * k], packed[4 * k + 1], packed[4 * k + 2], packed[4 * k + 3]);
    } else {
        while (n0 < lo + 64u) {
            word |= 1ull << (n0 - lo);
            n0 = ++p < b ? uint32_t(keys[p]) & 511u : 512u;
        }
        if (!L.isIndex) { // float leaves: 1 for an active voxel, 0 otherwise
            float4* dst = reinterpret_cast<float4*>(F + 96 + 256u * sub);
#pragma unroll
            for (uint32_t k = 0; k < 16u; ++k) {
                const uint32_t bits = uint32_t(word >> (4u * k)) & 15u;
                dst[k] = make_float4((bits & 1u) ? 1.0f : 0.0f, (bits & 2u) ? 1.0f : 0.0f, (bits & 4u) ? 1.0f : 0.0f,
                                     (bits & 8u) ? 1.0f : 0.0f);
            }
        }
    }
    *reinterpret_cast<uint64_t*>(F + 16 + 8u * sub) = word;
    // the group's words: x extent from which words are set, y and z extents from their union
    const uint32_t xbits = uint32_t(__ballot(word != 0ull) >> (lane & ~7u)) & 255u;
    uint64_t any = word;
    any |= __shfl_xor(any, 4, 64);
    any |= __shfl_xor(any, 2, 64);
    any |= __shfl_xor(any, 1, 64);
    // inclusive prefix of the per-word active counts (index grids)
    uint32_t pc = uint32_t(__popcll(word));
#pragma unroll
    for (uint32_t off = 1; off < 8u; off <<= 1) {
        const uint32_t t = __shfl_up(pc, off, 8);
        if (sub >= off) pc += t;
    }
    uint64_t prefix = 0;
#pragma unroll
    for (uint32_t j = 0; j < 7u; ++j) prefix |= uint64_t(__shfl(pc, (lane & ~7u) + j, 64)) << (9u * j);
    if (sub != 0u) return;
    const K k0 = keys[a];
    int32_t t[3];
    uint32_t n2;
    nodeOf(L, k0, t, n2);
    const uint32_t n1 = uint32_t(k0 >> 9) & 4095u;
    const int32_t origin[3] = {t[0] * 4096 + int32_t((n2 >> 10) & 31u) * 128 + int32_t((n1 >> 8) & 15u) * 8,
                               t[1] * 4096 + int32_t((n2 >> 5) & 31u) * 128 + int32_t((n1 >> 4) & 15u) * 8,
                               t[2] * 4096 + int32_t(n2 & 31u) * 128 + int32_t(n1 & 15u) * 8};
    const uint32_t xmin = uint32_t(__builtin_ctz(xbits)), xmax = 31u - uint32_t(__builtin_clz(xbits));
    const uint32_t ymin = uint32_t(__builtin_ctzll(any)) >> 3, ymax = (63u - uint32_t(__builtin_clzll(any))) >> 3;
    const uint32_t w32 = uint32_t(any) | uint32_t(any >> 32);
    const uint32_t w16 = (w32 & 0xffffu) | (w32 >> 16);
    const uint32_t byte = (w16 & 0xffu) | (w16 >> 8);
    const uint32_t zmin = uint32_t(__builtin_ctz(byte)), zmax = 31u - uint32_t(__builtin_clz(byte));
    int4 h;
    h.x = origin[0] + int32_t(xmin);
    h.y = origin[1] + int32_t(ymin);
    h.z = origin[2] + int32_t(zmin);
    h.w = int32_t((xmax - xmin) | ((ymax - ymin) << 8) | ((zmax - zmin) << 16) | (2u << 24)); // dif[3], flags HasBBox
    *reinterpret_cast<int4*>(F) = h;
    if (L.isIndex) {
        *reinterpret_cast<uint64_t*>(F + 80) = L.allIndex ? 1ull + 512ull * l : uint64_t(leafRank[l]);
        *reinterpret_cast<uint64_t*>(F + 88) = L.allIndex ? 0ull : prefix;
    } else if (!L.isPoint) { // float leaves: min/max/avg/std, zero as the builder leaves them
        *reinterpret_cast<uint4*>(F + 80) = make_uint4(0u, 0u, 0u, 0u);
    } else {
        *reinterpret_cast<uint64_t*>(F + 80) = a;
        *reinterpret_cast<uint64_t*>(F + 88) = b - a;
        if (b - a > 65535u) atomicMax(maxCount, b - a);
    }
}

__device__ __forceinline__ void waveBoxReduce(int32_t lo[3], int32_t hi[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            lo[a] = min(lo[a], __shfl_xor(lo[a], off, 64));
            hi[a] = max(hi[a], __shfl_xor(hi[a], off, 64));
        }
}

__device__ __forceinline__ void storeBox(uint8_t* node, const int32_t lo[3], const int32_t hi[3])
{
    *reinterpret_cast<int4*>(node) = make_int4(lo[0], lo[1], lo[2], hi[0]);
    *reinterpret_cast<int2*>(node + 16) = make_int2(hi[1], hi[2]);
}

// One wavefront per lower node: its bbox is the union of its leaves' (processBBox).
__global__ __launch_bounds__(kB) void k_v2g_lowers(const uint32_t* __restrict__ lowerFirst, uint32_t nLower, Layout L,
                                                   uint8_t* __restrict__ buf)
{
    const uint32_t lane = threadIdx.x & 63u, m = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (m >= nLower) return;
    int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (uint32_t l = lowerFirst[m] + lane; l < lowerFirst[m + 1]; l += 64) {
        const int4 h = *reinterpret_cast<const int4*>(buf + L.offLeaf + uint64_t(L.leafSize) * l);
        const uint32_t d = uint32_t(h.w);
        lo[0] = min(lo[0], h.x);
        lo[1] = min(lo[1], h.y);
        lo[2] = min(lo[2], h.z);
        hi[0] = max(hi[0], h.x + int32_t(d & 255u));
        hi[1] = max(hi[1], h.y + int32_t((d >> 8) & 255u));
        hi[2] = max(hi[2], h.z + int32_t((d >> 16) & 255u));
    }
    waveBoxReduce(lo, hi);
    if (lane == 0) storeBox(buf + L.offLower + kLowerBytes * m, lo, hi);
}

// One wavefront per upper node: union of its lower nodes' bboxes.
__global__ __launch_bounds__(kB) void k_v2g_uppers(const uint32_t* __restrict__ upperFirst, uint32_t nUpper, Layout L,
                                                   uint8_t* __restrict__ buf)
{
    const uint32_t lane = threadIdx.x & 63u, u = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (u >= nUpper) return;
    int32_t lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (uint32_t m = upperFirst[u] + lane; m < upperFirst[u + 1]; m += 64) {
        const int32_t* b = reinterpret_cast<const int32_t*>(buf + L.offLower + kLowerBytes * m);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], b[a]);
            hi[a] = max(hi[a], b[3 + a]);
        }
    }
    waveBoxReduce(lo, hi);
    if (lane == 0) storeBox(buf + L.offUpper + kUpperBytes * u, lo, hi);
}

template<typename T>
void put(uint8_t* b, uint64_t off, const T& v)
{
    static_assert(!std::is_pointer<T>::value, "put() copies values, not what a pointer points to");
    std::memcpy(b + off, &v, sizeof(T));
}

// Grid, tree and root headers (BuildGridTreeRootFunctor :877-1008 with GridData::init
// NanoVDB.h:1968-2000, Map(double) :1441-1451, UpdateRootWorldBBoxFunctor :1374-1384).
void writeHeaders(uint8_t* h, const Layout& L, const uint32_t counts[4], const int32_t box[6], double voxelSize,
                  const char* name, uint32_t gridType, uint32_t gridClass = 0, uint64_t points = 0)
{
    const uint64_t offTree = kTreeOff, offRoot = kRootOff;
    for (int a = 0; a < 6; ++a) put(h, offRoot + 4 * a, box[a]);
    put(h, offRoot + 24, counts[3]);
    put(h, offTree + 0, int64_t(L.offLeaf - offTree));
    put(h, offTree + 8, int64_t(L.offLower - offTree));
    put(h, offTree + 16, int64_t(L.offUpper - offTree));
    put(h, offTree + 24, int64_t(offRoot - offTree));
    put(h, offTree + 32, counts[1]);
    put(h, offTree + 36, counts[2]);
    put(h, offTree + 40, counts[3]);
    put(h, offTree + 56, uint64_t(counts[0]));

    put(h, 0, uint64_t(0x314244566f6e614eull)); // "NanoVDB1"
    put(h, 8, ~uint64_t(0));                    // checksum disabled
    put(h, 16, uint32_t((32u << 21) | (9u << 10) | 2u));
    put(h, 20, uint32_t(2u | 32u)); // HasBBox | IsBreadthFirst
    put(h, 28, uint32_t(1));
    put(h, 32, L.total);
    if (name) std::memcpy(h + 40, name, std::min<size_t>(std::strlen(name), 255));
    const float sF = float(voxelSize), invF = 1.0f / sF;
    for (int i = 0; i < 9; ++i) {
        const bool d = i % 4 == 0;
        put(h, 296 + 4 * i, d ? sF : 0.0f);
        put(h, 296 + 36 + 4 * i, d ? invF : 0.0f);
        put(h, 296 + 88 + 8 * i, d ? voxelSize : 0.0);
        put(h, 296 + 160 + 8 * i, d ? 1.0 / voxelSize : 0.0);
    }
    put(h, 296 + 84, 1.0f);
    put(h, 296 + 256, 1.0);
    // Map::applyMap (matMult, math/Math.h:1748-1753) of the 8 corners of (min, max + 1)
    auto apply = [&](const double x[3], double y[3]) {
        for (int r = 0; r < 3; ++r)
            y[r] = std::fma(x[0], r == 0 ? voxelSize : 0.0, std::fma(x[1], r == 1 ? voxelSize : 0.0, std::fma(x[2], r == 2 ? voxelSize : 0.0, 0.0)));
    };
    double wb[6];
    for (int c = 0; c < 8; ++c) {
        const double x[3] = {double(c & 1 ? box[3] + 1 : box[0]), double(c & 2 ? box[4] + 1 : box[1]),
                             double(c & 4 ? box[5] + 1 : box[2])};
        double y[3];
        apply(x, y);
        for (int a = 0; a < 3; ++a) {
            wb[a] = c == 0 ? y[a] : std::min(wb[a], y[a]);
            wb[3 + a] = c == 0 ? y[a] : std::max(wb[3 + a], y[a]);
        }
    }
    put(h, 560, wb);
    const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
    double p1[3], p0[3];
    apply(one, p1);
    apply(zero, p0);
    const double vs[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    put(h, 608, vs);
    put(h, 632, uint32_t(L.isIndex ? 8u : L.isPoint ? gridClass : 0u)); // IndexGrid / PointIndex|PointData / Unknown
    put(h, 636, gridType);
    put(h, 640, int64_t(L.offMeta)); // blind metadata: only point grids have some
    put(h, 648, uint32_t(L.isPoint ? 1u : 0u));
    put(h, 656, uint64_t(L.allIndex  ? 1u + 512u * uint64_t(counts[1])
                         : L.isIndex ? uint64_t(counts[0]) + 1u
                         : L.isPoint ? points : 0u)); // value count (index grids) or point count
}

// Per-device scratch of the builder, kept between builds and grown when a build needs more.  A
// build holds the device's lock from start to finish and waits for its stream before it returns,
// so no kernel of an earlier build still uses the memory when the next one starts.
struct Workspace {
    std::mutex lock;
    uint8_t* base = nullptr;
    size_t bytes = 0;
    // device memory of at least `need` bytes (the contents are not kept when it grows)
    uint8_t* reserve(size_t need)
    {
        if (need <= bytes) return base;
        if (base) (void)hipFree(base);
        base = nullptr;
        bytes = 0;
        const size_t grow = need + need / 4;
        if (hipMalloc(reinterpret_cast<void**>(&base), grow) != hipSuccess) return nullptr;
        bytes = grow;
        return base;
    }
};
constexpr int kMaxDevices = 64; // devices per process with a builder workspace
Workspace& workspaceOf(int device)
{
    static Workspace spaces[kMaxDevices];
    return spaces[device];
}

// Carves aligned arrays out of one block of device memory.
struct Carve {
    uint8_t* base;
    size_t used = 0;
    template<typename T>
    T* take(size_t count)
    {
        used = (used + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += count * sizeof(T);
        return p;
    }
};
template<typename T>
size_t carveBytes(size_t count)
{
    return ((count * sizeof(T) + 255) & ~size_t(255));
}

uint32_t blocksFor(uint64_t items, uint32_t perBlock) { return uint32_t((items + perBlock - 1) / perBlock); }

// Steps after the bbox: sort keys of width K over endBit bits, node numbering, node contents.
// The builder's key sort.  32-bit keys of 25-27 bits (single-tile inputs, e.g. a 512^3 volume)
// take three 9-bit onesweep passes instead of four 8-bit ones (1.11 vs 1.39 ms for 43.6 M keys
// on MI355X); other widths keep the library's tuned 8-bit passes.  The input index of each voxel
// rides along only when a per-voxel index or point data needs it.
using Onesweep9 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 9,
                                        rocprim::block_radix_rank_algorithm::match>>;
template<typename K>
hipError_t sortPairs(void* temp, size_t& bytes, const K* keys, K* keys2, const uint32_t* ids, uint32_t* ids2, uint32_t n,
                     int endBit, bool withIds, hipStream_t s)
{
    if constexpr (std::is_same<K, uint32_t>::value) {
        if ((endBit + 8) / 9 < (endBit + 7) / 8)
            return withIds ? rocprim::radix_sort_pairs<Onesweep9>(temp, bytes, keys, keys2, ids, ids2, size_t(n), 0u,
                                                                  unsigned(endBit), s)
                           : rocprim::radix_sort_keys<Onesweep9>(temp, bytes, keys, keys2, size_t(n), 0u,
                                                                 unsigned(endBit), s);
    }
    return withIds ? hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys, keys2, ids, ids2, int(n), 0, endBit, s)
                   : hipcub::DeviceRadixSort::SortKeys(temp, bytes, keys, keys2, int(n), 0, endBit, s);
}

template<typename K>
int buildSorted(Workspace& W, const int32_t* ijk, uint32_t n32, Layout L, int endBit, const int32_t box[6],
                double voxel_size, const char* name, uint32_t grid_type, uint64_t* index, const PointArgs* P,
                hipStream_t s, nvq_grid** out)
{
    // 2. keys, one radix sort, run-head totals per chunk and their scan
    const uint32_t chunks = uint32_t((uint64_t(n32) + kChunk - 1) / kChunk);
    size_t sortBytes = 0, scanBytes = 0;
    const bool withIds = index != nullptr || P != nullptr;
    if (sortPairs<K>(nullptr, sortBytes, static_cast<K*>(nullptr), static_cast<K*>(nullptr),
                     static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), n32, endBit, withIds, s) !=
            hipSuccess ||
        hipcub::DeviceScan::InclusiveScan(nullptr, scanBytes, static_cast<Cnt*>(nullptr), static_cast<Cnt*>(nullptr),
                                          CntSum(), int(chunks), s) != hipSuccess)
        return NVQ_ERR_HIP;
    const size_t tempBytes = std::max(sortBytes, scanBytes);
    // live until the end: sorted keys and positions, chunk counts; dead after the scan: the
    // unsorted keys and positions and the library temporaries (reused for the node starts)
    const size_t live = carveBytes<K>(n32) + carveBytes<uint32_t>(n32) + 2 * carveBytes<Cnt>(chunks);
    const size_t dead = carveBytes<K>(n32) + carveBytes<uint32_t>(n32) + carveBytes<uint8_t>(tempBytes);
    uint8_t* ws = W.reserve(live + dead);
    if (!ws) return NVQ_ERR_HIP;
    Carve c{ws};
    K* keys2 = c.take<K>(n32);
    uint32_t* ids2 = c.take<uint32_t>(n32);
    Cnt* chunkSums = c.take<Cnt>(chunks);
    Cnt* chunkIncl = c.take<Cnt>(chunks);
    const size_t deadStart = (c.used + 255) & ~size_t(255);
    K* keys = c.take<K>(n32);
    uint32_t* ids = c.take<uint32_t>(n32);
    void* temp = c.take<uint8_t>(tempBytes);
    k_v2g_keys<K><<<blocksFor(n32, kB), kB, 0, s>>>(ijk, n32, L, keys, ids);
    {
        ProfScope prof("v2g_sort", s);
        if (sortPairs<K>(temp, sortBytes, keys, keys2, ids, ids2, n32, endBit, withIds, s) != hipSuccess)
            return NVQ_ERR_HIP;
    }
    {
        ProfScope prof("v2g_scan", s);
        k_v2g_count<K><<<chunks, kB, 0, s>>>(keys2, n32, chunkSums);
        if (hipcub::DeviceScan::InclusiveScan(temp, scanBytes, chunkSums, chunkIncl, CntSum(), int(chunks), s) != hipSuccess)
            return NVQ_ERR_HIP;
    }
    Cnt tot;
    if (readBack(&tot, chunkIncl + (chunks - 1), sizeof(Cnt), s) != hipSuccess) return NVQ_ERR_HIP;

    // 3. layout (GridT::memUsage sums of PointsToGrid::getBuffer :848-856), zeroed buffer
    L.rootHdr = (L.isIndex || L.isPoint) ? 96u : 64u;
    L.leafSize = L.isIndex ? 96u : L.isPoint ? 1120u : 2144u;
    L.offUpper = kRootOff + L.rootHdr + 32ull * tot.u;
    L.offLower = L.offUpper + kUpperBytes * tot.u;
    L.offLeaf = L.offLower + kLowerBytes * tot.m;
    L.offMeta = L.offLeaf + uint64_t(L.leafSize) * tot.l;
    L.total = P ? L.offMeta + 288u + uint64_t(P->size) * n32 : L.offMeta;
    void* dbuf = nullptr;
    if (poolMalloc(&dbuf, L.total, s) != hipSuccess) return NVQ_ERR_HIP;
    uint8_t* buf = static_cast<uint8_t*>(dbuf);
    Header hdr{};
    hdr.bytes = uint32_t(kRootOff + L.rootHdr);
    const uint32_t counts[4] = {tot.v, tot.l, tot.m, tot.u};
    writeHeaders(reinterpret_cast<uint8_t*>(hdr.w), L, counts, box, voxel_size, name, grid_type, P ? P->gridClass : 0u,
                 n32);
    // node starts: in the dead region when they fit, else in a block of their own
    const size_t starts = carveBytes<uint32_t>(uint64_t(tot.l) + 1) + carveBytes<uint32_t>(tot.l) +
                          carveBytes<uint32_t>(uint64_t(tot.m) + 1) + carveBytes<uint32_t>(uint64_t(tot.u) + 1) +
                          carveBytes<uint32_t>(1);
    void* extra = nullptr;
    Carve st{ws + deadStart};
    if (starts > dead) {
        if (poolMalloc(&extra, starts, s) != hipSuccess) {
            (void)hipFreeAsync(dbuf, s);
            return NVQ_ERR_HIP;
        }
        st = Carve{static_cast<uint8_t*>(extra)};
    }
    uint32_t* leafFirst = st.take<uint32_t>(uint64_t(tot.l) + 1);
    uint32_t* leafRank = st.take<uint32_t>(tot.l);
    uint32_t* lowerFirst = st.take<uint32_t>(uint64_t(tot.m) + 1);
    uint32_t* upperFirst = st.take<uint32_t>(uint64_t(tot.u) + 1);
    uint32_t* maxCount = st.take<uint32_t>(1);
    int rc = NVQ_ERR_HIP;
    // zero the headers, root and internal nodes only: the leaf kernels write every byte of every
    // leaf, and the point kernels every byte of the blind data
    if (hipMemsetAsync(buf, 0, L.offLeaf, s) == hipSuccess) {
        k_v2g_header<<<1, 256, 0, s>>>(buf, hdr);
        // 4. nodes, then leaves, then bboxes bottom-up
        {
            ProfScope prof("k_v2g_heads", s);
            k_v2g_heads<K><<<chunks, kB, 0, s>>>(keys2, ids2, chunkIncl, n32, L, buf, leafFirst, leafRank, lowerFirst,
                                              upperFirst, index);
        }
        // fewer than 64 keys per leaf on average: eight lanes per leaf, else a wavefront per leaf
        const bool sparse = uint64_t(n32) < 64ull * tot.l;
        if (P) (void)hipMemsetAsync(maxCount, 0, sizeof(uint32_t), s);
        {
            ProfScope prof("k_v2g_leaves", s);
            if (sparse)
                k_v2g_leaves_sparse<K><<<blocksFor(uint64_t(tot.l) * 8u, kB), kB, 0, s>>>(keys2, leafRank, leafFirst, tot.l,
                                                                                      L, buf, maxCount);
            else
                k_v2g_leaves<K><<<blocksFor(tot.l, kWaves), kB, 0, s>>>(keys2, leafRank, leafFirst, tot.l, L, buf);
        }
        if (P) {
            // Point grids: leaf values, offsets and counts, the blind metadata and the points
            if (!sparse)
                k_p2g_leaf_values<K><<<blocksFor(tot.l, kWaves), kB, 0, s>>>(keys2, leafFirst, tot.l, L, buf, maxCount);
            Header meta{};
            meta.bytes = 288u;
            uint8_t* mb = reinterpret_cast<uint8_t*>(meta.w);
            put(mb, 0, int64_t(288));
            put(mb, 8, uint64_t(n32));
            put(mb, 16, P->size);
            put(mb, 20, P->semantic);
            put(mb, 24, P->dataClass);
            put(mb, 28, P->dataType);
            std::memcpy(mb + 32, P->name, std::min<size_t>(std::strlen(P->name), 255));
            k_v2g_header<<<1, 256, 0, s>>>(buf + L.offMeta, meta);
            if (P->f64)
                k_p2g_blind<double><<<blocksFor(n32, kB), kB, 0, s>>>(ids2, n32, *P, buf + L.offMeta + 288u);
            else
                k_p2g_blind<float><<<blocksFor(n32, kB), kB, 0, s>>>(ids2, n32, *P, buf + L.offMeta + 288u);
        }
        k_v2g_lowers<<<blocksFor(tot.m, kWaves), kB, 0, s>>>(lowerFirst, tot.m, L, buf);
        k_v2g_uppers<<<blocksFor(tot.u, kWaves), kB, 0, s>>>(upperFirst, tot.u, L, buf);
        // complete before returning, as PointsToGrid::getHandle does (:537): the grid may be used
        // on any stream next
        if (hipGetLastError() == hipSuccess)
            rc = makeGridHandle(reinterpret_cast<const uint8_t*>(hdr.w), L.total, dbuf, kOwnerPool, s, out);
    }
    uint32_t most = 0;
    if (rc == NVQ_OK && P &&
        (readBack(&most, maxCount, sizeof(most), s) != hipSuccess || most > 65535u)) {
        // a leaf with more points than its uint16 values can count (the reference throws)
        (void)nvq_grid_free(*out);
        *out = nullptr;
        rc = most > 65535u ? NVQ_ERR_INVALID_ARG : NVQ_ERR_HIP;
        dbuf = nullptr;
    }
    if (extra) (void)hipFreeAsync(extra, s);
    if (rc != NVQ_OK && dbuf) (void)hipFreeAsync(dbuf, s);
    // the workspace is free for the next build only once this stream has drained
    if (hipStreamSynchronize(s) != hipSuccess && rc == NVQ_OK) {
        (void)nvq_grid_free(*out);
        *out = nullptr;
        rc = NVQ_ERR_HIP;
    }
    return rc;
}

// From the input voxels on: bbox, key width, then buildSorted.  The caller holds W's lock.
// Dense numbering of the upper-node tiles (and, for one tile, the lower nodes) over an input bbox:
// the sort keys of the builder.
void keyLayout(const int32_t box[6], Layout& L, uint64_t& tiles, int& tileBits, int& lowerBits)
{
    tiles = 1;
    for (int a = 0; a < 3; ++a) {
        L.tlo[a] = box[a] >> 12;
        L.dim[a] = uint32_t((box[3 + a] >> 12) - L.tlo[a] + 1);
        tiles *= L.dim[a];
    }
    tileBits = tiles > 1 ? 64 - __builtin_clzll(tiles - 1) : 0;
    uint64_t lowers = 1;
    for (int a = 0; a < 3; ++a) {
        L.llo[a] = box[a] >> 7;
        L.ldim[a] = uint32_t((box[3 + a] >> 7) - L.llo[a] + 1);
        lowers *= L.ldim[a];
    }
    lowerBits = lowers > 1 ? 64 - __builtin_clzll(lowers - 1) : 0;
}

// Largest number of points in one voxel (the density of the voxel-size search): sort the voxel
// keys, run-length encode them and take the longest run (DeviceRunLengthEncode + DeviceReduce::Max
// over the sorted keys, as countNodes does, tools/cuda/PointsToGrid.cuh:722-740).
template<typename K>
int maxPerVoxelSorted(const int32_t* ijk, uint32_t n, const Layout& L, int endBit, hipStream_t s, uint32_t& most)
{
    size_t sortBytes = 0, rleBytes = 0, maxBytes = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, sortBytes, static_cast<K*>(nullptr), static_cast<K*>(nullptr), int(n), 0,
                                          endBit, s) != hipSuccess ||
        hipcub::DeviceRunLengthEncode::Encode(nullptr, rleBytes, static_cast<K*>(nullptr), static_cast<K*>(nullptr),
                                              static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr), int(n),
                                              s) != hipSuccess ||
        hipcub::DeviceReduce::Max(nullptr, maxBytes, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                  int(n), s) != hipSuccess)
        return NVQ_ERR_HIP;
    const size_t bytes = 3 * carveBytes<K>(n) + 2 * carveBytes<uint32_t>(n) + carveBytes<uint32_t>(2) +
                         carveBytes<uint8_t>(std::max({sortBytes, rleBytes, maxBytes}));
    void* mem = nullptr;
    if (poolMalloc(&mem, bytes, s) != hipSuccess) return NVQ_ERR_HIP;
    Carve c{static_cast<uint8_t*>(mem)};
    K* keys = c.take<K>(n);
    K* sorted = c.take<K>(n);
    K* unique = c.take<K>(n);
    uint32_t* ids = c.take<uint32_t>(n);
    uint32_t* counts = c.take<uint32_t>(n);
    uint32_t* small = c.take<uint32_t>(2); // [0] runs, [1] longest run
    void* temp = c.take<uint8_t>(std::max({sortBytes, rleBytes, maxBytes}));
    int rc = NVQ_ERR_HIP;
    uint32_t runs = 0;
    k_v2g_keys<K><<<blocksFor(n, kB), kB, 0, s>>>(ijk, n, L, keys, ids);
    if (hipGetLastError() == hipSuccess &&
        hipcub::DeviceRadixSort::SortKeys(temp, sortBytes, keys, sorted, int(n), 0, endBit, s) == hipSuccess &&
        hipcub::DeviceRunLengthEncode::Encode(temp, rleBytes, sorted, unique, counts, small, int(n), s) == hipSuccess &&
        readBack(&runs, small, sizeof(runs), s) == hipSuccess &&
        hipcub::DeviceReduce::Max(temp, maxBytes, counts, small + 1, int(runs), s) == hipSuccess &&
        readBack(&most, small + 1, sizeof(most), s) == hipSuccess)
        rc = NVQ_OK;
    (void)hipFreeAsync(mem, s);
    return rc;
}

int maxPerVoxel(Workspace& W, const int32_t* ijk, uint32_t n, hipStream_t s, uint32_t& most)
{
    uint8_t* ws = W.reserve(carveBytes<int32_t>(6 * kBoundBlocks) + carveBytes<int32_t>(6));
    if (!ws) return NVQ_ERR_HIP;
    Carve c{ws};
    int32_t* part = c.take<int32_t>(6 * kBoundBlocks);
    int32_t* bnd = c.take<int32_t>(6);
    const uint32_t nb = std::min<uint32_t>(blocksFor(n, kB), kBoundBlocks);
    k_v2g_bounds<<<nb, kB, 0, s>>>(ijk, n, part);
    k_v2g_bounds_final<<<1, kB, 0, s>>>(part, nb, bnd);
    int32_t box[6];
    if (readBack(box, bnd, sizeof(box), s) != hipSuccess) return NVQ_ERR_HIP;
    Layout L{};
    uint64_t tiles = 1;
    int tileBits = 0, lowerBits = 0;
    keyLayout(box, L, tiles, tileBits, lowerBits);
    if (tiles == 1 && 21 + lowerBits <= 32) return maxPerVoxelSorted<uint32_t>(ijk, n, L, 21 + lowerBits, s, most);
    if (36 + tileBits > 64) return NVQ_ERR_INVALID_ARG; // input spans more than 2^28 upper-node tiles
    return maxPerVoxelSorted<uint64_t>(ijk, n, L, 36 + tileBits, s, most);
}

int buildFromVoxels(Workspace& W, const int32_t* ijk, uint64_t n, uint32_t grid_type, double voxel_size,
                    const char* name, uint64_t* index, const PointArgs* P, hipStream_t s, nvq_grid** out)
{
    const uint32_t n32 = uint32_t(n);

    // 1. bbox of the input, read back to size the dense tile numbering
    uint8_t* ws = W.reserve(carveBytes<int32_t>(6 * kBoundBlocks) + carveBytes<int32_t>(6));
    if (!ws) return NVQ_ERR_HIP;
    Carve c{ws};
    int32_t* part = c.take<int32_t>(6 * kBoundBlocks);
    int32_t* bnd = c.take<int32_t>(6);
    const uint32_t nb = std::min<uint32_t>(blocksFor(n, kB), kBoundBlocks);
    k_v2g_bounds<<<nb, kB, 0, s>>>(ijk, n, part);
    k_v2g_bounds_final<<<1, kB, 0, s>>>(part, nb, bnd);
    int32_t box[6];
    if (readBack(box, bnd, sizeof(box), s) != hipSuccess) return NVQ_ERR_HIP;
    Layout L{};
    uint64_t tiles = 1;
    int tileBits = 0, lowerBits = 0;
    keyLayout(box, L, tiles, tileBits, lowerBits);
    L.isIndex = grid_type == GT_ONINDEX || grid_type == kGridTypeIndex;
    L.allIndex = grid_type == kGridTypeIndex;
    L.isPoint = P != nullptr;
    int rc;
    if (tiles == 1 && 21 + lowerBits <= 32)
        rc = buildSorted<uint32_t>(W, ijk, n32, L, 21 + lowerBits, box, voxel_size, name, grid_type, index, P, s, out);
    else if (36 + tileBits > 64)
        return NVQ_ERR_INVALID_ARG; // input spans more than 2^28 upper-node tiles
    else
        rc = buildSorted<uint64_t>(W, ijk, n32, L, 36 + tileBits, box, voxel_size, name, grid_type, index, P, s, out);
    if (rc != NVQ_OK) (void)hipStreamSynchronize(s); // no kernel may still use the workspace
    return rc;
}

// pointsToGrid for float or double world points (PtrT = Vec3f* or Vec3d*).
template<typename T>
int pointsToGrid(const T* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name, void* stream,
                 nvq_grid** out)
{
    if (!out) return NVQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!xyz || n == 0 || n > uint64_t(INT32_MAX) || !(voxel_size > 0.0) || !std::isfinite(voxel_size) ||
        point_type == 0 || point_type > 9)
        return NVQ_ERR_INVALID_ARG;
    // PointType -> blind data (processGridTreeRoot, tools/cuda/PointsToGrid.cuh:913-1001); Default
    // describes the input type
    static const struct {
        uint32_t size, semantic, dataClass, dataType, gridClass;
        const char* name;
    } kinds[10] = {{0, 0, 0, 0, 0, ""},
                   {4, 6, 1, 10, 4, "PointID: uint32_t indices to points"},
                   {24, 7, 2, 7, 5, "World64: Vec3<double> point coordinates in world space"},
                   {12, 7, 2, 6, 5, "World32: Vec3<float> point coordinates in world space"},
                   {24, 8, 2, 7, 5, "Grid64: Vec3<double> point coordinates in grid space"},
                   {12, 8, 2, 6, 5, "Grid32: Vec3<float> point coordinates in grid space"},
                   {12, 9, 2, 6, 5, "Voxel32: Vec3<float> point coordinates in voxel space"},
                   {6, 9, 2, 25, 5, "Voxel16: Vec3<uint16_t> point coordinates in voxel space"},
                   {3, 9, 2, 24, 5, "Voxel8: Vec3<uint8_t> point coordinates in voxel space"},
                   {12, 7, 2, 6, 5, "World32: Vec3<float> point coordinates in world space"}};
    constexpr bool kF64 = std::is_same<T, double>::value;
    const auto& k = kinds[point_type == 9 && kF64 ? 2 : point_type];
    PointArgs P{xyz, kF64 ? 1u : 0u, point_type, k.size, k.semantic, k.dataClass, k.dataType, k.gridClass, k.name,
                1.0f / float(voxel_size), 1.0 / voxel_size};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return NVQ_ERR_NO_DEVICE;
    Workspace& W = workspaceOf(device);
    std::lock_guard<std::mutex> hold(W.lock);
    int32_t* ijk = nullptr;
    if (poolMalloc(reinterpret_cast<void**>(&ijk), 12ull * n, s) != hipSuccess) return NVQ_ERR_HIP;
    k_p2g_voxels<T><<<blocksFor(n, kB), kB, 0, s>>>(xyz, uint32_t(n), P.invF, P.invD, ijk);
    const int rc = buildFromVoxels(W, ijk, n, 23u, voxel_size, name, nullptr, &P, s, out);
    (void)hipFreeAsync(ijk, s);
    (void)hipStreamSynchronize(s);
    return rc;
}

// The voxel-size search of PointsToGrid(maxPointsPerVoxel, tolerance, maxIterations)
// (tools/cuda/PointsToGrid.cuh:335-342, countNodes :630-774): start from Map(1.0); after each
// count of the largest points-per-voxel, stop when it is within the tolerance of the target (or
// after maxIterations, or for a single point), else bracket the target between the densest
// voxel size below it and the sparsest above it and interpolate (or scale, while no size above
// it is known).  The grid is then built with the last voxel size.
template<typename T>
int pointsToGridDensity(const T* xyz, uint64_t n, uint32_t target, int tolerance, int max_iterations, uint32_t point_type,
                        const char* name, void* stream, nvq_grid** out, double* voxel_size_out, uint32_t* density_out)
{
    if (!out) return NVQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!xyz || n == 0 || n > uint64_t(INT32_MAX) || target == 0 || point_type == 0 || point_type > 9)
        return NVQ_ERR_INVALID_ARG;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return NVQ_ERR_NO_DEVICE;
    target = uint32_t(std::min<uint64_t>(target, n));
    struct Sample { // voxel size and the largest points-per-voxel it gives (Foo, :634-638)
        double dx;
        uint32_t density;
        bool operator<(const Sample& r) const { return density < r.density || (density == r.density && dx < r.dx); }
    } lo{0.0, 1u}, hi{0.0, 0u};
    double dx = 1.0;
    uint32_t density = 0;
    {
        Workspace& W = workspaceOf(device);
        std::lock_guard<std::mutex> hold(W.lock);
        int32_t* ijk = nullptr;
        if (poolMalloc(reinterpret_cast<void**>(&ijk), 12ull * n, s) != hipSuccess) return NVQ_ERR_HIP;
        int rc = NVQ_OK;
        for (int iter = 0;;) {
            if (!(dx > 0.0) || !std::isfinite(dx)) {
                rc = NVQ_ERR_INVALID_ARG;
                break;
            }
            k_p2g_voxels<T><<<blocksFor(n, kB), kB, 0, s>>>(xyz, uint32_t(n), 1.0f / float(dx), 1.0 / dx, ijk);
            if ((rc = maxPerVoxel(W, ijk, uint32_t(n), s, density)) != NVQ_OK) break;
            if (++iter >= max_iterations || n == 1 || std::abs(int(density) - int(target)) <= tolerance) break;
            const Sample cur{dx, density};
            if (density < target) {
                if (lo < cur) lo = cur;
            } else if (hi.density == 0 || cur < hi) {
                hi = cur;
            }
            if (hi.density)
                dx = (lo.dx * double(hi.density - target) + hi.dx * double(target - lo.density)) / double(hi.density - lo.density);
            else if (density > 1u)
                dx *= (target - 1.0) / (density - 1.0);
            else
                dx *= 10.0;
        }
        (void)hipFreeAsync(ijk, s);
        (void)hipStreamSynchronize(s);
        if (rc != NVQ_OK) return rc;
    }
    if (voxel_size_out) *voxel_size_out = dx;
    if (density_out) *density_out = density;
    return pointsToGrid(xyz, n, dx, point_type, name, stream, out);
}

} // namespace

extern "C" {

int nvq_voxels_to_grid(const int32_t* ijk, uint64_t n, uint32_t grid_type, double voxel_size, const char* name,
                       uint64_t* index, void* stream, nvq_grid** out)
{
    if (!out) return NVQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!ijk || n == 0 || n > uint64_t(INT32_MAX) || !(voxel_size > 0.0) || !std::isfinite(voxel_size))
        return NVQ_ERR_INVALID_ARG;
    if (grid_type != GT_FLOAT && grid_type != GT_ONINDEX && grid_type != kGridTypeIndex) return NVQ_ERR_UNSUPPORTED_TYPE;
    if (index && grid_type == GT_FLOAT) return NVQ_ERR_INVALID_ARG;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return NVQ_ERR_NO_DEVICE;
    Workspace& W = workspaceOf(device);
    std::lock_guard<std::mutex> hold(W.lock);
    return buildFromVoxels(W, ijk, n, grid_type, voxel_size, name, index, nullptr, static_cast<hipStream_t>(stream), out);
}

int nvq_voxels_to_grid_fp(const void* xyz, int xyz_f64, uint64_t n, uint32_t grid_type, double voxel_size,
                          const char* name, uint64_t* index, void* stream, nvq_grid** out)
{
    if (!out) return NVQ_ERR_INVALID_ARG;
    *out = nullptr;
    if (!xyz || n == 0 || n > uint64_t(INT32_MAX) || !(voxel_size > 0.0) || !std::isfinite(voxel_size))
        return NVQ_ERR_INVALID_ARG;
    if (grid_type != GT_FLOAT && grid_type != GT_ONINDEX && grid_type != kGridTypeIndex) return NVQ_ERR_UNSUPPORTED_TYPE;
    if (index && grid_type == GT_FLOAT) return NVQ_ERR_INVALID_ARG;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return NVQ_ERR_NO_DEVICE;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    Workspace& W = workspaceOf(device);
    std::lock_guard<std::mutex> hold(W.lock);
    int32_t* ijk = nullptr;
    if (poolMalloc(reinterpret_cast<void**>(&ijk), 12ull * n, s) != hipSuccess) return NVQ_ERR_HIP;
    if (xyz_f64)
        k_round_voxels<double><<<blocksFor(n, kB), kB, 0, s>>>(static_cast<const double*>(xyz), uint32_t(n), ijk);
    else
        k_round_voxels<float><<<blocksFor(n, kB), kB, 0, s>>>(static_cast<const float*>(xyz), uint32_t(n), ijk);
    const int rc = buildFromVoxels(W, ijk, n, grid_type, voxel_size, name, index, nullptr, s, out);
    (void)hipFreeAsync(ijk, s);
    (void)hipStreamSynchronize(s);
    return rc;
}

int nvq_points_to_grid(const float* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                       void* stream, nvq_grid** out)
{
    return pointsToGrid(xyz, n, voxel_size, point_type, name, stream, out);
}

int nvq_points_to_grid_f64(const double* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                           void* stream, nvq_grid** out)
{
    return pointsToGrid(xyz, n, voxel_size, point_type, name, stream, out);
}

int nvq_points_to_grid_density(const void* xyz, int xyz_f64, uint64_t n, uint32_t max_points_per_voxel, int tolerance,
                               int max_iterations, uint32_t point_type, const char* name, void* stream, nvq_grid** out,
                               double* voxel_size, uint32_t* points_per_voxel)
{
    if (xyz_f64)
        return pointsToGridDensity(static_cast<const double*>(xyz), n, max_points_per_voxel, tolerance, max_iterations,
                                   point_type, name, stream, out, voxel_size, points_per_voxel);
    return pointsToGridDensity(static_cast<const float*>(xyz), n, max_points_per_voxel, tolerance, max_iterations,
                               point_type, name, stream, out, voxel_size, points_per_voxel);
}

int nvq_grid_download(const nvq_grid* grid, void* host_dst, uint64_t bytes, void* stream)
{
    nvq_grid_info info;
    if (!grid || !host_dst || nvq_grid_get_info(grid, &info) != NVQ_OK || bytes > info.bytes) return NVQ_ERR_INVALID_ARG;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(host_dst, info.device_buffer, bytes, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return NVQ_ERR_HIP;
    return NVQ_OK;
}

} // extern "C"
