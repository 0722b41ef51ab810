// Device-side grid statistics: tools::cuda::updateGridStats<float>(grid, StatsMode::Default)
// (tools/cuda/GridStats.cuh:343-362).  Bboxes, min / max / average / standard deviation of
// every node, render flags, the root and the grid's world bbox, recomputed in place.
//
// The merge order is the reference device code's, so the results are its bits:
//  - one 32-lane half of a wavefront per leaf (processLeaf :86-153): lane L folds voxels L,
//    L+32, ... then the lanes merge pairwise through shuffles (d = 16 ... 1);
//  - one 128-thread workgroup per internal node (processInternal :155-241): thread t folds
//    entries t, t+128, ..., then a pairwise tree in LDS (d = 64 ... 1);
//  - one thread for the root and the grid (processRootAndGrid :243-306).
// Statistics are FP64 Welford accumulators (Stats, tools/GridStats.h:244-344).  Each node's
// accumulator goes to a per-level array its parent reads by child index.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "../../include/nvq.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>

using namespace nvq;

namespace {

constexpr uint64_t kLeafBytes = 2144, kLowerBytes = 33856, kUpperBytes = 270400;

struct Acc {
    double avg, aux;
    unsigned long long size;
    float mn, mx;
};

__device__ __forceinline__ Acc accEmpty() { return Acc{0.0, 0.0, 0ull, FLT_MAX, -FLT_MAX}; }

__device__ __forceinline__ void extAdd(Acc& s, float v)
{
    if (v < s.mn) s.mn = v;
    if (v > s.mx) s.mx = v;
}
// Extrema::add(const Extrema&): min of the minima, max of the maxima
__device__ __forceinline__ void extMerge(Acc& s, const Acc& o)
{
    if (o.mn < s.mn) s.mn = o.mn;
    if (o.mx > s.mx) s.mx = o.mx;
}

// StatsMode (tools/GridStats.h:40-47): what a pass writes besides bboxes and render flags
enum StatsModeId : int { kModeBBox = 1, kModeMinMax = 2, kModeAll = 3 };
// Stats::add(value)
__device__ __forceinline__ void accAdd(Acc& s, float v)
{
    extAdd(s, v);
    s.size += 1ull;
    const double delta = double(v) - s.avg;
    s.avg += delta / double(s.size);
    s.aux += delta * (double(v) - s.avg);
}
// Stats::add(value, n)
__device__ __forceinline__ void accAddN(Acc& s, float v, unsigned long long n)
{
    const double denom = 1.0 / double(s.size + n);
    const double delta = double(v) - s.avg;
    s.avg += denom * delta * double(n);
    s.aux += denom * delta * delta * double(s.size) * double(n);
    extAdd(s, v);
    s.size += n;
}
// Stats::add(other)
__device__ __forceinline__ void accMerge(Acc& s, const Acc& o)
{
    if (o.size > 0ull) {
        const double denom = 1.0 / double(s.size + o.size);
        const double delta = o.avg - s.avg;
        s.avg += denom * delta * double(o.size);
        s.aux += o.aux + denom * delta * delta * double(s.size) * double(o.size);
        if (o.mn < s.mn) s.mn = o.mn;
        if (o.mx > s.mx) s.mx = o.mx;
        s.size += o.size;
    }
}
__device__ __forceinline__ float accStd(const Acc& s) { return float(sqrt(s.size < 2ull ? 0.0 : s.aux / double(s.size))); }

__device__ __forceinline__ void putStats(uint8_t* p, const Acc& s)
{
    float4 v;
    v.x = s.mn;
    v.y = s.mx;
    v.z = float(s.avg);
    v.w = accStd(s);
    *reinterpret_cast<float4*>(p) = v;
}

__device__ __forceinline__ void boxEmpty(int32_t b[6])
{
    b[0] = b[1] = b[2] = INT_MAX;
    b[3] = b[4] = b[5] = INT_MIN;
}
__device__ __forceinline__ void boxExpand(int32_t b[6], const int32_t o[6])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b[a] = min(b[a], o[a]);
        b[3 + a] = max(b[3 + a], o[3 + a]);
    }
}

// LeafNode::bbox (NanoVDB.h:4442-4453): empty unless the HasBBox flag bit is on
__device__ __forceinline__ void leafBox(const uint8_t* F, int32_t b[6])
{
    const int4 h = *reinterpret_cast<const int4*>(F);
    const uint32_t d = uint32_t(h.w);
    if (!((d >> 24) & 2u)) {
        boxEmpty(b);
        return;
    }
    b[0] = h.x;
    b[1] = h.y;
    b[2] = h.z;
    b[3] = h.x + int32_t(d & 255u);
    b[4] = h.y + int32_t((d >> 8) & 255u);
    b[5] = h.z + int32_t((d >> 16) & 255u);
}

// One 32-lane half-wavefront per leaf.
__global__ __launch_bounds__(256) void k_stats_leaf(uint8_t* __restrict__ leaf0, uint32_t count, Acc* __restrict__ acc,
                                                    int mode)
{
    const uint32_t lane = threadIdx.x & 31u, l = blockIdx.x * 8u + (threadIdx.x >> 5);
    const bool valid = l < count;
    uint8_t* F = leaf0 + kLeafBytes * (valid ? l : 0u);
    uint64_t w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = valid ? *reinterpret_cast<const uint64_t*>(F + 16 + 8 * j) : 0ull;
    Acc s = accEmpty();
    const float* v = reinterpret_cast<const float*>(F + 96);
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        const uint32_t i = lane + 32u * k;
        if ((w[k >> 1] >> (i & 63u)) & 1ull) accAdd(s, v[i]);
    }
#pragma unroll
    for (int d = 16; d >= 1; d >>= 1) {
        Acc o;
        o.avg = __shfl_down(s.avg, d, 32);
        o.aux = __shfl_down(s.aux, d, 32);
        o.size = __shfl_down(s.size, d, 32);
        o.mn = __shfl_down(s.mn, d, 32);
        o.mx = __shfl_down(s.mx, d, 32);
        if (lane < uint32_t(d)) accMerge(s, o);
    }
    if (lane != 0 || !valid) return;
    uint64_t word = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) word |= w[j];
    uint32_t flags = F[15];
    if (word) {
        // LeafNode::updateBBox (NanoVDB.h:4596-4624)
        uint32_t xmin = w[0] ? 0u : 8u, xmax = xmin;
#pragma unroll
        for (uint32_t j = 1; j < 8; ++j)
            if (w[j]) {
                if (xmin == 8u) xmin = j;
                xmax = j;
            }
        const uint32_t ymin = uint32_t(__builtin_ctzll(word)) >> 3, ymax = (63u - uint32_t(__builtin_clzll(word))) >> 3;
        const uint32_t w32 = uint32_t(word) | uint32_t(word >> 32), w16 = (w32 & 0xffffu) | (w32 >> 16);
        const uint32_t byte = (w16 & 0xffu) | (w16 >> 8);
        const uint32_t zmin = uint32_t(__builtin_ctz(byte)), zmax = 31u - uint32_t(__builtin_clz(byte));
        const int4 h0 = *reinterpret_cast<const int4*>(F);
        flags = (flags | 2u) & ~1u;
        int4 h;
        h.x = (h0.x & ~7) + int32_t(xmin);
        h.y = (h0.y & ~7) + int32_t(ymin);
        h.z = (h0.z & ~7) + int32_t(zmin);
        h.w = int32_t((xmax - xmin) | ((ymax - ymin) << 8) | ((zmax - zmin) << 16) | (flags << 24));
        *reinterpret_cast<int4*>(F) = h;
        if (mode == kModeAll) {
            putStats(F + 80, s);
        } else if (mode == kModeMinMax) { // Extrema::setStats: minimum and maximum only
            reinterpret_cast<float*>(F + 80)[0] = s.mn;
            reinterpret_cast<float*>(F + 80)[1] = s.mx;
        }
        acc[l] = s;
    } else {
        F[15] = uint8_t(flags & ~3u);
        acc[l] = accEmpty();
    }
}

// One 128-thread workgroup per internal node; LOG2C / LOG2N are the child's total and the node's
// own log2 dimensions (3 / 4 for lower nodes, 7 / 5 for upper nodes).
template<uint32_t LOG2C, uint32_t LOG2N, bool CHILD_IS_LEAF>
__global__ __launch_bounds__(128) void k_stats_internal(uint8_t* __restrict__ node0, const uint8_t* __restrict__ child0,
                                                        uint32_t count, const Acc* __restrict__ childAcc,
                                                        Acc* __restrict__ acc, int mode)
{
    constexpr uint32_t SIZE = 1u << (3 * LOG2N);
    constexpr uint64_t kNode = LOG2N == 4 ? kLowerBytes : kUpperBytes, kChild = CHILD_IS_LEAF ? kLeafBytes : kLowerBytes;
    constexpr uint32_t kChildMask = 32u + SIZE / 8u, kTable = kChildMask + SIZE / 8u + 32u, kStats = kTable - 32u;
    __shared__ Acc sAcc[128];
    __shared__ int32_t sBox[128][6];
    const uint32_t t = threadIdx.x, k = blockIdx.x;
    if (k >= count) return;
    uint8_t* N = node0 + kNode * k;
    const int32_t* b0 = reinterpret_cast<const int32_t*>(N);
    constexpr int32_t omask = ~int32_t((1u << (LOG2C + LOG2N)) - 1u);
    const int32_t ox = b0[0] & omask, oy = b0[1] & omask, oz = b0[2] & omask; // InternalNode::origin
    const uint64_t* vmask = reinterpret_cast<const uint64_t*>(N + 32);
    const uint64_t* cmask = reinterpret_cast<const uint64_t*>(N + kChildMask);
    Acc s = accEmpty();
    int32_t box[6];
    boxEmpty(box);
    for (uint32_t i = t; i < SIZE; i += 128u) {
        const uint64_t bit = 1ull << (i & 63u);
        const uint8_t* e = N + kTable + 8ull * i;
        if (cmask[i >> 6] & bit) {
            const uint8_t* C = N + *reinterpret_cast<const int64_t*>(e);
            const uint64_t ci = uint64_t(C - child0) / kChild;
            int32_t cb[6];
            if (CHILD_IS_LEAF) {
                leafBox(C, cb);
            } else {
#pragma unroll
                for (int a = 0; a < 6; ++a) cb[a] = reinterpret_cast<const int32_t*>(C)[a];
            }
            boxExpand(box, cb);
            if (mode == kModeAll) {
                accMerge(s, childAcc[ci]);
            } else if (mode == kModeMinMax) { // Extrema::add(child.minimum()), add(child.maximum())
                constexpr uint32_t kChildStats = CHILD_IS_LEAF ? 80u : 1056u;
                const float* cs = reinterpret_cast<const float*>(C + kChildStats);
                extAdd(s, cs[0]);
                extAdd(s, cs[1]);
            }
        } else if (vmask[i >> 6] & bit) {
            constexpr uint32_t m = (1u << LOG2N) - 1u;
            const int32_t c[3] = {ox + int32_t(((i >> (2 * LOG2N)) & m) << LOG2C), oy + int32_t(((i >> LOG2N) & m) << LOG2C),
                                  oz + int32_t((i & m) << LOG2C)};
            const int32_t tb[6] = {c[0], c[1], c[2], c[0] + int32_t(1u << LOG2C) - 1, c[1] + int32_t(1u << LOG2C) - 1,
                                   c[2] + int32_t(1u << LOG2C) - 1};
            boxExpand(box, tb);
            if (mode == kModeAll) accAddN(s, *reinterpret_cast<const float*>(e), 1ull << (3 * LOG2C));
            else if (mode == kModeMinMax) extAdd(s, *reinterpret_cast<const float*>(e));
        }
    }
    sAcc[t] = s;
#pragma unroll
    for (int a = 0; a < 6; ++a) sBox[t][a] = box[a];
    __syncthreads();
    for (uint32_t d = 64; d; d >>= 1) {
        if (t < d) {
            if (mode == kModeAll) accMerge(sAcc[t], sAcc[t + d]);
            else extMerge(sAcc[t], sAcc[t + d]);
            boxExpand(sBox[t], sBox[t + d]);
        }
        __syncthreads();
    }
    if (t == 0) {
        int32_t* bo = reinterpret_cast<int32_t*>(N);
#pragma unroll
        for (int a = 0; a < 6; ++a) bo[a] = sBox[0][a];
        if (mode == kModeAll) {
            putStats(N + kStats, sAcc[0]);
        } else if (mode == kModeMinMax) {
            reinterpret_cast<float*>(N + kStats)[0] = sAcc[0].mn;
            reinterpret_cast<float*>(N + kStats)[1] = sAcc[0].mx;
        }
        uint64_t* flags = reinterpret_cast<uint64_t*>(N + 24);
        *flags = *flags & ~1ull;
        acc[k] = sAcc[0];
    }
}

// Root (tiles in table order) and grid: one thread.
__global__ void k_stats_root(uint8_t* __restrict__ base, int64_t rootOff, const uint8_t* __restrict__ upper0,
                             const Acc* __restrict__ upperAcc, int mode)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint8_t* R = base + kGridSize + rootOff;
    const uint32_t tableSize = *reinterpret_cast<const uint32_t*>(R + 24);
    int32_t rb[6];
    boxEmpty(rb);
    float* st = reinterpret_cast<float*>(R + 32);
    if (tableSize == 0) {
        const float bg = *reinterpret_cast<const float*>(R + 28);
        st[0] = st[1] = bg;
        st[2] = st[3] = 0.0f;
    } else {
        Acc s = accEmpty();
        for (uint32_t i = 0; i < tableSize; ++i) {
            const uint8_t* tile = R + 64 + 32ull * i;
            const int64_t child = *reinterpret_cast<const int64_t*>(tile + 8);
            if (child) {
                const int32_t* cb = reinterpret_cast<const int32_t*>(R + child);
                boxExpand(rb, cb);
                if (mode == kModeAll) {
                    accMerge(s, upperAcc[uint64_t((R + child) - upper0) / kUpperBytes]);
                } else if (mode == kModeMinMax) {
                    const float* cs = reinterpret_cast<const float*>(R + child + 8224);
                    extAdd(s, cs[0]);
                    extAdd(s, cs[1]);
                }
            } else if (*reinterpret_cast<const uint32_t*>(tile + 16)) {
                // RootData::KeyToCoord (NanoVDB.h:2641-2648)
                const uint64_t key = *reinterpret_cast<const uint64_t*>(tile);
                const int32_t c[3] = {int32_t(uint32_t((key >> 42) & 0x1fffffu) << 12),
                                      int32_t(uint32_t((key >> 21) & 0x1fffffu) << 12), int32_t(uint32_t(key & 0x1fffffu) << 12)};
                const int32_t tb[6] = {c[0], c[1], c[2], c[0] + 4095, c[1] + 4095, c[2] + 4095};
                boxExpand(rb, tb);
                if (mode == kModeAll) accAddN(s, *reinterpret_cast<const float*>(tile + 20), 1ull << 36);
                else if (mode == kModeMinMax) extAdd(s, *reinterpret_cast<const float*>(tile + 20));
            }
        }
        if (mode >= kModeMinMax) {
            st[0] = s.mn;
            st[1] = s.mx;
        }
        if (mode == kModeAll) {
            st[2] = float(s.avg);
            st[3] = accStd(s);
        }
    }
    int32_t* rbo = reinterpret_cast<int32_t*>(R);
    for (int a = 0; a < 6; ++a) rbo[a] = rb[a];
    uint32_t* gflags = reinterpret_cast<uint32_t*>(base + 20);
    double* wb = reinterpret_cast<double*>(base + 560);
    uint32_t flags = *gflags;
    if (rb[0] > rb[3] || rb[1] > rb[4] || rb[2] > rb[5]) {
        for (int a = 0; a < 3; ++a) {
            wb[a] = DBL_MAX;
            wb[3 + a] = -DBL_MAX;
        }
        flags &= ~2u;
    } else {
        // CoordBBox(min, max + 1).transform(map) (math/Math.h:2070-2083), Map::applyMap
        const double* mat = reinterpret_cast<const double*>(base + kGridMap + 88);
        const double* vec = reinterpret_cast<const double*>(base + kGridMap + 232);
        double out[6];
        for (int c = 0; c < 8; ++c) {
            const double x[3] = {double(c & 1 ? rb[3] + 1 : rb[0]), double(c & 2 ? rb[4] + 1 : rb[1]),
                                 double(c & 4 ? rb[5] + 1 : rb[2])};
            for (int r = 0; r < 3; ++r) {
                const double y = __builtin_fma(x[0], mat[3 * r], __builtin_fma(x[1], mat[3 * r + 1], __builtin_fma(x[2], mat[3 * r + 2], vec[r])));
                out[r] = c == 0 ? y : (y < out[r] ? y : out[r]);
                out[3 + r] = c == 0 ? y : (y > out[3 + r] ? y : out[3 + r]);
            }
        }
        for (int a = 0; a < 6; ++a) wb[a] = out[a];
        flags |= 2u;
    }
    // HasMinMax | HasAverage | HasStdDeviation as the mode computes them (:299-301)
    flags &= ~(4u | 8u | 16u);
    if (mode >= kModeMinMax) flags |= 4u;
    if (mode == kModeAll) flags |= 8u | 16u;
    *gflags = flags;
}

} // namespace

extern "C" {

int nvq_update_grid_stats_mode(nvq_grid* grid, int mode, void* stream)
{
    // StatsMode::Disable with a grid is an error in the reference (:346-358)
    if (!grid || mode < kModeBBox || mode > kModeAll) return NVQ_ERR_INVALID_ARG;
    const GridLayout L = gridLayout(grid);
    if (L.type != GT_FLOAT) return NVQ_ERR_UNSUPPORTED_TYPE;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const uint64_t slots = uint64_t(L.count[0]) + L.count[1] + L.count[2] + 1;
    Acc* acc = nullptr;
    if (poolMalloc(reinterpret_cast<void**>(&acc), slots * sizeof(Acc), s) != hipSuccess) return NVQ_ERR_HIP;
    Acc* leafAcc = acc;
    Acc* lowerAcc = leafAcc + L.count[0];
    Acc* upperAcc = lowerAcc + L.count[1];
    uint8_t* tree = L.base + kGridSize;
    uint8_t* leaf0 = tree + L.nodeOff[0];
    uint8_t* lower0 = tree + L.nodeOff[1];
    uint8_t* upper0 = tree + L.nodeOff[2];
    if (L.count[0]) k_stats_leaf<<<(L.count[0] + 7u) / 8u, 256, 0, s>>>(leaf0, L.count[0], leafAcc, mode);
    if (L.count[1])
        k_stats_internal<3, 4, true><<<L.count[1], 128, 0, s>>>(lower0, leaf0, L.count[1], leafAcc, lowerAcc, mode);
    if (L.count[2])
        k_stats_internal<7, 5, false><<<L.count[2], 128, 0, s>>>(upper0, lower0, L.count[2], lowerAcc, upperAcc, mode);
    k_stats_root<<<1, 64, 0, s>>>(L.base, L.nodeOff[3], upper0, upperAcc, mode);
    const bool launched = hipGetLastError() == hipSuccess;
    (void)hipFreeAsync(acc, s);
    int rc = (launched && hipStreamSynchronize(s) == hipSuccess) ? NVQ_OK : NVQ_ERR_HIP;
    if (rc == NVQ_OK) rc = gridRefreshBBox(grid, s);
    return rc;
}

int nvq_update_grid_stats(nvq_grid* grid, void* stream) { return nvq_update_grid_stats_mode(grid, kModeAll, stream); }

} // extern "C"
