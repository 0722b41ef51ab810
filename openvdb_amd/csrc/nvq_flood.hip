// Device-side signed flood fill: tools::cuda::signedFloodFill<float>(grid)
// (tools/cuda/SignedFloodFill.cuh:240-253), in place on a float level set's device buffer.
//  - leaves (processLeaf :172-185): one wavefront per leaf, 8 voxels per lane; an inactive
//    voxel takes +-background with the sign of the next active voxel in the leaf, else the
//    previous one, else voxel 0;
//  - lower, then upper nodes (processNode :149-170): one thread per table entry; an entry that
//    is not a child takes +-background with the sign of the next child's first value, else the
//    previous child's last value, else entry 0;
//  - root (processRoot :91-145, host code in the reference): one thread per root child finds
//    its successor on the same z-scanline (the next entry of the lexicographic order) and, when
//    both facing values are <= 0, makes the inactive tiles between them -background.
// Writes that race with reads in these passes only ever replace a value by one of the same sign,
// so the result is deterministic.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "../../include/nvq.h"

#include <hip/hip_runtime.h>

#include <cstring>

using namespace nvq;

namespace {

constexpr uint64_t kLeafBytes = 2144, kLowerBytes = 33856, kUpperBytes = 270400;

// Mask::findNext<true> / findPrev<true> (NanoVDB.h:1385-1409) over `words` 64-bit words
__device__ __forceinline__ uint32_t findNextOn(const uint64_t* w, uint32_t words, uint32_t start)
{
    uint32_t n = start >> 6;
    if (n >= words) return words * 64u;
    uint64_t b = w[n];
    if (b & (1ull << (start & 63u))) return start;
    b &= ~0ull << (start & 63u);
    while (!b && ++n < words) b = w[n];
    return b ? (n << 6) + uint32_t(__builtin_ctzll(b)) : words * 64u;
}
__device__ __forceinline__ uint32_t findPrevOn(const uint64_t* w, uint32_t words, uint32_t start)
{
    uint32_t n = start >> 6;
    if (n >= words) return words * 64u;
    uint64_t b = w[n];
    if (b & (1ull << (start & 63u))) return start;
    b &= (1ull << (start & 63u)) - 1ull;
    while (!b && n) b = w[--n];
    return b ? (n << 6) + 63u - uint32_t(__builtin_clzll(b)) : words * 64u;
}

__global__ __launch_bounds__(256) void k_sff_leaf(uint8_t* __restrict__ leaf0, uint32_t count, const float* __restrict__ bgp)
{
    const float bg = *bgp;
    const uint32_t lane = threadIdx.x & 63u, l = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (l >= count) return;
    uint8_t* F = leaf0 + kLeafBytes * l;
    const uint64_t* m = reinterpret_cast<const uint64_t*>(F + 16);
    float* v = reinterpret_cast<float*>(F + 96);
    const uint64_t word = m[lane >> 3];
    for (uint32_t e = 0; e < 8; ++e) {
        const uint32_t n = lane * 8u + e;
        if ((word >> (n & 63u)) & 1ull) continue;
        uint32_t k = findNextOn(m, 8, n);
        if (k == 512u && (k = findPrevOn(m, 8, n)) == 512u) k = 0;
        v[n] = v[k] < 0.0f ? -bg : bg;
    }
}

// InternalNode::getFirstValue / getLastValue (NanoVDB.h:3529-3541) of a lower node, and
// LeafNode::getFirstValue / getLastValue (:4476-4478)
__device__ __forceinline__ float leafEdge(const uint8_t* F, bool last)
{
    return reinterpret_cast<const float*>(F + 96)[last ? 511 : 0];
}
__device__ __forceinline__ float lowerEdge(const uint8_t* N, bool last)
{
    const uint32_t n = last ? 4095u : 0u;
    const uint64_t cm = reinterpret_cast<const uint64_t*>(N + 544)[n >> 6];
    const int64_t e = reinterpret_cast<const int64_t*>(N + 1088)[n];
    if ((cm >> (n & 63u)) & 1ull) return leafEdge(N + e, last);
    return *reinterpret_cast<const float*>(N + 1088 + 8ull * n);
}
__device__ __forceinline__ float upperEdge(const uint8_t* N, bool last)
{
    const uint32_t n = last ? 32767u : 0u;
    const uint64_t cm = reinterpret_cast<const uint64_t*>(N + 4128)[n >> 6];
    const int64_t e = reinterpret_cast<const int64_t*>(N + 8256)[n];
    if ((cm >> (n & 63u)) & 1ull) return lowerEdge(N + e, last);
    return *reinterpret_cast<const float*>(N + 8256 + 8ull * n);
}

// One thread per table entry of the nodes of one level (LOG2N 4: lower, 5: upper).
template<uint32_t LOG2N>
__global__ __launch_bounds__(256) void k_sff_node(uint8_t* __restrict__ node0, uint32_t count, const float* __restrict__ bgp)
{
    const float bg = *bgp;
    constexpr uint32_t SIZE = 1u << (3 * LOG2N), WORDS = SIZE / 64u;
    constexpr uint64_t kNode = LOG2N == 4 ? kLowerBytes : kUpperBytes;
    constexpr uint32_t kChildMask = 32u + SIZE / 8u, kTable = kChildMask + SIZE / 8u + 32u;
    const uint64_t tid = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (tid >= uint64_t(count) * SIZE) return;
    const uint32_t v = uint32_t(tid & (SIZE - 1u));
    uint8_t* N = node0 + kNode * (tid >> (3 * LOG2N));
    const uint64_t* cm = reinterpret_cast<const uint64_t*>(N + kChildMask);
    if ((cm[v >> 6] >> (v & 63u)) & 1ull) return;
    const int64_t* table = reinterpret_cast<const int64_t*>(N + kTable);
    auto edge = [&](uint32_t n, bool last) {
        const uint8_t* C = N + table[n];
        return LOG2N == 4 ? leafEdge(C, last) : lowerEdge(C, last);
    };
    float value = bg;
    uint32_t n = findNextOn(cm, WORDS, v);
    if (n < SIZE) {
        if (edge(n, false) < 0.0f) value = -bg;
    } else if ((n = findPrevOn(cm, WORDS, v)) < SIZE) {
        if (edge(n, true) < 0.0f) value = -bg;
    } else if (*reinterpret_cast<const float*>(N + kTable) < 0.0f) {
        value = -bg;
    }
    *reinterpret_cast<float*>(N + kTable + 8ull * v) = value;
}

__device__ __forceinline__ void keyToCoord(uint64_t key, int32_t c[3])
{
    c[0] = int32_t(uint32_t((key >> 42) & 0x1fffffu) << 12);
    c[1] = int32_t(uint32_t((key >> 21) & 0x1fffffu) << 12);
    c[2] = int32_t(uint32_t(key & 0x1fffffu) << 12);
}

// One thread per root tile; tiles that are children look for their successor on their z-scanline.
__global__ __launch_bounds__(256) void k_sff_root(uint8_t* __restrict__ root, uint32_t tableSize, const float* __restrict__ bgp,
                                                  int* __restrict__ missing)
{
    const float bg = *bgp;
    const uint32_t a = blockIdx.x * 256u + threadIdx.x;
    if (a >= tableSize) return;
    const uint8_t* ta = root + 64 + 32ull * a;
    const int64_t childA = *reinterpret_cast<const int64_t*>(ta + 8);
    if (!childA) return;
    int32_t ca[3];
    keyToCoord(*reinterpret_cast<const uint64_t*>(ta), ca);
    // successor in lexicographic order among the children on the same (x, y) line
    uint32_t best = tableSize;
    int32_t bestZ = 0;
    for (uint32_t j = 0; j < tableSize; ++j) {
        const uint8_t* tj = root + 64 + 32ull * j;
        if (!*reinterpret_cast<const int64_t*>(tj + 8)) continue;
        int32_t c[3];
        keyToCoord(*reinterpret_cast<const uint64_t*>(tj), c);
        if (c[0] != ca[0] || c[1] != ca[1] || c[2] <= ca[2]) continue;
        if (best == tableSize || c[2] < bestZ) {
            best = j;
            bestZ = c[2];
        }
    }
    if (best == tableSize || bestZ - ca[2] == 4096) return;
    const uint8_t* tb = root + 64 + 32ull * best;
    if (upperEdge(root + childA, true) > 0.0f || upperEdge(root + *reinterpret_cast<const int64_t*>(tb + 8), false) > 0.0f)
        return;
    for (int32_t z = ca[2] + 4096; z != bestZ; z += 4096) {
        // RootData::CoordToKey (NanoVDB.h:2633-2640), then RootNode::probeTile
        const uint64_t key = uint64_t(uint32_t(z) >> 12) | (uint64_t(uint32_t(ca[1]) >> 12) << 21) |
                             (uint64_t(uint32_t(ca[0]) >> 12) << 42);
        uint32_t t = 0;
        while (t < tableSize && *reinterpret_cast<const uint64_t*>(root + 64 + 32ull * t) != key) ++t;
        if (t == tableSize) {
            atomicOr(missing, 1);
            return;
        }
        uint8_t* tile = root + 64 + 32ull * t;
        *reinterpret_cast<int64_t*>(tile + 8) = 0;    // Tile::setValue(c, false, -background)
        *reinterpret_cast<uint32_t*>(tile + 16) = 0u;
        *reinterpret_cast<float*>(tile + 20) = -bg;
    }
}

__global__ void k_sff_flag_clear(int* flag)
{
    if (threadIdx.x == 0) *flag = 0;
}

} // namespace

extern "C" {

int nvq_signed_flood_fill(nvq_grid* grid, void* stream)
{
    if (!grid) return NVQ_ERR_INVALID_ARG;
    const GridLayout L = gridLayout(grid);
    if (L.type != GT_FLOAT) return NVQ_ERR_UNSUPPORTED_TYPE;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    uint8_t* tree = L.base + kGridSize;
    uint8_t* root = tree + L.nodeOff[3];
    // RootData::mBackground, read by the kernels on the device
    const float* bg = reinterpret_cast<const float*>(root + kRootBackground);
    int* missing = nullptr;
    if (poolMalloc(reinterpret_cast<void**>(&missing), sizeof(int), s) != hipSuccess) return NVQ_ERR_HIP;
    k_sff_flag_clear<<<1, 64, 0, s>>>(missing);
    if (L.count[0]) k_sff_leaf<<<(L.count[0] + 3u) / 4u, 256, 0, s>>>(tree + L.nodeOff[0], L.count[0], bg);
    if (L.count[1]) k_sff_node<4><<<uint32_t((uint64_t(L.count[1]) * 4096u + 255u) / 256u), 256, 0, s>>>(tree + L.nodeOff[1], L.count[1], bg);
    if (L.count[2]) k_sff_node<5><<<uint32_t((uint64_t(L.count[2]) * 32768u + 255u) / 256u), 256, 0, s>>>(tree + L.nodeOff[2], L.count[2], bg);
    if (L.tableSize) k_sff_root<<<(L.tableSize + 255u) / 256u, 256, 0, s>>>(root, L.tableSize, bg, missing);
    int flag = 0;
    int rc = NVQ_ERR_HIP;
    const void* st = hipGetLastError() == hipSuccess ? stageCopy(missing, sizeof(int), s) : nullptr;
    (void)hipFreeAsync(missing, s);
    if (st && hipStreamSynchronize(s) == hipSuccess) {
        std::memcpy(&flag, st, sizeof(int));
        rc = flag ? NVQ_ERR_INVALID_ARG : NVQ_OK;
    }
    return rc;
}

} // extern "C"
