// Large-batch point sampling: spatial binning + LDS-staged voxel bricks, and its adjoint.
//
// Random query points touch a dozen distinct cache lines each over a grid that does not fit in
// L2, so a lane-per-query kernel is bound by line traffic from the last-level cache.  For large
// batches (forward):
//   k_slot_table   : every 8^3 slot of the binned region: leaf bit (a 1-bit map) and the tile or
//                    background value of slots without a leaf
//   k_bin_classify : per query, whether a slot of its [cell-1, cell+2]^3 block holds a leaf (map
//                    staged in LDS); marks those queries per chunk of 2048 and counts them
//   scan + compact : the marked queries' bin keys (16^3 bins over the grid's active bbox grown by
//                    2 voxels; points outside are clamped into the border bins) and points
//                    {index-space xyz, original index}, in chunk order; the call waits once, to
//                    size the sort
//   radix sort     : (key, point) pairs of the marked queries (rocPRIM/hipCUB DeviceRadixSort)
//   k_bin_bounds   : [begin, end) of every non-empty bin in the sorted order
//   k_bin_sample   : one 512-thread workgroup per bin stages the 20^3 brick of voxel values around the bin in
//                    LDS (ReadAccessor::getValue semantics, so results are bit-identical) and
//                    evaluates the bin's queries from it; each query's results go to its original
//                    index as one 64-byte record, written by the four lanes of a quad
//   k_sample_list  : queries whose stencil leaves the brick run through the lane accessor
//   k_bin_finish   : the outputs in caller order: marked queries copy their record, the others
//                    (their block has no leaf) are evaluated from the slot table's values
// On MI355X a random store is cheap only when one instruction writes a whole aligned 64-byte block
// (2^26 random 32-byte records: 2.9 ms; 64-byte records written by lane quads: 1.3 ms), which is
// why records are 64 bytes and the compaction writes in item/wave/lane order.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "nvq_sample.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>

namespace nvq {
namespace {

constexpr int kThreads = 256;
constexpr int kBinThreads = 512;        // one bin per workgroup: 8 waves share the 32 KB brick
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

// One query's results at its original index: 64 bytes although 20 are used.  On MI355X a random
// write is cheap only when one store instruction covers a whole aligned 64-byte block: 2^26 random
// records take 1.3 ms written that way, against 2.9 ms for 32-byte records (a partial block is
// read, merged and written back) and 2.8 ms for 64-byte records one lane writes with four stores.
// So the four lanes of a quad write the four 16-byte quarters of one record per instruction
// (storeRecQuad).  k_unpack reads only the first 32 bytes.
struct alignas(64) Rec {
    float4 a;    // nn, val, grad x, grad y
    float4 b;    // grad z, unused
    float4 c, d; // unused
};

// Wave-wide record store: every lane of the wave must call it (inactive queries pass valid =
// false).  Four instructions; in instruction j the quad's lanes h = 0..3 write quarter h of the
// record of the quad's lane j.
__device__ __forceinline__ void storeRecQuad(Rec* __restrict__ rec, bool valid, uint32_t id, float nn, float val,
                                             const float g[3])
{
    const int lane = int(threadIdx.x & 63u), h = lane & 3, quad = lane & ~3;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int src = quad + j;
        const int ok = __shfl(int(valid), src, 64);
        const uint32_t d = __shfl(id, src, 64);
        const float a = __shfl(nn, src, 64), b = __shfl(val, src, 64), c = __shfl(g[0], src, 64),
                    e = __shfl(g[1], src, 64), f = __shfl(g[2], src, 64);
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (h == 0) v = make_float4(a, b, c, e);
        if (h == 1) v.x = f;
        if (ok) reinterpret_cast<float4*>(rec + d)[h] = v;
    }
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

// The 8^3 slots that [cell-1, cell+2]^3 overlaps (1 or 2 per axis), with the value every voxel
// of a slot without a leaf reads (its tile or background value, ReadAccessor::getValue semantics).
struct SlotBlock {
    int32_t s0[3], s1[3];
    Cell v; // v.v[a][b][e]: slot (s0 + (a,b,e)); unused where s1 == s0
};

// Every 8^3 slot of a box covering the binned region: whether a leaf holds it, and otherwise the
// value all its voxels read (tile or background, ReadAccessor::getValue semantics).  Built per
// call by one tree walk per slot; the classify pass then reads it with independent loads instead
// of walking the tree per query.
struct SlotTable {
    int32_t lo[3];   // first slot per axis
    uint32_t n[3];   // slots per axis
    const float* value;
    const uint32_t* leaf; // one bit per slot (bit i%32 of word i/32): a leaf holds the slot
};

// Bitmaps of at most this many 32-bit words (48 KB, 384K slots) are staged in LDS by the classify
// pass, which takes exactly the bitmap's size of dynamic LDS; larger ones are read from global
// memory.  The torus (300, 80) needs 6.5K words, the 512^3 fog 10.3K.
constexpr uint32_t kLdsLeafWords = 12288;

// One thread per slot; the bitmap words are assembled with a wave ballot (64 slots per wave, so
// every word is written whole by one lane).  The launch covers whole waves: count rounded up to 64.
template<uint32_t GT>
__global__ __launch_bounds__(kThreads) void k_slot_table(GridView g, SlotTable T, float* __restrict__ value,
                                                         uint32_t* __restrict__ leaf)
{
    Accessor<GT> acc(g);
    const uint64_t count = uint64_t(T.n[0]) * T.n[1] * T.n[2];
    const uint64_t padded = (count + 63) & ~uint64_t(63);
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < padded; i += uint64_t(gridDim.x) * kThreads) {
        bool isLeaf = false;
        if (i < count) {
            const uint32_t z = uint32_t(i % T.n[2]), y = uint32_t((i / T.n[2]) % T.n[1]),
                           x = uint32_t(i / (uint64_t(T.n[2]) * T.n[1]));
            const Probe pr = acc.probe((T.lo[0] + int32_t(x)) << 3, (T.lo[1] + int32_t(y)) << 3,
                                       (T.lo[2] + int32_t(z)) << 3);
            value[i] = pr.value;
            isLeaf = pr.leaf != nullptr;
        }
        const uint64_t b = __ballot(isLeaf);
        const uint32_t lane = threadIdx.x & 63u;
        if (lane == 0) leaf[i >> 5] = uint32_t(b);
        if (lane == 32) leaf[i >> 5] = uint32_t(b >> 32);
    }
}

// True when a slot of the block around p's cell holds a leaf or lies outside the table (the query
// then goes through the sort).  `bits` is the leaf bitmap (in LDS or global memory).
__device__ __forceinline__ bool touchesLeaf(const SlotTable& T, const uint32_t* bits, const float p[3], SlotBlock& B)
{
    uint32_t r0[3], r1[3];
    bool outside = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t c = int32_t(floorf(p[a]));
        B.s0[a] = (c - 1) >> 3;
        B.s1[a] = (c + 2) >> 3;
        r0[a] = uint32_t(B.s0[a] - T.lo[a]);
        r1[a] = uint32_t(B.s1[a] - T.lo[a]);
        outside = outside || r0[a] >= T.n[a] || r1[a] >= T.n[a];
    }
    if (outside) return true;
    const bool two[3] = {r1[0] != r0[0], r1[1] != r0[1], r1[2] != r0[2]};
    bool leaf = false;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if ((a && !two[0]) || (b && !two[1]) || (e && !two[2])) continue; // a repeated slot
                const uint64_t i = (uint64_t(a ? r1[0] : r0[0]) * T.n[1] + (b ? r1[1] : r0[1])) * T.n[2] + (e ? r1[2] : r0[2]);
                leaf = leaf || ((bits[i >> 5] >> (i & 31u)) & 1u);
            }
    return leaf;
}

// The slot values of the block around p's cell (a query touchesLeaf found inside the table and
// without leaves): distinct slots are loaded, repeated ones copied.
__device__ __forceinline__ void loadSlots(const SlotTable& T, const float p[3], SlotBlock& B)
{
    uint32_t r[2][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int32_t c = int32_t(floorf(p[a]));
        B.s0[a] = (c - 1) >> 3;
        B.s1[a] = (c + 2) >> 3;
        r[0][a] = uint32_t(B.s0[a] - T.lo[a]);
        r[1][a] = uint32_t(B.s1[a] - T.lo[a]);
    }
    const bool two[3] = {r[1][0] != r[0][0], r[1][1] != r[0][1], r[1][2] != r[0][2]};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int sa = two[0] ? a : 0, sb = two[1] ? b : 0, se = two[2] ? e : 0;
                if ((a && !two[0]) || (b && !two[1]) || (e && !two[2])) {
                    B.v.v[a][b][e] = pick(B.v, sa, sb, se);
                    continue;
                }
                B.v.v[a][b][e] = T.value[(uint64_t(r[a][0]) * T.n[1] + r[b][1]) * T.n[2] + r[e][2]];
            }
}

// Value of voxel (x, y, z) of the block (no leaf in it): the value of its slot
__device__ __forceinline__ float slotValue(const SlotBlock& B, int32_t x, int32_t y, int32_t z)
{
    return pick(B.v, (x >> 3) != B.s0[0], (y >> 3) != B.s0[1], (z >> 3) != B.s0[2]);
}

// The query's outputs from the slot values (bit-identical to sampleEval: same cells, fractions and
// operation order).  Requires the FD6 sample cells to be the neighbouring ones (adjointInBlock).
__device__ __forceinline__ SampleOut sampleSlots(const SlotBlock& B, const float p[3], bool wantNN, bool wantVal,
                                                 int mode, const SampleConsts& k)
{
    SampleOut r{0.0f, 0.0f, {0.0f, 0.0f, 0.0f}};
    const float V = B.v.v[0][0][0];
    bool uniform = isfinite(V) && __float_as_uint(V) != 0x80000000u;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) uniform = uniform && __float_as_uint(B.v.v[a][b][e]) == __float_as_uint(V);
    if (uniform) {
        // every value the query reads is V (finite, not -0): lerps give V, differences +0
        r.nn = V;
        r.val = V;
        return r;
    }
    auto cellAt = [&](int32_t i, int32_t j, int32_t l, Cell& C) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) C.v[a][b][e] = slotValue(B, i + a, j + b, l + e);
    };
    auto tri = [&](float qx, float qy, float qz) {
        const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
        Cell C;
        cellAt(int32_t(fx), int32_t(fy), int32_t(fz), C);
        return trilerp(C, qx - fx, qy - fy, qz - fz);
    };
    const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
    const int32_t i0 = int32_t(fx), j0 = int32_t(fy), k0 = int32_t(fz);
    if (wantNN) r.nn = slotValue(B, int32_t(rintf(p[0])), int32_t(rintf(p[1])), int32_t(rintf(p[2])));
    if (wantVal) r.val = tri(p[0], p[1], p[2]);
    if (mode == 1) {
        r.g[0] = (tri(p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f) - tri(p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f)) * k.fdx;
        r.g[1] = (tri(p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f) - tri(p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f)) * k.fdy;
        r.g[2] = (tri(p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f) - tri(p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f)) * k.fdz;
    } else if (mode == 2) {
        r.g[0] = k.gsInv2dx * (slotValue(B, i0 + 1, j0, k0) - slotValue(B, i0 - 1, j0, k0));
        r.g[1] = k.gsInv2dx * (slotValue(B, i0, j0 + 1, k0) - slotValue(B, i0, j0 - 1, k0));
        r.g[2] = k.gsInv2dx * (slotValue(B, i0, j0, k0 + 1) - slotValue(B, i0, j0, k0 - 1));
    } else if (mode == 3) {
        Cell C;
        cellAt(i0, j0, k0, C);
        trilinearGradient(C, p[0] - fx, p[1] - fy, p[2] - fz, r.g);
    }
    return r;
}

constexpr int kClassifyItems = 8; // points per thread per block iteration

constexpr uint32_t kChunk = kThreads * kClassifyItems; // queries per block iteration

// Exclusive prefix over the block of one count per thread (waves in order, lanes in order); `total`
// receives the block's sum.  Uses `waveSum`, which the caller must not reuse before a barrier.
__device__ __forceinline__ uint32_t blockPrefix(uint32_t cnt, uint32_t* waveSum, uint32_t& total)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d, 64);
        if (lane >= uint32_t(d)) incl += v;
    }
    if (lane == 63) waveSum[wave] = incl;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        off += uint32_t(w) < wave ? waveSum[w] : 0u;
        total += waveSum[w];
    }
    return off + incl - cnt;
}

// Classify: chunk c holds queries c*kChunk + j*kThreads + t (item j of thread t), so loads are
// coalesced.  Queries that touch a leaf (or whose FD6 samples round into a farther cell, or that
// lie outside the slot table) are marked in flags[c*kThreads + t] (bit j) and counted per chunk in
// counts[c]; k_bin_compact places them after a scan of the counts (no atomics) in chunk, item,
// wave and lane order.  The others are answered from their slot values by k_bin_finish.
__global__ __launch_bounds__(kThreads) void k_bin_classify(GridView g, SlotTable T, const float* __restrict__ xyz,
                                                           uint32_t n, uint32_t chunks, int world, int mode,
                                                           uint32_t ldsWords, uint8_t* __restrict__ flagsOut,
                                                           uint32_t* __restrict__ counts)
{
    __shared__ uint32_t waveSum[kThreads / 64];
    extern __shared__ uint32_t ldsLeaf[]; // ldsWords words: the whole bitmap, or none
    const uint32_t* bits = T.leaf;
    if (ldsWords) {
        for (uint32_t w = threadIdx.x; w < ldsWords; w += kThreads) ldsLeaf[w] = T.leaf[w];
        __syncthreads();
        bits = ldsLeaf;
    }
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint64_t c0 = uint64_t(c) * kChunk;
        uint32_t flags = 0;
        // the chunk's points of this thread are loaded before the first test, so their loads
        // overlap instead of each waiting for the previous test
        float P[kClassifyItems][3];
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            const uint64_t i = c0 + uint64_t(j) * kThreads + threadIdx.x;
            P[j][0] = P[j][1] = P[j][2] = 0.0f;
            if (i < n) loadPoint(g, xyz, i, world, P[j]);
        }
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            const uint64_t i = c0 + uint64_t(j) * kThreads + threadIdx.x;
            if (i >= n) continue;
            SlotBlock B;
            // FD6 samples that round into a farther cell read outside the block's slots: such
            // queries go through the sort with the others
            if (touchesLeaf(T, bits, P[j], B) || !adjointInBlock(P[j], mode)) flags |= 1u << j;
        }
        flagsOut[uint64_t(c) * kThreads + threadIdx.x] = uint8_t(flags);
        uint32_t total;
        (void)blockPrefix(__popc(flags), waveSum, total);
        if (threadIdx.x == 0) counts[c] = total;
        __syncthreads(); // waveSum is reused by the next chunk
    }
}

// Compaction of the marked queries: chunk c's go to offsets[c] on, ordered by item, wave and lane,
// so each store instruction writes one contiguous run (whole 64-byte blocks of keys and points).
__global__ __launch_bounds__(kThreads) void k_bin_compact(GridView g, BinLayout L, const float* __restrict__ xyz,
                                                          uint32_t chunks, int world,
                                                          const uint8_t* __restrict__ flagsIn,
                                                          const uint32_t* __restrict__ offsets,
                                                          uint32_t* __restrict__ keys, Pt* __restrict__ pts)
{
    constexpr int kWaves = kThreads / 64;
    __shared__ uint32_t runBase[kClassifyItems][kWaves];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint32_t flags = flagsIn[uint64_t(c) * kThreads + threadIdx.x];
        uint64_t ball[kClassifyItems];
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            ball[j] = __ballot((flags >> j) & 1u);
            if (lane == 0) runBase[j][wave] = uint32_t(__popcll(ball[j]));
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = offsets[c];
            for (int j = 0; j < kClassifyItems; ++j)
                for (int w = 0; w < kWaves; ++w) {
                    const uint32_t v = runBase[j][w];
                    runBase[j][w] = acc;
                    acc += v;
                }
        }
        __syncthreads();
        const uint64_t below = (uint64_t(1) << lane) - 1u;
        // the marked points are loaded before the first store (the loads overlap)
        float P[kClassifyItems][3];
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            P[j][0] = P[j][1] = P[j][2] = 0.0f;
            if ((flags >> j) & 1u) loadPoint(g, xyz, uint64_t(c) * kChunk + uint64_t(j) * kThreads + threadIdx.x, world, P[j]);
        }
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            if (!((flags >> j) & 1u)) continue;
            const uint32_t pos = runBase[j][wave] + uint32_t(__popcll(ball[j] & below));
            const uint64_t i = uint64_t(c) * kChunk + uint64_t(j) * kThreads + threadIdx.x;
            const float* p = P[j];
            keys[pos] = binOf(L, p);
            Pt t;
            t.x = p[0];
            t.y = p[1];
            t.z = p[2];
            t.id = uint32_t(i);
            pts[pos] = t;
        }
        __syncthreads(); // runBase is reused by the next chunk
    }
}

// Four consecutive sorted keys per thread: one 16-byte load (`keys` is 256-byte aligned) and the
// keys on either side, so all three loads are in flight at once.  Bin keys are below the bin
// count, so ~0u stands for "no key" before the first and after the last.
__global__ __launch_bounds__(kThreads) void k_bin_bounds(const uint32_t* __restrict__ keys, uint32_t n,
                                                         uint32_t* __restrict__ begin, uint32_t* __restrict__ end)
{
    const uint64_t i0 = (uint64_t(blockIdx.x) * kThreads + threadIdx.x) * 4u;
    if (i0 >= n) return;
    uint32_t k[6]; // keys[i0 - 1 .. i0 + 4]
    if (i0 + 4u <= n) {
        const uint4 v = *reinterpret_cast<const uint4*>(keys + i0);
        k[1] = v.x;
        k[2] = v.y;
        k[3] = v.z;
        k[4] = v.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) k[1 + j] = i0 + j < n ? keys[i0 + j] : ~0u;
    }
    k[0] = i0 ? keys[i0 - 1] : ~0u;
    k[5] = i0 + 4u < n ? keys[i0 + 4] : ~0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = uint32_t(i0) + uint32_t(j);
        if (i >= n) break;
        if (k[j] != k[j + 1]) begin[k[j + 1]] = i;
        if (k[j + 2] != k[j + 1]) end[k[j + 1]] = i + 1;
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
    // wave-uniform loop (the record store is a wave-wide operation): 64 consecutive queries per wave;
    // the next iteration's point is loaded before this iteration's evaluation
    uint32_t base = q0 + (threadIdx.x & ~63u);
    Pt next = sorted[base + (threadIdx.x & 63u) < q1 ? base + (threadIdx.x & 63u) : q0];
    for (; base < q1; base += kBinThreads) {
        const uint32_t q = base + (threadIdx.x & 63u);
        bool valid = q < q1;
        const Pt t = next;
        if (base + kBinThreads < q1) next = sorted[q + kBinThreads < q1 ? q + kBinThreads : q0];
        const uint32_t id = t.id;
        const float p[3] = {t.x, t.y, t.z};
        const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
        const int32_t ci = int32_t(fx) - o[0], cj = int32_t(fy) - o[1], ck = int32_t(fz) - o[2];
        // every FD6 sample cell lies in [cell-1, cell+2] (p +- 1 rounds at most one voxel further up),
        // so the whole stencil is inside the brick iff the cell is in [1, kBrick-4]
        if (valid && (uint32_t(ci - 1) > uint32_t(kBrick - 5) || uint32_t(cj - 1) > uint32_t(kBrick - 5) ||
                      uint32_t(ck - 1) > uint32_t(kBrick - 5))) {
            spill[atomicAdd(spillCount, 1u)] = id;
            valid = false;
        }
        float nnv = 0.0f, vv = 0.0f;
        float g3[3] = {0.0f, 0.0f, 0.0f};
        if (valid) {
            auto load = [&](int32_t cb, Cell& c) {
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                        for (int e = 0; e < 2; ++e) c.v[a][bb][e] = brick[cb + (a * kBrick + bb) * kBrick + e];
            };
            auto cell = [&](float qx, float qy, float qz) -> float {
                const float gx = floorf(qx), gy = floorf(qy), gz = floorf(qz);
                const int32_t cb = ((int32_t(gx) - o[0]) * kBrick + (int32_t(gy) - o[1])) * kBrick + (int32_t(gz) - o[2]);
                Cell c;
                load(cb, c);
                return trilerp(c, qx - gx, qy - gy, qz - gz);
            };
            const int32_t c0 = (ci * kBrick + cj) * kBrick + ck;
            if (wantNN) {
                const int32_t ri = int32_t(rintf(p[0])) - o[0], rj = int32_t(rintf(p[1])) - o[1], rk = int32_t(rintf(p[2])) - o[2];
                nnv = brick[(ri * kBrick + rj) * kBrick + rk];
            }
            if (wantVal) vv = cell(p[0], p[1], p[2]);
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
                load(c0, C);
                trilinearGradient(C, p[0] - fx, p[1] - fy, p[2] - fz, g3);
            }
        }
        storeRecQuad(rec, valid, id, nnv, vv, g3);
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
    for (uint32_t base = blockIdx.x * kThreads + (threadIdx.x & ~63u); base < m; base += gridDim.x * kThreads) {
        const uint32_t j = base + (threadIdx.x & 63u);
        const bool valid = j < m;
        SampleOut r{0.0f, 0.0f, {0.0f, 0.0f, 0.0f}};
        uint32_t id = 0;
        if (valid) {
            id = list[j];
            float p[3];
            loadPoint(g, xyz, id, world, p);
            r = sampleEval(acc, p, wantNN, wantVal, mode, k);
        }
        storeRecQuad(rec, valid, id, r.nn, r.val, r.g);
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

// One 12-byte store per gradient, so that a wavefront's store covers one contiguous run
__device__ __forceinline__ void storeGrad(float* __restrict__ grad, uint64_t i, float x, float y, float z)
{
    typedef float v3f __attribute__((ext_vector_type(3)));
    v3f g3;
    g3.x = x;
    g3.y = y;
    g3.z = z;
    __builtin_memcpy(grad + 3 * i, &g3, 12);
}

// Outputs of a classified call, in the classify pass's chunk order: a marked query copies its record,
// the others are answered from their slot values (sampleSlots, bit-identical to the lane accessor).
// 146 VGPRs uncapped (3 wavefronts per SIMD); capped at 4 wavefronts (128, a 60-byte spill) it runs
// 1.30 -> 1.21 ms per 2^26-query call; at 5 (96, 208 bytes) 1.81 ms.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_bin_finish(GridView g, SlotTable T, const float* __restrict__ xyz,
                                                         uint32_t n, uint32_t chunks, int world, bool wantNN,
                                                         bool wantVal, int mode, SampleConsts k,
                                                         const uint8_t* __restrict__ flagsIn,
                                                         const Rec* __restrict__ rec, float* __restrict__ nn,
                                                         float* __restrict__ val, float* __restrict__ grad)
{
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint32_t flags = flagsIn[uint64_t(c) * kThreads + threadIdx.x];
        // the marked queries' records, all loads issued before the first use
        float4 ra[kClassifyItems];
        float rb[kClassifyItems];
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            const uint64_t i = uint64_t(c) * kChunk + uint64_t(j) * kThreads + threadIdx.x;
            const bool marked = ((flags >> j) & 1u) != 0; // only queries below n are marked
            ra[j] = marked ? rec[i].a : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            rb[j] = marked && grad ? rec[i].b.x : 0.0f;
        }
        // and the points of the other queries, answered from the slot table
        float P[kClassifyItems][3];
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            const uint64_t i = uint64_t(c) * kChunk + uint64_t(j) * kThreads + threadIdx.x;
            P[j][0] = P[j][1] = P[j][2] = 0.0f;
            if (i < n && !((flags >> j) & 1u)) loadPoint(g, xyz, i, world, P[j]);
        }
        // one store per output for every query of the chunk, whether its result comes from a
        // record or from the slot table: stores that covered the marked and the other queries of a
        // line separately wrote each output line twice
#pragma unroll
        for (int j = 0; j < kClassifyItems; ++j) {
            const uint64_t i = uint64_t(c) * kChunk + uint64_t(j) * kThreads + threadIdx.x;
            if (i >= n) continue;
            float o[5] = {ra[j].x, ra[j].y, ra[j].z, ra[j].w, rb[j]};
            if (!((flags >> j) & 1u)) {
                const float* p = P[j];
                SlotBlock B;
                loadSlots(T, p, B);
                const SampleOut r = sampleSlots(B, p, wantNN, wantVal, mode, k);
                o[0] = r.nn;
                o[1] = r.val;
                o[2] = r.g[0];
                o[3] = r.g[1];
                o[4] = r.g[2];
            }
            if (nn) nn[i] = o[0];
            if (val) val[i] = o[1];
            if (grad) storeGrad(grad, i, o[2], o[3], o[4]);
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
// | forward only: rec[n], classify marks, counts and offsets, scan temporaries | sort temporaries,
// in the stream's workspace.
struct Bins {
    BinLayout L;
    uint32_t n = 0;
    uint32_t m = 0; // queries in the sort (all n, or those the forward's classify pass kept)
    WorkspaceLease lease;
    char* base = nullptr;
    // keys/idx: the unsorted (key, point) pairs; after the sort, keys2/idx2 name the sorted ones
    uint32_t *keys = nullptr, *keys2 = nullptr, *begin = nullptr, *end = nullptr, *spillCount = nullptr,
             *flag = nullptr, *spill = nullptr;
    uint32_t chunks = 0;                                 // classify chunks (forward only; 0: not classified)
    SlotTable T{};                                       // the classify pass's slot table
    uint8_t* cflags = nullptr;                           // [chunks * kThreads] per-thread marks
    uint32_t *counts = nullptr, *offsets = nullptr;      // [chunks + 1] marked per chunk, their scan
    void* scanTemp = nullptr;
    size_t scanBytes = 0;
    Pt *idx = nullptr, *idx2 = nullptr;
    Rec* rec = nullptr;
};

// What the forward's classify pass evaluates for the queries it answers itself.
struct Classify {
    bool wantNN, wantVal;
    int mode;
};

constexpr uint64_t kMaxTableSlots = uint64_t(1) << 24; // larger regions sort every query

// The slot table's box: the bins' region (2 voxels around the grid's bbox, rounded up to bins)
// plus one slot on each side.
bool slotBox(const BinLayout& L, SlotTable& T)
{
    uint64_t count = 1;
    for (int a = 0; a < 3; ++a) {
        const int64_t lo = (int64_t(L.lo[a]) >> 3) - 1, hi = ((int64_t(L.lo[a]) + (int64_t(L.nb[a]) << kBinLog2) - 1) >> 3) + 1;
        T.lo[a] = int32_t(lo);
        T.n[a] = uint32_t(hi - lo + 1);
        count *= T.n[a];
    }
    return count <= kMaxTableSlots;
}

template<uint32_t GT>
void launchClassify(const GridView& g, const Bins& B, const SlotTable& T, float* tv, uint32_t* tl, const float* xyz,
                    int world, const Classify& c, uint32_t blocks, hipStream_t s)
{
    const uint64_t count = uint64_t(T.n[0]) * T.n[1] * T.n[2];
    hipLaunchKernelGGL(k_slot_table<GT>, dim3(uint32_t(std::min<uint64_t>((count + kThreads - 1) / kThreads, 1u << 16))),
                       dim3(kThreads), 0, s, g, T, tv, tl);
    ProfScope prof("k_bin_classify", s);
    const uint64_t words = (count + 31) / 32;
    const uint32_t ldsWords = words <= kLdsLeafWords ? uint32_t(words) : 0u;
    hipLaunchKernelGGL(k_bin_classify, dim3(blocks), dim3(kThreads), size_t(ldsWords) * 4u, s, g, T, xyz, B.n, B.chunks,
                       world, c.mode, ldsWords, B.cflags, B.counts);
}

// Allocate, key, sort and bound the bins of n points.  With `cls`, only the queries whose block
// touches a leaf (marked by the classify pass, whose slot table and marks stay in B) are sorted;
// the forward answers the others in k_bin_finish, the adjoint of a Float grid drops them.  The call
// then waits for the stream once, to size the sort.  0, or -1 (no layout), -2 (allocation), -3
// (grid type), -4 (sort or HIP error: nothing was launched that reads the sorted arrays; the
// scratch is already freed).
int binPoints(const GridView& g, const float* xyz, uint32_t n, int world, bool withRecords, Bins& B, hipStream_t s,
              const Classify* cls = nullptr)
{
    if (!makeLayout(g, B.L)) return -1;
    const uint32_t m = B.L.nbins;
    B.n = n;
    // The sort ping-pongs between keys/keys2 and idx/idx2, so its own temporaries are only its
    // histograms and look-back state (the form with separate outputs adds a 20-byte-per-point buffer).
    size_t sortBytes = 0;
    {
        hipcub::DoubleBuffer<uint32_t> dk(nullptr, nullptr);
        hipcub::DoubleBuffer<Pt> dv(nullptr, nullptr);
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, dk, dv, n, 0, int(B.L.bits), s) != hipSuccess)
            return -4;
    }
    SlotTable T{};
    if (cls && !slotBox(B.L, T)) cls = nullptr; // too large a region for the slot table: sort every query
    const uint64_t slots = cls ? uint64_t(T.n[0]) * T.n[1] * T.n[2] : 0;
    const uint32_t chunks = cls ? uint32_t((uint64_t(n) + kChunk - 1) / kChunk) : 0u;
    size_t scanBytes = 0;
    if (cls && hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, static_cast<uint32_t*>(nullptr),
                                                static_cast<uint32_t*>(nullptr), chunks + 1, s) != hipSuccess)
        return -4;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t oKeys = 0, oKeys2 = up(oKeys + 4ull * n), oIdx = up(oKeys2 + 4ull * n),
                 oIdx2 = up(oIdx + sizeof(Pt) * size_t(n)), oBegin = up(oIdx2 + sizeof(Pt) * size_t(n)),
                 oEnd = up(oBegin + 4ull * m), oCount = up(oEnd + 4ull * m), oSpill = up(oCount + 8),
                 oRec = up(oSpill + 4ull * n), oFlags = up(oRec + (withRecords ? sizeof(Rec) * size_t(n) : 0)),
                 oCounts = up(oFlags + size_t(chunks) * kThreads), oOffsets = up(oCounts + 4ull * (chunks + 1)),
                 oScan = up(oOffsets + 4ull * (chunks + 1)), oTabV = up(oScan + scanBytes),
                 oTabL = up(oTabV + 4 * slots), oTemp = up(oTabL + 4 * ((slots + 63) / 64) * 2),
                 total = oTemp + up(sortBytes);
    B.base = B.lease.acquire(total, s);
    if (!B.base) return -2;
    char* base = B.base;
    B.keys = reinterpret_cast<uint32_t*>(base + oKeys);
    B.keys2 = reinterpret_cast<uint32_t*>(base + oKeys2);
    B.idx = reinterpret_cast<Pt*>(base + oIdx);
    B.idx2 = reinterpret_cast<Pt*>(base + oIdx2);
    B.begin = reinterpret_cast<uint32_t*>(base + oBegin);
    B.end = reinterpret_cast<uint32_t*>(base + oEnd);
    B.spillCount = reinterpret_cast<uint32_t*>(base + oCount);
    B.flag = B.spillCount + 1;
    B.chunks = chunks;
    B.cflags = reinterpret_cast<uint8_t*>(base + oFlags);
    B.counts = reinterpret_cast<uint32_t*>(base + oCounts);
    B.offsets = reinterpret_cast<uint32_t*>(base + oOffsets);
    B.scanTemp = base + oScan;
    B.scanBytes = scanBytes;
    B.spill = reinterpret_cast<uint32_t*>(base + oSpill);
    B.rec = withRecords ? reinterpret_cast<Rec*>(base + oRec) : nullptr;
    void* temp = base + oTemp;
    const uint32_t gridN = std::min<uint32_t>((n + kThreads - 1) / kThreads, 1u << 16);
    const uint32_t zeroWords = uint32_t((oCount + 8) / 4 - oBegin / 4); // begin[], end[], spillCount and flag
    hipLaunchKernelGGL(k_fill_u32, dim3((zeroWords + kThreads - 1) / kThreads), dim3(kThreads), 0, s, B.begin,
                       zeroWords, 0u);
    B.m = n;
    if (cls) {
        // classify: enough blocks to fill the chip, each staging the leaf bitmap once
        const uint32_t blocks = std::min<uint32_t>(chunks, 4096u);
        // counts[chunks] stays 0, so the exclusive scan's last entry is the number of marked queries
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kThreads), 0, s, B.counts + chunks, 1u, 0u);
        float* tv = reinterpret_cast<float*>(base + oTabV);
        uint32_t* tl = reinterpret_cast<uint32_t*>(base + oTabL);
        T.value = tv;
        T.leaf = tl;
        B.T = T;
        switch (g.type) {
        case GT_FLOAT: launchClassify<GT_FLOAT>(g, B, T, tv, tl, xyz, world, *cls, blocks, s); break;
        case GT_FP4: launchClassify<GT_FP4>(g, B, T, tv, tl, xyz, world, *cls, blocks, s); break;
        case GT_FP8: launchClassify<GT_FP8>(g, B, T, tv, tl, xyz, world, *cls, blocks, s); break;
        case GT_FP16: launchClassify<GT_FP16>(g, B, T, tv, tl, xyz, world, *cls, blocks, s); break;
        case GT_FPN: launchClassify<GT_FPN>(g, B, T, tv, tl, xyz, world, *cls, blocks, s); break;
        case GT_ONINDEX: launchClassify<GT_ONINDEX>(g, B, T, tv, tl, xyz, world, *cls, blocks, s); break;
        default: return -3;
        }
        uint32_t m = 0;
        if (hipGetLastError() != hipSuccess ||
            hipcub::DeviceScan::ExclusiveSum(B.scanTemp, B.scanBytes, B.counts, B.offsets, chunks + 1, s) !=
                hipSuccess ||
            readBack(&m, B.offsets + chunks, sizeof(m), s) != hipSuccess || m > n) {
            (void)hipGetLastError();
            return -4;
        }
        B.m = m;
        if (m)
            hipLaunchKernelGGL(k_bin_compact, dim3(blocks), dim3(kThreads), 0, s, g, B.L, xyz, chunks, world, B.cflags,
                               B.offsets, B.keys, B.idx);
    } else {
        hipLaunchKernelGGL(k_bin_keys, dim3(gridN), dim3(kThreads), 0, s, g, B.L, xyz, n, world, B.keys, B.idx);
    }
    if (B.m == 0) return 0;
    hipError_t sorted;
    {
        ProfScope prof("bin_sort", s);
        hipcub::DoubleBuffer<uint32_t> dk(B.keys, B.keys2);
        hipcub::DoubleBuffer<Pt> dv(B.idx, B.idx2);
        sorted = hipcub::DeviceRadixSort::SortPairs(temp, sortBytes, dk, dv, B.m, 0, int(B.L.bits), s);
        // the sorted pairs end in either buffer, depending on the number of digit passes
        B.keys2 = dk.Current();
        B.idx2 = dv.Current();
    }
    if (sorted != hipSuccess) {
        // the sorted arrays were never written: launch nothing that reads them
        (void)hipGetLastError();
        return -4;
    }
    // four keys per thread, one pass over the grid
    const uint32_t gridM = uint32_t((uint64_t(B.m) + 4u * kThreads - 1) / (4u * kThreads));
    hipLaunchKernelGGL(k_bin_bounds, dim3(gridM), dim3(kThreads), 0, s, B.keys2, B.m, B.begin, B.end);
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
    const bool wantNN = nn != nullptr, wantVal = val != nullptr;
    const int gm = grad ? mode : 0;
    const Classify cls{wantNN, wantVal, gm};
    const int bs = binPoints(g, xyz, uint32_t(n64), world, true, B, s, &cls);
    if (bs) return bs;
    const uint32_t n = B.n;
    int st = 0;
    if (B.m) switch (g.type) {
    case GT_FLOAT: launchSample<GT_FLOAT>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FP4: launchSample<GT_FP4>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FP8: launchSample<GT_FP8>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FP16: launchSample<GT_FP16>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_FPN: launchSample<GT_FPN>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    case GT_ONINDEX: launchSample<GT_ONINDEX>(g, B.L, xyz, world, B.idx2, B.begin, B.end, B.rec, wantNN, wantVal, gm, B.spill, B.spillCount, s); break;
    default: st = -3;
    }
    if (st == 0 && B.chunks) {
        ProfScope prof("k_bin_finish", s);
        hipLaunchKernelGGL(k_bin_finish, dim3(std::min<uint32_t>(B.chunks, 4096u)), dim3(kThreads), 0, s, g, B.T, xyz, n,
                           B.chunks, world, wantNN, wantVal, gm, sampleConsts(g), B.cflags, B.rec, nn, val, grad);
    } else if (st == 0) {
        ProfScope prof("k_unpack", s);
        hipLaunchKernelGGL(k_unpack, dim3(std::min<uint32_t>((n + kThreads - 1) / kThreads, 1u << 14)), dim3(kThreads), 0,
                           s, B.rec, n, nn, val, grad);
    }
    if (hipGetLastError() != hipSuccess) st = -4;
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
        // Float grids: a query whose block touches no leaf reads only tile and background values,
        // which receive no gradient, so only the queries the classify pass marks are sorted
        const Classify cls{false, false, gm};
        const int bs = binPoints(g, xyz, uint32_t(n64), world, false, B, s, g.type == GT_FLOAT ? &cls : nullptr);
        if (bs) return bs;
        if (g.type == GT_FLOAT) launchSampleBwd<GT_FLOAT>(g, B, xyz, world, gnn, gval, ggrad, gm, scale, acc, s);
        else launchSampleBwd<GT_ONINDEX>(g, B, xyz, world, gnn, gval, ggrad, gm, scale, acc, s);
        bool ok = hipGetLastError() == hipSuccess;
        const void* st = stageCopy(B.flag, sizeof(over), s);
        ok = st && ok;
        ok = hipStreamSynchronize(s) == hipSuccess && ok;
        if (!ok) return -4;
        std::memcpy(&over, st, sizeof(over));
        return over ? -6 : 0;
    }
    uint32_t* flag = nullptr;
    if (poolMalloc(reinterpret_cast<void**>(&flag), sizeof(uint32_t), s) != hipSuccess) return -2;
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
    const void* st = stageCopy(flag, sizeof(over), s);
    ok = st && ok;
    ok = hipFreeAsync(flag, s) == hipSuccess && ok;
    ok = hipStreamSynchronize(s) == hipSuccess && ok;
    if (!ok) return -4;
    std::memcpy(&over, st, sizeof(over));
    return over ? -6 : 0;
}

} // namespace nvq
