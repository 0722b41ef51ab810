// Device-side read path over a flat NanoVDB buffer for gfx950.
//
// The buffer is read by byte offset (NanoVDB.h is never compiled for the
// device).  Layout: nanovdb/nanovdb/NanoVDB.h GridData 1944-2135, Map
// 1418-1587, TreeData 2393-2423, RootData 2621-2840, InternalData 3181-3250,
// LeafData 3671-3973.  Every float expression keeps the reference's operation
// order and the file is compiled with -ffp-contract=off; fmaf is written out
// exactly where the reference writes fmaf, so hit voxels, active states and
// lookups are bit-identical to the reference host code.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <stdint.h>

namespace nvq {

// ------------------------------------------------------------------ layout
constexpr uint32_t kGridMap = 296, kGridVoxelSize = 608, kGridClass = 632, kGridType = 636, kGridSize = 672;
constexpr uint32_t kMapInvMatF = 36, kMapVecF = 72, kMapInvMatD = 160, kMapVecD = 232;
constexpr uint32_t kRootBBox = 0, kRootTableSize = 24, kRootBackground = 28, kRootSize = 64, kTileSize = 32;
constexpr uint32_t kUpFlags = 24, kUpVMask = 32, kUpCMask = 4128, kUpTable = 8256, kUpSize = 270400;
constexpr uint32_t kLoFlags = 24, kLoVMask = 32, kLoCMask = 544, kLoTable = 1088, kLoSize = 33856;
constexpr uint32_t kLeafFlags = 15, kLeafVMask = 16, kLeafMin = 80, kLeafQuantum = 84, kLeafData = 96;
enum GridTypeId : uint32_t { GT_FLOAT = 1, GT_FP4 = 13, GT_FP8 = 14, GT_FP16 = 15, GT_FPN = 16, GT_ONINDEX = 20 };
// ValueOnIndex grids (NanoVDB.h:4147-4171): the root is 96 B (uint64 background index at 32),
// root tiles carry a uint64 index at 24, node tables hold uint64 indices, and a leaf holds an
// offset (80) and a prefix sum of 9 bits per mask word (88).  Values live in a float channel.
constexpr uint32_t kIdxRootSize = 96, kIdxRootBackground = 32, kIdxTileValue = 24, kIdxLeafOffset = 80,
                   kIdxLeafPrefix = 88;
__host__ __device__ constexpr uint32_t rootHeaderSize(uint32_t gt) { return gt == GT_ONINDEX ? kIdxRootSize : kRootSize; }

// Host FLT_MAX: the reference's Maximum<float> on the host (math/Math.h:118-124);
// the reference's own HIP branch would use 1e38 (:112-115), which would change
// which rays are clipped, so the host value is used.
constexpr float kFltMax = 3.40282346638528859812e+38F;

// Kernel-argument view of an uploaded grid (all pointers are device pointers).
struct GridView {
    const uint8_t* root;  // RootData
    const uint8_t* leaf0; // first leaf node
    const float* channel; // index grids: values by index (blind channel 0 or a caller array)
    uint32_t tableSize;
    uint32_t type;
    float background;
    int32_t bbox[6];      // RootData::mBBox
    float invMatF[9], vecF[3];
    double invMatD[9], vecD[3];
    double voxelSize[3];
    uint32_t leafCount;
    uint32_t leafSize;    // 0 for FpN (variable)
    uint32_t rayMapF64;   // ray queries map world rays with the FP64 Ray::applyInverseMap
    uint32_t pad;
};

__device__ __forceinline__ uint64_t ld_u64(const uint8_t* p) { return *reinterpret_cast<const uint64_t*>(p); }
__device__ __forceinline__ int64_t ld_i64(const uint8_t* p) { return *reinterpret_cast<const int64_t*>(p); }
__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) { return *reinterpret_cast<const uint32_t*>(p); }
__device__ __forceinline__ float ld_f32(const uint8_t* p) { return *reinterpret_cast<const float*>(p); }
__device__ __forceinline__ uint32_t ld_u16(const uint8_t* p) { return *reinterpret_cast<const uint16_t*>(p); }
__device__ __forceinline__ uint32_t ld_u8(const uint8_t* p) { return *p; }

__device__ __forceinline__ bool maskOn(const uint8_t* mask, uint32_t n)
{
    return (ld_u32(mask + 4 * (n >> 5)) >> (n & 31)) & 1u;
}

// ------------------------------------------------------------------ leaf codecs
// float NanoVDB.h:3703; Fp4 :3839-3847; Fp8 :3875-3878; Fp16 :3906-3909; FpN :3937-3951.
// Quantised values are float(code)*quantum + minimum with no fused multiply-add.
template<uint32_t GT> struct Codec;
template<> struct Codec<GT_FLOAT> {
    static __device__ __forceinline__ float get(const uint8_t* leaf, uint32_t n, const float*)
    {
        return ld_f32(leaf + kLeafData + 4 * n);
    }
};
__device__ __forceinline__ float dequant(const uint8_t* leaf, uint32_t code)
{
    return __fadd_rn(__fmul_rn(float(code), ld_f32(leaf + kLeafQuantum)), ld_f32(leaf + kLeafMin));
}
template<> struct Codec<GT_FP4> {
    static __device__ __forceinline__ float get(const uint8_t* leaf, uint32_t n, const float*)
    {
        return dequant(leaf, (ld_u8(leaf + kLeafData + (n >> 1)) >> ((n & 1) << 2)) & 15u);
    }
};
template<> struct Codec<GT_FP8> {
    static __device__ __forceinline__ float get(const uint8_t* leaf, uint32_t n, const float*)
    {
        return dequant(leaf, ld_u8(leaf + kLeafData + n));
    }
};
template<> struct Codec<GT_FP16> {
    static __device__ __forceinline__ float get(const uint8_t* leaf, uint32_t n, const float*)
    {
        return dequant(leaf, ld_u16(leaf + kLeafData + 2 * n));
    }
};
template<> struct Codec<GT_FPN> {
    static __device__ __forceinline__ float get(const uint8_t* leaf, uint32_t n, const float*)
    {
        const uint32_t b = ld_u8(leaf + kLeafFlags) >> 5; // log2 of the bit width
        uint32_t code = ld_u32(leaf + kLeafData + 4 * (n >> (5 - b)));
        code >>= (n & ((32u >> b) - 1u)) << b;
        code &= (1u << (1u << b)) - 1u;
        return dequant(leaf, code);
    }
};
// LeafData<ValueOnIndex>::getValue (NanoVDB.h:4160-4170): index 0 (the background) for an inactive
// voxel, else offset + the active voxels before n (popcount within its mask word + the 9-bit
// prefix sum of the words before); the value is channel[index] (ChannelAccessor, :5790-5792).
template<> struct Codec<GT_ONINDEX> {
    static __device__ __forceinline__ uint64_t index(const uint8_t* leaf, uint32_t n)
    {
        const uint32_t w = n >> 6;
        const uint64_t word = ld_u64(leaf + kLeafVMask + 8 * w), bit = uint64_t(1) << (n & 63u);
        if (!(word & bit)) return 0;
        uint64_t sum = ld_u64(leaf + kIdxLeafOffset) + uint64_t(__popcll(word & (bit - 1u)));
        if (w) sum += (ld_u64(leaf + kIdxLeafPrefix) >> (9u * (w - 1u))) & 511u;
        return sum;
    }
    static __device__ __forceinline__ float get(const uint8_t* leaf, uint32_t n, const float* ch)
    {
        return ch[index(leaf, n)];
    }
};

// ------------------------------------------------------------------ accessor
// Returned values and getDim results do not depend on the cache state
// (NanoVDB.h:5379-5403), so a single-key three-level cache (the
// NANOVDB_USE_SINGLE_ACCESSOR_KEY form, NanoVDB.h:5329-5348) is used: one
// Coord key plus three node pointers per lane.
struct Probe {
    float value;
    uint32_t dim;       // ReadAccessor::getDim semantics
    const uint8_t* leaf; // leaf holding the voxel or nullptr
    uint32_t n;         // voxel offset in the leaf
    bool active;
    uint64_t index;     // index grids: the channel index the value comes from
};

template<uint32_t GT>
struct Accessor {
    static constexpr uint32_t kType = GT;
    const uint8_t* root;
    const float* channel;
    uint32_t tableSize;
    float background;
    int32_t kx, ky, kz;
    const uint8_t* upper;
    const uint8_t* lower;
    const uint8_t* leaf;

    __device__ __forceinline__ explicit Accessor(const GridView& grid)
        : root(grid.root), channel(grid.channel), tableSize(grid.tableSize),
          background(GT == GT_ONINDEX ? grid.channel[ld_u64(grid.root + kIdxRootBackground)] : grid.background),
          kx(0x7fffffff), ky(0x7fffffff), kz(0x7fffffff), upper(nullptr), lower(nullptr), leaf(nullptr) {}

    // value of a node-table entry or root tile: a float in the low word, or an index into the channel
    __device__ __forceinline__ float entryValue(uint64_t entry) const
    {
        if constexpr (GT == GT_ONINDEX) return channel[entry];
        else return __uint_as_float(uint32_t(entry));
    }

    __device__ __forceinline__ static uint32_t leafOffset(int32_t x, int32_t y, int32_t z)
    {
        return ((uint32_t(x) & 7u) << 6) | ((uint32_t(y) & 7u) << 3) | (uint32_t(z) & 7u);
    }

    __device__ __forceinline__ void fromLeaf(Probe& r, int32_t x, int32_t y, int32_t z) const
    {
        const uint32_t n = leafOffset(x, y, z);
        r.leaf = leaf;
        r.n = n;
        if constexpr (GT == GT_ONINDEX) {
            r.index = Codec<GT_ONINDEX>::index(leaf, n);
            r.value = channel[r.index];
        } else {
            r.value = Codec<GT>::get(leaf, n, channel);
        }
        r.active = maskOn(leaf + kLeafVMask, n);
        r.dim = (ld_u8(leaf + kLeafFlags) & 1u) ? 8u : 1u;
    }

    // lower node: InternalNode<Leaf,4>::getAndCache / getDimAndCache, NanoVDB.h:3619-3647
    __device__ __forceinline__ void fromLower(Probe& r, int32_t x, int32_t y, int32_t z)
    {
        const uint32_t n = (((uint32_t(x) & 127u) >> 3) << 8) | (((uint32_t(y) & 127u) >> 3) << 4) | ((uint32_t(z) & 127u) >> 3);
        const uint64_t entry = ld_u64(lower + kLoTable + 8 * n);
        if (maskOn(lower + kLoCMask, n)) {
            leaf = lower + int64_t(entry);
            kx = x; ky = y; kz = z;
            fromLeaf(r, x, y, z);
            if (ld_u64(lower + kLoFlags) & 1u) {
                // render-skip flag: getDimAndCache stops at this node (NanoVDB.h:3622); keep the
                // leaf out of the cache so later getDim calls see the flag again
                r.dim = 128u;
                leaf = nullptr;
            }
        } else {
            leaf = nullptr;
            r.leaf = nullptr;
            r.n = 0;
            r.value = entryValue(entry);
            if constexpr (GT == GT_ONINDEX) r.index = entry;
            r.active = maskOn(lower + kLoVMask, n);
            r.dim = (ld_u64(lower + kLoFlags) & 1u) ? 128u : 8u;
        }
    }

    // upper node: InternalNode<Lower,5>, NanoVDB.h:3619-3647
    __device__ __forceinline__ void fromUpper(Probe& r, int32_t x, int32_t y, int32_t z)
    {
        const uint32_t n = (((uint32_t(x) & 4095u) >> 7) << 10) | (((uint32_t(y) & 4095u) >> 7) << 5) | ((uint32_t(z) & 4095u) >> 7);
        const uint64_t entry = ld_u64(upper + kUpTable + 8 * n);
        if (maskOn(upper + kUpCMask, n)) {
            lower = upper + int64_t(entry);
            kx = x; ky = y; kz = z;
            fromLower(r, x, y, z);
        } else {
            lower = nullptr;
            leaf = nullptr;
            r.leaf = nullptr;
            r.n = 0;
            r.value = entryValue(entry);
            if constexpr (GT == GT_ONINDEX) r.index = entry;
            r.active = maskOn(upper + kUpVMask, n);
            r.dim = 128u;
        }
        if (ld_u64(upper + kUpFlags) & 1u) {
            r.dim = 4096u;
            lower = nullptr;
            leaf = nullptr;
        }
    }

    // root: CoordToKey (single 64-bit key, NanoVDB.h:2633-2640) + linear tile probe (:2785-2799)
    __device__ __forceinline__ void fromRoot(Probe& r, int32_t x, int32_t y, int32_t z)
    {
        const uint64_t key = uint64_t(uint32_t(z) >> 12) | (uint64_t(uint32_t(y) >> 12) << 21) |
                             (uint64_t(uint32_t(x) >> 12) << 42);
        const uint8_t* tile = nullptr;
        for (uint32_t i = 0; i < tableSize; ++i) {
            const uint8_t* t = root + rootHeaderSize(GT) + kTileSize * i;
            if (ld_u64(t) == key) { tile = t; break; }
        }
        upper = nullptr;
        lower = nullptr;
        leaf = nullptr;
        r.leaf = nullptr;
        r.n = 0;
        r.dim = 4096u;
        if (!tile) {
            r.value = background;
            r.active = false;
            if constexpr (GT == GT_ONINDEX) r.index = ld_u64(root + kIdxRootBackground);
            return;
        }
        const int64_t child = ld_i64(tile + 8);
        if (child == 0) {
            if constexpr (GT == GT_ONINDEX) {
                r.index = ld_u64(tile + kIdxTileValue);
                r.value = channel[r.index];
            } else {
                r.value = ld_f32(tile + 20);
            }
            r.active = ld_u32(tile + 16) > 0u;
            return;
        }
        upper = root + child;
        kx = x; ky = y; kz = z;
        fromUpper(r, x, y, z);
    }

    __device__ __forceinline__ Probe probe(int32_t x, int32_t y, int32_t z)
    {
        Probe r;
        const uint32_t dirty = uint32_t((x ^ kx) | (y ^ ky) | (z ^ kz));
        if (leaf && !(dirty & ~7u)) fromLeaf(r, x, y, z);
        else if (lower && !(dirty & ~127u)) fromLower(r, x, y, z);
        else if (upper && !(dirty & ~4095u)) fromUpper(r, x, y, z);
        else fromRoot(r, x, y, z);
        return r;
    }

    __device__ __forceinline__ float getValue(int32_t x, int32_t y, int32_t z)
    {
        const uint32_t dirty = uint32_t((x ^ kx) | (y ^ ky) | (z ^ kz));
        if (leaf && !(dirty & ~7u)) return Codec<GT>::get(leaf, leafOffset(x, y, z), channel);
        return probe(x, y, z).value;
    }
};

// ------------------------------------------------------------------ samplers
// lerp a + w*(b-a), math/SampleFromVoxels.h:423-464
__device__ __forceinline__ float lerpf(float a, float b, float w) { return a + w * (b - a); }

struct Cell { float v[2][2][2]; };

template<class AccT>
__device__ __forceinline__ void stencil8(AccT& acc, int32_t i, int32_t j, int32_t k, Cell& c)
{
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) c.v[a][b][e] = acc.getValue(i + a, j + b, k + e);
}

// c.v[a][b][e] for a,b,e in {0,1} without dynamic register indexing.  The empty asm keeps the
// eight corners as register values: otherwise the optimiser rewrites the selects of elements into
// one load through a selected address, which forces the cell out of registers into memory.
__device__ __forceinline__ float pick(const Cell& c, int32_t a, int32_t b, int32_t e)
{
    float v000 = c.v[0][0][0], v001 = c.v[0][0][1], v010 = c.v[0][1][0], v011 = c.v[0][1][1];
    float v100 = c.v[1][0][0], v101 = c.v[1][0][1], v110 = c.v[1][1][0], v111 = c.v[1][1][1];
    asm("" : "+v"(v000), "+v"(v001), "+v"(v010), "+v"(v011), "+v"(v100), "+v"(v101), "+v"(v110), "+v"(v111));
    const float x0 = b ? (e ? v011 : v010) : (e ? v001 : v000);
    const float x1 = b ? (e ? v111 : v110) : (e ? v101 : v100);
    return a ? x1 : x0;
}

__device__ __forceinline__ float trilerp(const Cell& s, float u, float v, float w)
{
    return lerpf(lerpf(lerpf(s.v[0][0][0], s.v[0][0][1], w), lerpf(s.v[0][1][0], s.v[0][1][1], w), v),
                 lerpf(lerpf(s.v[1][0][0], s.v[1][0][1], w), lerpf(s.v[1][1][0], s.v[1][1][1], w), v), u);
}

// Floor<Coord>(Vec3f&) + fused trilinear, math/SampleFromVoxels.h:61-69, 423-464
template<class AccT>
__device__ __forceinline__ float trilinear(AccT& acc, float x, float y, float z)
{
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    Cell c;
    stencil8(acc, int32_t(fx), int32_t(fy), int32_t(fz), c);
    return trilerp(c, x - fx, y - fy, z - fz);
}

// ------------------------------------------------------------------ sliding corner window
// The eight corner values of the current trilinear cell, kept across the steps of a march.  A
// move to a cell whose offset is -1, 0 or +1 on every axis keeps the shared corners and looks
// up only the new ones; larger jumps reload all eight.  Values are ReadAccessor::getValue
// results, so the sample does not depend on the window.  With KEYS, every corner also carries
// its leaf-value index (leaf index * 512 + voxel offset, -1 outside leaves; Float leaves) and a
// partial sum that is handed to `sink(key, sum)` exactly once, when the corner leaves the window:
// the cells sharing a corner voxel form a convex 2x2x2 block, so a ray visits them in one run.
struct NoSink {
    __device__ __forceinline__ void operator()(int32_t, float) const {}
};

// f(std::integral_constant<int, I>) for I = 0..N-1, so register arrays are indexed by constants
template<class F, int... I>
__device__ __forceinline__ void staticForImpl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template<int N, class F>
__device__ __forceinline__ void staticFor(F&& f)
{
    staticForImpl(f, std::make_integer_sequence<int, N>{});
}

// slot k (0..3) among those whose `bit` is clear (slot s = a*4 + b*2 + e)
__host__ __device__ constexpr int clearSlot(int k, int bit)
{
    return bit == 4 ? k : (bit == 2 ? (((k & 2) << 1) | (k & 1)) : (k << 1));
}

template<bool KEYS>
struct CellWindow {
    int32_t ci, cj, ck;
    Cell c;
    int32_t key[KEYS ? 8 : 1];
    float sum[KEYS ? 8 : 1];

    __device__ __forceinline__ CellWindow() : ci(0x7fffffff), cj(0), ck(0)
    {
        if constexpr (KEYS) {
            staticFor<8>([&](auto S) {
                key[S] = -1;
                sum[S] = 0.0f;
            });
        }
    }

    template<int S, class Sink>
    __device__ __forceinline__ void drop(Sink& sink)
    {
        if constexpr (KEYS) {
            if (key[S] >= 0 && sum[S] != 0.0f) sink(key[S], sum[S]);
            key[S] = -1;
            sum[S] = 0.0f;
        }
    }

    // shift the window by d in {-1, 0, +1} along the axis of slot bit BIT (4 = x, 2 = y, 1 = z).
    // Both slots of a pair are written unconditionally with selected values (the kept corner; the
    // vacated slot is reloaded afterwards), so every register index stays a constant.
    template<int BIT, class Sink>
    __device__ __forceinline__ void shift(int32_t d, Sink& sink)
    {
        if (d == 0) return;
        const bool up = d > 0;
        staticFor<4>([&](auto K) {
            constexpr int lo = clearSlot(decltype(K)::value, BIT), hi = lo | BIT;
            constexpr int la = lo >> 2, lb = (lo >> 1) & 1, le = lo & 1;
            constexpr int ha = hi >> 2, hb = (hi >> 1) & 1, he = hi & 1;
            const float kept = up ? c.v[ha][hb][he] : c.v[la][lb][le];
            c.v[la][lb][le] = kept;
            c.v[ha][hb][he] = kept;
            if constexpr (KEYS) {
                const int32_t kOut = up ? key[lo] : key[hi], kKeep = up ? key[hi] : key[lo];
                const float sOut = up ? sum[lo] : sum[hi], sKeep = up ? sum[hi] : sum[lo];
                if (kOut >= 0 && sOut != 0.0f) sink(kOut, sOut);
                key[lo] = up ? kKeep : -1;
                sum[lo] = up ? sKeep : 0.0f;
                key[hi] = up ? -1 : kKeep;
                sum[hi] = up ? 0.0f : sKeep;
            }
        });
    }

    template<class AccT, class Sink>
    __device__ __forceinline__ void moveTo(AccT& acc, const GridView& g, int32_t ni, int32_t nj, int32_t nk, Sink& sink)
    {
        const int32_t dx = ni - ci, dy = nj - cj, dz = nk - ck;
        if ((dx | dy | dz) == 0) return;
        uint32_t need = 0xffu;
        if (uint32_t(dx + 1) > 2u || uint32_t(dy + 1) > 2u || uint32_t(dz + 1) > 2u) {
            staticFor<8>([&](auto S) { drop<decltype(S)::value>(sink); });
        } else {
            shift<4>(dx, sink);
            shift<2>(dy, sink);
            shift<1>(dz, sink);
            // new corner (a,b,e) was an old corner iff a+dx, b+dy and e+dz all lie in {0,1}
            // (a = 1 is slots 0xf0, b = 1 is 0xcc, e = 1 is 0xaa)
            const uint32_t mx = dx == 0 ? 0u : (dx > 0 ? 0xf0u : 0x0fu);
            const uint32_t my = dy == 0 ? 0u : (dy > 0 ? 0xccu : 0x33u);
            const uint32_t mz = dz == 0 ? 0u : (dz > 0 ? 0xaau : 0x55u);
            need = mx | my | mz;
        }
        ci = ni;
        cj = nj;
        ck = nk;
        staticFor<8>([&](auto S) {
            constexpr int s = decltype(S)::value, a = s >> 2, b = (s >> 1) & 1, e = s & 1;
            if (!(need & (1u << s))) return;
            if constexpr (KEYS) {
                const Probe p = acc.probe(ni + a, nj + b, nk + e);
                c.v[a][b][e] = p.value;
                if constexpr (AccT::kType == GT_ONINDEX) {
                    key[s] = int32_t(p.index); // the channel entry the corner reads (< 2^31 values)
                } else {
                    // leaf index = byte offset / 2144 (Float leaf size) = (offset / 32) / 67, exact and
                    // in 32 bits for the <= 2^22 leaves the adjoint accepts
                    key[s] = p.leaf ? int32_t((uint32_t(uint64_t(p.leaf - g.leaf0) >> 5) / 67u) * 512u + p.n) : -1;
                }
                sum[s] = 0.0f;
            } else {
                c.v[a][b][e] = acc.getValue(ni + a, nj + b, nk + e);
            }
        });
    }

    template<class Sink>
    __device__ __forceinline__ void flush(Sink& sink)
    {
        staticFor<8>([&](auto S) { drop<decltype(S)::value>(sink); });
    }
};

// TrilinearSampler::gradient, math/SampleFromVoxels.h:232-271 (index space)
__device__ __forceinline__ void trilinearGradient(const Cell& s, float u, float v, float w, float out[3])
{
    float D[4] = {s.v[0][0][1] - s.v[0][0][0], s.v[0][1][1] - s.v[0][1][0], s.v[1][0][1] - s.v[1][0][0],
                  s.v[1][1][1] - s.v[1][1][0]};
    out[2] = lerpf(lerpf(D[0], D[1], v), lerpf(D[2], D[3], v), u);
    D[0] = s.v[0][0][0] + D[0] * w;
    D[1] = s.v[0][1][0] + D[1] * w;
    D[2] = s.v[1][0][0] + D[2] * w;
    D[3] = s.v[1][1][0] + D[3] * w;
    out[0] = lerpf(D[2], D[3], v) - lerpf(D[0], D[1], v);
    out[1] = lerpf(D[1] - D[0], D[3] - D[2], u);
}

// ------------------------------------------------------------------ Ray / HDDA
// math/Ray.h:64-74, 154-163, 258-268, 398-424; math/HDDA.h:28-179
struct Ray {
    float eye[3], dir[3], inv[3], t0, t1;
};

__device__ __forceinline__ void rayAt(const Ray& r, float t, float p[3])
{
    p[0] = __builtin_fmaf(t, r.dir[0], r.eye[0]);
    p[1] = __builtin_fmaf(t, r.dir[1], r.eye[1]);
    p[2] = __builtin_fmaf(t, r.dir[2], r.eye[2]);
}

__device__ __forceinline__ Ray makeRay(const float e[3], const float d[3], float t0, float t1)
{
    Ray r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.eye[a] = e[a];
        r.dir[a] = d[a];
        r.inv[a] = 1.0f / d[a];
    }
    r.t0 = t0;
    r.t1 = t1;
    return r;
}

// World ray -> index ray.  Ray::worldToIndexF (math/Ray.h:258-268): Map::applyInverseMapF /
// applyInverseJacobianF (fmaf form, math/Math.h:1704-1709), t1 left at FLT_MAX.  With
// g.rayMapF64, Ray::applyInverseMap (math/Ray.h:223-230): the FP64 Map::applyInverseMap /
// applyInverseJacobian (float(xyz - vecD), then a double fma chain rounded to float,
// NanoVDB.h:1507-1510, math/Math.h:1718-1723) and both times scaled by the length.
__device__ __forceinline__ Ray worldToIndexRay(const GridView& g, const float e[3], const float d[3])
{
    float eye[3], dir[3];
    if (g.rayMapF64) {
        const float t[3] = {float(double(e[0]) - g.vecD[0]), float(double(e[1]) - g.vecD[1]),
                            float(double(e[2]) - g.vecD[2])};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const double* m = g.invMatD + 3 * r;
            eye[r] = float(__builtin_fma(double(t[0]), m[0], __builtin_fma(double(t[1]), m[1], double(t[2]) * m[2])));
            dir[r] = float(__builtin_fma(double(d[0]), m[0], __builtin_fma(double(d[1]), m[1], double(d[2]) * m[2])));
        }
    } else {
        const float t[3] = {e[0] - g.vecF[0], e[1] - g.vecF[1], e[2] - g.vecF[2]};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            eye[r] = __builtin_fmaf(t[0], g.invMatF[3 * r], __builtin_fmaf(t[1], g.invMatF[3 * r + 1], t[2] * g.invMatF[3 * r + 2]));
            dir[r] = __builtin_fmaf(d[0], g.invMatF[3 * r], __builtin_fmaf(d[1], g.invMatF[3 * r + 1], d[2] * g.invMatF[3 * r + 2]));
        }
    }
    const float length = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const float invLength = 1.0f / length;
    const float nd[3] = {invLength * dir[0], invLength * dir[1], invLength * dir[2]};
    return makeRay(eye, nd, length * 1.0e-5f, g.rayMapF64 ? length * kFltMax : kFltMax);
}

// Ray::clip(CoordBBox) with the +1 max padding
__device__ __forceinline__ bool rayClip(Ray& r, const int32_t* bbox)
{
    float t0 = r.t0, t1 = r.t1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float a = float(bbox[i]), b = float(bbox[3 + i] + 1);
        if (a >= b) return false;
        a = (a - r.eye[i]) * r.inv[i];
        b = (b - r.eye[i]) * r.inv[i];
        if (a > b) { const float tmp = a; a = b; b = tmp; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return false;
    }
    r.t0 = t0;
    r.t1 = t1;
    return true;
}

__device__ __forceinline__ int32_t ifloor(float x) { return int32_t(floorf(x)); }

// Per-axis state is kept in scalar members: selecting among array elements by the minimum axis
// would be folded by the compiler into a dynamically indexed array and spilled to scratch.
struct HDDA {
    int32_t dim;
    float t0, t1;
    int32_t vx, vy, vz;    // current voxel / node origin
    int32_t sx, sy, sz;    // step direction per axis
    float dx, dy, dz;      // delta time per axis
    float nx, ny, nz;      // next crossing time per axis

    __device__ __forceinline__ static void initAxis(float dir, float inv, float pos, int32_t vox, int32_t d, float tStart,
                                                    int32_t& st, float& de, float& nxt)
    {
        de = 0.0f;
        if (dir == 0.0f) {
            nxt = kFltMax;
            st = 0;
        } else if (inv > 0.0f) {
            st = 1;
            nxt = tStart + (float(vox + d) - pos) * inv;
            de = inv;
        } else {
            st = -1;
            nxt = tStart + (float(vox) - pos) * inv;
            de = -inv;
        }
    }

    __device__ __forceinline__ void init(const Ray& r, float startTime, float maxTime, int32_t d)
    {
        dim = d;
        t0 = startTime;
        t1 = maxTime;
        float pos[3];
        rayAt(r, t0, pos);
        vx = ifloor(pos[0]) & ~(d - 1);
        vy = ifloor(pos[1]) & ~(d - 1);
        vz = ifloor(pos[2]) & ~(d - 1);
        initAxis(r.dir[0], r.inv[0], pos[0], vx, d, t0, sx, dx, nx);
        initAxis(r.dir[1], r.inv[1], pos[1], vy, d, t0, sy, dy, ny);
        initAxis(r.dir[2], r.inv[2], pos[2], vz, d, t0, sz, dz, nz);
    }

    __device__ __forceinline__ static int32_t clampVox(float pos, int32_t vox, int32_t oldDim, int32_t d)
    {
        const int32_t vmax = (vox + (oldDim - 1)) & ~(d - 1), vmin = vox & ~(d - 1);
        int32_t v = ifloor(pos) & ~(d - 1);
        v = v < vmax ? v : vmax;
        return v > vmin ? v : vmin;
    }

    __device__ __forceinline__ static void nextAxis(int32_t st, int32_t vox, float pos, float inv, int32_t d, float tStart,
                                                    float& nxt)
    {
        if (st == 0) return;
        nxt = tStart + (float(vox) - pos) * inv;
        if (st > 0) nxt += float(d) * inv;
    }

    __device__ __forceinline__ bool update(const Ray& r, int32_t d)
    {
        if (dim == d) return false;
        float pos[3];
        rayAt(r, t0, pos);
        vx = clampVox(pos[0], vx, dim, d);
        vy = clampVox(pos[1], vy, dim, d);
        vz = clampVox(pos[2], vz, dim, d);
        dim = d;
        nextAxis(sx, vx, pos[0], r.inv[0], d, t0, nx);
        nextAxis(sy, vy, pos[1], r.inv[1], d, t0, ny);
        nextAxis(sz, vz, pos[2], r.inv[2], d, t0, nz);
        return true;
    }

    __device__ __forceinline__ bool advance(float& nxt, float de, int32_t& vox, int32_t st)
    {
        if (nxt <= t0) nxt += t0 - 0.999999f * nxt + 1.0e-6f; // ENFORCE_FORWARD_STEPPING, math/HDDA.h:163-169
        t0 = nxt;
        nxt += float(dim) * de;
        vox += dim * st;
        return t0 <= t1;
    }

    // MinIndex (math/Math.h:294-309) then step along that axis
    __device__ __forceinline__ bool stepOnce()
    {
        if (nx < ny && nx < nz) return advance(nx, dx, vx, sx);
        if (ny < nz) return advance(ny, dy, vy, sy);
        return advance(nz, dz, vz, sz);
    }
};

// ------------------------------------------------------------------ isoCrossing / zeroCrossing
// math/HDDA.h:225-248 (isoCrossing) over the device accessor on an index-space ray.  zeroCrossing
// (math/HDDA.h:189-214) is the same loop with iso = 0: value - 0.0f is the value, bit for bit.
// Returns the first voxel whose value - iso changes sign against the value at the entry voxel.
template<uint32_t GT>
__device__ __forceinline__ bool isoCrossingCore(Accessor<GT>& acc, const GridView& g, Ray& ray, float iso,
                                                int32_t c[3], float& v, float& t)
{
    if (!rayClip(ray, g.bbox) || double(ray.t1) > 1e20) return false;
    const float Delta = 1.0001f;
    float p[3];
    rayAt(ray, ray.t0, p);
    c[0] = ifloor(p[0]);
    c[1] = ifloor(p[1]);
    c[2] = ifloor(p[2]);
    const Probe p0 = acc.probe(c[0], c[1], c[2]);
    HDDA h;
    h.init(ray, ray.t0, ray.t1, int32_t(p0.dim));
    const float v0 = p0.value - iso;
    // The reference nests a voxel-by-voxel loop through the narrow band inside the node-level
    // loop.  Here both are one loop with a mode flag: every lane performs the same sequence of
    // HDDA steps and probes as the nested form (a step that fails ends the march in both, since
    // HDDA time only grows), but a wavefront no longer runs one lane's band walk while the others
    // wait at their own outer step, which made rays at the silhouette cost the sum of their
    // lanes' walks.
    bool band = false;
    while (h.stepOnce()) {
        int32_t x, y, z;
        if (band) {
            x = h.vx;
            y = h.vy;
            z = h.vz;
        } else {
            rayAt(ray, h.t0 + Delta, p);
            x = ifloor(p[0]);
            y = ifloor(p[1]);
            z = ifloor(p[2]);
        }
        const Probe q = acc.probe(x, y, z);
        if (band) { // in the narrow band
            if (!q.active) {
                band = false;
                continue;
            }
            v = q.value - iso;
            if (v * v0 < 0) {
                c[0] = x;
                c[1] = y;
                c[2] = z;
                t = h.t0;
                return true;
            }
        } else {
            c[0] = x;
            c[1] = y;
            c[2] = z;
            h.update(ray, int32_t(q.dim));
            band = !(h.dim > 1) && q.active;
        }
    }
    return false;
}

// ------------------------------------------------------------------ fog-march step absorption
// a = exp(-sigma * rho * dt) of one march step, as exp2 of the exponent times log2(e) on the
// hardware exp2 (1 ulp): two instructions instead of the fourteen of the range-reduced library expf,
// within two ulp of it for the small exponents of a march step.  The forward march and its adjoint
// both call this, so the adjoint's re-march reproduces the forward's transmittance bit for bit.
__device__ __forceinline__ float stepAbsorption(float sigma, float rho, float dt)
{
    return __expf(-sigma * rho * dt);
}

// ------------------------------------------------------------------ fog-march step structure
// HDDA over cells of size >= 8 (the leaf level, as TreeMarcher, math/HDDA.h:450-536, with
// Eps = 1e-6 and Delta = 0.01); a cell is marched when it lies in a leaf (getDim == 1) or in an
// active tile; steps are globally aligned at t = k*dt with ta <= t < tb inside each marched
// cell.  onStep(x, y, z) gets every step's index-space position, in order; the forward march and its adjoint share
// this routine, so both see the same steps bit for bit.
template<class AccT, class F>
__device__ __forceinline__ void marchSteps(AccT& acc, const GridView& g, const float e[3], const float d[3], float dt,
                                           F&& onStep)
{
    Ray R = worldToIndexRay(g, e, d);
    if (!rayClip(R, g.bbox) || double(R.t1) > 1e20) return;
    const float Eps = 0.000001f, Delta = 0.01f;
    const float t0 = R.t0 + Eps, t1 = R.t1 - Eps;
    if (t0 > t1) return;
    float p[3];
    rayAt(R, t0, p);
    Probe l = acc.probe(ifloor(p[0]), ifloor(p[1]), ifloor(p[2]));
    HDDA h;
    h.init(R, t0, t1, int32_t(l.dim > 8u ? l.dim : 8u));
    bool on = (l.dim == 1u) || l.active;
    for (;;) {
        const float ta = h.t0;
        const bool valid = h.stepOnce();
        const float tb = fminf(h.t0, t1);
        if (on && ta < tb) {
            float kf = ceilf(ta / dt);
            for (float t = kf * dt; t < tb; t = kf * dt) {
                const float x0 = __builtin_fmaf(t, R.dir[0], R.eye[0]);
                const float x1 = __builtin_fmaf(t, R.dir[1], R.eye[1]);
                const float x2 = __builtin_fmaf(t, R.dir[2], R.eye[2]);
                onStep(x0, x1, x2);
                const float kn = kf + 1.0f;
                if (kn == kf) break; // step index beyond 2^24: float cannot advance it
                kf = kn;
            }
        }
        if (!valid) break;
        rayAt(R, h.t0 + Delta, p);
        l = acc.probe(ifloor(p[0]), ifloor(p[1]), ifloor(p[2]));
        h.update(R, int32_t(l.dim > 8u ? l.dim : 8u));
        on = (l.dim == 1u) || l.active;
    }
}

} // namespace nvq
