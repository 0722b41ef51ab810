// Host-side NanoVDB grid builder of libnvq.so (C ABI in include/nvq_build.h).
//
// A sparse build tree (root map -> 32^3 upper -> 16^3 lower -> 8^3 leaf) is
// filled the way the reference's build::Grid is filled, then serialised
// breadth-first into the flat NanoVDB layout with per-node statistics.  The
// steps that decide bytes follow the reference so the buffers agree:
//   voxel values of the primitives     tools/CreatePrimitives.h:598-738
//   signed flood fill                  tools/GridBuilder.h:575-606, 1055-1088, 1636-1666, 2147-2164
//   level set -> fog                   tools/GridBuilder.h:2166-2238
//   functor fill + constant pruning    tools/GridBuilder.h:1927-2019
//   serialisation order and codecs     tools/CreateNanoGrid.h:887-959, 1009-1092, 1185-1474, 1481-1611, 1651-1700
//   statistics, bboxes and node flags  tools/GridStats.h:230-330, 580-756
// The checksum field is written disabled (all bits set).
#include "../../include/nvq.h"
#include "../../include/nvq_build.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

// ------------------------------------------------------------------ basics
struct Coord {
    int32_t v[3];
    Coord() : v{0, 0, 0} {}
    Coord(int32_t x, int32_t y, int32_t z) : v{x, y, z} {}
    int32_t& operator[](int i) { return v[i]; }
    int32_t operator[](int i) const { return v[i]; }
    bool operator<(const Coord& o) const
    {
        return v[0] < o.v[0] ? true : v[0] > o.v[0] ? false : v[1] < o.v[1] ? true : v[1] > o.v[1] ? false : v[2] < o.v[2];
    }
    bool operator==(const Coord& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
    Coord masked(int32_t m) const { return Coord(v[0] & ~m, v[1] & ~m, v[2] & ~m); }
};

struct CoordBox {
    Coord lo{std::numeric_limits<int32_t>::max(), std::numeric_limits<int32_t>::max(), std::numeric_limits<int32_t>::max()};
    Coord hi{std::numeric_limits<int32_t>::min(), std::numeric_limits<int32_t>::min(), std::numeric_limits<int32_t>::min()};
    bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
    void expand(const Coord& a, const Coord& b)
    {
        for (int i = 0; i < 3; ++i) {
            lo[i] = std::min(lo[i], a[i]);
            hi[i] = std::max(hi[i], b[i]);
        }
    }
};

template<int WORDS>
struct Mask {
    uint64_t w[WORDS];
    void clear() { std::memset(w, 0, sizeof(w)); }
    void fill() { std::memset(w, 0xff, sizeof(w)); }
    bool on(uint32_t n) const { return (w[n >> 6] >> (n & 63)) & 1u; }
    void set(uint32_t n, bool s)
    {
        if (s) w[n >> 6] |= uint64_t(1) << (n & 63);
        else w[n >> 6] &= ~(uint64_t(1) << (n & 63));
    }
    bool isOff() const
    {
        for (int i = 0; i < WORDS; ++i)
            if (w[i]) return false;
        return true;
    }
    bool isOn() const
    {
        for (int i = 0; i < WORDS; ++i)
            if (~w[i]) return false;
        return true;
    }
    uint32_t count() const
    {
        uint32_t c = 0;
        for (int i = 0; i < WORDS; ++i) c += uint32_t(__builtin_popcountll(w[i]));
        return c;
    }
    uint32_t first() const
    {
        for (int i = 0; i < WORDS; ++i)
            if (w[i]) return uint32_t(i * 64 + __builtin_ctzll(w[i]));
        return WORDS * 64;
    }
};

// ------------------------------------------------------------------ build tree
struct LeafT {
    static constexpr int32_t TOTAL = 3;
    Coord origin;
    Mask<8> mask;
    bool prune = false;
    float v[512];
    LeafT(const Coord& o, float value, bool state) : origin(o.masked(7))
    {
        for (float& x : v) x = value;
        if (state) mask.fill();
        else mask.clear();
    }
    float first() const { return v[0]; }
    float last() const { return v[511]; }
};

template<typename ChildT, int LOG2>
struct Internal {
    using Child = ChildT;
    static constexpr uint32_t SIZE = 1u << (3 * LOG2);
    static constexpr int32_t TOTAL = LOG2 + ChildT::TOTAL;
    static constexpr int32_t DIM = 1 << TOTAL;
    Coord origin;
    Mask<SIZE / 64> cmask, vmask;
    std::vector<float> tile;
    std::vector<ChildT*> child;
    Internal(const Coord& o, float value, bool state) : origin(o.masked(DIM - 1)), tile(SIZE, value), child(SIZE, nullptr)
    {
        cmask.clear();
        if (state) vmask.fill();
        else vmask.clear();
    }
    ~Internal()
    {
        for (auto* c : child) delete c;
    }
    Internal(const Internal&) = delete;
    Internal& operator=(const Internal&) = delete;
    static uint32_t offset(const Coord& ijk)
    {
        return (((uint32_t(ijk[0]) & (DIM - 1)) >> ChildT::TOTAL) << (2 * LOG2)) |
               (((uint32_t(ijk[1]) & (DIM - 1)) >> ChildT::TOTAL) << LOG2) | ((uint32_t(ijk[2]) & (DIM - 1)) >> ChildT::TOTAL);
    }
    Coord childOrigin(uint32_t n) const
    {
        const uint32_t m = n & ((1u << (2 * LOG2)) - 1);
        return Coord(origin[0] + (int32_t(n >> (2 * LOG2)) << ChildT::TOTAL),
                     origin[1] + (int32_t(m >> LOG2) << ChildT::TOTAL),
                     origin[2] + (int32_t(m & ((1u << LOG2) - 1)) << ChildT::TOTAL));
    }
    float first() const { return cmask.on(0) ? child[0]->first() : tile[0]; }
    float last() const { return cmask.on(SIZE - 1) ? child[SIZE - 1]->last() : tile[SIZE - 1]; }
    void detach(uint32_t n)
    {
        delete child[n];
        child[n] = nullptr;
        cmask.set(n, false);
    }
};
using Lower = Internal<LeafT, 4>;
using Upper = Internal<Lower, 5>;

struct RootTile {
    Upper* child = nullptr;
    float value = 0.0f;
    bool state = false;
};

struct Tree {
    float background;
    std::map<Coord, RootTile> root;
    explicit Tree(float bg) : background(bg) {}
    ~Tree()
    {
        for (auto& kv : root) delete kv.second.child;
    }
    Tree(const Tree&) = delete;
    Tree& operator=(const Tree&) = delete;

    Upper* touchUpper(const Coord& ijk)
    {
        const Coord key = ijk.masked(4095);
        auto it = root.find(key);
        if (it == root.end()) {
            RootTile t;
            t.child = new Upper(key, background, false);
            root.emplace(key, t);
            return root[key].child;
        }
        if (!it->second.child) it->second.child = new Upper(key, it->second.value, it->second.state);
        return it->second.child;
    }
    Lower* touchLower(const Coord& ijk)
    {
        Upper* u = touchUpper(ijk);
        const uint32_t n = Upper::offset(ijk);
        if (!u->cmask.on(n)) {
            u->child[n] = new Lower(ijk, u->tile[n], u->vmask.on(n));
            u->cmask.set(n, true);
        }
        return u->child[n];
    }
    // build::Tree::addNode for a leaf (takes ownership)
    void addLeaf(LeafT* leaf)
    {
        Lower* l = touchLower(leaf->origin);
        const uint32_t n = Lower::offset(leaf->origin);
        if (l->cmask.on(n)) delete l->child[n];
        l->child[n] = leaf;
        l->cmask.set(n, true);
    }
    template<typename F>
    void forEachUpper(F&& f)
    {
        for (auto& kv : root)
            if (kv.second.child) f(*kv.second.child);
    }
    std::vector<Lower*> lowers()
    {
        std::vector<Lower*> out;
        forEachUpper([&](Upper& u) {
            for (uint32_t n = 0; n < Upper::SIZE; ++n)
                if (u.cmask.on(n)) out.push_back(u.child[n]);
        });
        return out;
    }
    std::vector<LeafT*> leaves()
    {
        std::vector<LeafT*> out;
        for (Lower* l : lowers())
            for (uint32_t n = 0; n < Lower::SIZE; ++n)
                if (l->cmask.on(n)) out.push_back(l->child[n]);
        return out;
    }
};

template<typename F>
void parallelFor(size_t n, F&& f)
{
    unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    if (n < 64 || nt == 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    nt = unsigned(std::min<size_t>(nt, n));
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto& th : pool) th.join();
}

// ------------------------------------------------------------------ primitives
// initSphere / initTorus: per-voxel signed distance in float with the reference's leapfrogging.
// Voxels are produced per x-slab of 8 (leaf-aligned), so every leaf is complete in one slab.
struct Voxel {
    Coord ijk;
    float v;
};

template<typename SlabFn>
void fillBySlabs(Tree& tree, int imin, int imax, SlabFn&& slab)
{
    const int s0 = imin & ~7, s1 = imax & ~7;
    const size_t nslab = size_t((s1 - s0) / 8 + 1);
    std::vector<std::map<Coord, LeafT*>> parts(nslab);
    parallelFor(nslab, [&](size_t s) {
        const int ib = std::max(imin, s0 + int(s) * 8), ie = std::min(imax, s0 + int(s) * 8 + 7);
        auto& leaves = parts[s];
        LeafT* cur = nullptr;
        slab(ib, ie, [&](const Coord& ijk, float v) {
            const Coord key = ijk.masked(7);
            if (!cur || !(cur->origin == key)) {
                auto it = leaves.find(key);
                if (it == leaves.end()) it = leaves.emplace(key, new LeafT(key, tree.background, false)).first;
                cur = it->second;
            }
            const uint32_t n = ((uint32_t(ijk[0]) & 7u) << 6) | ((uint32_t(ijk[1]) & 7u) << 3) | (uint32_t(ijk[2]) & 7u);
            cur->v[n] = v;
            cur->mask.set(n, true);
        });
    });
    for (auto& part : parts)
        for (auto& kv : part) tree.addLeaf(kv.second);
}

float fPow2(float x) { return x * x; }

void initSphere(Tree& tree, double radius, const double center[3], double voxelSize, double halfWidth)
{
    using ValueT = float;
    const ValueT r0 = ValueT(radius / voxelSize), rmax = r0 + ValueT(halfWidth);
    if (r0 < ValueT(1.5f)) return;
    const ValueT c[3] = {ValueT(center[0] - 0.0) / ValueT(voxelSize), ValueT(center[1] - 0.0) / ValueT(voxelSize),
                         ValueT(center[2] - 0.0) / ValueT(voxelSize)};
    const int imin = int(std::floor(c[0] - rmax)), imax = int(std::ceil(c[0] + rmax));
    const int jmin = int(std::floor(c[1] - rmax)), jmax = int(std::ceil(c[1] + rmax));
    const int kmin = int(std::floor(c[2] - rmax)), kmax = int(std::ceil(c[2] + rmax));
    fillBySlabs(tree, imin, imax, [&](int ib, int ie, auto&& emit) {
        for (int i = ib; i <= ie; ++i) {
            const ValueT x2 = fPow2(ValueT(i) - c[0]);
            for (int j = jmin; j <= jmax; ++j) {
                const ValueT x2y2 = fPow2(ValueT(j) - c[1]) + x2;
                for (int k = kmin, m = 1; k <= kmax; k += m) {
                    m = 1;
                    const ValueT v = std::sqrt(x2y2 + fPow2(ValueT(k) - c[2])) - r0;
                    const ValueT d = v < 0 ? -v : v;
                    if (d < halfWidth) emit(Coord(i, j, k), ValueT(voxelSize) * v);
                    else m += int(std::floor(d - halfWidth));
                }
            }
        }
    });
}

void initTorus(Tree& tree, double radius1, double radius2, const double center[3], double voxelSize, double halfWidth)
{
    using ValueT = float;
    const ValueT r1 = ValueT(radius1 / voxelSize), r2 = ValueT(radius2 / voxelSize),
                 rmax1 = r1 + r2 + ValueT(halfWidth), rmax2 = r2 + ValueT(halfWidth);
    if (r2 < ValueT(1.5)) return;
    const ValueT c[3] = {ValueT(center[0] - 0.0) / ValueT(voxelSize), ValueT(center[1] - 0.0) / ValueT(voxelSize),
                         ValueT(center[2] - 0.0) / ValueT(voxelSize)};
    const int imin = int(std::floor(c[0] - rmax1)), imax = int(std::ceil(c[0] + rmax1));
    const int jmin = int(std::floor(c[1] - rmax2)), jmax = int(std::ceil(c[1] + rmax2));
    const int kmin = int(std::floor(c[2] - rmax1)), kmax = int(std::ceil(c[2] + rmax1));
    fillBySlabs(tree, imin, imax, [&](int ib, int ie, auto&& emit) {
        for (int i = ib; i <= ie; ++i) {
            const ValueT x2 = fPow2(ValueT(i) - c[0]);
            for (int k = kmin; k <= kmax; ++k) {
                const ValueT x2z2 = fPow2(std::sqrt(fPow2(ValueT(k) - c[2]) + x2) - r1);
                for (int j = jmin, m = 1; j <= jmax; j += m) {
                    m = 1;
                    const ValueT v = std::sqrt(x2z2 + fPow2(ValueT(j) - c[1])) - r2;
                    const ValueT d = v < 0 ? -v : v;
                    if (d < halfWidth) emit(Coord(i, j, k), ValueT(voxelSize) * v);
                    else m += int(std::floor(d - halfWidth));
                }
            }
        }
    });
}

// ------------------------------------------------------------------ flood fill / fog
void floodLeaf(LeafT& L, float outside)
{
    const uint32_t first = L.mask.first();
    if (first >= 512) return;
    bool xIn = L.v[first] < 0, yIn = xIn, zIn = xIn;
    for (uint32_t x = 0; x != 8; ++x) {
        const uint32_t x00 = x << 6;
        if (L.mask.on(x00)) xIn = L.v[x00] < 0;
        yIn = xIn;
        for (uint32_t y = 0; y != 8; ++y) {
            const uint32_t xy0 = x00 + (y << 3);
            if (L.mask.on(xy0)) yIn = L.v[xy0] < 0;
            zIn = yIn;
            for (uint32_t z = 0; z != 8; ++z) {
                const uint32_t xyz = xy0 + z;
                if (L.mask.on(xyz)) zIn = L.v[xyz] < 0;
                else L.v[xyz] = zIn ? -outside : outside;
            }
        }
    }
}

template<typename NodeT, int LOG2>
void floodInternal(NodeT& N, float outside)
{
    const uint32_t first = N.cmask.first();
    if (first >= NodeT::SIZE) return;
    bool xIn = N.child[first]->first() < 0, yIn = xIn, zIn = xIn;
    for (uint32_t x = 0; x != (1u << LOG2); ++x) {
        const uint32_t x00 = x << (2 * LOG2);
        if (N.cmask.on(x00)) xIn = N.child[x00]->last() < 0;
        yIn = xIn;
        for (uint32_t y = 0; y != (1u << LOG2); ++y) {
            const uint32_t xy0 = x00 + (y << LOG2);
            if (N.cmask.on(xy0)) yIn = N.child[xy0]->last() < 0;
            zIn = yIn;
            for (uint32_t z = 0; z != (1u << LOG2); ++z) {
                const uint32_t xyz = xy0 + z;
                if (N.cmask.on(xyz)) zIn = N.child[xyz]->last() < 0;
                else N.tile[xyz] = zIn ? -outside : outside;
            }
        }
    }
}

void floodRoot(Tree& t, float outside)
{
    std::vector<std::pair<Coord, Upper*>> keys;
    for (auto& kv : t.root)
        if (kv.second.child) keys.emplace_back(kv.first, kv.second.child);
    if (keys.empty()) return;
    for (size_t i = 0; i + 1 < keys.size(); ++i) {
        const auto& a = keys[i];
        const auto& b = keys[i + 1];
        const Coord d(b.first[0] - a.first[0], b.first[1] - a.first[1], b.first[2] - a.first[2]);
        if (d[0] != 0 || d[1] != 0 || d[2] == 4096) continue;
        if (!(a.second->last() < 0) || !(b.second->first() < 0)) continue;
        for (Coord c(a.first[0], a.first[1], a.first[2] + 4096); c[2] != b.first[2]; c[2] += 4096) {
            RootTile tile;
            tile.value = -outside;
            tile.state = false;
            t.root[c] = tile;
        }
    }
}

void sdfToLevelSet(Tree& t)
{
    const float outside = t.background;
    auto leaves = t.leaves();
    parallelFor(leaves.size(), [&](size_t i) { floodLeaf(*leaves[i], outside); });
    auto lowers = t.lowers();
    parallelFor(lowers.size(), [&](size_t i) { floodInternal<Lower, 4>(*lowers[i], outside); });
    t.forEachUpper([&](Upper& u) { floodInternal<Upper, 5>(u, outside); });
    floodRoot(t, outside);
}

void levelSetToFog(Tree& t)
{
    const float d = -t.background, w = 1.0f / d;
    auto op = [&](float& v) -> bool {
        if (v > 0.0f) {
            v = 0.0f;
            return false;
        }
        v = v > d ? v * w : 1.0f;
        return true;
    };
    auto leaves = t.leaves();
    parallelFor(leaves.size(), [&](size_t i) {
        for (uint32_t n = 0; n < 512u; ++n) leaves[i]->mask.set(n, op(leaves[i]->v[n]));
    });
    auto lowers = t.lowers();
    parallelFor(lowers.size(), [&](size_t i) {
        Lower& node = *lowers[i];
        for (uint32_t n = 0; n < Lower::SIZE; ++n) {
            if (node.cmask.on(n)) {
                LeafT* leaf = node.child[n];
                if (leaf->mask.isOff()) {
                    node.tile[n] = leaf->first();
                    node.detach(n);
                }
            } else {
                node.vmask.set(n, op(node.tile[n]));
            }
        }
    });
    t.forEachUpper([&](Upper& node) {
        for (uint32_t n = 0; n < Upper::SIZE; ++n) {
            if (node.cmask.on(n)) {
                Lower* c = node.child[n];
                if (c->cmask.isOff() && c->vmask.isOff()) {
                    node.tile[n] = c->first();
                    node.detach(n);
                }
            } else {
                node.vmask.set(n, op(node.tile[n]));
            }
        }
    });
    for (auto& kv : t.root) {
        RootTile& tile = kv.second;
        if (!tile.child) {
            tile.state = op(tile.value);
        } else if (tile.child->cmask.isOff() && tile.child->vmask.isOff()) {
            tile.value = tile.child->first();
            tile.state = false;
            delete tile.child;
            tile.child = nullptr;
        }
    }
}

// build::Grid<float>::operator()(func, bbox): leaf-by-leaf evaluation, background voxels stay
// inactive, fully active constant leaves and then constant all-tile nodes are pruned.
template<typename F>
void fillFromFunction(Tree& t, const int32_t* bbox, F&& func)
{
    const Coord lo(bbox[0] >> 3, bbox[1] >> 3, bbox[2] >> 3), hi(bbox[3] >> 3, bbox[4] >> 3, bbox[5] >> 3);
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    std::vector<LeafT*> made(size_t(nx) * ny * nz, nullptr);
    parallelFor(made.size(), [&](size_t idx) {
        const int lx = lo[0] + int(idx / (size_t(ny) * nz)), ly = lo[1] + int((idx / nz) % ny), lz = lo[2] + int(idx % nz);
        const Coord mn(lx << 3, ly << 3, lz << 3);
        const Coord bmin(std::max(mn[0], bbox[0]), std::max(mn[1], bbox[1]), std::max(mn[2], bbox[2]));
        const Coord bmax(std::min(mn[0] + 7, bbox[3]), std::min(mn[1] + 7, bbox[4]), std::min(mn[2] + 7, bbox[5]));
        LeafT* leaf = new LeafT(bmin, t.background, false);
        for (int i = bmin[0]; i <= bmax[0]; ++i)
            for (int j = bmin[1]; j <= bmax[1]; ++j)
                for (int k = bmin[2]; k <= bmax[2]; ++k) {
                    const float v = func(i, j, k);
                    if (v != t.background) {
                        const uint32_t n = ((uint32_t(i) & 7u) << 6) | ((uint32_t(j) & 7u) << 3) | (uint32_t(k) & 7u);
                        leaf->v[n] = v;
                        leaf->mask.set(n, true);
                    }
                }
        if (leaf->mask.isOff()) {
            delete leaf;
            return;
        }
        if (leaf->mask.isOn()) {
            const float first = leaf->first();
            int n = 1;
            while (n < 512) {
                if (leaf->v[n++] != first) break;
            }
            if (n == 512) leaf->prune = true;
        }
        made[idx] = leaf;
    });
    for (LeafT* l : made)
        if (l) t.addLeaf(l);
    for (auto& kv : t.root) {
        Upper* upper = kv.second.child;
        if (!upper) continue;
        for (uint32_t n1 = 0; n1 < Upper::SIZE; ++n1) {
            if (!upper->cmask.on(n1)) continue;
            Lower* lower = upper->child[n1];
            for (uint32_t n0 = 0; n0 < Lower::SIZE; ++n0) {
                if (!lower->cmask.on(n0)) continue;
                LeafT* leaf = lower->child[n0];
                if (leaf->prune) {
                    lower->tile[n0] = leaf->first();
                    lower->detach(n0);
                    lower->vmask.set(n0, true);
                }
            }
            if (lower->cmask.isOff()) {
                const float first = lower->first();
                uint32_t n = 1;
                while (n < Lower::SIZE) {
                    if (lower->tile[n++] != first) break;
                }
                if (n == Lower::SIZE) {
                    upper->tile[n1] = first;
                    upper->detach(n1);
                    upper->vmask.set(n1, true);
                }
            }
        }
        if (upper->cmask.isOff()) {
            const float first = upper->first();
            uint32_t n = 1;
            while (n < Upper::SIZE) {
                if (upper->tile[n++] != first) break;
            }
            if (n == Upper::SIZE) {
                kv.second.value = first;
                kv.second.state = upper->vmask.isOn();
                delete upper;
                kv.second.child = nullptr;
            }
        }
    }
}

// ------------------------------------------------------------------ improved Perlin noise + fBm
struct Perlin {
    int p[512];
    explicit Perlin(uint64_t seed)
    {
        int perm[256];
        for (int i = 0; i < 256; ++i) perm[i] = i;
        uint64_t s = seed;
        auto next = [&]() { // SplitMix64
            uint64_t z = (s += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            return z ^ (z >> 31);
        };
        for (int i = 255; i > 0; --i) { // Fisher-Yates
            const int j = int(next() % uint64_t(i + 1));
            std::swap(perm[i], perm[j]);
        }
        for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
    }
    static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static double lerp(double t, double a, double b) { return a + t * (b - a); }
    static double grad(int hash, double x, double y, double z)
    {
        const int h = hash & 15;
        const double u = h < 8 ? x : y, v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
        return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
    }
    double noise(double x, double y, double z) const
    {
        const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
        const int X = int(int64_t(fx) & 255), Y = int(int64_t(fy) & 255), Z = int(int64_t(fz) & 255);
        x -= fx;
        y -= fy;
        z -= fz;
        const double u = fade(x), v = fade(y), w = fade(z);
        const int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z, B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
        return lerp(w,
                    lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                         lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                    lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                         lerp(u, grad(p[AB + 1], x, y - 1, z - 1), grad(p[BB + 1], x - 1, y - 1, z - 1))));
    }
};

struct Fbm {
    Perlin perlin;
    uint32_t octaves;
    double freq;
    float amplitude, threshold;
    Fbm(uint64_t seed, uint32_t oct, double f, float amp, float thr) : perlin(seed), octaves(oct), freq(f), amplitude(amp), threshold(thr) {}
    float operator()(int i, int j, int k) const
    {
        double sum = 0.0, a = 1.0, f = freq;
        for (uint32_t o = 0; o < octaves; ++o) {
            sum += a * perlin.noise(double(i) * f, double(j) * f, double(k) * f);
            a *= 0.5;
            f *= 2.0;
        }
        float d = float(sum) * amplitude;
        d = d < 0.0f ? 0.0f : (d > 1.0f ? 1.0f : d);
        return d < threshold ? 0.0f : d;
    }
};

// ------------------------------------------------------------------ statistics (tools/GridStats.h:230-330)
struct Stats {
    float mn = std::numeric_limits<float>::max(), mx = std::numeric_limits<float>::lowest();
    size_t size = 0;
    double avg = 0.0, aux = 0.0;
    void ext(float v)
    {
        if (v < mn) mn = v;
        if (mx < v) mx = v;
    }
    void add(float val)
    {
        ext(val);
        size += 1;
        const double delta = double(val) - avg;
        avg += delta / double(size);
        aux += delta * (double(val) - avg);
    }
    void add(float val, uint64_t n)
    {
        const double denom = 1.0 / double(size + n);
        const double delta = double(val) - avg;
        avg += denom * delta * double(n);
        aux += denom * delta * delta * double(size) * double(n);
        ext(val);
        size += n;
    }
    void add(const Stats& o)
    {
        if (o.size > 0) {
            const double denom = 1.0 / double(size + o.size);
            const double delta = o.avg - avg;
            avg += denom * delta * double(o.size);
            aux += o.aux + denom * delta * delta * double(size) * double(o.size);
            if (o.mn < mn) mn = o.mn;
            if (mx < o.mx) mx = o.mx;
            size += o.size;
        }
    }
    double var() const { return size < 2 ? 0.0 : aux / double(size); }
    double stdev() const { return std::sqrt(var()); }
};
struct NodeStats {
    Stats stats;
    CoordBox bbox;
    void add(const NodeStats& o)
    {
        stats.add(o.stats);
        bbox.expand(o.bbox.lo, o.bbox.hi);
    }
};

// ------------------------------------------------------------------ serialisation
constexpr uint32_t kGT_FLOAT = 1, kGT_FP4 = 13, kGT_FP8 = 14, kGT_FP16 = 15, kGT_FPN = 16;

struct Writer {
    uint8_t* b;
    template<typename T> void put(uint64_t off, const T& v) { std::memcpy(b + off, &v, sizeof(T)); }
    template<typename T> T get(uint64_t off) const
    {
        T v;
        std::memcpy(&v, b + off, sizeof(T));
        return v;
    }
};

struct MapD {
    double mat[9], inv[9], vec[3];
    float matF[9], invF[9], vecF[3];
    explicit MapD(double dx)
    {
        // Map::set(dx, translation=0), NanoVDB.h:1574-1587 (matrices stored transposed)
        const double idx = 1.0 / dx;
        for (int i = 0; i < 9; ++i) {
            mat[i] = (i % 4 == 0) ? dx : 0.0;
            inv[i] = (i % 4 == 0) ? idx : 0.0;
            matF[i] = float(mat[i]);
            invF[i] = float(inv[i]);
        }
        for (int i = 0; i < 3; ++i) {
            vec[i] = 0.0;
            vecF[i] = 0.0f;
        }
    }
    void apply(const double x[3], double out[3]) const
    {
        for (int r = 0; r < 3; ++r) out[r] = std::fma(x[0], mat[3 * r], std::fma(x[1], mat[3 * r + 1], std::fma(x[2], mat[3 * r + 2], vec[r])));
    }
};

uint64_t rootKey(const Coord& c)
{
    return uint64_t(uint32_t(c[2]) >> 12) | (uint64_t(uint32_t(c[1]) >> 12) << 21) | (uint64_t(uint32_t(c[0]) >> 12) << 42);
}

struct LeafCodec {
    float min = 0.0f, max = 0.0f;
    uint8_t log2 = 0;
    uint64_t offset = 0; // byte offset of the leaf in the buffer
};

uint32_t leafBytes(uint32_t gt, uint8_t log2)
{
    switch (gt) {
    case kGT_FLOAT: return 2144u;
    case kGT_FP4: return 352u;
    case kGT_FP8: return 608u;
    case kGT_FP16: return 1120u;
    default: return 96u + (1u << log2) * 64u;
    }
}

// LeafFnBase setters (NanoVDB.h:3790-3810); float->integer conversions of non-finite values
// give 0 in the low bits on x86, which is what the reference writes for degenerate leaves.
uint16_t toU16(float x) { return std::isfinite(x) && x >= 0.0f && x < 65536.0f ? uint16_t(x) : uint16_t(0); }

int serialise(Tree& tree, const char* name, uint32_t gridClass, double voxelSize, uint32_t gt, float tol, void** out,
              uint64_t* bytesOut)
{
    // breadth-first node lists (NodeManager order)
    std::vector<Upper*> uppers;
    tree.forEachUpper([&](Upper& u) { uppers.push_back(&u); });
    std::vector<Lower*> lowers = tree.lowers();
    std::vector<LeafT*> leaves = tree.leaves();
    const uint32_t tableSize = uint32_t(tree.root.size());
    const uint64_t offTree = 672, offRoot = offTree + 64, offUpper = offRoot + 64 + 32ull * tableSize;
    const uint64_t offLower = offUpper + 270400ull * uppers.size(), offLeaf = offLower + 33856ull * lowers.size();

    // leaf codecs: FpN bit-width selection (tools/CreateNanoGrid.h:1009-1067)
    std::vector<LeafCodec> codec(leaves.size());
    parallelFor(leaves.size(), [&](size_t i) {
        const LeafT& L = *leaves[i];
        float mn = std::numeric_limits<float>::max(), mx = -mn;
        for (int j = 0; j < 512; ++j) {
            const float v = L.v[j];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        codec[i].min = mn;
        codec[i].max = mx;
        if (gt != kGT_FPN) return;
        const float range = mx - mn;
        uint8_t lb = 0;
        while (range > 0.0f && lb < 4u) {
            const uint32_t mask = (uint32_t(1) << (uint32_t(1) << lb)) - 1u;
            const float encode = float(mask) / range, decode = range / float(mask);
            int j = 0;
            do {
                const float exact = L.v[j];
                const uint32_t code = uint32_t(encode * (exact - mn) + 0.5f);
                const float approx = float(code) * decode + mn;
                j += std::fabs(exact - approx) <= tol ? 1 : 513;
            } while (j < 512);
            if (j == 512) break;
            ++lb;
        }
        codec[i].log2 = lb;
    });
    uint64_t cur = offLeaf;
    for (auto& c : codec) {
        c.offset = cur;
        cur += leafBytes(gt, c.log2);
    }
    // FpN: the reference's preProcess<FpN> returns the absolute end offset of the last leaf as the
    // leaf-section *size* and initHandle adds it to the leaf offset again (tools/CreateNanoGrid.h:1055-1067,
    // 937), so FpN buffers carry an unused tail as long as the leaf offset; mGridSize includes it.
    const uint64_t gridBytes = gt == kGT_FPN && !leaves.empty() ? cur + offLeaf : cur;
    const uint64_t alloc = (gridBytes + 31) & ~uint64_t(31);
    uint8_t* buf = static_cast<uint8_t*>(std::aligned_alloc(32, alloc));
    if (!buf) return NVQ_ERR_INVALID_ARG;
    std::memset(buf, 0, alloc);
    Writer W{buf};

    // child pointers -> buffer offsets
    std::map<const void*, uint64_t> where;
    for (size_t i = 0; i < uppers.size(); ++i) where[uppers[i]] = offUpper + 270400ull * i;
    for (size_t i = 0; i < lowers.size(); ++i) where[lowers[i]] = offLower + 33856ull * i;
    for (size_t i = 0; i < leaves.size(); ++i) where[leaves[i]] = codec[i].offset;

    // ---- leaves: values / codes, bbox, flags, stats (decoded active values)
    std::vector<NodeStats> leafStats(leaves.size());
    parallelFor(leaves.size(), [&](size_t i) {
        const LeafT& L = *leaves[i];
        const uint64_t o = codec[i].offset;
        W.put(o + 16, L.mask.w);
        const float mn = codec[i].min, mx = codec[i].max;
        uint8_t flags = 0;
        auto decodeAt = [&](uint32_t n) -> float {
            if (gt == kGT_FLOAT) return W.get<float>(o + 96 + 4 * n);
            const float q = W.get<float>(o + 84), m = W.get<float>(o + 80);
            uint32_t code;
            if (gt == kGT_FP4) code = (buf[o + 96 + (n >> 1)] >> ((n & 1) << 2)) & 15u;
            else if (gt == kGT_FP8) code = buf[o + 96 + n];
            else if (gt == kGT_FP16) code = W.get<uint16_t>(o + 96 + 2 * n);
            else {
                const uint32_t b = codec[i].log2;
                code = W.get<uint32_t>(o + 96 + 4 * (n >> (5 - b)));
                code >>= (n & ((32u >> b) - 1u)) << b;
                code &= (1u << (1u << b)) - 1u;
            }
            return float(code) * q + m;
        };
        if (gt == kGT_FLOAT) {
            std::memcpy(buf + o + 96, L.v, 2048);
        } else if (gt == kGT_FP4 || gt == kGT_FP8 || gt == kGT_FP16) {
            const int bw = gt == kGT_FP4 ? 4 : gt == kGT_FP8 ? 8 : 16;
            W.put(o + 80, mn);
            W.put(o + 84, (mx - mn) / float((1 << bw) - 1));
            if (bw == 16) {
                const double encode = 65535.0 / double(mx - mn);
                for (uint32_t j = 0; j < 512; ++j) {
                    const double x = encode * double(L.v[j] - mn) + double(0.5f);
                    W.put(o + 96 + 2 * j, uint16_t(std::isfinite(x) ? uint16_t(x) : 0));
                }
            } else {
                const float encode = float((1 << bw) - 1) / (mx - mn);
                auto q = [&](uint32_t j) -> uint8_t {
                    const float x = encode * (L.v[j] - mn) + 0.5f;
                    return std::isfinite(x) ? uint8_t(x) : uint8_t(0);
                };
                if (bw == 8) {
                    for (uint32_t j = 0; j < 512; ++j) buf[o + 96 + j] = q(j);
                } else {
                    for (uint32_t j = 0; j < 256; ++j) buf[o + 96 + j] = uint8_t((q(2 * j + 1) << 4) | q(2 * j));
                }
            }
        } else { // FpN, tools/CreateNanoGrid.h:1394-1474
            const uint8_t lb = codec[i].log2;
            flags = uint8_t(lb << 5);
            W.put(o + 80, mn);
            W.put(o + 84, (mx - mn) / float((1 << (1 << lb)) - 1));
            uint8_t* dst = buf + o + 96;
            if (lb == 4) {
                const double encode = 65535.0 / double(mx - mn);
                for (uint32_t j = 0; j < 512; ++j) {
                    const double x = encode * double(L.v[j] - mn) + double(0.5f);
                    W.put(o + 96 + 2 * j, uint16_t(std::isfinite(x) ? uint16_t(x) : 0));
                }
            } else {
                const float encode = float((1u << (1u << lb)) - 1u) / (mx - mn);
                auto q = [&](uint32_t j) -> uint8_t {
                    const float x = encode * (L.v[j] - mn) + 0.5f;
                    return std::isfinite(x) ? uint8_t(x) : uint8_t(0);
                };
                const uint32_t bits = 1u << lb, per = 8u / bits;
                for (uint32_t byte = 0; byte < 64u * bits; ++byte) {
                    uint8_t a = 0;
                    for (uint32_t k = 0; k < per; ++k) a = uint8_t(a | (q(byte * per + k) << (k * bits)));
                    dst[byte] = a;
                }
            }
        }
        // LeafNode::updateBBox (NanoVDB.h:4596-4624) + GridStats::process(Leaf)
        NodeStats ns;
        Coord bmin = L.origin;
        if (!L.mask.isOff()) {
            int lo[3] = {8, 8, 8}, hi[3] = {-1, -1, -1};
            for (uint32_t n = 0; n < 512; ++n)
                if (L.mask.on(n)) {
                    const int c[3] = {int(n >> 6), int((n >> 3) & 7), int(n & 7)};
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], c[a]);
                        hi[a] = std::max(hi[a], c[a]);
                    }
                }
            for (int a = 0; a < 3; ++a) {
                bmin[a] = L.origin[a] + lo[a];
                buf[o + 12 + a] = uint8_t(hi[a] - lo[a]);
            }
            flags = uint8_t(flags | 2u);
            ns.bbox.expand(bmin, Coord(bmin[0] + buf[o + 12], bmin[1] + buf[o + 13], bmin[2] + buf[o + 14]));
            W.put(o + 15, flags);
            for (uint32_t n = 0; n < 512; ++n)
                if (L.mask.on(n)) ns.stats.add(decodeAt(n));
            if (gt == kGT_FLOAT) {
                W.put(o + 80, ns.stats.mn);
                W.put(o + 84, ns.stats.mx);
                W.put(o + 88, float(ns.stats.avg));
                W.put(o + 92, float(ns.stats.stdev()));
            } else {
                const float q = W.get<float>(o + 84), m = W.get<float>(o + 80);
                W.put(o + 88, toU16((ns.stats.mn - m) / q + 0.5f));
                W.put(o + 90, toU16((ns.stats.mx - m) / q + 0.5f));
                W.put(o + 92, toU16((float(ns.stats.avg) - m) / q + 0.5f));
                W.put(o + 94, toU16(float(ns.stats.stdev()) / q + 0.5f));
            }
        } else {
            W.put(o + 15, flags);
        }
        W.put(o + 0, bmin);
        leafStats[i] = ns;
    });

    // ---- lower and upper nodes (GridStats::process(InternalNode): tiles first, then children)
    std::map<const void*, NodeStats> childStats;
    for (size_t i = 0; i < leaves.size(); ++i) childStats[leaves[i]] = leafStats[i];
    auto writeInternal = [&](auto& node, uint64_t o, uint32_t tableOff, uint32_t cmaskOff, uint32_t statsOff) {
        using NodeT = std::remove_reference_t<decltype(node)>;
        using ChildT = typename NodeT::Child;
        constexpr int32_t childDim = 1 << ChildT::TOTAL;
        constexpr uint64_t childVoxels = uint64_t(1) << (3 * ChildT::TOTAL); // ChildT::NUM_VALUES
        NodeStats total;
        for (uint32_t n = 0; n < NodeT::SIZE; ++n)
            if (node.vmask.on(n)) {
                total.stats.add(node.tile[n], childVoxels);
                const Coord c = node.childOrigin(n);
                total.bbox.expand(c, Coord(c[0] + childDim - 1, c[1] + childDim - 1, c[2] + childDim - 1));
            }
        for (uint32_t n = 0; n < NodeT::SIZE; ++n)
            if (node.cmask.on(n)) total.add(childStats.at(node.child[n]));
        W.put(o + 32, node.vmask.w);
        W.put(o + cmaskOff, node.cmask.w);
        for (uint32_t n = 0; n < NodeT::SIZE; ++n) {
            if (node.cmask.on(n)) W.put(o + tableOff + 8ull * n, int64_t(where.at(node.child[n])) - int64_t(o));
            else W.put(o + tableOff + 8ull * n, node.tile[n]);
        }
        uint64_t flags = 0;
        W.put(o, total.bbox.lo);
        W.put(o + 12, total.bbox.hi);
        if (total.bbox.empty()) {
            flags |= 1u; // no active values: render-skip on, has-active off
        } else {
            flags |= 2u;
            W.put(o + statsOff, total.stats.mn);
            W.put(o + statsOff + 4, total.stats.mx);
            W.put(o + statsOff + 8, float(total.stats.avg));
            W.put(o + statsOff + 12, float(total.stats.stdev()));
        }
        W.put(o + 24, flags);
        return total;
    };
    std::vector<NodeStats> lowerStats(lowers.size());
    parallelFor(lowers.size(), [&](size_t i) {
        // childStats is only read here
        lowerStats[i] = writeInternal(*lowers[i], where.at(lowers[i]), 1088u, 544u, 1056u);
    });
    for (size_t i = 0; i < lowers.size(); ++i) childStats[lowers[i]] = lowerStats[i];
    for (Upper* u : uppers) childStats[u] = writeInternal(*u, where.at(u), 8256u, 4128u, 8224u);

    // ---- root (tiles in key order) and its statistics
    NodeStats rootTotal;
    uint32_t ti = 0, activeRootTiles = 0;
    for (auto& kv : tree.root) {
        const uint64_t t = offRoot + 64 + 32ull * ti++;
        W.put(t, rootKey(kv.first));
        if (kv.second.child) {
            W.put(t + 8, int64_t(where.at(kv.second.child)) - int64_t(offRoot));
            W.put(t + 16, uint32_t(0));
            rootTotal.add(childStats.at(kv.second.child));
        } else {
            W.put(t + 8, int64_t(0));
            W.put(t + 16, uint32_t(kv.second.state ? 1 : 0));
            W.put(t + 20, kv.second.value);
            if (kv.second.state) {
                ++activeRootTiles;
                rootTotal.bbox.expand(kv.first, Coord(kv.first[0] + 4095, kv.first[1] + 4095, kv.first[2] + 4095));
                rootTotal.stats.add(kv.second.value, uint64_t(1) << 36);
            }
        }
    }
    W.put(offRoot, rootTotal.bbox.lo);
    W.put(offRoot + 12, rootTotal.bbox.hi);
    W.put(offRoot + 24, tableSize);
    W.put(offRoot + 28, tree.background);
    if (tableSize == 0) {
        W.put(offRoot + 32, tree.background);
        W.put(offRoot + 36, tree.background);
    } else {
        W.put(offRoot + 32, rootTotal.stats.mn);
        W.put(offRoot + 36, rootTotal.stats.mx);
        W.put(offRoot + 40, float(rootTotal.stats.avg));
        W.put(offRoot + 44, float(rootTotal.stats.stdev()));
    }

    // ---- tree (CreateNanoGrid::processTree, tools/CreateNanoGrid.h:1651-1700)
    W.put(offTree + 0, int64_t(leaves.empty() ? 0 : offLeaf - offTree));
    W.put(offTree + 8, int64_t(lowers.empty() ? 0 : offLower - offTree));
    W.put(offTree + 16, int64_t(uppers.empty() ? 0 : offUpper - offTree));
    W.put(offTree + 24, int64_t(offRoot - offTree));
    W.put(offTree + 32, uint32_t(leaves.size()));
    W.put(offTree + 36, uint32_t(lowers.size()));
    W.put(offTree + 40, uint32_t(uppers.size()));
    uint32_t tiles0 = 0, tiles1 = 0;
    uint64_t voxels = 0;
    for (Lower* l : lowers) tiles0 += l->vmask.count();
    for (Upper* u : uppers) tiles1 += u->vmask.count();
    for (LeafT* l : leaves) voxels += l->mask.count();
    W.put(offTree + 44, tiles0);
    W.put(offTree + 48, tiles1);
    W.put(offTree + 52, activeRootTiles);
    voxels += uint64_t(tiles0) << 9;
    voxels += uint64_t(tiles1) << 21;
    voxels += uint64_t(activeRootTiles) << 36;
    W.put(offTree + 56, voxels);

    // ---- grid (GridData::init + GridStats::process(Grid))
    const MapD map(voxelSize);
    W.put(0, uint64_t(0x314244566f6e614eull)); // "NanoVDB1"
    W.put(8, ~uint64_t(0));                    // checksum disabled
    W.put(16, uint32_t((32u << 21) | (9u << 10) | 2u));
    uint32_t gflags = 1u << 5; // IsBreadthFirst
    gflags |= (1u << 2) | (1u << 3) | (1u << 4); // min/max, average, std deviation
    W.put(24, uint32_t(0));
    W.put(28, uint32_t(1));
    W.put(32, gridBytes);
    std::strncpy(reinterpret_cast<char*>(buf + 40), name, 255);
    W.put(296, map.matF);
    W.put(296 + 36, map.invF);
    W.put(296 + 72, map.vecF);
    W.put(296 + 84, 1.0f);
    W.put(296 + 88, map.mat);
    W.put(296 + 160, map.inv);
    W.put(296 + 232, map.vec);
    W.put(296 + 256, 1.0);
    double wb[6] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    -std::numeric_limits<double>::max(), -std::numeric_limits<double>::max(), -std::numeric_limits<double>::max()};
    if (!rootTotal.bbox.empty()) {
        gflags |= 1u << 1; // HasBBox
        for (int c = 0; c < 8; ++c) {
            const double x[3] = {double(c & 1 ? rootTotal.bbox.hi[0] + 1 : rootTotal.bbox.lo[0]),
                                 double(c & 2 ? rootTotal.bbox.hi[1] + 1 : rootTotal.bbox.lo[1]),
                                 double(c & 4 ? rootTotal.bbox.hi[2] + 1 : rootTotal.bbox.lo[2])};
            double y[3];
            map.apply(x, y);
            for (int a = 0; a < 3; ++a) {
                wb[a] = std::min(wb[a], y[a]);
                wb[3 + a] = std::max(wb[3 + a], y[a]);
            }
        }
    }
    W.put(560, wb);
    {
        const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
        double a[3], z[3];
        map.apply(one, a);
        map.apply(zero, z);
        const double vs[3] = {a[0] - z[0], a[1] - z[1], a[2] - z[2]};
        W.put(608, vs);
    }
    W.put(20, gflags);
    W.put(632, gridClass);
    W.put(636, gt);
    W.put(640, int64_t(gridBytes)); // blind metadata offset = grid size (none)
    W.put(648, uint32_t(0));
    *out = buf;
    *bytesOut = gridBytes;
    return NVQ_OK;
}

bool supported(uint32_t gt) { return gt == kGT_FLOAT || gt == kGT_FP4 || gt == kGT_FP8 || gt == kGT_FP16 || gt == kGT_FPN; }

const char* defaultName(const char* base, uint32_t gt)
{
    // the reference's default grid names (tools/CreatePrimitives.h:53-200)
    if (gt == kGT_FPN) {
        if (!std::strcmp(base, "sphere_ls")) return "sphere_ls_FpN";
    }
    return base;
}

// Read an FP32 NanoVDB buffer back into a build tree (for re-encoding).
int loadFloatTree(const uint8_t* b, uint64_t bytes, Tree*& out, uint32_t& gridClass, double& voxelSize, std::string& name)
{
    if (!b || bytes < 800) return NVQ_ERR_INVALID_ARG;
    uint32_t type;
    std::memcpy(&type, b + 636, 4);
    if (type != kGT_FLOAT) return NVQ_ERR_UNSUPPORTED_TYPE;
    std::memcpy(&gridClass, b + 632, 4);
    std::memcpy(&voxelSize, b + 608, 8);
    name.assign(reinterpret_cast<const char*>(b + 40), strnlen(reinterpret_cast<const char*>(b + 40), 256));
    const uint8_t* tree = b + 672;
    int64_t offRoot;
    std::memcpy(&offRoot, tree + 24, 8);
    const uint8_t* root = tree + offRoot;
    float bg;
    std::memcpy(&bg, root + 28, 4);
    uint32_t ts;
    std::memcpy(&ts, root + 24, 4);
    std::unique_ptr<Tree> t(new Tree(bg));
    auto rd64 = [](const uint8_t* p) { uint64_t v; std::memcpy(&v, p, 8); return v; };
    auto rdf = [](const uint8_t* p) { float v; std::memcpy(&v, p, 4); return v; };
    for (uint32_t i = 0; i < ts; ++i) {
        const uint8_t* tile = root + 64 + 32ull * i;
        const uint64_t key = rd64(tile);
        const Coord c(int32_t(uint32_t((key >> 42) & 0x1fffff) << 12), int32_t(uint32_t((key >> 21) & 0x1fffff) << 12),
                      int32_t(uint32_t(key & 0x1fffff) << 12));
        int64_t child;
        std::memcpy(&child, tile + 8, 8);
        RootTile rt;
        if (child == 0) {
            uint32_t st;
            std::memcpy(&st, tile + 16, 4);
            rt.state = st != 0;
            rt.value = rdf(tile + 20);
            t->root[c] = rt;
            continue;
        }
        const uint8_t* up = root + child;
        Upper* U = new Upper(c, bg, false);
        std::memcpy(U->vmask.w, up + 32, 4096);
        std::memcpy(U->cmask.w, up + 4128, 4096);
        for (uint32_t n = 0; n < Upper::SIZE; ++n) {
            if (!U->cmask.on(n)) {
                U->tile[n] = rdf(up + 8256 + 8ull * n);
                continue;
            }
            int64_t lo;
            std::memcpy(&lo, up + 8256 + 8ull * n, 8);
            const uint8_t* L = up + lo;
            Lower* Lw = new Lower(U->childOrigin(n), bg, false);
            std::memcpy(Lw->vmask.w, L + 32, 512);
            std::memcpy(Lw->cmask.w, L + 544, 512);
            for (uint32_t m = 0; m < Lower::SIZE; ++m) {
                if (!Lw->cmask.on(m)) {
                    Lw->tile[m] = rdf(L + 1088 + 8ull * m);
                    continue;
                }
                int64_t lf;
                std::memcpy(&lf, L + 1088 + 8ull * m, 8);
                const uint8_t* F = L + lf;
                LeafT* leaf = new LeafT(Lw->childOrigin(m), bg, false);
                std::memcpy(leaf->mask.w, F + 16, 64);
                std::memcpy(leaf->v, F + 96, 2048);
                Lw->child[m] = leaf;
            }
            U->child[n] = Lw;
        }
        rt.child = U;
        t->root[c] = rt;
    }
    out = t.release();
    return NVQ_OK;
}

} // namespace

// ======================================================================= C ABI
extern "C" {

void nvqb_free(void* p) { std::free(p); }

int nvqb_level_set_sphere(double radius, const double center[3], double voxel_size, double half_width,
                          uint32_t grid_type, float tol, void** out, uint64_t* bytes)
{
    if (!out || !bytes || !center || !(radius > 0) || !(voxel_size > 0) || !(half_width > 0) || !supported(grid_type))
        return !supported(grid_type) ? NVQ_ERR_UNSUPPORTED_TYPE : NVQ_ERR_INVALID_ARG;
    Tree t(float(half_width * voxel_size));
    initSphere(t, radius, center, voxel_size, half_width);
    sdfToLevelSet(t);
    return serialise(t, defaultName("sphere_ls", grid_type), 1u, voxel_size, grid_type, tol, out, bytes);
}

int nvqb_level_set_torus(double major_radius, double minor_radius, const double center[3], double voxel_size,
                         double half_width, uint32_t grid_type, float tol, void** out, uint64_t* bytes)
{
    if (!out || !bytes || !center || !(minor_radius > 0) || !(major_radius > minor_radius) || !(voxel_size > 0) ||
        !(half_width > 0) || !supported(grid_type))
        return !supported(grid_type) ? NVQ_ERR_UNSUPPORTED_TYPE : NVQ_ERR_INVALID_ARG;
    Tree t(float(half_width * voxel_size));
    initTorus(t, major_radius, minor_radius, center, voxel_size, half_width);
    sdfToLevelSet(t);
    return serialise(t, "torus_ls", 1u, voxel_size, grid_type, tol, out, bytes);
}

int nvqb_fog_volume_sphere(double radius, const double center[3], double voxel_size, double half_width, void** out,
                           uint64_t* bytes)
{
    if (!out || !bytes || !center || !(radius > 0) || !(voxel_size > 0) || !(half_width > 0)) return NVQ_ERR_INVALID_ARG;
    Tree t(float(half_width * voxel_size));
    initSphere(t, radius, center, voxel_size, half_width);
    sdfToLevelSet(t);
    levelSetToFog(t);
    return serialise(t, "sphere_fog", 2u, voxel_size, kGT_FLOAT, 0.0f, out, bytes);
}

int nvqb_fog_fbm(const int32_t bbox[6], uint32_t octaves, double base_freq, uint64_t seed, float amplitude,
                 float threshold, uint32_t grid_type, float tol, void** out, uint64_t* bytes)
{
    if (!out || !bytes || !bbox || octaves == 0 || octaves > 16 || !supported(grid_type)) return NVQ_ERR_INVALID_ARG;
    for (int a = 0; a < 3; ++a)
        if (bbox[a] > bbox[3 + a]) return NVQ_ERR_INVALID_ARG;
    const Fbm f(seed, octaves, base_freq, amplitude, threshold);
    Tree t(0.0f);
    fillFromFunction(t, bbox, [&](int i, int j, int k) { return f(i, j, k); });
    return serialise(t, "fog_fbm", 2u, 1.0, grid_type, tol, out, bytes);
}

int nvqb_fbm_density(const int32_t* ijk, uint64_t n, uint32_t octaves, double base_freq, uint64_t seed,
                     float amplitude, float threshold, float* out)
{
    if ((n && (!ijk || !out)) || octaves == 0 || octaves > 16) return NVQ_ERR_INVALID_ARG;
    const Fbm f(seed, octaves, base_freq, amplitude, threshold);
    for (uint64_t i = 0; i < n; ++i) out[i] = f(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
    return NVQ_OK;
}

int nvqb_quantize(const void* float_buf, uint64_t bytes_in, uint32_t grid_type, float tol, void** out, uint64_t* bytes)
{
    if (!out || !bytes || !supported(grid_type)) return NVQ_ERR_INVALID_ARG;
    Tree* t = nullptr;
    uint32_t cls = 0;
    double vs = 1.0;
    std::string name;
    const int st = loadFloatTree(static_cast<const uint8_t*>(float_buf), bytes_in, t, cls, vs, name);
    if (st != NVQ_OK) return st;
    std::unique_ptr<Tree> own(t);
    return serialise(*t, name.c_str(), cls, vs, grid_type, tol, out, bytes);
}

int nvqb_camera_rays(uint32_t width, uint32_t height, const float eye[3], const float look_at[3], const float up[3],
                     float fov_deg, float* out_eye, float* out_dir)
{
    if (!width || !height || !eye || !look_at || !up || !out_eye || !out_dir) return NVQ_ERR_INVALID_ARG;
    // camera frame: forward f = normalize(look_at - eye), right r = normalize(f x up), true up u = r x f
    float f[3] = {look_at[0] - eye[0], look_at[1] - eye[1], look_at[2] - eye[2]};
    const float fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (!(fl > 0.0f)) return NVQ_ERR_INVALID_ARG;
    for (float& x : f) x /= fl;
    float r[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    const float rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(rl > 0.0f)) return NVQ_ERR_INVALID_ARG;
    for (float& x : r) x /= rl;
    const float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    const float tanHalf = std::tan(fov_deg / 2 * 3.14159265358979323846f / 180.f);
    const float aspect = float(width) / float(height);
    for (uint64_t i = 0; i < uint64_t(width) * height; ++i) {
        const uint32_t x = uint32_t(i % width), y = uint32_t(i / width);
        const float uu = (float(x) + 0.5f) / float(width);
        const float vv = (float(y) + 0.5f) / float(height);
        const float Px = (2.f * uu - 1.f) * tanHalf * aspect;
        const float Py = (2.f * vv - 1.f) * tanHalf;
        float d[3];
        for (int a = 0; a < 3; ++a) d[a] = Px * r[a] + Py * u[a] + f[a];
        const float len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const float inv = 1.0f / len;
        for (int a = 0; a < 3; ++a) {
            out_eye[3 * i + a] = eye[a];
            out_dir[3 * i + a] = d[a] * inv;
        }
    }
    return NVQ_OK;
}

int nvqb_iso_camera(const void* grid_buf, uint64_t bytes, int32_t width, int32_t height, nvq_iso_camera* out)
{
    if (!grid_buf || bytes < 672 || width <= 0 || height <= 0 || !out) return NVQ_ERR_INVALID_ARG;
    const uint8_t* b = static_cast<const uint8_t*>(grid_buf);
    if (std::memcmp(b, "NanoVDB", 7) != 0) return NVQ_ERR_BAD_MAGIC;
    double lo[3], hi[3], vs[3]; // GridData::mWorldBBox at 560, mVoxelSize at 608 (NanoVDB.h:1956-1957)
    std::memcpy(lo, b + 560, sizeof(lo));
    std::memcpy(hi, b + 584, sizeof(hi));
    std::memcpy(vs, b + 608, sizeof(vs));
    // BBox<Vec3d>::dim(): max - min when min < max on every axis, else 0 (math/Math.h:1907-1910)
    const bool valid = lo[0] < hi[0] && lo[1] < hi[1] && lo[2] < hi[2];
    double dim[3];
    for (int a = 0; a < 3; ++a) dim[a] = valid ? hi[a] - lo[a] : 0.0;
    out->width = width;
    out->height = height;
    out->voxel_size = float(vs[0]);
    out->iso = out->voxel_size;
    out->wbbox_dim_z = float(dim[2]) * 2;
    const float off[3] = {0.0f, 0.0f, out->wbbox_dim_z};
    for (int a = 0; a < 3; ++a) out->origin[a] = float(lo[a] + 0.5 * dim[a]) + off[a];
    const float fov = 45.f;
    out->tan_half_fov = std::tan(fov / 2 * 3.14159265358979323846f / 180.f);
    return NVQ_OK;
}

} // extern "C"
