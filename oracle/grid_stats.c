/* Oracle (test infrastructure only, never part of the product): CPU restatement of
 * tools/cuda/GridStats.cuh updateGridStats<float>(grid, StatsMode::Default) in place on a
 * host copy of a float grid buffer, with the device code's merge order:
 *  - processLeaf (:86-153): 32 lanes, lane L folds voxels L, L+32, ... (active only) with
 *    Stats::add(value), then lanes merge pairwise (lane l += lane l+d for d = 16, 8, 4, 2, 1);
 *    LeafNode::updateBBox (NanoVDB.h:4596-4624); render flag bit 0 cleared;
 *  - processInternal (:155-241): 128 threads, thread t folds entries t, t+128, ... (a child's
 *    bbox and accumulator, or an active tile's extent and Stats::add(value, childVoxels)), then a
 *    pairwise tree over the 128 partials (d = 64 ... 1);
 *  - processRootAndGrid (:243-306): root tiles in table order, world bbox of (min, max + 1)
 *    through the grid's map, grid flags HasBBox | HasMinMax | HasAverage | HasStdDeviation.
 * Stats / Extrema arithmetic: tools/GridStats.h:72-140, :244-344 (FP64 mean and M2, FP32
 * extrema; avg and std stored as float).  Leaves without active voxels keep their statistics
 * and contribute nothing (their accumulator is empty).
 * The other modes (updateGridStats :344-362): StatsMode::MinMax (Extrema) writes only the
 * extrema, and a parent merges the minimum and maximum its children store (Extrema::add of
 * child.minimum() and child.maximum(), :193-196, :265-268); StatsMode::BBox (NoopStats) writes
 * no statistics.  The grid's HasMinMax / HasAverage / HasStdDeviation flags follow the mode
 * (:299-301); an empty root gets the background as minimum and maximum in every mode. */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float mn, mx;
    uint64_t size;
    double avg, aux;
} St;

static St st_empty(void)
{
    St s = {FLT_MAX, -FLT_MAX, 0, 0.0, 0.0};
    return s;
}
static void ext_add(St* s, float v)
{
    if (v < s->mn) s->mn = v;
    if (v > s->mx) s->mx = v;
}
/* Extrema::add(const Extrema&) (tools/GridStats.h:117-121): min of the minima, max of the maxima
 * (an empty side, FLT_MAX / -FLT_MAX, is neutral) */
static void ext_merge(St* s, const St* o)
{
    if (o->mn < s->mn) s->mn = o->mn;
    if (o->mx > s->mx) s->mx = o->mx;
}
static void st_add(St* s, float v)
{
    ext_add(s, v);
    s->size += 1;
    const double delta = (double)v - s->avg;
    s->avg += delta / (double)s->size;
    s->aux += delta * ((double)v - s->avg);
}
static void st_add_n(St* s, float v, uint64_t n)
{
    const double denom = 1.0 / (double)(s->size + n);
    const double delta = (double)v - s->avg;
    s->avg += denom * delta * (double)n;
    s->aux += denom * delta * delta * (double)s->size * (double)n;
    ext_add(s, v);
    s->size += n;
}
static void st_merge(St* s, const St* o)
{
    if (o->size > 0) {
        const double denom = 1.0 / (double)(s->size + o->size);
        const double delta = o->avg - s->avg;
        s->avg += denom * delta * (double)o->size;
        s->aux += o->aux + denom * delta * delta * (double)s->size * (double)o->size;
        if (o->mn < s->mn) s->mn = o->mn;
        if (o->mx > s->mx) s->mx = o->mx;
        s->size += o->size;
    }
}
static float st_std(const St* s) { return (float)sqrt(s->size < 2 ? 0.0 : s->aux / (double)s->size); }

static uint64_t rd64(const uint8_t* p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
static int32_t rdi(const uint8_t* p)
{
    int32_t v;
    memcpy(&v, p, 4);
    return v;
}
static float rdf(const uint8_t* p)
{
    float v;
    memcpy(&v, p, 4);
    return v;
}
static void wri(uint8_t* p, int32_t v) { memcpy(p, &v, 4); }
static void wrf(uint8_t* p, float v) { memcpy(p, &v, 4); }
static void wr64(uint8_t* p, uint64_t v) { memcpy(p, &v, 8); }
static void wrd(uint8_t* p, double v) { memcpy(p, &v, 8); }

static void box_empty(int32_t b[6])
{
    b[0] = b[1] = b[2] = INT_MAX;
    b[3] = b[4] = b[5] = INT_MIN;
}
static void box_expand(int32_t b[6], const int32_t o[6])
{
    for (int a = 0; a < 3; ++a) {
        if (o[a] < b[a]) b[a] = o[a];
        if (o[3 + a] > b[3 + a]) b[3 + a] = o[3 + a];
    }
}

/* node sizes and per-level layout of a float grid */
enum { LEAF = 2144, LOWER = 33856, UPPER = 270400 };

/* LeafNode::bbox (NanoVDB.h:4442-4453): empty unless the HasBBox flag bit is on */
static void leaf_box(const uint8_t* L, int32_t b[6])
{
    if (!(L[15] & 2u)) {
        box_empty(b);
        return;
    }
    for (int a = 0; a < 3; ++a) {
        b[a] = rdi(L + 4 * a);
        b[3 + a] = b[a] + (int32_t)L[12 + a];
    }
}

int or_update_grid_stats_mode(void* buf_, int mode)
{
    uint8_t* b = (uint8_t*)buf_;
    if (mode < 1 || mode > 3) return -1; /* Disable with a grid throws in the reference (:357-358) */
    uint32_t type;
    memcpy(&type, b + 636, 4);
    if (type != 1u) return -4;
    const uint8_t* T = b + 672;
    const int64_t offLeaf = (int64_t)rd64(T), offLower = (int64_t)rd64(T + 8), offUpper = (int64_t)rd64(T + 16),
                  offRoot = (int64_t)rd64(T + 24);
    uint32_t cnt[3];
    memcpy(cnt, T + 32, 12);
    uint8_t* leaf0 = b + 672 + offLeaf;
    uint8_t* lower0 = b + 672 + offLower;
    uint8_t* upper0 = b + 672 + offUpper;
    uint8_t* root = b + 672 + offRoot;

    St* leafAcc = (St*)malloc(sizeof(St) * (cnt[0] ? cnt[0] : 1));
    St* lowerAcc = (St*)malloc(sizeof(St) * (cnt[1] ? cnt[1] : 1));
    St* upperAcc = (St*)malloc(sizeof(St) * (cnt[2] ? cnt[2] : 1));
    if (!leafAcc || !lowerAcc || !upperAcc) {
        free(leafAcc);
        free(lowerAcc);
        free(upperAcc);
        return -2;
    }

    /* leaves */
    for (uint32_t l = 0; l < cnt[0]; ++l) {
        uint8_t* L = leaf0 + (uint64_t)LEAF * l;
        uint64_t w[8];
        for (int j = 0; j < 8; ++j) w[j] = rd64(L + 16 + 8 * j);
        uint64_t word = 0;
        for (int j = 0; j < 8; ++j) word |= w[j];
        leafAcc[l] = st_empty();
        if (word) {
            uint32_t xmin = w[0] ? 0u : 8u, xmax = xmin;
            for (uint32_t j = 1; j < 8; ++j)
                if (w[j]) {
                    if (xmin == 8) xmin = j;
                    xmax = j;
                }
            const uint32_t ymin = (uint32_t)__builtin_ctzll(word) >> 3, ymax = (63u - (uint32_t)__builtin_clzll(word)) >> 3;
            const uint32_t w32 = (uint32_t)word | (uint32_t)(word >> 32), w16 = (w32 & 0xffffu) | (w32 >> 16);
            const uint32_t byte = (w16 & 0xffu) | (w16 >> 8);
            const uint32_t mn[3] = {xmin, ymin, (uint32_t)__builtin_ctz(byte)};
            const uint32_t mx[3] = {xmax, ymax, 31u - (uint32_t)__builtin_clz(byte)};
            for (int a = 0; a < 3; ++a) {
                wri(L + 4 * a, (rdi(L + 4 * a) & ~7) + (int32_t)mn[a]);
                L[12 + a] = (uint8_t)(mx[a] - mn[a]);
            }
            L[15] |= 2u;
            if (mode == 2) { /* Extrema over the active values (order-free) */
                St e = st_empty();
                for (uint32_t i = 0; i < 512; ++i)
                    if ((w[i >> 6] >> (i & 63)) & 1u) ext_add(&e, rdf(L + 96 + 4 * i));
                wrf(L + 80, e.mn);
                wrf(L + 84, e.mx);
            }
            St part[32];
            for (int lane = 0; lane < 32 && mode == 3; ++lane) {
                part[lane] = st_empty();
                for (uint32_t i = (uint32_t)lane; i < 512; i += 32)
                    if ((w[i >> 6] >> (i & 63)) & 1u) st_add(&part[lane], rdf(L + 96 + 4 * i));
            }
            if (mode == 3) {
                for (int d = 16; d; d >>= 1)
                    for (int lane = 0; lane < d; ++lane) st_merge(&part[lane], &part[lane + d]);
                leafAcc[l] = part[0];
                wrf(L + 80, part[0].mn);
                wrf(L + 84, part[0].mx);
                wrf(L + 88, (float)part[0].avg);
                wrf(L + 92, st_std(&part[0]));
            }
        } else {
            L[15] &= (uint8_t)~2u;
        }
        L[15] &= (uint8_t)~1u;
    }

    /* internal nodes: level 1 (children are leaves), then level 2 (children are lower nodes) */
    for (int level = 1; level <= 2; ++level) {
        const uint32_t count = cnt[level], size = level == 1 ? 4096u : 32768u;
        const uint64_t nodeBytes = level == 1 ? LOWER : UPPER, childBytes = level == 1 ? LEAF : LOWER;
        const uint32_t vmOff = 32, cmOff = level == 1 ? 544u : 4128u, tabOff = level == 1 ? 1088u : 8256u;
        const uint32_t stOff = level == 1 ? 1056u : 8224u, log2c = level == 1 ? 3u : 7u, log2n = level == 1 ? 4u : 5u;
        const uint64_t childVoxels = (uint64_t)1 << (3 * log2c);
        uint8_t* node0 = level == 1 ? lower0 : upper0;
        uint8_t* child0 = level == 1 ? leaf0 : lower0;
        St* childAcc = level == 1 ? leafAcc : lowerAcc;
        St* acc = level == 1 ? lowerAcc : upperAcc;
        for (uint32_t k = 0; k < count; ++k) {
            uint8_t* N = node0 + nodeBytes * k;
            St part[128];
            int32_t pbox[128][6];
            const int32_t ox = rdi(N) & ~((1 << (log2c + log2n)) - 1), oy = rdi(N + 4) & ~((1 << (log2c + log2n)) - 1),
                          oz = rdi(N + 8) & ~((1 << (log2c + log2n)) - 1);
            for (int t = 0; t < 128; ++t) {
                part[t] = st_empty();
                box_empty(pbox[t]);
                for (uint32_t i = (uint32_t)t; i < size; i += 128) {
                    const int childOn = (rd64(N + cmOff + 8 * (i >> 6)) >> (i & 63)) & 1u;
                    if (childOn) {
                        const int64_t off = (int64_t)rd64(N + tabOff + 8ull * i);
                        const uint8_t* C = N + off;
                        const uint64_t ci = (uint64_t)(C - child0) / childBytes;
                        int32_t cb[6];
                        if (level == 1) leaf_box(C, cb);
                        else
                            for (int a = 0; a < 6; ++a) cb[a] = rdi(C + 4 * a);
                        box_expand(pbox[t], cb);
                        if (mode == 3) {
                            st_merge(&part[t], &childAcc[ci]);
                        } else if (mode == 2) { /* the child's stored extrema */
                            const uint32_t cst = level == 1 ? 80u : 1056u;
                            ext_add(&part[t], rdf(C + cst));
                            ext_add(&part[t], rdf(C + cst + 4));
                        }
                    } else if ((rd64(N + vmOff + 8 * (i >> 6)) >> (i & 63)) & 1u) {
                        const uint32_t m = (1u << log2n) - 1u;
                        const int32_t c[3] = {ox + (int32_t)(((i >> (2 * log2n)) & m) << log2c),
                                              oy + (int32_t)(((i >> log2n) & m) << log2c), oz + (int32_t)((i & m) << log2c)};
                        const int32_t tb[6] = {c[0], c[1], c[2], c[0] + (1 << log2c) - 1, c[1] + (1 << log2c) - 1,
                                               c[2] + (1 << log2c) - 1};
                        box_expand(pbox[t], tb);
                        if (mode == 3) st_add_n(&part[t], rdf(N + tabOff + 8ull * i), childVoxels);
                        else if (mode == 2) ext_add(&part[t], rdf(N + tabOff + 8ull * i));
                    }
                }
            }
            for (int d = 64; d; d >>= 1)
                for (int t = 0; t < d; ++t) {
                    if (mode == 2) ext_merge(&part[t], &part[t + d]);
                    else st_merge(&part[t], &part[t + d]);
                    box_expand(pbox[t], pbox[t + d]);
                }
            for (int a = 0; a < 6; ++a) wri(N + 4 * a, pbox[0][a]);
            if (mode >= 2) {
                wrf(N + stOff, part[0].mn);
                wrf(N + stOff + 4, part[0].mx);
            }
            if (mode == 3) {
                wrf(N + stOff + 8, (float)part[0].avg);
                wrf(N + stOff + 12, st_std(&part[0]));
            }
            wr64(N + 24, rd64(N + 24) & ~(uint64_t)1);
            acc[k] = part[0];
        }
    }

    /* root and grid */
    uint32_t tableSize;
    memcpy(&tableSize, root + 24, 4);
    int32_t rbox[6];
    box_empty(rbox);
    if (tableSize == 0) {
        const float bg = rdf(root + 28);
        wrf(root + 32, bg);
        wrf(root + 36, bg);
        wrf(root + 40, 0.0f);
        wrf(root + 44, 0.0f);
    } else {
        St s = st_empty();
        for (uint32_t i = 0; i < tableSize; ++i) {
            const uint8_t* tile = root + 64 + 32ull * i;
            const int64_t child = (int64_t)rd64(tile + 8);
            if (child) {
                const uint8_t* C = root + child;
                const uint64_t ci = (uint64_t)(C - upper0) / UPPER;
                int32_t cb[6];
                for (int a = 0; a < 6; ++a) cb[a] = rdi(C + 4 * a);
                box_expand(rbox, cb);
                if (mode == 3) {
                    st_merge(&s, &upperAcc[ci]);
                } else if (mode == 2) {
                    ext_add(&s, rdf(C + 8224));
                    ext_add(&s, rdf(C + 8228));
                }
            } else {
                uint32_t state;
                memcpy(&state, tile + 16, 4);
                if (state) {
                    /* RootData::KeyToCoord of the tile key */
                    const uint64_t key = rd64(tile);
                    const int32_t c[3] = {(int32_t)((uint32_t)((key >> 42) & 0x1fffffu) << 12),
                                          (int32_t)((uint32_t)((key >> 21) & 0x1fffffu) << 12),
                                          (int32_t)((uint32_t)(key & 0x1fffffu) << 12)};
                    const int32_t tb[6] = {c[0], c[1], c[2], c[0] + 4095, c[1] + 4095, c[2] + 4095};
                    box_expand(rbox, tb);
                    if (mode == 3) st_add_n(&s, rdf(tile + 20), (uint64_t)1 << 36);
                    else if (mode == 2) ext_add(&s, rdf(tile + 20));
                }
            }
        }
        if (mode >= 2) {
            wrf(root + 32, s.mn);
            wrf(root + 36, s.mx);
        }
        if (mode == 3) {
            wrf(root + 40, (float)s.avg);
            wrf(root + 44, st_std(&s));
        }
    }
    for (int a = 0; a < 6; ++a) wri(root + 4 * a, rbox[a]);

    uint32_t flags;
    memcpy(&flags, b + 20, 4);
    if (rbox[0] > rbox[3] || rbox[1] > rbox[4] || rbox[2] > rbox[5]) {
        for (int a = 0; a < 3; ++a) {
            wrd(b + 560 + 8 * a, DBL_MAX);
            wrd(b + 584 + 8 * a, -DBL_MAX);
        }
        flags &= ~2u;
    } else {
        double mat[9], vec[3], wb[6];
        memcpy(mat, b + 296 + 88, 72);
        memcpy(vec, b + 296 + 232, 24);
        for (int c = 0; c < 8; ++c) {
            const double x[3] = {(double)(c & 1 ? rbox[3] + 1 : rbox[0]), (double)(c & 2 ? rbox[4] + 1 : rbox[1]),
                                 (double)(c & 4 ? rbox[5] + 1 : rbox[2])};
            for (int r = 0; r < 3; ++r) {
                const double y = fma(x[0], mat[3 * r], fma(x[1], mat[3 * r + 1], fma(x[2], mat[3 * r + 2], vec[r])));
                if (c == 0 || y < wb[r]) wb[r] = y;
                if (c == 0 || y > wb[3 + r]) wb[3 + r] = y;
            }
        }
        for (int a = 0; a < 6; ++a) wrd(b + 560 + 8 * a, wb[a]);
        flags |= 2u;
    }
    flags &= ~(4u | 8u | 16u); /* HasMinMax, HasAverage, HasStdDeviation as the mode has them */
    if (mode >= 2) flags |= 4u;
    if (mode == 3) flags |= 8u | 16u;
    memcpy(b + 20, &flags, 4);
    free(leafAcc);
    free(lowerAcc);
    free(upperAcc);
    return 0;
}

int or_update_grid_stats(void* buf) { return or_update_grid_stats_mode(buf, 3); }
