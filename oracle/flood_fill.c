/* Oracle (test infrastructure only, never part of the product): CPU restatement of
 * tools/cuda/SignedFloodFill.cuh signedFloodFill<float>(grid), in place on a host copy of a
 * float grid buffer:
 *  - processLeaf (:172-185): every inactive voxel becomes +-background with the sign of the
 *    next active voxel (Mask::findNext<true>), else the previous one (findPrev<true>), else
 *    voxel 0;
 *  - processNode<1>, then <2> (:149-170): every table entry that is not a child becomes
 *    +-background with the sign of the next child's first value, else the previous child's last
 *    value, else entry 0 (active tiles included, as the reference writes them);
 *  - processRoot (:91-145): root children in lexicographic order; inactive tiles strictly
 *    between two children on one z-scanline whose facing values are both <= 0 become
 *    inactive tiles of value -background.  A missing tile there is an error (-3).
 * Mask::findNext / findPrev: NanoVDB.h:1385-1409; getFirstValue / getLastValue:
 * NanoVDB.h:3529-3541, :4476-4478.  Concurrent updates in the device code only ever rewrite
 * a value with one of the same sign, so this sequential order gives the same bytes. */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

static uint64_t r64(const uint8_t* p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
static float rf(const uint8_t* p)
{
    float v;
    memcpy(&v, p, 4);
    return v;
}
static void wf(uint8_t* p, float v) { memcpy(p, &v, 4); }

/* first on bit >= start, or size */
static uint32_t find_next(const uint8_t* words, uint32_t nwords, uint32_t start)
{
    uint32_t n = start >> 6;
    if (n >= nwords) return nwords * 64;
    uint64_t b = r64(words + 8 * n);
    if (b & ((uint64_t)1 << (start & 63))) return start;
    b &= ~(uint64_t)0 << (start & 63);
    while (!b && ++n < nwords) b = r64(words + 8 * n);
    return b ? (n << 6) + (uint32_t)__builtin_ctzll(b) : nwords * 64;
}
/* last on bit <= start, or size */
static uint32_t find_prev(const uint8_t* words, uint32_t nwords, uint32_t start)
{
    uint32_t n = start >> 6;
    if (n >= nwords) return nwords * 64;
    uint64_t b = r64(words + 8 * n);
    if (b & ((uint64_t)1 << (start & 63))) return start;
    b &= ((uint64_t)1 << (start & 63)) - 1;
    while (!b && n) b = r64(words + 8 * --n);
    return b ? (n << 6) + 63u - (uint32_t)__builtin_clzll(b) : nwords * 64;
}

/* node levels: 0 leaf (2144 B), 1 lower (33856 B), 2 upper (270400 B) */
static const uint32_t kCmask[3] = {16, 544, 4128}, kTable[3] = {96, 1088, 8256}, kWords[3] = {8, 64, 512};

static float edge_value(const uint8_t* node, int level, int last)
{
    if (level == 0) return rf(node + 96 + 4 * (last ? 511 : 0));
    const uint32_t size = kWords[level] * 64, n = last ? size - 1 : 0;
    if ((r64(node + kCmask[level] + 8 * (n >> 6)) >> (n & 63)) & 1u)
        return edge_value(node + (int64_t)r64(node + kTable[level] + 8ull * n), level - 1, last);
    return rf(node + kTable[level] + 8ull * n);
}

typedef struct {
    int32_t c[3];
    uint32_t tile;
} RootChild;

static int cmp_child(const void* pa, const void* pb)
{
    const RootChild* a = (const RootChild*)pa;
    const RootChild* b = (const RootChild*)pb;
    for (int i = 0; i < 3; ++i)
        if (a->c[i] != b->c[i]) return a->c[i] < b->c[i] ? -1 : 1;
    return 0;
}

static uint64_t coord_key(const int32_t c[3])
{
    return (uint64_t)((uint32_t)c[2] >> 12) | ((uint64_t)((uint32_t)c[1] >> 12) << 21) | ((uint64_t)((uint32_t)c[0] >> 12) << 42);
}

int or_signed_flood_fill(void* buf_)
{
    uint8_t* b = (uint8_t*)buf_;
    uint32_t type;
    memcpy(&type, b + 636, 4);
    if (type != 1u) return -4;
    const uint8_t* T = b + 672;
    int64_t off[4];
    for (int i = 0; i < 4; ++i) off[i] = (int64_t)r64(T + 8 * i);
    uint32_t cnt[3];
    memcpy(cnt, T + 32, 12);
    uint8_t* root = b + 672 + off[3];
    const float bg = rf(root + 28);

    /* leaves */
    for (uint32_t l = 0; l < cnt[0]; ++l) {
        uint8_t* L = b + 672 + off[0] + 2144ull * l;
        for (uint32_t v = 0; v < 512; ++v) {
            if ((r64(L + 16 + 8 * (v >> 6)) >> (v & 63)) & 1u) continue;
            uint32_t n = find_next(L + 16, 8, v);
            if (n == 512 && (n = find_prev(L + 16, 8, v)) == 512) n = 0;
            wf(L + 96 + 4 * v, rf(L + 96 + 4 * n) < 0.0f ? -bg : bg);
        }
    }
    /* lower, then upper nodes */
    for (int level = 1; level <= 2; ++level) {
        const uint64_t bytes = level == 1 ? 33856u : 270400u;
        const uint32_t size = kWords[level] * 64;
        for (uint32_t k = 0; k < cnt[level]; ++k) {
            uint8_t* N = b + 672 + off[level] + bytes * k;
            const uint8_t* cm = N + kCmask[level];
            for (uint32_t v = 0; v < size; ++v) {
                if ((r64(cm + 8 * (v >> 6)) >> (v & 63)) & 1u) continue;
                float value = bg;
                uint32_t n = find_next(cm, kWords[level], v);
                if (n < size) {
                    if (edge_value(N + (int64_t)r64(N + kTable[level] + 8ull * n), level - 1, 0) < 0.0f) value = -bg;
                } else if ((n = find_prev(cm, kWords[level], v)) < size) {
                    if (edge_value(N + (int64_t)r64(N + kTable[level] + 8ull * n), level - 1, 1) < 0.0f) value = -bg;
                } else if (rf(N + kTable[level]) < 0.0f) {
                    value = -bg;
                }
                wf(N + kTable[level] + 8ull * v, value);
            }
        }
    }
    /* root */
    uint32_t tableSize;
    memcpy(&tableSize, root + 24, 4);
    if (tableSize == 0) return 0;
    RootChild* ch = (RootChild*)malloc(sizeof(RootChild) * tableSize);
    if (!ch) return -2;
    uint32_t nc = 0;
    for (uint32_t i = 0; i < tableSize; ++i) {
        const uint8_t* tile = root + 64 + 32ull * i;
        if (!r64(tile + 8)) continue;
        const uint64_t key = r64(tile);
        ch[nc].c[0] = (int32_t)((uint32_t)((key >> 42) & 0x1fffffu) << 12);
        ch[nc].c[1] = (int32_t)((uint32_t)((key >> 21) & 0x1fffffu) << 12);
        ch[nc].c[2] = (int32_t)((uint32_t)(key & 0x1fffffu) << 12);
        ch[nc].tile = i;
        ++nc;
    }
    int rc = 0;
    if (nc >= 2) {
        qsort(ch, nc, sizeof(RootChild), cmp_child);
        for (uint32_t i = 0; i + 1 < nc && rc == 0; ++i) {
            const RootChild* a = &ch[i];
            const RootChild* c = &ch[i + 1];
            if (a->c[0] != c->c[0] || a->c[1] != c->c[1] || c->c[2] - a->c[2] == 4096) continue;
            const uint8_t* ua = root + (int64_t)r64(root + 64 + 32ull * a->tile + 8);
            const uint8_t* uc = root + (int64_t)r64(root + 64 + 32ull * c->tile + 8);
            if (edge_value(ua, 2, 1) > 0.0f || edge_value(uc, 2, 0) > 0.0f) continue;
            for (int32_t z = a->c[2] + 4096; z != c->c[2] && rc == 0; z += 4096) {
                const int32_t q[3] = {a->c[0], a->c[1], z};
                const uint64_t key = coord_key(q);
                uint32_t t = 0;
                while (t < tableSize && r64(root + 64 + 32ull * t) != key) ++t;
                if (t == tableSize) {
                    rc = -3;
                    break;
                }
                uint8_t* tile = root + 64 + 32ull * t;
                const uint32_t state = 0;
                const int64_t child = 0;
                memcpy(tile + 8, &child, 8);
                memcpy(tile + 16, &state, 4);
                wf(tile + 20, -bg);
            }
        }
    }
    free(ch);
    return rc;
}
