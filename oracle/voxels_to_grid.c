/* Oracle (test infrastructure only, never part of the product): CPU restatement of
 * tools/cuda/PointsToGrid.cuh voxelsToGrid<BuildT> (:1417-1424) for BuildT = float and
 * ValueOnIndex: the NanoVDB buffer the reference's device builder writes for a list of
 * integer voxels, byte for byte.
 *
 * The reference orders voxels by a 63-bit tile key ((x+2^31)>>12 | ... , TileKeyFunctor
 * :553-580), then by the 36-bit in-tile key upperOffset<<21 | lowerOffset<<9 | leafOffset
 * (VoxelKeyFunctor :582-598); here one comparison does both (signed tile coordinates in
 * x, y, z order, then the in-tile key).  Node lists are the runs of that order at 36, 21
 * and 9 bits (countNodes :619-826); node contents follow BuildGridTreeRootFunctor
 * (:877-1008), BuildUpperNodesFunctor (:1034-1061), BuildLowerNodesFunctor (:1085-1104),
 * ProcessLeafMetaDataFunctor (:1126-1154), SetLeafActiveVoxelStateAndValuesFunctor
 * (:1156-1174), leafPrefixSumKernel (:459-479), processBBox (:1386-1414) and
 * LeafNode::updateBBox (NanoVDB.h:4596-4624).  The buffer starts zeroed (getBuffer :866-871).
 */
#include "oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

enum { GT_FLOAT_ = 1, GT_INDEX_ = 19, GT_ONINDEX_ = 20 };

typedef struct {
    int32_t t[3];   /* tile coordinates (x>>12, arithmetic) */
    uint64_t key;   /* upper<<21 | lower<<9 | leaf offsets */
    uint64_t id;    /* input position */
} VKey;

static int cmp_vkey(const void* pa, const void* pb)
{
    const VKey* a = (const VKey*)pa;
    const VKey* b = (const VKey*)pb;
    for (int i = 0; i < 3; ++i)
        if (a->t[i] != b->t[i]) return a->t[i] < b->t[i] ? -1 : 1;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->id < b->id ? -1 : (a->id > b->id);
}

static void put_u32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }
static void put_u64(uint8_t* p, uint64_t v) { memcpy(p, &v, 8); }
static void put_i32(uint8_t* p, int32_t v) { memcpy(p, &v, 4); }
static void put_i64(uint8_t* p, int64_t v) { memcpy(p, &v, 8); }
static void put_f32(uint8_t* p, float v) { memcpy(p, &v, 4); }
static void put_f64(uint8_t* p, double v) { memcpy(p, &v, 8); }
static uint64_t get_u64(const uint8_t* p)
{
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
static int32_t get_i32(const uint8_t* p)
{
    int32_t v;
    memcpy(&v, p, 4);
    return v;
}

/* Map::applyMap with mMatD = s*I, mVecD = 0 (matMult, math/Math.h:1748-1753) */
static void apply_map(double s, const double x[3], double y[3])
{
    for (int r = 0; r < 3; ++r) {
        const double m[3] = {r == 0 ? s : 0.0, r == 1 ? s : 0.0, r == 2 ? s : 0.0};
        y[r] = fma(x[0], m[0], fma(x[1], m[1], fma(x[2], m[2], 0.0)));
    }
}

static void bbox_empty(int32_t b[6])
{
    b[0] = b[1] = b[2] = INT_MAX;
    b[3] = b[4] = b[5] = INT_MIN;
}
static void bbox_expand(int32_t b[6], const int32_t lo[3], const int32_t hi[3])
{
    for (int a = 0; a < 3; ++a) {
        if (lo[a] < b[a]) b[a] = lo[a];
        if (hi[a] > b[3 + a]) b[3 + a] = hi[a];
    }
}

static void put_u16(uint8_t* p, uint16_t v) { memcpy(p, &v, 2); }

/* Blind data of a point grid: PointType (NanoVDB.h:395-405) -> value size, GridBlindDataSemantic,
 * GridBlindDataClass, GridType of the values and the name processGridTreeRoot writes
 * (tools/cuda/PointsToGrid.cuh:913-1001, float input points). */
typedef struct {
    uint32_t size, semantic, dataClass, dataType, gridClass;
    const char* name;
} PointKind;
static int point_kind(uint32_t t, PointKind* k)
{
    static const PointKind kinds[10] = {
        {0, 0, 0, 0, 0, ""},
        {4, 6, 1, 10, 4, "PointID: uint32_t indices to points"},
        {24, 7, 2, 7, 5, "World64: Vec3<double> point coordinates in world space"},
        {12, 7, 2, 6, 5, "World32: Vec3<float> point coordinates in world space"},
        {24, 8, 2, 7, 5, "Grid64: Vec3<double> point coordinates in grid space"},
        {12, 8, 2, 6, 5, "Grid32: Vec3<float> point coordinates in grid space"},
        {12, 9, 2, 6, 5, "Voxel32: Vec3<float> point coordinates in voxel space"},
        {6, 9, 2, 25, 5, "Voxel16: Vec3<uint16_t> point coordinates in voxel space"},
        {3, 9, 2, 24, 5, "Voxel8: Vec3<uint8_t> point coordinates in voxel space"},
        {12, 7, 2, 6, 5, "World32: Vec3<float> point coordinates in world space"}};
    if (t == 0 || t > 9) return -1;
    *k = kinds[t];
    return 0;
}

/* Map(dx)::applyInverseMapF of a float point (matMult with mInvMatF, math/Math.h:1704-1709) */
static void inv_map_f(float invF, const float p[3], float q[3])
{
    for (int r = 0; r < 3; ++r) {
        const float m[3] = {r == 0 ? invF : 0.0f, r == 1 ? invF : 0.0f, r == 2 ? invF : 0.0f};
        q[r] = fmaf(p[0], m[0], fmaf(p[1], m[1], p[2] * m[2]));
    }
}
/* Map(dx)::applyInverseMap of a float point: FP64 matMult (math/Math.h:1716-1722), returned as
 * the input's float type */
static void inv_map_d(double inv, const float p[3], float q[3])
{
    for (int r = 0; r < 3; ++r) {
        const double m[3] = {r == 0 ? inv : 0.0, r == 1 ? inv : 0.0, r == 2 ? inv : 0.0};
        q[r] = (float)fma((double)p[0], m[0], fma((double)p[1], m[1], (double)p[2] * m[2]));
    }
}

/* Map(dx)::applyInverseMap of a double point: FP64 matMult (math/Math.h:1716-1722) */
static void inv_map_dd(double inv, const double p[3], double q[3])
{
    for (int r = 0; r < 3; ++r) {
        const double m[3] = {r == 0 ? inv : 0.0, r == 1 ? inv : 0.0, r == 2 ? inv : 0.0};
        q[r] = fma(p[0], m[0], fma(p[1], m[1], p[2] * m[2]));
    }
}

/* mode 1 float, 20 ValueOnIndex, 23 Point (pts or ptsd, float or double points, and point_type
 * then describe the points) */
static int build_grid(const int32_t* ijk, uint64_t n, uint32_t grid_type, double voxel_size, const char* name,
                      const float* pts, const double* ptsd, uint32_t point_type, void** out, uint64_t* bytes_out,
                      uint64_t* out_index)
{
    *out = NULL;
    *bytes_out = 0;
    const int pt = grid_type == 23u;
    PointKind kind = {0, 0, 0, 0, 0, ""};
    if (pt && ((!pts && !ptsd) || point_kind(point_type, &kind) != 0)) return -1;
    if (pt && ptsd && point_type == 9) { /* Default stores the input type: Vec3<double> */
        kind.size = 24;
        kind.dataType = 7;
        kind.name = "World64: Vec3<double> point coordinates in world space";
    }
    VKey* k = (VKey*)malloc(n * sizeof(VKey));
    if (!k) return -2;
    for (uint64_t i = 0; i < n; ++i) {
        const int32_t x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
        k[i].t[0] = x >> 12;
        k[i].t[1] = y >> 12;
        k[i].t[2] = z >> 12;
        /* InternalNode::CoordToOffset (NanoVDB.h:3560-3566), LeafNode::CoordToOffset (:4516-4519) */
        const uint64_t up = ((uint64_t)((x & 4095) >> 7) << 10) | ((uint64_t)((y & 4095) >> 7) << 5) | (uint64_t)((z & 4095) >> 7);
        const uint64_t lo = ((uint64_t)((x & 127) >> 3) << 8) | ((uint64_t)((y & 127) >> 3) << 4) | (uint64_t)((z & 127) >> 3);
        const uint64_t lf = ((uint64_t)(x & 7) << 6) | ((uint64_t)(y & 7) << 3) | (uint64_t)(z & 7);
        k[i].key = (up << 21) | (lo << 9) | lf;
        k[i].id = i;
    }
    qsort(k, n, sizeof(VKey), cmp_vkey);

    /* run heads at the voxel, leaf, lower and upper level */
    uint64_t nVox = 0, nLeaf = 0, nLower = 0, nUpper = 0;
    for (uint64_t p = 0; p < n; ++p) {
        const int newTile = p == 0 || memcmp(k[p].t, k[p - 1].t, sizeof(k[p].t)) != 0;
        nUpper += newTile;
        nLower += newTile || (k[p].key >> 21) != (k[p - 1].key >> 21);
        nLeaf += newTile || (k[p].key >> 9) != (k[p - 1].key >> 9);
        nVox += newTile || k[p].key != k[p - 1].key;
    }
    const int idx = grid_type == GT_ONINDEX_ || grid_type == GT_INDEX_;
    const int all = grid_type == GT_INDEX_; /* ValueIndex: every leaf value has an index */
    const uint64_t rootHdr = (idx || pt) ? 96u : 64u, leafSize = idx ? 96u : pt ? 1120u : 2144u;
    const uint64_t offTree = 672, offRoot = offTree + 64, offUpper = offRoot + rootHdr + 32u * nUpper;
    const uint64_t offLower = offUpper + 270400u * nUpper, offLeaf = offLower + 33856u * nLower;
    const uint64_t offMeta = offLeaf + leafSize * nLeaf;
    const uint64_t total = pt ? offMeta + 288u + (uint64_t)kind.size * n : offMeta;
    uint8_t* b = (uint8_t*)aligned_alloc(32, (total + 31) & ~(uint64_t)31);
    if (!b) {
        free(k);
        return -2;
    }
    memset(b, 0, total);

    /* node contents in list order */
    int64_t u = -1, m = -1, l = -1, v = -1;
    int32_t rootBox[6];
    bbox_empty(rootBox);
    int32_t* lowerBox = (int32_t*)malloc(6 * sizeof(int32_t) * nLower);
    int32_t* upperBox = (int32_t*)malloc(6 * sizeof(int32_t) * nUpper);
    int64_t* lowerParent = (int64_t*)malloc(sizeof(int64_t) * nLower);
    int64_t* leafParent = (int64_t*)malloc(sizeof(int64_t) * nLeaf);
    uint64_t* leafStart = (uint64_t*)malloc(sizeof(uint64_t) * (nLeaf + 1));
    if (!lowerBox || !upperBox || !lowerParent || !leafParent || !leafStart) {
        free(leafStart);
        free(k);
        free(b);
        free(lowerBox);
        free(upperBox);
        free(lowerParent);
        free(leafParent);
        return -2;
    }
    for (uint64_t p = 0; p < n; ++p) {
        const int newTile = p == 0 || memcmp(k[p].t, k[p - 1].t, sizeof(k[p].t)) != 0;
        const int newLower = newTile || (k[p].key >> 21) != (k[p - 1].key >> 21);
        const int newLeaf = newTile || (k[p].key >> 9) != (k[p - 1].key >> 9);
        const int newVox = newTile || k[p].key != k[p - 1].key;
        const int32_t* t = k[p].t;
        const uint32_t n2 = (uint32_t)(k[p].key >> 21), n1 = (uint32_t)(k[p].key >> 9) & 4095u, n0 = (uint32_t)k[p].key & 511u;
        if (newTile) {
            ++u;
            const uint64_t oU = offUpper + 270400u * (uint64_t)u, tile = offRoot + rootHdr + 32u * (uint64_t)u;
            const int32_t o[3] = {t[0] * 4096, t[1] * 4096, t[2] * 4096};
            /* RootData::CoordToKey (NanoVDB.h:2633-2640) + Tile::setChild */
            put_u64(b + tile, (uint64_t)((uint32_t)o[2] >> 12) | ((uint64_t)((uint32_t)o[1] >> 12) << 21) |
                                  ((uint64_t)((uint32_t)o[0] >> 12) << 42));
            put_i64(b + tile + 8, (int64_t)(oU - offRoot));
            put_u64(b + oU + 24, 2u); /* GridFlags::HasBBox */
            bbox_empty(upperBox + 6 * u);
        }
        if (newLower) {
            ++m;
            const uint64_t oU = offUpper + 270400u * (uint64_t)u, oL = offLower + 33856u * (uint64_t)m;
            uint8_t* cm = b + oU + 4128 + 8 * (n2 >> 6);
            put_u64(cm, get_u64(cm) | ((uint64_t)1 << (n2 & 63)));
            put_i64(b + oU + 8256 + 8ull * n2, (int64_t)(oL - oU));
            put_u64(b + oL + 24, 2u);
            bbox_empty(lowerBox + 6 * m);
            lowerParent[m] = u;
        }
        if (newLeaf) {
            ++l;
            const uint64_t oL = offLower + 33856u * (uint64_t)m, oF = offLeaf + leafSize * (uint64_t)l;
            uint8_t* cm = b + oL + 544 + 8 * (n1 >> 6);
            put_u64(cm, get_u64(cm) | ((uint64_t)1 << (n1 & 63)));
            put_i64(b + oL + 1088 + 8ull * n1, (int64_t)(oF - oL));
            /* lower->offsetToGlobalCoord(n1): the leaf origin */
            const int32_t ox = t[0] * 4096 + (int32_t)((n2 >> 10) & 31) * 128 + (int32_t)((n1 >> 8) & 15) * 8;
            const int32_t oy = t[1] * 4096 + (int32_t)((n2 >> 5) & 31) * 128 + (int32_t)((n1 >> 4) & 15) * 8;
            const int32_t oz = t[2] * 4096 + (int32_t)(n2 & 31) * 128 + (int32_t)(n1 & 15) * 8;
            put_i32(b + oF, ox);
            put_i32(b + oF + 4, oy);
            put_i32(b + oF + 8, oz);
            b[oF + 15] = 2u; /* HasBBox */
            /* 1 + values before this leaf: active voxels (ValueOnIndex), all leaf values (ValueIndex,
             * tools/cuda/PointsToGrid.cuh:1147) */
            if (idx) put_u64(b + oF + 80, all ? 1u + 512u * (uint64_t)l : (uint64_t)(v + 1) + 1u);
            if (pt) put_u64(b + oF + 80, p);                       /* points before this leaf */
            leafStart[l] = p;
            leafParent[l] = m;
        }
        if (newVox) {
            ++v;
            const uint64_t oF = offLeaf + leafSize * (uint64_t)l;
            uint8_t* w = b + oF + 16 + 8 * (n0 >> 6);
            put_u64(w, get_u64(w) | ((uint64_t)1 << (n0 & 63)));
            if (!idx && !pt) put_f32(b + oF + 96 + 4 * n0, 1.0f);
        }
        /* Point leaves: the points of the leaf up to and including this voxel
         * (SetLeafActiveVoxelStateAndValuesFunctor, tools/cuda/PointsToGrid.cuh:1156-1174) */
        if (pt && (p + 1 == n || k[p + 1].key != k[p].key || memcmp(k[p + 1].t, k[p].t, sizeof(k[p].t)) != 0))
            put_u16(b + offLeaf + leafSize * (uint64_t)l + 96 + 2 * n0, (uint16_t)(p + 1 - leafStart[l]));
        if (out_index) out_index[k[p].id] = all ? 1u + 512u * (uint64_t)l + n0 : (uint64_t)v + 1u;
    }

    /* leaves: updateBBox, prefix sums; bboxes up the tree (processBBox) */
    leafStart[nLeaf] = n;
    for (uint64_t i = 0; i < nLeaf; ++i) {
        uint8_t* L = b + offLeaf + leafSize * i;
        uint64_t w[8];
        for (int j = 0; j < 8; ++j) w[j] = get_u64(L + 16 + 8 * j);
        if (pt) {
            if (leafStart[i + 1] - leafStart[i] > 65535u) { /* the reference throws: too many points per leaf */
                free(k);
                free(b);
                free(lowerBox);
                free(upperBox);
                free(lowerParent);
                free(leafParent);
                free(leafStart);
                return -3;
            }
            put_u64(L + 88, leafStart[i + 1] - leafStart[i]);
            /* inactive voxels: the value of the previous active voxel, else 0
             * (SetLeafInactiveVoxelValuesFunctor, :1176-1191) */
            uint16_t last = 0;
            for (uint32_t q = 0; q < 512; ++q) {
                if ((w[q >> 6] >> (q & 63)) & 1u) memcpy(&last, L + 96 + 2 * q, 2);
                else put_u16(L + 96 + 2 * q, last);
            }
        }
        uint64_t word = w[0];
        uint32_t xmin = w[0] ? 0u : 8u, xmax = xmin;
        for (int j = 1; j < 8; ++j)
            if (w[j]) {
                word |= w[j];
                if (xmin == 8) xmin = (uint32_t)j;
                xmax = (uint32_t)j;
            }
        const uint32_t ymin = (uint32_t)__builtin_ctzll(word) >> 3, ymax = (63u - (uint32_t)__builtin_clzll(word)) >> 3;
        const uint32_t w32 = (uint32_t)word | (uint32_t)(word >> 32);
        const uint32_t w16 = (w32 & 0xffffu) | (w32 >> 16);
        const uint32_t byte = (w16 & 0xffu) | (w16 >> 8);
        const uint32_t zmin = (uint32_t)__builtin_ctz(byte), zmax = 31u - (uint32_t)__builtin_clz(byte);
        const uint32_t mn[3] = {xmin, ymin, zmin}, mx[3] = {xmax, ymax, zmax};
        int32_t lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = (get_i32(L + 4 * a) & ~7) + (int32_t)mn[a];
            hi[a] = lo[a] + (int32_t)(mx[a] - mn[a]);
            put_i32(L + 4 * a, lo[a]);
            L[12 + a] = (uint8_t)(mx[a] - mn[a]);
        }
        if (idx && !all) { /* ValueIndex leaves keep a zero prefix sum (:1148) */
            uint64_t sum = (uint64_t)__builtin_popcountll(w[0]), prefix = sum;
            for (int j = 1; j < 7; ++j) {
                sum += (uint64_t)__builtin_popcountll(w[j]);
                prefix |= sum << (9 * j);
            }
            put_u64(L + 88, prefix);
        }
        bbox_expand(lowerBox + 6 * leafParent[i], lo, hi);
    }
    for (uint64_t i = 0; i < nLower; ++i) {
        uint8_t* N = b + offLower + 33856u * i;
        const int32_t* bx = lowerBox + 6 * i;
        for (int a = 0; a < 6; ++a) put_i32(N + 4 * a, bx[a]);
        bbox_expand(upperBox + 6 * lowerParent[i], bx, bx + 3);
    }
    for (uint64_t i = 0; i < nUpper; ++i) {
        uint8_t* N = b + offUpper + 270400u * i;
        const int32_t* bx = upperBox + 6 * i;
        for (int a = 0; a < 6; ++a) put_i32(N + 4 * a, bx[a]);
        bbox_expand(rootBox, bx, bx + 3);
    }

    /* root, tree and grid headers */
    uint8_t* R = b + offRoot;
    for (int a = 0; a < 6; ++a) put_i32(R + 4 * a, rootBox[a]);
    put_u32(R + 24, (uint32_t)nUpper);
    uint8_t* T = b + offTree;
    put_i64(T + 0, (int64_t)(offLeaf - offTree));
    put_i64(T + 8, (int64_t)(offLower - offTree));
    put_i64(T + 16, (int64_t)(offUpper - offTree));
    put_i64(T + 24, (int64_t)(offRoot - offTree));
    put_u32(T + 32, (uint32_t)nLeaf);
    put_u32(T + 36, (uint32_t)nLower);
    put_u32(T + 40, (uint32_t)nUpper);
    put_u64(T + 56, nVox);

    put_u64(b + 0, 0x314244566f6e614eull); /* NANOVDB_MAGIC_GRID */
    put_u64(b + 8, ~(uint64_t)0);          /* checksum disabled */
    put_u32(b + 16, (32u << 21) | (9u << 10) | 2u);
    put_u32(b + 20, 2u | 32u); /* HasBBox | IsBreadthFirst */
    put_u32(b + 24, 0u);
    put_u32(b + 28, 1u);
    put_u64(b + 32, total);
    if (name) {
        size_t len = strlen(name);
        if (len > 255) len = 255;
        memcpy(b + 40, name, len);
    }
    /* Map(voxelSize), NanoVDB.h:1441-1451 */
    const float sF = (float)voxel_size, invF = 1.0f / sF;
    for (int i = 0; i < 9; ++i) {
        put_f32(b + 296 + 4 * i, i % 4 == 0 ? sF : 0.0f);
        put_f32(b + 296 + 36 + 4 * i, i % 4 == 0 ? invF : 0.0f);
        put_f64(b + 296 + 88 + 8 * i, i % 4 == 0 ? voxel_size : 0.0);
        put_f64(b + 296 + 160 + 8 * i, i % 4 == 0 ? 1.0 / voxel_size : 0.0);
    }
    put_f32(b + 296 + 84, 1.0f);
    put_f64(b + 296 + 256, 1.0);
    /* world bbox: CoordBBox(min, max + 1).transform(map), math/Math.h:2070-2083 */
    double wb[6];
    for (int c = 0; c < 8; ++c) {
        const double x[3] = {(double)(c & 1 ? rootBox[3] + 1 : rootBox[0]), (double)(c & 2 ? rootBox[4] + 1 : rootBox[1]),
                             (double)(c & 4 ? rootBox[5] + 1 : rootBox[2])};
        double y[3];
        apply_map(voxel_size, x, y);
        for (int a = 0; a < 3; ++a) {
            if (c == 0 || y[a] < wb[a]) wb[a] = y[a];
            if (c == 0 || y[a] > wb[3 + a]) wb[3 + a] = y[a];
        }
    }
    for (int a = 0; a < 6; ++a) put_f64(b + 560 + 8 * a, wb[a]);
    {
        const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
        double p1[3], p0[3];
        apply_map(voxel_size, one, p1);
        apply_map(voxel_size, zero, p0);
        for (int a = 0; a < 3; ++a) put_f64(b + 608 + 8 * a, p1[a] - p0[a]);
    }
    put_u32(b + 632, idx ? 8u : pt ? kind.gridClass : 0u); /* GridClass IndexGrid / Point* / Unknown */
    put_u32(b + 636, grid_type);
    put_i64(b + 640, (int64_t)offMeta); /* blind metadata: none follows unless this is a point grid */
    put_u32(b + 648, pt ? 1u : 0u);
    put_u64(b + 656, all ? 1u + 512u * (uint64_t)nLeaf : idx ? nVox + 1u : pt ? n : 0u);
    if (pt) {
        /* GridBlindMetaData (NanoVDB.h:1591-1601), then the points in sorted order (encodePoints,
         * tools/cuda/PointsToGrid.cuh:1262-1317) */
        uint8_t* M = b + offMeta;
        put_i64(M, 288);
        put_u64(M + 8, n);
        put_u32(M + 16, kind.size);
        put_u32(M + 20, kind.semantic);
        put_u32(M + 24, kind.dataClass);
        put_u32(M + 28, kind.dataType);
        memcpy(M + 32, kind.name, strlen(kind.name));
        uint8_t* D = M + 288;
        const float invF = 1.0f / (float)voxel_size;
        const double invD = 1.0 / voxel_size;
        for (uint64_t q = 0; q < n; ++q) {
            const uint64_t id = k[q].id;
            uint8_t* o = D + (uint64_t)kind.size * q;
            if (ptsd) { /* double points: the same encodings computed from Vec3<double> */
                const double* wd = ptsd + 3 * id;
                double gd[3];
                float wf[3], gf[3];
                switch (point_type) {
                case 1: put_u32(o, (uint32_t)id); break;
                case 2:
                case 9:
                    for (int a = 0; a < 3; ++a) put_f64(o + 8 * a, wd[a]);
                    break;
                case 3:
                    for (int a = 0; a < 3; ++a) put_f32(o + 4 * a, (float)wd[a]);
                    break;
                case 4:
                    inv_map_dd(invD, wd, gd);
                    for (int a = 0; a < 3; ++a) put_f64(o + 8 * a, gd[a]);
                    break;
                case 5: /* applyInverseMapF casts each coordinate to float first (math/Math.h:1704-1709) */
                    for (int a = 0; a < 3; ++a) wf[a] = (float)wd[a];
                    inv_map_f(invF, wf, gf);
                    memcpy(o, gf, 12);
                    break;
                default:
                    inv_map_dd(invD, wd, gd);
                    for (int a = 0; a < 3; ++a) {
                        const double f = gd[a] - (double)(int32_t)floor(gd[a] + 0.5);
                        if (point_type == 6) put_f32(o + 4 * a, (float)f);
                        else if (point_type == 7) put_u16(o + 2 * a, (uint16_t)(65535.0 * (f + 0.5)));
                        else o[a] = (uint8_t)(255.0 * (f + 0.5));
                    }
                }
                continue;
            }
            const float* w = pts + 3 * id;
            float g[3];
            switch (point_type) {
            case 1: put_u32(o, (uint32_t)id); break;
            case 2:
                for (int a = 0; a < 3; ++a) put_f64(o + 8 * a, (double)w[a]);
                break;
            case 3:
            case 9: memcpy(o, w, 12); break;
            case 4:
                inv_map_d(invD, w, g);
                for (int a = 0; a < 3; ++a) put_f64(o + 8 * a, (double)g[a]);
                break;
            case 5:
                inv_map_f(invF, w, g);
                memcpy(o, g, 12);
                break;
            default: { /* worldToVoxel (:197-236): FP64 offsets from the voxel centre */
                inv_map_d(invD, w, g);
                for (int a = 0; a < 3; ++a) {
                    const double x = (double)g[a];
                    const double f = x - (double)(int32_t)floor(x + 0.5);
                    if (point_type == 6) put_f32(o + 4 * a, (float)f);
                    else if (point_type == 7) put_u16(o + 2 * a, (uint16_t)(65535.0 * (f + 0.5)));
                    else o[a] = (uint8_t)(255.0 * (f + 0.5));
                }
            }
            }
        }
    }

    free(k);
    free(lowerBox);
    free(upperBox);
    free(lowerParent);
    free(leafParent);
    free(leafStart);
    *out = b;
    *bytes_out = total;
    return 0;
}

void or_free(void* p) { free(p); }

int or_voxels_to_grid(const int32_t* ijk, uint64_t n, uint32_t grid_type, double voxel_size, const char* name,
                      void** out, uint64_t* bytes_out, uint64_t* out_index)
{
    *out = NULL;
    *bytes_out = 0;
    if (!ijk || n == 0 || (grid_type != GT_FLOAT_ && grid_type != GT_ONINDEX_ && grid_type != GT_INDEX_)) return -1;
    if (out_index && grid_type == GT_FLOAT_) return -1;
    return build_grid(ijk, n, grid_type, voxel_size, name, NULL, NULL, 0, out, bytes_out, out_index);
}

/* tools/cuda/PointsToGrid.cuh pointsToGrid(d_points, n, voxelSize, type) (:57-64) for float world
 * points: each point's voxel is Round(Map(voxelSize).applyInverseMapF(p)) (TileKeyFunctor,
 * :553-580; Vec3<float>::round, math/Math.h:1515-1524). */
int or_points_to_grid(const float* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                      void** out, uint64_t* bytes_out)
{
    *out = NULL;
    *bytes_out = 0;
    if (!xyz || n == 0) return -1;
    int32_t* ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * n);
    if (!ijk) return -2;
    const float invF = 1.0f / (float)voxel_size;
    for (uint64_t i = 0; i < n; ++i) {
        float q[3];
        inv_map_f(invF, xyz + 3 * i, q);
        for (int a = 0; a < 3; ++a) ijk[3 * i + a] = (int32_t)floorf(q[a] + 0.5f);
    }
    const int rc = build_grid(ijk, n, 23u, voxel_size, name, xyz, NULL, point_type, out, bytes_out, NULL);
    free(ijk);
    return rc;
}

/* The same for double world points (PtrT = Vec3d*): each point's voxel is
 * Round(Map(voxelSize).applyInverseMap(p)) in FP64 (TileKeyFunctor :571-572; Vec3<double>::round,
 * math/Math.h:1515-1524), and the encodings start from the double coordinates. */
int or_points_to_grid_f64(const double* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                          void** out, uint64_t* bytes_out)
{
    *out = NULL;
    *bytes_out = 0;
    if (!xyz || n == 0) return -1;
    int32_t* ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * n);
    if (!ijk) return -2;
    const double invD = 1.0 / voxel_size;
    for (uint64_t i = 0; i < n; ++i) {
        double q[3];
        inv_map_dd(invD, xyz + 3 * i, q);
        for (int a = 0; a < 3; ++a) ijk[3 * i + a] = (int32_t)floor(q[a] + 0.5);
    }
    const int rc = build_grid(ijk, n, 23u, voxel_size, name, NULL, xyz, point_type, out, bytes_out, NULL);
    free(ijk);
    return rc;
}

/* ---- voxel size for a target point density ---------------------------------------------------
 * PointsToGrid(maxPointsPerVoxel, tolerance, maxIterations) (tools/cuda/PointsToGrid.cuh:335-342)
 * restated: countNodes (:630-774) starts from Map(1.0), counts the largest number of points in one
 * voxel at the current voxel size dx, and stops when that count is within the tolerance of the
 * target (clamped to the point count), after maxIterations counts, or for a single point.
 * Otherwise the sample (dx, count) becomes the bracket's lower end (count below the target: the
 * largest such sample by count, then dx) or upper end (the smallest sample at or above it), and
 * dx is interpolated linearly between the two ends; while no upper end is known dx is scaled by
 * (target - 1) / (count - 1), or by 10 at a count of 1.  The grid is then built with the last dx. */
static int cmp_ijk(const void* a, const void* b)
{
    const int32_t* p = (const int32_t*)a;
    const int32_t* q = (const int32_t*)b;
    for (int i = 0; i < 3; ++i)
        if (p[i] != q[i]) return p[i] < q[i] ? -1 : 1;
    return 0;
}

/* largest number of points that round to one voxel at voxel size dx */
static uint32_t max_per_voxel(const float* xf, const double* xd, uint64_t n, double dx, int32_t* ijk)
{
    const float invF = 1.0f / (float)dx;
    const double invD = 1.0 / dx;
    for (uint64_t i = 0; i < n; ++i) {
        if (xd) {
            double q[3];
            inv_map_dd(invD, xd + 3 * i, q);
            for (int a = 0; a < 3; ++a) ijk[3 * i + a] = (int32_t)floor(q[a] + 0.5);
        } else {
            float q[3];
            inv_map_f(invF, xf + 3 * i, q);
            for (int a = 0; a < 3; ++a) ijk[3 * i + a] = (int32_t)floorf(q[a] + 0.5f);
        }
    }
    qsort(ijk, n, 12, cmp_ijk);
    uint32_t most = 0, run = 0;
    for (uint64_t i = 0; i < n; ++i) {
        run = (i > 0 && cmp_ijk(ijk + 3 * i, ijk + 3 * (i - 1)) == 0) ? run + 1 : 1;
        if (run > most) most = run;
    }
    return most;
}

typedef struct {
    double dx;
    uint32_t density;
} DensitySample;
static int sample_less(DensitySample a, DensitySample b)
{
    return a.density < b.density || (a.density == b.density && a.dx < b.dx);
}

int or_points_to_grid_density(const void* xyz, int f64, uint64_t n, uint32_t target, int tolerance, int max_iterations,
                              uint32_t point_type, const char* name, void** out, uint64_t* bytes_out,
                              double* dx_out, uint32_t* density_out)
{
    *out = NULL;
    *bytes_out = 0;
    if (!xyz || n == 0 || target == 0) return -1;
    const float* xf = f64 ? NULL : (const float*)xyz;
    const double* xd = f64 ? (const double*)xyz : NULL;
    if ((uint64_t)target > n) target = (uint32_t)n;
    int32_t* ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * n);
    if (!ijk) return -2;
    DensitySample lo = {0.0, 1u}, hi = {0.0, 0u};
    double dx = 1.0;
    uint32_t density = 0;
    for (int iter = 0;;) {
        if (!(dx > 0.0) || !isfinite(dx)) { /* the update left the positive reals (a target of 1) */
            free(ijk);
            return -1;
        }
        density = max_per_voxel(xf, xd, n, dx, ijk);
        if (++iter >= max_iterations || n == 1 || abs((int)density - (int)target) <= tolerance) break;
        const DensitySample cur = {dx, density};
        if (density < target) {
            if (sample_less(lo, cur)) lo = cur;
        } else if (hi.density == 0 || sample_less(cur, hi)) {
            hi = cur;
        }
        if (hi.density)
            dx = (lo.dx * (double)(hi.density - target) + hi.dx * (double)(target - lo.density)) /
                 (double)(hi.density - lo.density);
        else if (density > 1u)
            dx *= (target - 1.0) / (density - 1.0);
        else
            dx *= 10.0;
    }
    free(ijk);
    *dx_out = dx;
    *density_out = density;
    return f64 ? or_points_to_grid_f64(xd, n, dx, point_type, name, out, bytes_out)
               : or_points_to_grid(xf, n, dx, point_type, name, out, bytes_out);
}
