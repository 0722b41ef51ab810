/* Oracle: CPU restatement of NanoVDB's read-only query path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py as the checker; never linked into, loaded by or
 * called from the product library (openvdb_amd/).
 *
 * The flat buffer is read by byte offsets.  Layout and algorithms follow the
 * reference files cited at each function (paths relative to the reference
 * root).  Every FP32 routine keeps the reference's operation order; the file is
 * compiled with -ffp-contract=off and fmaf appears only where the reference
 * writes fmaf/fma.
 *
 * Pinning (what checks each part; the tests under tests/ hold them):
 *   lookup/getValue/isActive/getDim, codecs  : compiled reference (oracle/_ref) bit-exact, closed-form sphere,
 *                                              RandomReadAccessorGet known answers
 *   trilinear / FD6 / GradStencil / gradient : compiled reference bit-exact, linear-field closed forms,
 *                                              sampleFromVoxels golden values of the reference Python tests
 *   Ray clip / HDDA / zeroCrossing           : reference unit-test known answers, compiled reference bit-exact
 *   fog march (FP64)                         : compiled-reference march over the reference HDDA/accessor/sampler
 *                                              (same contract), constant-density closed form
 *   fog-march adjoint (FP64)                 : central finite differences of or_fog_march; no reference
 *                                              implementation exists (parity vs the reference is unpinned by
 *                                              construction, the adjoint is pinned by calculus only)
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ layout
 * nanovdb/nanovdb/NanoVDB.h: GridData 1944-2135 (672 B), Map 1418-1587,
 * TreeData 2393-2423, RootData 2621-2840 (+32 B tiles), InternalData 3181-3250,
 * LeafData 3671-3973.  Byte offsets as pinned by unittest/TestNanoVDB.cc:4938+. */
enum {
    GRID_MAGIC = 0, GRID_VERSION = 16, GRID_MAP = 296, GRID_VOXELSIZE = 608, GRID_CLASS = 632, GRID_TYPE = 636,
    GRID_SIZE = 672,
    MAP_INVMATF = 36, MAP_VECF = 72, MAP_INVMATD = 160, MAP_VECD = 232,
    TREE_NODEOFF = 0, TREE_NODECNT = 32,
    ROOT_BBOX = 0, ROOT_TABLESIZE = 24, ROOT_BACKGROUND = 28, ROOT_SIZE = 64, TILE_SIZE = 32,
    UP_FLAGS = 24, UP_VMASK = 32, UP_CMASK = 4128, UP_TABLE = 8256,
    LO_FLAGS = 24, LO_VMASK = 32, LO_CMASK = 544, LO_TABLE = 1088,
    LEAF_FLAGS = 15, LEAF_VMASK = 16, LEAF_MIN = 80, LEAF_QUANTUM = 84, LEAF_DATA = 96
};
enum { GT_FLOAT = 1, GT_FP4 = 13, GT_FP8 = 14, GT_FP16 = 15, GT_FPN = 16, GT_ONINDEX = 20 };
/* ValueOnIndex grids (NanoVDB.h:4147-4171): RootData is 96 B with a uint64 background index and
 * tiles {key, child, state, uint64 index}; node tables hold uint64 indices; leaves hold an offset
 * (80) and a 9-bit-per-word prefix sum (88). */
enum { IDX_ROOT_SIZE = 96, IDX_ROOT_BACKGROUND = 32, IDX_TILE_VALUE = 24, IDX_LEAF_OFFSET = 80, IDX_LEAF_PREFIX = 88,
       GRID_BLIND_OFFSET = 640, GRID_BLIND_COUNT = 648, BLIND_DATATYPE = 28 };
static const uint64_t MAGIC_GRID = 0x314244566f6e614eULL; /* "NanoVDB1", NanoVDB.h:140 */

static inline uint64_t rd_u64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }
static inline int64_t rd_i64(const uint8_t* p) { int64_t v; memcpy(&v, p, 8); return v; }
static inline uint32_t rd_u32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static inline int32_t rd_i32(const uint8_t* p) { int32_t v; memcpy(&v, p, 4); return v; }
static inline uint16_t rd_u16(const uint8_t* p) { uint16_t v; memcpy(&v, p, 2); return v; }
static inline float rd_f32(const uint8_t* p) { float v; memcpy(&v, p, 4); return v; }
static inline double rd_f64(const uint8_t* p) { double v; memcpy(&v, p, 8); return v; }

typedef struct {
    const uint8_t* buf;
    const uint8_t* tree;
    const uint8_t* root;
    const uint8_t* leaf0;
    uint32_t type, leafCount, tableSize, leafSize, rootSize;
    float background;
    int32_t bbox[6];
    const float* channel; /* index grids: values by index (blind channel 0) */
} G;

static uint32_t leaf_size_of(uint32_t type)
{
    switch (type) {
    case GT_FLOAT: return 2144u;
    case GT_FP4: return 352u;
    case GT_FP8: return 608u;
    case GT_FP16: return 1120u;
    case GT_ONINDEX: return 96u;
    default: return 0u; /* FpN: variable */
    }
}

static G open_grid(const void* p)
{
    G g;
    g.buf = (const uint8_t*)p;
    g.tree = g.buf + GRID_SIZE;
    g.type = rd_u32(g.buf + GRID_TYPE);
    g.root = g.tree + rd_i64(g.tree + TREE_NODEOFF + 8 * 3);
    g.leaf0 = g.tree + rd_i64(g.tree + TREE_NODEOFF);
    g.leafCount = rd_u32(g.tree + TREE_NODECNT);
    g.tableSize = rd_u32(g.root + ROOT_TABLESIZE);
    g.leafSize = leaf_size_of(g.type);
    g.channel = NULL;
    g.rootSize = ROOT_SIZE;
    if (g.type == GT_ONINDEX) {
        /* ChannelAccessor(grid, 0): blind data #0 when it holds floats (NanoVDB.h:5737-5790,
         * GridBlindMetaData :1591-1600, offsets relative to the metadata entry) */
        if (rd_u32(g.buf + GRID_BLIND_COUNT) > 0) {
            const uint8_t* m = g.buf + rd_i64(g.buf + GRID_BLIND_OFFSET);
            if (rd_u32(m + BLIND_DATATYPE) == GT_FLOAT) g.channel = (const float*)(m + rd_i64(m));
        }
        g.rootSize = IDX_ROOT_SIZE;
        g.background = g.channel ? g.channel[rd_u64(g.root + IDX_ROOT_BACKGROUND)] : 0.0f;
    } else {
        g.background = rd_f32(g.root + ROOT_BACKGROUND);
    }
    for (int i = 0; i < 6; ++i) g.bbox[i] = rd_i32(g.root + ROOT_BBOX + 4 * i);
    return g;
}

int or_validate(const void* buf, uint64_t bytes)
{
    if (!buf || bytes < GRID_SIZE + 64) return -1;
    const uint8_t* b = (const uint8_t*)buf;
    if ((rd_u64(b + GRID_MAGIC) & 0x00FFFFFFFFFFFFFFULL) != (MAGIC_GRID & 0x00FFFFFFFFFFFFFFULL)) return -2;
    if ((rd_u32(b + GRID_VERSION) >> 21) != 32u) return -3; /* Version: major<<21 | minor<<10 | patch */
    const uint32_t t = rd_u32(b + GRID_TYPE);
    if (t != GT_FLOAT && t != GT_FP4 && t != GT_FP8 && t != GT_FP16 && t != GT_FPN && t != GT_ONINDEX) return -4;
    return 0;
}
uint32_t or_grid_type(const void* buf) { return rd_u32((const uint8_t*)buf + GRID_TYPE); }
uint32_t or_leaf_count(const void* buf) { return open_grid(buf).leafCount; }
/* index grids: GridData::mData1, the number of channel values */
uint64_t or_value_count(const void* buf) { return rd_u64((const uint8_t*)buf + 656); }

/* ------------------------------------------------------------------ leaf decode
 * float NanoVDB.h:3703; Fp4 :3839-3847; Fp8 :3875-3878; Fp16 :3906-3909;
 * FpN branchless :3937-3951.  value = float(code)*quantum + minimum (no fma). */
static uint32_t leaf_code(const G* g, const uint8_t* leaf, uint32_t n)
{
    const uint8_t* d = leaf + LEAF_DATA;
    switch (g->type) {
    case GT_FP4: return (d[n >> 1] >> ((n & 1) << 2)) & 15u;
    case GT_FP8: return d[n];
    case GT_FP16: return rd_u16(d + 2 * n);
    default: {
        const int b = leaf[LEAF_FLAGS] >> 5;
        uint32_t code = rd_u32(d + 4 * (n >> (5 - b)));
        code >>= (n & ((32u >> b) - 1u)) << b;
        return code & ((1u << (1u << b)) - 1u);
    }
    }
}

/* LeafData<ValueOnIndex>::getValue, NanoVDB.h:4160-4170: 0 (the background) for an inactive voxel,
 * else offset + active voxels before n (popcount of the word below n + 9-bit prefix of the words
 * before it) */
static uint64_t leaf_index(const uint8_t* leaf, uint32_t n)
{
    const uint32_t w = n >> 6;
    const uint64_t word = rd_u64(leaf + LEAF_VMASK + 8 * w), bit = (uint64_t)1 << (n & 63u);
    if (!(word & bit)) return 0;
    uint64_t sum = rd_u64(leaf + IDX_LEAF_OFFSET) + (uint64_t)__builtin_popcountll(word & (bit - 1u));
    if (w) sum += (rd_u64(leaf + IDX_LEAF_PREFIX) >> (9u * (w - 1u))) & 511u;
    return sum;
}

/* value of an internal-node table entry or a root tile: a float, or an index into the channel */
static float entry_value(const G* g, const uint8_t* entry)
{
    return g->type == GT_ONINDEX ? g->channel[rd_u64(entry)] : rd_f32(entry);
}

static float leaf_value(const G* g, const uint8_t* leaf, uint32_t n)
{
    if (g->type == GT_FLOAT) return rd_f32(leaf + LEAF_DATA + 4 * n);
    if (g->type == GT_ONINDEX) return g->channel[leaf_index(leaf, n)];
    return (float)leaf_code(g, leaf, n) * rd_f32(leaf + LEAF_QUANTUM) + rd_f32(leaf + LEAF_MIN);
}

/* same code, FP64 reconstruction code*quantum + minimum */
static double leaf_value_f64(const G* g, const uint8_t* leaf, uint32_t n)
{
    if (g->type == GT_FLOAT) return (double)rd_f32(leaf + LEAF_DATA + 4 * n);
    if (g->type == GT_ONINDEX) return (double)g->channel[leaf_index(leaf, n)];
    return (double)leaf_code(g, leaf, n) * (double)rd_f32(leaf + LEAF_QUANTUM) + (double)rd_f32(leaf + LEAF_MIN);
}

static inline int mask_on(const uint8_t* mask, uint32_t n) { return (int)((rd_u64(mask + 8 * (n >> 6)) >> (n & 63)) & 1u); }

/* ------------------------------------------------------------------ traversal
 * RootData::CoordToKey (NANOVDB_USE_SINGLE_ROOT_KEY) NanoVDB.h:2633-2640; linear probe :2785-2799;
 * InternalNode::CoordToOffset :3560-3565; LeafNode::CoordToOffset :4516-4519;
 * getDimAndCache root :3126-3137, internal :3619-3633, leaf :4565-4573;
 * GetValue/GetState :6196-6236. */
typedef struct {
    float value;
    int active;
    uint32_t dim;
    const uint8_t* leaf; /* leaf node holding the voxel, or NULL */
    uint32_t n;          /* voxel offset in that leaf */
    uint64_t index;      /* index grids: the channel index the value comes from */
    const uint8_t* lower; /* lower / upper node the lookup passed through, or NULL */
    const uint8_t* upper;
} Look;

static Look lookup(const G* g, int32_t x, int32_t y, int32_t z)
{
    Look r;
    r.leaf = NULL;
    r.n = 0;
    r.index = 0;
    r.lower = NULL;
    r.upper = NULL;
    const int idx = g->type == GT_ONINDEX;
    const uint64_t key = ((uint64_t)((uint32_t)z >> 12)) | ((uint64_t)((uint32_t)y >> 12) << 21) |
                         ((uint64_t)((uint32_t)x >> 12) << 42);
    const uint8_t* tile = NULL;
    for (uint32_t i = 0; i < g->tableSize; ++i) {
        const uint8_t* t = g->root + g->rootSize + (uint64_t)TILE_SIZE * i;
        if (rd_u64(t) == key) { tile = t; break; }
    }
    if (!tile) {
        r.value = g->background;
        r.active = 0;
        r.dim = 4096u;
        if (idx) r.index = rd_u64(g->root + IDX_ROOT_BACKGROUND);
        return r;
    }
    const int64_t child = rd_i64(tile + 8);
    if (child == 0) {
        r.value = entry_value(g, tile + (idx ? IDX_TILE_VALUE : 20));
        if (idx) r.index = rd_u64(tile + IDX_TILE_VALUE);
        r.active = rd_u32(tile + 16) > 0u;
        r.dim = 4096u;
        return r;
    }
    const uint8_t* up = g->root + child;
    r.upper = up;
    const int upSkip = (int)(rd_u64(up + UP_FLAGS) & 1u);
    const uint32_t nu = ((((uint32_t)x & 4095u) >> 7) << 10) | ((((uint32_t)y & 4095u) >> 7) << 5) | (((uint32_t)z & 4095u) >> 7);
    if (!mask_on(up + UP_CMASK, nu)) {
        r.value = entry_value(g, up + UP_TABLE + 8 * nu);
        if (idx) r.index = rd_u64(up + UP_TABLE + 8 * nu);
        r.active = mask_on(up + UP_VMASK, nu);
        r.dim = upSkip ? 4096u : 128u;
        return r;
    }
    const uint8_t* lo = up + rd_i64(up + UP_TABLE + 8 * nu);
    r.lower = lo;
    const int loSkip = (int)(rd_u64(lo + LO_FLAGS) & 1u);
    const uint32_t nl = ((((uint32_t)x & 127u) >> 3) << 8) | ((((uint32_t)y & 127u) >> 3) << 4) | (((uint32_t)z & 127u) >> 3);
    if (!mask_on(lo + LO_CMASK, nl)) {
        r.value = entry_value(g, lo + LO_TABLE + 8 * nl);
        if (idx) r.index = rd_u64(lo + LO_TABLE + 8 * nl);
        r.active = mask_on(lo + LO_VMASK, nl);
        r.dim = upSkip ? 4096u : (loSkip ? 128u : 8u);
        return r;
    }
    const uint8_t* lf = lo + rd_i64(lo + LO_TABLE + 8 * nl);
    const uint32_t n = (((uint32_t)x & 7u) << 6) | (((uint32_t)y & 7u) << 3) | ((uint32_t)z & 7u);
    r.leaf = lf;
    r.n = n;
    r.value = leaf_value(g, lf, n);
    if (idx) r.index = leaf_index(lf, n);
    r.active = mask_on(lf + LEAF_VMASK, n);
    r.dim = upSkip ? 4096u : (loSkip ? 128u : ((lf[LEAF_FLAGS] & 1u) ? 8u : 1u));
    return r;
}

static inline float get_value(const G* g, int32_t x, int32_t y, int32_t z) { return lookup(g, x, y, z).value; }

void or_get_value(const void* buf, const int32_t* ijk, uint64_t n, float* val, uint8_t* active, uint32_t* dim)
{
    const G g = open_grid(buf);
    for (uint64_t i = 0; i < n; ++i) {
        const Look l = lookup(&g, ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
        if (val) val[i] = l.value;
        if (active) active[i] = (uint8_t)l.active;
        if (dim) dim[i] = l.dim;
    }
}

/* ------------------------------------------------------------------ samplers
 * Floor<Coord>(Vec3f&) math/SampleFromVoxels.h:61-69; Round (rintf) math/Math.h:255-261;
 * fused non-cached trilinear :423-464 (lerp a + w*(b-a)); TrilinearSampler::gradient :232-271. */
static inline float lerpf(float a, float b, float w) { return a + w * (b - a); }

static void stencil8(const G* g, int32_t i, int32_t j, int32_t k, float v[2][2][2])
{
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) v[a][b][c] = get_value(g, i + a, j + b, k + c);
}

static float trilinear(const G* g, float x, float y, float z)
{
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const float u = x - fx, v = y - fy, w = z - fz;
    float s[2][2][2];
    stencil8(g, (int32_t)fx, (int32_t)fy, (int32_t)fz, s);
    return lerpf(lerpf(lerpf(s[0][0][0], s[0][0][1], w), lerpf(s[0][1][0], s[0][1][1], w), v),
                 lerpf(lerpf(s[1][0][0], s[1][0][1], w), lerpf(s[1][1][0], s[1][1][1], w), v), u);
}

static void trilinear_grad(const G* g, float x, float y, float z, float out[3])
{
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const float u = x - fx, v = y - fy, w = z - fz;
    float s[2][2][2];
    stencil8(g, (int32_t)fx, (int32_t)fy, (int32_t)fz, s);
    float D[4] = {s[0][0][1] - s[0][0][0], s[0][1][1] - s[0][1][0], s[1][0][1] - s[1][0][0], s[1][1][1] - s[1][1][0]};
    out[2] = lerpf(lerpf(D[0], D[1], v), lerpf(D[2], D[3], v), u);
    D[0] = s[0][0][0] + D[0] * w;
    D[1] = s[0][1][0] + D[1] * w;
    D[2] = s[1][0][0] + D[2] * w;
    D[3] = s[1][1][0] + D[3] * w;
    out[0] = lerpf(D[2], D[3], v) - lerpf(D[0], D[1], v);
    out[1] = lerpf(D[1] - D[0], D[3] - D[2], u);
}

/* Grid::worldToIndex -> Map::applyInverseMap with Vec3f: translation subtracted in FP64 and
 * rounded to float, matMult(const double*, Vec3f) with fma, result rounded to float
 * (NanoVDB.h:1507-1510, math/Math.h:1718-1723). */
static void world_to_index_f64map(const G* g, const float in[3], float out[3])
{
    const uint8_t* m = g->buf + GRID_MAP;
    float t[3];
    for (int a = 0; a < 3; ++a) t[a] = (float)((double)in[a] - rd_f64(m + MAP_VECD + 8 * a));
    for (int r = 0; r < 3; ++r) {
        const double m0 = rd_f64(m + MAP_INVMATD + 8 * (3 * r)), m1 = rd_f64(m + MAP_INVMATD + 8 * (3 * r + 1)),
                     m2 = rd_f64(m + MAP_INVMATD + 8 * (3 * r + 2));
        out[r] = (float)fma((double)t[0], m0, fma((double)t[1], m1, (double)t[2] * m2));
    }
}

typedef struct { float inv2dx, invdx2; } GradC;
static GradC grad_consts(const G* g)
{
    GradC c; /* GradStencil ctor, math/Stencils.h:477-481 */
    c.inv2dx = (float)(0.5 / rd_f64(g->buf + GRID_VOXELSIZE));
    c.invdx2 = (float)(4.0 * (double)c.inv2dx * (double)c.inv2dx);
    return c;
}

void or_sample(const void* buf, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
               int grad_mode)
{
    const G g = open_grid(buf);
    /* FD6 scale (BuildT).5/voxelSize() computed in FP64 then rounded, PySampleFromVoxels.cu:42 */
    float inv2dx[3];
    for (int a = 0; a < 3; ++a) inv2dx[a] = (float)(0.5 / rd_f64(g.buf + GRID_VOXELSIZE + 8 * a));
    const GradC gc = grad_consts(&g);
    for (uint64_t i = 0; i < n; ++i) {
        float p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        if (world) {
            float q[3];
            world_to_index_f64map(&g, p, q);
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        }
        if (nn) nn[i] = get_value(&g, (int32_t)rintf(p[0]), (int32_t)rintf(p[1]), (int32_t)rintf(p[2]));
        if (val) val[i] = trilinear(&g, p[0], p[1], p[2]);
        if (grad) {
            float o[3] = {0.0f, 0.0f, 0.0f};
            if (grad_mode == 1) {
                const float dx = trilinear(&g, p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f) -
                                 trilinear(&g, p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f);
                const float dy = trilinear(&g, p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f) -
                                 trilinear(&g, p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f);
                const float dz = trilinear(&g, p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f) -
                                 trilinear(&g, p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f);
                o[0] = dx * inv2dx[0];
                o[1] = dy * inv2dx[1];
                o[2] = dz * inv2dx[2];
            } else if (grad_mode == 2) {
                /* GradStencil::moveTo(RoundDown(xyz)) + gradient(), math/Stencils.h:148-152, 502-507 */
                const int32_t a = (int32_t)floorf(p[0]), b = (int32_t)floorf(p[1]), c = (int32_t)floorf(p[2]);
                o[0] = gc.inv2dx * (get_value(&g, a + 1, b, c) - get_value(&g, a - 1, b, c));
                o[1] = gc.inv2dx * (get_value(&g, a, b + 1, c) - get_value(&g, a, b - 1, c));
                o[2] = gc.inv2dx * (get_value(&g, a, b, c + 1) - get_value(&g, a, b, c - 1));
            } else if (grad_mode == 3) {
                trilinear_grad(&g, p[0], p[1], p[2], o);
            }
            grad[3 * i] = o[0];
            grad[3 * i + 1] = o[1];
            grad[3 * i + 2] = o[2];
        }
    }
}

static double trilinear_f64(const G* g, float x, float y, float z)
{
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const double u = (double)(x - fx), v = (double)(y - fy), w = (double)(z - fz);
    const int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    double s[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) {
                const Look l = lookup(g, i + a, j + b, k + c);
                s[a][b][c] = l.leaf ? leaf_value_f64(g, l.leaf, l.n) : (double)l.value;
            }
#define LERPD(a, b, t) ((a) + (t) * ((b) - (a)))
    return LERPD(LERPD(LERPD(s[0][0][0], s[0][0][1], w), LERPD(s[0][1][0], s[0][1][1], w), v),
                 LERPD(LERPD(s[1][0][0], s[1][0][1], w), LERPD(s[1][1][0], s[1][1][1], w), v), u);
}

void or_sample_f64(const void* buf, const float* xyz, uint64_t n, double* val, double* grad_fd6)
{
    const G g = open_grid(buf);
    double inv2dx[3];
    for (int a = 0; a < 3; ++a) inv2dx[a] = 0.5 / rd_f64(g.buf + GRID_VOXELSIZE + 8 * a);
    for (uint64_t i = 0; i < n; ++i) {
        const float* p = xyz + 3 * i;
        if (val) val[i] = trilinear_f64(&g, p[0], p[1], p[2]);
        if (grad_fd6) {
            grad_fd6[3 * i] = (trilinear_f64(&g, p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f) -
                               trilinear_f64(&g, p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f)) * inv2dx[0];
            grad_fd6[3 * i + 1] = (trilinear_f64(&g, p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f) -
                                   trilinear_f64(&g, p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f)) * inv2dx[1];
            grad_fd6[3 * i + 2] = (trilinear_f64(&g, p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f) -
                                   trilinear_f64(&g, p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f)) * inv2dx[2];
        }
    }
}

/* ------------------------------------------------------------------ adjoint of the point queries
 * The reference sampler has no backward: parity against the reference is unpinned by
 * construction.  Two restatements of dLoss/dvalue of or_sample's outputs:
 *  - or_sample_bwd_fixed: the FP32 / int64 fixed-point contract the device adjoint implements
 *    (per query and voxel, the FP32 sum in term order of ((s*A)*B)*C, converted once with
 *    round-to-nearest-even and saturation at +-2^62, integer-summed with wrap-around);
 *  - or_sample_bwd_f64: the exact derivative in FP64 (weights from FP64 fractions, FD6 cells at
 *    floor(p) +- 1), pinned by central differences of or_sample_loss_f64, an FP64 forward of the
 *    same outputs with one value perturbed.
 * Keys: Float grids leafIndex*512 + n for voxels in leaves (others receive nothing), index grids
 * the channel index every lookup reads.  Forward semantics: PySampleFromVoxels.cu:42-47 (FD6),
 * math/SampleFromVoxels.h:149-156 (nearest), :232-271 (trilinear gradient), :423-476
 * (trilinear), math/Stencils.h:502-507 (GradStencil). */
typedef struct {
    int32_t o[3];
    float s;
    float A[2], B[2], C[2];
    int single;
} ATerm;

static int64_t adj_key(const G* g, const Look* l)
{
    if (g->type == GT_ONINDEX) return (int64_t)l->index;
    if (!l->leaf) return -1;
    return (int64_t)((uint64_t)(l->leaf - g->leaf0) / g->leafSize) * 512 + l->n;
}

static void adj_cell(ATerm* t, int32_t x, int32_t y, int32_t z, float s, float a0, float a1, float b0, float b1,
                     float c0, float c1)
{
    t->o[0] = x; t->o[1] = y; t->o[2] = z;
    t->s = s;
    t->A[0] = a0; t->A[1] = a1; t->B[0] = b0; t->B[1] = b1; t->C[0] = c0; t->C[1] = c1;
    t->single = 0;
}

static void adj_single(ATerm* t, int32_t x, int32_t y, int32_t z, float s)
{
    t->o[0] = x; t->o[1] = y; t->o[2] = z;
    t->s = s;
    t->single = 1;
}

static int adj_terms(const float p[3], float gn, float gv, const float gg[3], int mode, const float fd[3],
                     float gs_inv2dx, ATerm T[8])
{
    int nt = 0;
    const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
    const int32_t i0 = (int32_t)fx, j0 = (int32_t)fy, k0 = (int32_t)fz;
    const float u = p[0] - fx, v = p[1] - fy, w = p[2] - fz;
    if (gv != 0.0f) adj_cell(&T[nt++], i0, j0, k0, gv, 1.0f - u, u, 1.0f - v, v, 1.0f - w, w);
    if (mode == 1) {
        for (int t = 0; t < 6; ++t) {
            const int ax = t >> 1, plus = (t & 1) == 0;
            float q[3];
            for (int b = 0; b < 3; ++b) {
                const float d = b == ax ? 1.0f : 0.0f;
                q[b] = plus ? p[b] + d : p[b] - d;
            }
            const float s0 = gg[ax] * fd[ax];
            const float s = plus ? s0 : -s0;
            if (s == 0.0f) continue;
            const float gx = floorf(q[0]), gy = floorf(q[1]), gz = floorf(q[2]);
            const float a = q[0] - gx, b = q[1] - gy, c = q[2] - gz;
            adj_cell(&T[nt++], (int32_t)gx, (int32_t)gy, (int32_t)gz, s, 1.0f - a, a, 1.0f - b, b, 1.0f - c, c);
        }
    } else if (mode == 2) {
        for (int ax = 0; ax < 3; ++ax) {
            const float s = gs_inv2dx * gg[ax];
            if (s == 0.0f) continue;
            const int32_t dx = ax == 0, dy = ax == 1, dz = ax == 2;
            adj_single(&T[nt++], i0 + dx, j0 + dy, k0 + dz, s);
            adj_single(&T[nt++], i0 - dx, j0 - dy, k0 - dz, -s);
        }
    } else if (mode == 3) {
        if (gg[0] != 0.0f) adj_cell(&T[nt++], i0, j0, k0, gg[0], -1.0f, 1.0f, 1.0f - v, v, 1.0f - w, w);
        if (gg[1] != 0.0f) adj_cell(&T[nt++], i0, j0, k0, gg[1], 1.0f - u, u, -1.0f, 1.0f, 1.0f - w, w);
        if (gg[2] != 0.0f) adj_cell(&T[nt++], i0, j0, k0, gg[2], 1.0f - u, u, 1.0f - v, v, -1.0f, 1.0f);
    }
    if (gn != 0.0f) adj_single(&T[nt++], (int32_t)rintf(p[0]), (int32_t)rintf(p[1]), (int32_t)rintf(p[2]), gn);
    return nt;
}

static inline int adj_corner_is(const ATerm* t, int c, int32_t x, int32_t y, int32_t z)
{
    return t->o[0] + (c >> 2) == x && t->o[1] + ((c >> 1) & 1) == y && t->o[2] + (c & 1) == z;
}

void or_sample_bwd_fixed(const void* buf, const float* xyz, uint64_t n, int world, const float* gnn,
                         const float* gval, const float* ggrad, int grad_mode, int shift, int64_t* accum,
                         int* overflow)
{
    const G g = open_grid(buf);
    float fd[3];
    for (int a = 0; a < 3; ++a) fd[a] = (float)(0.5 / rd_f64(g.buf + GRID_VOXELSIZE + 8 * a));
    const float gs = grad_consts(&g).inv2dx;
    const float scale = ldexpf(1.0f, shift);
    const float lim = 4611686018427387904.0f; /* 2^62 */
    int over = 0;
    for (uint64_t i = 0; i < n; ++i) {
        float p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        if (world) {
            float q[3];
            world_to_index_f64map(&g, p, q);
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        }
        const float gg[3] = {ggrad ? ggrad[3 * i] : 0.0f, ggrad ? ggrad[3 * i + 1] : 0.0f,
                             ggrad ? ggrad[3 * i + 2] : 0.0f};
        ATerm T[8];
        const int nt = adj_terms(p, gnn ? gnn[i] : 0.0f, gval ? gval[i] : 0.0f, gg, ggrad ? grad_mode : 0, fd, gs, T);
        for (int t = 0; t < nt; ++t) {
            const int nc = T[t].single ? 1 : 8;
            for (int c = 0; c < nc; ++c) {
                const int32_t X = T[t].o[0] + (c >> 2), Y = T[t].o[1] + ((c >> 1) & 1), Z = T[t].o[2] + (c & 1);
                int seen = 0;
                for (int t2 = 0; t2 < t && !seen; ++t2)
                    for (int c2 = 0; c2 < (T[t2].single ? 1 : 8); ++c2)
                        if (adj_corner_is(&T[t2], c2, X, Y, Z)) { seen = 1; break; }
                if (seen) continue;
                float sum = 0.0f;
                for (int t2 = t; t2 < nt; ++t2)
                    for (int c2 = 0; c2 < (T[t2].single ? 1 : 8); ++c2)
                        if (adj_corner_is(&T[t2], c2, X, Y, Z))
                            sum += T[t2].single ? T[t2].s
                                                : ((T[t2].s * T[t2].A[c2 >> 2]) * T[t2].B[(c2 >> 1) & 1]) * T[t2].C[c2 & 1];
                const float x = sum * scale;
                long long q;
                if (fabsf(x) < lim) q = llrintf(x);
                else {
                    over = 1;
                    q = (x != x) ? 0 : (x > 0.0f ? (1ll << 62) : -(1ll << 62));
                }
                if (q == 0) continue;
                const Look l = lookup(&g, X, Y, Z);
                const int64_t key = adj_key(&g, &l);
                if (key >= 0) accum[key] = (int64_t)((uint64_t)accum[key] + (uint64_t)q);
            }
        }
    }
    if (overflow) *overflow = over;
}

/* FP64: the value a lookup returns, plus delta where its key is `key` */
static double adj_value_f64(const G* g, int32_t x, int32_t y, int32_t z, int64_t key, double delta)
{
    const Look l = lookup(g, x, y, z);
    double v = l.leaf ? leaf_value_f64(g, l.leaf, l.n) : (double)l.value;
    if (key >= 0 && adj_key(g, &l) == key) v += delta;
    return v;
}

static double trilinear_f64d(const G* g, double x, double y, double z, int64_t key, double delta, double grad[3])
{
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const double u = x - fx, v = y - fy, w = z - fz;
    const int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    double s[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) s[a][b][c] = adj_value_f64(g, i + a, j + b, k + c, key, delta);
    if (grad) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    const double wa = a ? u : 1.0 - u, wb = b ? v : 1.0 - v, wc = c ? w : 1.0 - w;
                    gx += (a ? 1.0 : -1.0) * wb * wc * s[a][b][c];
                    gy += wa * (b ? 1.0 : -1.0) * wc * s[a][b][c];
                    gz += wa * wb * (c ? 1.0 : -1.0) * s[a][b][c];
                }
        grad[0] = gx; grad[1] = gy; grad[2] = gz;
    }
    double r = 0.0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) r += (a ? u : 1.0 - u) * (b ? v : 1.0 - v) * (c ? w : 1.0 - w) * s[a][b][c];
    return r;
}

/* Loss = sum_i gnn_i nn_i + gval_i val_i + ggrad_i . grad_i with every output in FP64 and the
 * value at `key` raised by delta (key -1: none).  Index-space points (no map). */
double or_sample_loss_f64(const void* buf, const float* xyz, uint64_t n, const double* gnn, const double* gval,
                          const double* ggrad, int grad_mode, int64_t key, double delta)
{
    const G g = open_grid(buf);
    double fd[3];
    for (int a = 0; a < 3; ++a) fd[a] = 0.5 / rd_f64(g.buf + GRID_VOXELSIZE + 8 * a);
    const double gs = 0.5 / rd_f64(g.buf + GRID_VOXELSIZE);
    double L = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        if (gnn) L += gnn[i] * adj_value_f64(&g, (int32_t)rint(p[0]), (int32_t)rint(p[1]), (int32_t)rint(p[2]), key, delta);
        if (gval) L += gval[i] * trilinear_f64d(&g, p[0], p[1], p[2], key, delta, NULL);
        if (!ggrad) continue;
        double o[3] = {0.0, 0.0, 0.0};
        if (grad_mode == 1) {
            for (int ax = 0; ax < 3; ++ax) {
                double qp[3] = {p[0], p[1], p[2]}, qm[3] = {p[0], p[1], p[2]};
                qp[ax] += 1.0;
                qm[ax] -= 1.0;
                o[ax] = (trilinear_f64d(&g, qp[0], qp[1], qp[2], key, delta, NULL) -
                         trilinear_f64d(&g, qm[0], qm[1], qm[2], key, delta, NULL)) * fd[ax];
            }
        } else if (grad_mode == 2) {
            const int32_t a = (int32_t)floor(p[0]), b = (int32_t)floor(p[1]), c = (int32_t)floor(p[2]);
            o[0] = gs * (adj_value_f64(&g, a + 1, b, c, key, delta) - adj_value_f64(&g, a - 1, b, c, key, delta));
            o[1] = gs * (adj_value_f64(&g, a, b + 1, c, key, delta) - adj_value_f64(&g, a, b - 1, c, key, delta));
            o[2] = gs * (adj_value_f64(&g, a, b, c + 1, key, delta) - adj_value_f64(&g, a, b, c - 1, key, delta));
        } else if (grad_mode == 3) {
            trilinear_f64d(&g, p[0], p[1], p[2], key, delta, o);
        }
        L += ggrad[3 * i] * o[0] + ggrad[3 * i + 1] * o[1] + ggrad[3 * i + 2] * o[2];
    }
    return L;
}

/* Exact FP64 derivative of or_sample_loss_f64 in every value, added into grad (keys as above). */
void or_sample_bwd_f64(const void* buf, const float* xyz, uint64_t n, const double* gnn, const double* gval,
                       const double* ggrad, int grad_mode, double* grad)
{
    const G g = open_grid(buf);
    double fd[3];
    for (int a = 0; a < 3; ++a) fd[a] = 0.5 / rd_f64(g.buf + GRID_VOXELSIZE + 8 * a);
    const double gs = 0.5 / rd_f64(g.buf + GRID_VOXELSIZE);
#define ADJ_ADD(X, Y, Z, W)                                                                                            \
    do {                                                                                                               \
        const Look l_ = lookup(&g, (X), (Y), (Z));                                                                     \
        const int64_t k_ = adj_key(&g, &l_);                                                                           \
        if (k_ >= 0) grad[k_] += (W);                                                                                  \
    } while (0)
    for (uint64_t i = 0; i < n; ++i) {
        const double p[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        const double f[3] = {floor(p[0]), floor(p[1]), floor(p[2])};
        const int32_t c0[3] = {(int32_t)f[0], (int32_t)f[1], (int32_t)f[2]};
        const double u[3] = {p[0] - f[0], p[1] - f[1], p[2] - f[2]};
        if (gnn && gnn[i] != 0.0) ADJ_ADD((int32_t)rint(p[0]), (int32_t)rint(p[1]), (int32_t)rint(p[2]), gnn[i]);
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    const double wa = a ? u[0] : 1.0 - u[0], wb = b ? u[1] : 1.0 - u[1], wc = c ? u[2] : 1.0 - u[2];
                    const int32_t X = c0[0] + a, Y = c0[1] + b, Z = c0[2] + c;
                    if (gval && gval[i] != 0.0) ADJ_ADD(X, Y, Z, gval[i] * wa * wb * wc);
                    if (!ggrad) continue;
                    const double* d = ggrad + 3 * i;
                    if (grad_mode == 3)
                        ADJ_ADD(X, Y, Z, d[0] * (a ? 1.0 : -1.0) * wb * wc + d[1] * wa * (b ? 1.0 : -1.0) * wc +
                                              d[2] * wa * wb * (c ? 1.0 : -1.0));
                    if (grad_mode == 1)
                        /* sample at p +- e_ax: cell c0 +- e_ax, same fractions */
                        for (int ax = 0; ax < 3; ++ax) {
                            const double wgt = d[ax] * fd[ax] * wa * wb * wc;
                            ADJ_ADD(X + (ax == 0), Y + (ax == 1), Z + (ax == 2), wgt);
                            ADJ_ADD(X - (ax == 0), Y - (ax == 1), Z - (ax == 2), -wgt);
                        }
                }
        if (ggrad && grad_mode == 2)
            for (int ax = 0; ax < 3; ++ax) {
                const double s = gs * ggrad[3 * i + ax];
                ADJ_ADD(c0[0] + (ax == 0), c0[1] + (ax == 1), c0[2] + (ax == 2), s);
                ADJ_ADD(c0[0] - (ax == 0), c0[1] - (ax == 1), c0[2] - (ax == 2), -s);
            }
    }
#undef ADJ_ADD
}

/* GradStencil, math/Stencils.h:455-576; GodunovsNormSqrd :62-80 */
static inline float sqf(float x) { return x * x; }
void or_grad_stencil(const void* buf, const int32_t* ijk, uint64_t n, float* grad, float* normsq, float* lap,
                     uint8_t* zc)
{
    const G g = open_grid(buf);
    const GradC gc = grad_consts(&g);
    for (uint64_t i = 0; i < n; ++i) {
        const int32_t a = ijk[3 * i], b = ijk[3 * i + 1], c = ijk[3 * i + 2];
        float v[7];
        v[0] = get_value(&g, a, b, c);
        v[1] = get_value(&g, a - 1, b, c);
        v[2] = get_value(&g, a + 1, b, c);
        v[3] = get_value(&g, a, b - 1, c);
        v[4] = get_value(&g, a, b + 1, c);
        v[5] = get_value(&g, a, b, c - 1);
        v[6] = get_value(&g, a, b, c + 1);
        if (grad) {
            grad[3 * i] = gc.inv2dx * (v[2] - v[1]);
            grad[3 * i + 1] = gc.inv2dx * (v[4] - v[3]);
            grad[3 * i + 2] = gc.inv2dx * (v[6] - v[5]);
        }
        if (normsq) {
            const float xm = v[0] - v[1], xp = v[2] - v[0], ym = v[0] - v[3], yp = v[4] - v[0], zm = v[0] - v[5],
                        zp = v[6] - v[0];
            float d;
            if (v[0] > 0.0f) {
                d = fmaxf(sqf(fmaxf(xm, 0.0f)), sqf(fminf(xp, 0.0f)));
                d += fmaxf(sqf(fmaxf(ym, 0.0f)), sqf(fminf(yp, 0.0f)));
                d += fmaxf(sqf(fmaxf(zm, 0.0f)), sqf(fminf(zp, 0.0f)));
            } else {
                d = fmaxf(sqf(fminf(xm, 0.0f)), sqf(fmaxf(xp, 0.0f)));
                d += fmaxf(sqf(fminf(ym, 0.0f)), sqf(fmaxf(yp, 0.0f)));
                d += fmaxf(sqf(fminf(zm, 0.0f)), sqf(fmaxf(zp, 0.0f)));
            }
            normsq[i] = gc.invdx2 * d;
        }
        if (lap) lap[i] = gc.invdx2 * (v[1] + v[2] + v[3] + v[4] + v[5] + v[6] - 6.0f * v[0]);
        if (zc)
            zc[i] = (uint8_t)(v[0] > 0.0f ? (v[1] < 0 || v[2] < 0 || v[3] < 0 || v[4] < 0 || v[5] < 0 || v[6] < 0)
                                          : (v[1] > 0 || v[2] > 0 || v[3] > 0 || v[4] > 0 || v[5] > 0 || v[6] > 0));
    }
}

/* ------------------------------------------------------------------ Ray
 * math/Ray.h: ctor :64-74 (invDir = 1/dir), operator() via fmaf :154-163,
 * worldToIndexF :258-268, intersects(CoordBBox) :398-424 (max padded by +1), clip :506-515.
 * Maximum<float> is FLT_MAX on the host (math/Math.h:118-124). */
typedef struct { float eye[3], dir[3], inv[3], t0, t1; } Ray;

static Ray make_ray(const float e[3], const float d[3], float t0, float t1)
{
    Ray r;
    for (int a = 0; a < 3; ++a) {
        r.eye[a] = e[a];
        r.dir[a] = d[a];
        r.inv[a] = 1.0f / d[a];
    }
    r.t0 = t0;
    r.t1 = t1;
    return r;
}

static inline void ray_at(const Ray* r, float t, float p[3])
{
    p[0] = fmaf(t, r->dir[0], r->eye[0]);
    p[1] = fmaf(t, r->dir[1], r->eye[1]);
    p[2] = fmaf(t, r->dir[2], r->eye[2]);
}

static Ray world_to_index_ray(const G* g, const float e[3], const float d[3])
{
    /* Map::applyInverseMapF / applyInverseJacobianF: matMult(const float*, Vec3f) with fmaf,
     * NanoVDB.h:1518-1539, math/Math.h:1704-1709 */
    const uint8_t* m = g->buf + GRID_MAP;
    float M[9], V[3];
    for (int a = 0; a < 9; ++a) M[a] = rd_f32(m + MAP_INVMATF + 4 * a);
    for (int a = 0; a < 3; ++a) V[a] = rd_f32(m + MAP_VECF + 4 * a);
    float t[3] = {e[0] - V[0], e[1] - V[1], e[2] - V[2]}, eye[3], dir[3];
    for (int r = 0; r < 3; ++r) {
        eye[r] = fmaf(t[0], M[3 * r], fmaf(t[1], M[3 * r + 1], t[2] * M[3 * r + 2]));
        dir[r] = fmaf(d[0], M[3 * r], fmaf(d[1], M[3 * r + 1], d[2] * M[3 * r + 2]));
    }
    const float length = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const float invLength = 1.0f / length;
    const float nd[3] = {invLength * dir[0], invLength * dir[1], invLength * dir[2]};
    const float t0 = 1.0e-5f, t1 = FLT_MAX; /* Delta<float>, Maximum<float> (host) */
    return make_ray(eye, nd, length * t0, t1);
}

/* Ray::applyInverseMap(map), math/Ray.h:223-230: eye through the FP64 Map::applyInverseMap
 * (translation subtracted in FP64 and rounded to float, then the double fma chain rounded to
 * float, NanoVDB.h:1507-1510) and the direction through the FP64 applyInverseJacobian (:1530);
 * the direction is normalised in FP32 and both times are scaled by the length (t1 = FLT_MAX times
 * a length above 1 is inf, as in the reference). */
static Ray world_to_index_ray_f64(const G* g, const float e[3], const float d[3])
{
    const uint8_t* m = g->buf + GRID_MAP;
    double M[9];
    for (int a = 0; a < 9; ++a) M[a] = rd_f64(m + MAP_INVMATD + 8 * a);
    float eye[3], dir[3];
    world_to_index_f64map(g, e, eye);
    for (int r = 0; r < 3; ++r)
        dir[r] = (float)fma((double)d[0], M[3 * r], fma((double)d[1], M[3 * r + 1], (double)d[2] * M[3 * r + 2]));
    const float length = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    const float invLength = 1.0f / length;
    const float nd[3] = {invLength * dir[0], invLength * dir[1], invLength * dir[2]};
    return make_ray(eye, nd, length * 1.0e-5f, length * FLT_MAX);
}

static int ray_clip(Ray* r, const int32_t* bbox)
{
    float t0 = r->t0, t1 = r->t1;
    for (int i = 0; i < 3; ++i) {
        float a = (float)bbox[i], b = (float)(bbox[3 + i] + 1);
        if (a >= b) return 0;
        a = (a - r->eye[i]) * r->inv[i];
        b = (b - r->eye[i]) * r->inv[i];
        if (a > b) { const float tmp = a; a = b; b = tmp; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return 0;
    }
    r->t0 = t0;
    r->t1 = t1;
    return 1;
}

int or_ray_clip(const float* eye, const float* dir, float t0, float t1, const int32_t* bbox, float* out)
{
    Ray r = make_ray(eye, dir, t0, t1);
    const int hit = ray_clip(&r, bbox);
    out[0] = r.t0;
    out[1] = r.t1;
    return hit;
}

/* ------------------------------------------------------------------ HDDA, math/HDDA.h:28-179 */
typedef struct {
    int32_t dim;
    float t0, t1;
    int32_t voxel[3], step[3];
    float delta[3], next[3];
} HDDA;

static inline int32_t ifloor(float x) { return (int32_t)floorf(x); }

static void hdda_init(HDDA* h, const Ray* r, float startTime, float maxTime, int dim)
{
    h->dim = dim;
    h->t0 = startTime;
    h->t1 = maxTime;
    float pos[3];
    ray_at(r, h->t0, pos);
    for (int a = 0; a < 3; ++a) h->voxel[a] = ifloor(pos[a]) & ~(dim - 1);
    for (int a = 0; a < 3; ++a) {
        h->delta[a] = 0.0f;
        if (r->dir[a] == 0.0f) {
            h->next[a] = FLT_MAX;
            h->step[a] = 0;
        } else if (r->inv[a] > 0.0f) {
            h->step[a] = 1;
            h->next[a] = h->t0 + ((float)(h->voxel[a] + dim) - pos[a]) * r->inv[a];
            h->delta[a] = r->inv[a];
        } else {
            h->step[a] = -1;
            h->next[a] = h->t0 + ((float)h->voxel[a] - pos[a]) * r->inv[a];
            h->delta[a] = -r->inv[a];
        }
    }
}

static int hdda_update(HDDA* h, const Ray* r, int dim)
{
    if (h->dim == dim) return 0;
    int32_t vmax[3], vmin[3];
    for (int a = 0; a < 3; ++a) {
        vmax[a] = (h->voxel[a] + (h->dim - 1)) & ~(dim - 1);
        vmin[a] = h->voxel[a] & ~(dim - 1);
    }
    h->dim = dim;
    float pos[3];
    ray_at(r, h->t0, pos);
    for (int a = 0; a < 3; ++a) {
        int32_t v = ifloor(pos[a]) & ~(dim - 1);
        v = v < vmax[a] ? v : vmax[a];
        v = v > vmin[a] ? v : vmin[a];
        h->voxel[a] = v;
    }
    for (int a = 0; a < 3; ++a) {
        if (h->step[a] == 0) continue;
        h->next[a] = h->t0 + ((float)h->voxel[a] - pos[a]) * r->inv[a];
        if (h->step[a] > 0) h->next[a] += (float)dim * r->inv[a];
    }
    return 1;
}

static inline int min_index(const float v[3])
{
    if (v[0] < v[1] && v[0] < v[2]) return 0;
    return v[1] < v[2] ? 1 : 2;
}

static int hdda_step(HDDA* h)
{
    const int a = min_index(h->next);
    if (h->next[a] <= h->t0) h->next[a] += h->t0 - 0.999999f * h->next[a] + 1.0e-6f; /* ENFORCE_FORWARD_STEPPING */
    h->t0 = h->next[a];
    h->next[a] += (float)h->dim * h->delta[a];
    h->voxel[a] += h->dim * h->step[a];
    return h->t0 <= h->t1;
}

static inline float hdda_next(const HDDA* h) { return fminf(h->t1, fminf(h->next[0], fminf(h->next[1], h->next[2]))); }

void or_hdda_trace(const float* eye, const float* dir, float t0, float t1, int dim, int nsteps, float* times,
                   float* nexts, int32_t* voxels, uint8_t* ok)
{
    const Ray r = make_ray(eye, dir, t0, t1);
    HDDA h;
    hdda_init(&h, &r, r.t0, r.t1, dim);
    for (int s = 0; s <= nsteps; ++s) {
        ok[s] = s > 0 ? (uint8_t)hdda_step(&h) : 1;
        times[s] = h.t0;
        nexts[s] = hdda_next(&h);
        for (int a = 0; a < 3; ++a) voxels[3 * s + a] = h.voxel[a];
    }
}

/* ------------------------------------------------------------------ zeroCrossing, math/HDDA.h:189-214 */
/* math::isoCrossing, math/HDDA.h:225-248 (zeroCrossing, :189-214, is the iso = 0 case:
 * value - 0.0f is the value bit for bit) */
static int iso_crossing(const G* g, Ray* ray, float iso, int32_t ijk[3], float* v, float* t)
{
    if (!ray_clip(ray, g->bbox) || (double)ray->t1 > 1e20) return 0;
    const float Delta = 1.0001f;
    float p[3];
    ray_at(ray, ray->t0, p);
    int32_t c[3] = {ifloor(p[0]), ifloor(p[1]), ifloor(p[2])};
    HDDA h;
    hdda_init(&h, ray, ray->t0, ray->t1, (int)lookup(g, c[0], c[1], c[2]).dim);
    const float v0 = get_value(g, c[0], c[1], c[2]) - iso;
    while (hdda_step(&h)) {
        ray_at(ray, h.t0 + Delta, p);
        c[0] = ifloor(p[0]);
        c[1] = ifloor(p[1]);
        c[2] = ifloor(p[2]);
        const Look l = lookup(g, c[0], c[1], c[2]);
        hdda_update(&h, ray, (int)l.dim);
        if (h.dim > 1 || !l.active) continue;
        while (hdda_step(&h)) {
            const Look lv = lookup(g, h.voxel[0], h.voxel[1], h.voxel[2]);
            if (!lv.active) break;
            *v = lv.value - iso;
            if (*v * v0 < 0) {
                ijk[0] = h.voxel[0];
                ijk[1] = h.voxel[1];
                ijk[2] = h.voxel[2];
                *t = h.t0;
                return 1;
            }
        }
    }
    return 0;
}

void or_zero_crossing(const void* buf, const float* eye, const float* dir, uint64_t n, uint8_t* hit, int32_t* ijk,
                      float* t, float* v, float* normal)
{
    or_iso_crossing(buf, eye, dir, n, 0.0f, hit, ijk, t, v, normal);
}

void or_iso_crossing(const void* buf, const float* eye, const float* dir, uint64_t n, float iso, uint8_t* hit,
                     int32_t* ijk, float* t, float* v, float* normal)
{
    or_iso_crossing_map(buf, eye, dir, n, iso, 0, hit, ijk, t, v, normal);
}

void or_iso_crossing_map(const void* buf, const float* eye, const float* dir, uint64_t n, float iso, int f64map,
                         uint8_t* hit, int32_t* ijk, float* t, float* v, float* normal)
{
    const G g = open_grid(buf);
    const uint8_t* m = g.buf + GRID_MAP;
    float IM[9];
    for (int a = 0; a < 9; ++a) IM[a] = rd_f32(m + MAP_INVMATF + 4 * a);
    for (uint64_t i = 0; i < n; ++i) {
        Ray r = f64map ? world_to_index_ray_f64(&g, eye + 3 * i, dir + 3 * i)
                       : world_to_index_ray(&g, eye + 3 * i, dir + 3 * i);
        int32_t c[3] = {0, 0, 0};
        float vv = 0.0f, tt = 0.0f;
        const int h = iso_crossing(&g, &r, iso, c, &vv, &tt);
        if (hit) hit[i] = (uint8_t)h;
        if (ijk) for (int a = 0; a < 3; ++a) ijk[3 * i + a] = h ? c[a] : 0;
        if (t) t[i] = h ? tt : 0.0f;
        if (v) v[i] = h ? vv : 0.0f;
        if (normal) {
            float nr[3] = {0.0f, 0.0f, 0.0f};
            if (h) {
                float p[3], gi[3], gw[3];
                ray_at(&r, tt, p);
                trilinear_grad(&g, p[0], p[1], p[2], gi);
                /* Grid::indexToWorldGradF -> Map::applyIJTF -> matMultT(const float*, Vec3f), math/Math.h:1758-1763 */
                for (int a = 0; a < 3; ++a) gw[a] = fmaf(gi[0], IM[a], fmaf(gi[1], IM[3 + a], gi[2] * IM[6 + a]));
                const float len = sqrtf(gw[0] * gw[0] + gw[1] * gw[1] + gw[2] * gw[2]);
                if (len > 0.0f) {
                    const float s = 1.0f / len;
                    nr[0] = s * gw[0];
                    nr[1] = s * gw[1];
                    nr[2] = s * gw[2];
                }
            }
            for (int a = 0; a < 3; ++a) normal[3 * i + a] = nr[a];
        }
    }
}

/* ------------------------------------------------------------------ iso-surface renderer
 * RenderOp of examples/ex_raytrace_iso_surface/common.h: getIndexRay (:130-145), isoCrossing,
 * composite (:147-155).  cam = {width, height, iso, dx, wbboxDimZ, tanHalfFov, origin[3]}. */
void or_iso_render(const void* buf, int32_t width, int32_t height, float iso, float dx, float wbbox_dim_z,
                   float tan_half_fov, const float origin[3], float* image, uint8_t* hit, float* t)
{
    const G g = open_grid(buf);
    for (uint64_t i = 0; i < (uint64_t)width * (uint64_t)height; ++i) {
        const uint32_t x = (uint32_t)(i % (uint32_t)width), y = (uint32_t)(i / (uint32_t)width);
        const float u = ((float)x + 0.5f) / (float)width;
        const float v = ((float)y + 0.5f) / (float)height;
        const float aspect = (float)width / (float)height;
        const float px = (2.f * u - 1.f) * tan_half_fov * aspect;
        const float py = (2.f * v - 1.f) * tan_half_fov;
        const float pz = -1.f;
        /* Vec3::normalize: *this *= 1 / length(), length = Sqrt(x*x + y*y + z*z), math/Math.h:1423-1473 */
        const float inv = 1.0f / sqrtf(px * px + py * py + pz * pz);
        const float d[3] = {px * inv, py * inv, pz * inv};
        Ray r = world_to_index_ray(&g, origin, d);
        int32_t c[3] = {0, 0, 0};
        float vv = 0.0f, tt = 0.0f;
        const int h = iso_crossing(&g, &r, iso, c, &vv, &tt);
        const float value = h ? (tt * dx) / (wbbox_dim_z * 2) : 0.0f;
        const float alpha = h ? 1.0f : 0.0f;
        const uint32_t mask = 1u << 7;
        const float bg = ((x & mask) ^ (y & mask)) ? 1.0f : 0.5f;
        image[i] = alpha * value + (1.0f - alpha) * bg;
        if (hit) hit[i] = (uint8_t)h;
        if (t) t[i] = h ? tt : 0.0f;
    }
}

/* ------------------------------------------------------------------ fog march
 * Emission-absorption march over HDDA cells of size >= 8 (the leaf level, as TreeMarcher,
 * math/HDDA.h:450-536, with Eps = 1e-6 and Delta = 0.01).  A cell is marched when it lies in a
 * leaf (getDim == 1) or an active tile.  Steps are globally aligned at t = k*dt, ta <= t < tb.
 * Per step: rho = trilinear(x(t)), a = exp(-sigma*rho*dt), L += T*(1-a)*rho, T *= a
 * (emission-absorption as openvdb/openvdb/tools/RayTracer.h:993-1064 with density emission).
 * The step structure is FP32 (identical to the GPU path); the arithmetic here is FP64. */
typedef struct {
    float u, v, w;     /* fractional position */
    int32_t c[3];      /* base voxel */
} Step;

typedef void (*step_fn)(const G* g, const Ray* r, float t, void* ctx);

static void march_structure(const G* g, const float* eye, const float* dir, float dt, step_fn fn, void* ctx)
{
    Ray R = world_to_index_ray(g, eye, dir);
    if (!ray_clip(&R, g->bbox) || (double)R.t1 > 1e20) return;
    const float Eps = 0.000001f, Delta = 0.01f;
    const float t0 = R.t0 + Eps, t1 = R.t1 - Eps;
    if (t0 > t1) return;
    float p[3];
    ray_at(&R, t0, p);
    Look l = lookup(g, ifloor(p[0]), ifloor(p[1]), ifloor(p[2]));
    HDDA h;
    hdda_init(&h, &R, t0, t1, (int)(l.dim > 8u ? l.dim : 8u));
    int on = (l.dim == 1u) || l.active;
    for (;;) {
        const float ta = h.t0;
        const int valid = hdda_step(&h);
        const float tb = fminf(h.t0, t1);
        if (on && ta < tb) {
            float kf = ceilf(ta / dt);
            for (float t = kf * dt; t < tb; t = kf * dt) {
                fn(g, &R, t, ctx);
                const float kn = kf + 1.0f;
                if (kn == kf) break; /* step index beyond 2^24: float cannot advance it */
                kf = kn;
            }
        }
        if (!valid) break;
        ray_at(&R, h.t0 + Delta, p);
        l = lookup(g, ifloor(p[0]), ifloor(p[1]), ifloor(p[2]));
        hdda_update(&h, &R, (int)(l.dim > 8u ? l.dim : 8u));
        on = (l.dim == 1u) || l.active;
    }
}

typedef struct {
    double sigma_dt, L, T;
    uint32_t steps;
    /* adjoint recording */
    int record;
    size_t cap, count;
    double* rho;
    double* a;
    double* Tk;
    Step* st;
} MarchCtx;

static void fwd_step(const G* g, const Ray* r, float t, void* vctx)
{
    MarchCtx* m = (MarchCtx*)vctx;
    float p[3];
    ray_at(r, t, p);
    const double rho = trilinear_f64(g, p[0], p[1], p[2]);
    const double a = exp(-m->sigma_dt * rho);
    if (m->record) {
        if (m->count == m->cap) {
            m->cap = m->cap ? 2 * m->cap : 1024;
            m->rho = (double*)realloc(m->rho, m->cap * sizeof(double));
            m->a = (double*)realloc(m->a, m->cap * sizeof(double));
            m->Tk = (double*)realloc(m->Tk, m->cap * sizeof(double));
            m->st = (Step*)realloc(m->st, m->cap * sizeof(Step));
        }
        const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
        Step s;
        s.u = p[0] - fx; s.v = p[1] - fy; s.w = p[2] - fz;
        s.c[0] = (int32_t)fx; s.c[1] = (int32_t)fy; s.c[2] = (int32_t)fz;
        m->rho[m->count] = rho;
        m->a[m->count] = a;
        m->Tk[m->count] = m->T;
        m->st[m->count] = s;
        ++m->count;
    }
    m->L += m->T * (1.0 - a) * rho;
    m->T *= a;
    ++m->steps;
}

void or_fog_march(const void* buf, const float* eye, const float* dir, uint64_t n, float dt, float sigma, double* L,
                  double* T, uint32_t* steps)
{
    const G g = open_grid(buf);
    for (uint64_t i = 0; i < n; ++i) {
        MarchCtx m;
        memset(&m, 0, sizeof(m));
        m.sigma_dt = (double)sigma * (double)dt;
        m.T = 1.0;
        march_structure(&g, eye + 3 * i, dir + 3 * i, dt, fwd_step, &m);
        if (L) L[i] = m.L;
        if (T) T[i] = m.T;
        if (steps) steps[i] = m.steps;
    }
}

/* dLoss/drho_k = dL*(T_k[(1-a_k) + rho_k*s*a_k] - s*R_{>k}) - dT*s*T_N, with s = sigma*dt and
 * R_{>k} = sum_{j>k} T_j (1-a_j) rho_j; scattered with the trilinear weights of step k onto
 * the corner voxels that live in leaf nodes (grad index = leafIndex*512 + n), or, for index grids,
 * onto the channel entry every corner reads (grad index = channel index). */
void or_fog_march_adjoint(const void* buf, const float* eye, const float* dir, uint64_t n, float dt, float sigma,
                          const double* dL, const double* dT, double* grad)
{
    const G g = open_grid(buf);
    if (g.leafSize == 0) return;
    const int idx = g.type == GT_ONINDEX;
    MarchCtx m;
    memset(&m, 0, sizeof(m));
    m.record = 1;
    for (uint64_t i = 0; i < n; ++i) {
        m.count = 0;
        m.L = 0.0;
        m.T = 1.0;
        m.steps = 0;
        m.sigma_dt = (double)sigma * (double)dt;
        march_structure(&g, eye + 3 * i, dir + 3 * i, dt, fwd_step, &m);
        const double s = m.sigma_dt, TN = m.T, gl = dL ? dL[i] : 0.0, gt = dT ? dT[i] : 0.0;
        double R = 0.0; /* suffix radiance after step k */
        for (size_t k = m.count; k-- > 0;) {
            const double rho = m.rho[k], a = m.a[k], Tk = m.Tk[k];
            const double gk = gl * (Tk * ((1.0 - a) + rho * s * a) - s * R) - gt * s * TN;
            R += Tk * (1.0 - a) * rho;
            const Step st = m.st[k];
            const double wx[2] = {1.0 - st.u, st.u}, wy[2] = {1.0 - st.v, st.v}, wz[2] = {1.0 - st.w, st.w};
            for (int a0 = 0; a0 < 2; ++a0)
                for (int b0 = 0; b0 < 2; ++b0)
                    for (int c0 = 0; c0 < 2; ++c0) {
                        const Look l = lookup(&g, st.c[0] + a0, st.c[1] + b0, st.c[2] + c0);
                        if (idx) {
                            grad[l.index] += gk * wx[a0] * wy[b0] * wz[c0];
                            continue;
                        }
                        if (!l.leaf) continue;
                        const uint64_t li = (uint64_t)(l.leaf - g.leaf0) / g.leafSize;
                        grad[li * 512u + l.n] += gk * wx[a0] * wy[b0] * wz[c0];
                    }
        }
    }
    free(m.rho);
    free(m.a);
    free(m.Tk);
    free(m.st);
}

/* ------------------------------------------------------------------ node spans and first hits
 * TreeMarcher<NodeT, RayT, AccT>::init / step, math/HDDA.h:450-536.  level 0 = LeafNode (edge 8,
 * children 1), 1 = lower InternalNode (128, 8), 2 = upper InternalNode (4096, 128).  The HDDA
 * steps at max(getDim, edge); a span starts where the lookup at the span's start lands in a
 * child of NodeT (getDim <= child edge), its node is the NodeT that lookup passed through
 * (ReadAccessor::getNode<NodeT>, the node cached by that getDim), and it ends at the next HDDA
 * crossing (clamped to the clipped t1).  Spans with t0 >= t1 are skipped. */
static const uint8_t* look_node(const Look* l, int level)
{
    return level == 0 ? l->leaf : (level == 1 ? l->lower : l->upper);
}

static uint32_t tree_march_ray(const G* g, const float* e, const float* d, int level, uint32_t max_spans, float* T0,
                               float* T1, int64_t* node)
{
    const uint32_t DIM = level == 0 ? 8u : (level == 1 ? 128u : 4096u);
    const uint32_t CHILD = level == 0 ? 1u : (level == 1 ? 8u : 128u);
    Ray R = world_to_index_ray(g, e, d);
    if (!ray_clip(&R, g->bbox)) return 0;
    const float Eps = 0.000001f, Delta = 0.01f;
    const float t0 = R.t0 + Eps, t1 = R.t1 - Eps;
    if (t0 > t1) return 0;
    float p[3];
    ray_at(&R, t0, p);
    Look l = lookup(g, ifloor(p[0]), ifloor(p[1]), ifloor(p[2]));
    HDDA h;
    hdda_init(&h, &R, t0, t1, (int)(l.dim > DIM ? l.dim : DIM));
    float mT0 = l.dim <= CHILD ? h.t0 : -1.0f;
    const uint8_t* cached = look_node(&l, level);
    uint32_t count = 0;
    for (;;) {
        const float s0 = mT0;
        const uint8_t* current = cached;
        const int valid = hdda_step(&h);
        ray_at(&R, h.t0 + Delta, p);
        l = lookup(g, ifloor(p[0]), ifloor(p[1]), ifloor(p[2]));
        cached = look_node(&l, level);
        hdda_update(&h, &R, (int)(l.dim > DIM ? l.dim : DIM));
        mT0 = l.dim <= CHILD ? h.t0 : -1.0f;
        if (s0 >= 0.0f) {
            const float s1 = fminf(t1, h.t0);
            if (s0 < s1 && current) {
                if (count < max_spans) {
                    T0[count] = s0;
                    T1[count] = s1;
                    node[count] = (int64_t)(current - g->buf);
                }
                ++count;
            }
        }
        if (!valid) break;
    }
    return count;
}

void or_tree_march(const void* buf, const float* eye, const float* dir, uint64_t n, int level, uint32_t max_spans,
                   uint32_t* count, float* t0, float* t1, int64_t* node)
{
    const G g = open_grid(buf);
    for (uint64_t i = 0; i < n; ++i)
        count[i] = tree_march_ray(&g, eye + 3 * i, dir + 3 * i, level, max_spans, t0 + i * max_spans,
                                  t1 + i * max_spans, node + i * max_spans);
}

/* firstActive(ray, acc, ijk, t), math/HDDA.h:429-444: clip to the root bbox, then HDDA steps with
 * a 1.0001 nudge until the lookup at RoundDown(ray(t)) is active (a voxel or a tile). */
static int first_active(const G* g, Ray* ray, int32_t ijk[3], float* t)
{
    if (!ray_clip(ray, g->bbox) || (double)ray->t1 > 1e20) return 0;
    const float Delta = 1.0001f;
    float p[3];
    *t = ray->t0;
    ray_at(ray, ray->t0, p);
    for (int a = 0; a < 3; ++a) ijk[a] = ifloor(p[a]);
    Look l = lookup(g, ijk[0], ijk[1], ijk[2]);
    HDDA h;
    hdda_init(&h, ray, ray->t0, ray->t1, (int)l.dim);
    while (!l.active) {
        if (!hdda_step(&h)) return 0;
        *t = h.t0 + Delta;
        ray_at(ray, *t, p);
        for (int a = 0; a < 3; ++a) ijk[a] = ifloor(p[a]);
        l = lookup(g, ijk[0], ijk[1], ijk[2]);
        hdda_update(&h, ray, (int)l.dim);
    }
    return 1;
}

void or_first_active(const void* buf, const float* eye, const float* dir, uint64_t n, uint8_t* hit, int32_t* ijk,
                     float* t)
{
    const G g = open_grid(buf);
    for (uint64_t i = 0; i < n; ++i) {
        Ray r = world_to_index_ray(&g, eye + 3 * i, dir + 3 * i);
        int32_t c[3] = {0, 0, 0};
        float tt = 0.0f;
        const int h = first_active(&g, &r, c, &tt);
        hit[i] = (uint8_t)h;
        for (int a = 0; a < 3; ++a) ijk[3 * i + a] = h ? c[a] : 0;
        t[i] = h ? tt : 0.0f;
    }
}

/* DDA<RayT, Coord, Dim>, math/HDDA.h:258-379: voxel = RoundDown(ray(t0)) & ~(Dim-1); a step moves
 * the voxel by +-Dim and the axis' next time by |1/dir| (one voxel's time, as the reference
 * writes it), with the same forward-stepping fix as the HDDA. */
void or_dda_trace(const float* eye, const float* dir, float t0, float t1, int dim, int nsteps, float* times,
                  float* nexts, int32_t* voxels, uint8_t* ok)
{
    const Ray r = make_ray(eye, dir, t0, t1);
    float T0 = r.t0, next[3], delta[3], pos[3];
    int32_t voxel[3], step[3];
    ray_at(&r, T0, pos);
    for (int a = 0; a < 3; ++a) {
        voxel[a] = ifloor(pos[a]) & ~(dim - 1);
        delta[a] = 0.0f;
        if (r.dir[a] == 0.0f) {
            next[a] = FLT_MAX;
            step[a] = 0;
        } else if (r.inv[a] > 0.0f) {
            step[a] = dim;
            next[a] = T0 + ((float)(voxel[a] + dim) - pos[a]) * r.inv[a];
            delta[a] = r.inv[a];
        } else {
            step[a] = -dim;
            next[a] = T0 + ((float)voxel[a] - pos[a]) * r.inv[a];
            delta[a] = -r.inv[a];
        }
    }
    for (int s = 0; s <= nsteps; ++s) {
        if (s == 0) {
            ok[0] = 1;
        } else {
            const int a = min_index(next);
            if (next[a] <= T0) next[a] += T0 - 0.999999f * next[a] + 1.0e-6f;
            T0 = next[a];
            next[a] += delta[a];
            voxel[a] += step[a];
            ok[s] = (uint8_t)(T0 <= r.t1);
        }
        times[s] = T0;
        nexts[s] = fminf(r.t1, fminf(next[0], fminf(next[1], next[2])));
        for (int a = 0; a < 3; ++a) voxels[3 * s + a] = voxel[a];
    }
}

/* ------------------------------------------------------------------ BoxStencil, math/Stencils.h:306-432
 * moveTo(Vec3f) fills the 8 corners of the cell at RoundDown(xyz) in the order 000, 001, 011, 010,
 * 100, 101, 111, 110; interpolation nests z, y, x as A + (B - A)*w; gradient returns the partials
 * mapped by the FP64 Map::applyIJT (matMultT(mInvMatD, g), math/Math.h:1766-1776) and rounded to
 * float.  Float grids. */
void or_box_stencil(const void* buf, const float* xyz, uint64_t n, float* interp, float* grad)
{
    const G g = open_grid(buf);
    const uint8_t* m = g.buf + GRID_MAP;
    double IM[9];
    for (int a = 0; a < 9; ++a) IM[a] = rd_f64(m + MAP_INVMATD + 8 * a);
    for (uint64_t i = 0; i < n; ++i) {
        const float* p = xyz + 3 * i;
        const int32_t c[3] = {ifloor(p[0]), ifloor(p[1]), ifloor(p[2])};
        static const int off[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0},
                                      {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
        float V[8];
        for (int k = 0; k < 8; ++k) V[k] = get_value(&g, c[0] + off[k][0], c[1] + off[k][1], c[2] + off[k][2]);
        const float u = p[0] - (float)c[0], v = p[1] - (float)c[1], w = p[2] - (float)c[2];
        if (interp) {
            float A = V[0] + (V[1] - V[0]) * w;
            float B = V[3] + (V[2] - V[3]) * w;
            const float C = A + (B - A) * v;
            A = V[4] + (V[5] - V[4]) * w;
            B = V[7] + (V[6] - V[7]) * w;
            const float D = A + (B - A) * v;
            interp[i] = C + (D - C) * u;
        }
        if (grad) {
            float D[4] = {V[1] - V[0], V[2] - V[3], V[5] - V[4], V[6] - V[7]};
            float A = D[0] + (D[1] - D[0]) * v;
            float B = D[2] + (D[3] - D[2]) * v;
            float gi[3];
            gi[2] = A + (B - A) * u;
            D[0] = V[0] + D[0] * w;
            D[1] = V[3] + D[1] * w;
            D[2] = V[4] + D[2] * w;
            D[3] = V[7] + D[3] * w;
            A = D[0] + (D[1] - D[0]) * v;
            B = D[2] + (D[3] - D[2]) * v;
            gi[0] = B - A;
            A = D[1] - D[0];
            B = D[3] - D[2];
            gi[1] = A + (B - A) * u;
            for (int a = 0; a < 3; ++a)
                grad[3 * i + a] =
                    (float)fma((double)gi[0], IM[a], fma((double)gi[1], IM[3 + a], (double)gi[2] * IM[6 + a]));
        }
    }
}

/* GradStencil extras at integer voxels (math/Stencils.h:512-518, 544-554): the first-order upwind
 * gradient along V[i] (V > 0 takes the backward difference) times 2*mInv2Dx, and the
 * closest-point transform ijk - d*(central differences), d = float(v0 * 0.5 * mInvDx2).  Float
 * grids. */
void or_grad_stencil_ext(const void* buf, const int32_t* ijk, uint64_t n, const float* V, float* upwind, float* cpt)
{
    const G g = open_grid(buf);
    const GradC gc = grad_consts(&g);
    for (uint64_t i = 0; i < n; ++i) {
        const int32_t a = ijk[3 * i], b = ijk[3 * i + 1], c = ijk[3 * i + 2];
        float v[7];
        v[0] = get_value(&g, a, b, c);
        v[1] = get_value(&g, a - 1, b, c);
        v[2] = get_value(&g, a + 1, b, c);
        v[3] = get_value(&g, a, b - 1, c);
        v[4] = get_value(&g, a, b + 1, c);
        v[5] = get_value(&g, a, b, c - 1);
        v[6] = get_value(&g, a, b, c + 1);
        if (upwind) {
            const float* d = V + 3 * i;
            const float s[3] = {d[0] > 0 ? v[0] - v[1] : v[2] - v[0], d[1] > 0 ? v[0] - v[3] : v[4] - v[0],
                                d[2] > 0 ? v[0] - v[5] : v[6] - v[0]};
            for (int k = 0; k < 3; ++k) upwind[3 * i + k] = (s[k] * 2.0f) * gc.inv2dx;
        }
        if (cpt) {
            const float d = (float)(v[0] * 0.5 * gc.invdx2);
            cpt[3 * i] = (float)a - d * (v[2] - v[1]);
            cpt[3 * i + 1] = (float)b - d * (v[4] - v[3]);
            cpt[3 * i + 2] = (float)c - d * (v[6] - v[5]);
        }
    }
}

/* TrilinearSampler::zeroCrossing of the cell at Floor(xyz), math/SampleFromVoxels.h:274-288 */
void or_sample_zero_crossing(const void* buf, const float* xyz, uint64_t n, uint8_t* zc)
{
    const G g = open_grid(buf);
    for (uint64_t i = 0; i < n; ++i) {
        const float* p = xyz + 3 * i;
        float s[2][2][2];
        stencil8(&g, ifloor(p[0]), ifloor(p[1]), ifloor(p[2]), s);
        const int less = s[0][0][0] < 0.0f;
        zc[i] = (uint8_t)((less ^ (s[0][0][1] < 0.0f)) || (less ^ (s[0][1][1] < 0.0f)) ||
                          (less ^ (s[0][1][0] < 0.0f)) || (less ^ (s[1][0][0] < 0.0f)) ||
                          (less ^ (s[1][0][1] < 0.0f)) || (less ^ (s[1][1][1] < 0.0f)) ||
                          (less ^ (s[1][1][0] < 0.0f)));
    }
}
