// Test infrastructure only (never linked into the product library).
//
// C ABI over the reference NanoVDB host code, compiled directly from the
// reference's own headers (nanovdb/nanovdb/*.h) by oracle/Makefile into
// oracle/_ref/libnanoref.so.  The tests use it to pin the oracle restatement
// and the product's host grid builder, and bench.py's CPU-baseline leg times
// the reference query calls through it on the host cores.
//
// Every per-item function here is a thin loop around one reference call:
//   ReadAccessor::getValue / isActive / getDim      nanovdb/nanovdb/NanoVDB.h:5370-5447
//   SampleFromVoxels<Acc,0|1,false>                 nanovdb/nanovdb/math/SampleFromVoxels.h:149-156, 423-476
//   FD6 gradient of the CUDA sampler op             nanovdb/nanovdb/python/cuda/PySampleFromVoxels.cu:31-52
//   GradStencil                                     nanovdb/nanovdb/math/Stencils.h:455-576
//   Ray::worldToIndexF + math::zeroCrossing         nanovdb/nanovdb/math/Ray.h:258-268, math/HDDA.h:189-214
//   HDDA (fog-march spans)                          nanovdb/nanovdb/math/HDDA.h:28-179
//   RayGenOp (pinhole camera)                       nanovdb/nanovdb/examples/ex_raytrace_level_set/common.h:91-125
//   createLevelSetSphere / Torus, build::Grid::operator()(func,bbox), createNanoGrid
//                                                   nanovdb/nanovdb/tools/{CreatePrimitives,GridBuilder,CreateNanoGrid}.h
#include <nanovdb/NanoVDB.h>
#include <nanovdb/GridHandle.h>
#include <nanovdb/HostBuffer.h>
#include <nanovdb/math/Ray.h>
#include <nanovdb/math/HDDA.h>
#include <nanovdb/math/SampleFromVoxels.h>
#include <nanovdb/math/Stencils.h>
#include <nanovdb/tools/CreatePrimitives.h>
#include <nanovdb/tools/CreateNanoGrid.h>
#include <nanovdb/tools/GridBuilder.h>
#include <nanovdb/tools/GridValidator.h>
#include <nanovdb/tools/GridStats.h>
#include <nanovdb/io/IO.h>
#include "common.h" // examples/ex_raytrace_level_set/common.h (RayGenOp)

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

using namespace nanovdb;

namespace {

enum BuildKind { kFloat = 0, kFp4 = 1, kFp8 = 2, kFp16 = 3, kFpN = 4 };

void* copyOut(const GridHandle<HostBuffer>& h, uint64_t* bytes)
{
    *bytes = h.bufferSize();
    void* p = std::aligned_alloc(32, (h.bufferSize() + 31) & ~uint64_t(31));
    std::memcpy(p, h.data(), h.bufferSize());
    return p;
}

// dynamic-chunk parallel loop over [0,n) on nthreads host threads
template<typename F>
void parallelFor(uint64_t n, int nthreads, F&& f)
{
    if (nthreads <= 1 || n < 2048) {
        f(uint64_t(0), n);
        return;
    }
    std::atomic<uint64_t> next{0};
    const uint64_t chunk = 1024;
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                const uint64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                f(b, std::min(n, b + chunk));
            }
        });
    }
    for (auto& th : pool) th.join();
}



template<typename BuildT>
const NanoGrid<BuildT>* asGrid(const void* buf) { return reinterpret_cast<const NanoGrid<BuildT>*>(buf); }

// Float-valued accessor of any grid: ReadAccessor for value grids, ChannelAccessor over blind
// channel 0 for index grids (getAccessor<GridT, float>, NanoVDB.h:5814-5826).
template<typename GridT>
auto makeAcc(const GridT* grid)
{
    return nanovdb::getAccessor<GridT, float>(*grid);
}

// Dispatch a functor templated on the build type of the buffer.
template<typename F>
int dispatch(const void* buf, F&& f)
{
    switch (reinterpret_cast<const GridData*>(buf)->mGridType) {
    case GridType::Float: return f(asGrid<float>(buf));
    case GridType::Fp4: return f(asGrid<Fp4>(buf));
    case GridType::Fp8: return f(asGrid<Fp8>(buf));
    case GridType::Fp16: return f(asGrid<Fp16>(buf));
    case GridType::FpN: return f(asGrid<FpN>(buf));
    case GridType::OnIndex: return f(asGrid<ValueOnIndex>(buf));
    default: return -4;
    }
}

// The fog-march contract shared with the product kernels and the oracle
// restatement: HDDA over the tree at node granularity >= 8 (leaf level, as in
// TreeMarcher, math/HDDA.h:450-536); a cell is marched when it lies in a leaf
// (getDim == 1) or in an active tile; steps are globally aligned at
// t = k*dt with ta <= t < tb inside each marched cell; the per-step density is
// the trilinear sample at the index-space position; emission-absorption with
// a = exp(-sigma*rho*dt), L += T*(1-a)*rho, T *= a.
template<typename GridT, typename RealT>
void fogMarch(const GridT* grid, const math::Ray<float>& wRay, float dt, float sigma, RealT& L, RealT& T, uint32_t& nsteps)
{
    using RayT = math::Ray<float>;
    auto acc = makeAcc(grid);
    L = RealT(0);
    T = RealT(1);
    nsteps = 0;
    RayT R = wRay.worldToIndexF(*grid);
    if (!R.clip(acc.root().bbox()) || R.t1() > 1e20f) return;
    const float Eps = 0.000001f, Delta = 0.01f;
    const float t0 = R.t0() + Eps, t1 = R.t1() - Eps;
    if (t0 > t1) return;
    Coord ijk = math::RoundDown<Coord>(R(t0));
    uint32_t dim = acc.getDim(ijk, R);
    math::HDDA<RayT, Coord> hdda;
    hdda.init(R, t0, t1, int(math::Max(dim, 8u)));
    bool on = (dim == 1u) || acc.isActive(ijk);
    auto sampler = math::createSampler<1, typename GridT::AccessorType, false>(acc);
    for (;;) {
        const float ta = hdda.time();
        const bool valid = hdda.step();
        const float tb = math::Min(hdda.time(), t1);
        if (on && ta < tb) {
            float kf = std::ceil(ta / dt);
            for (float t = kf * dt; t < tb; t = kf * dt) {
                const Vec3f x = R(t);
                RealT rho;
                if constexpr (sizeof(RealT) == 4) {
                    rho = sampler(x);
                } else {
                    // FP64 arithmetic over the same FP32 position and stencil values
                    Vec3f uvw = x;
                    Coord c = math::Floor<Coord>(uvw);
                    double v[2][2][2];
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b)
                            for (int e = 0; e < 2; ++e) v[a][b][e] = acc.getValue(c.offsetBy(a, b, e));
                    auto lerp = [](double p, double q, double w) { return p + w * (q - p); };
                    rho = lerp(lerp(lerp(v[0][0][0], v[0][0][1], uvw[2]), lerp(v[0][1][0], v[0][1][1], uvw[2]), uvw[1]),
                               lerp(lerp(v[1][0][0], v[1][0][1], uvw[2]), lerp(v[1][1][0], v[1][1][1], uvw[2]), uvw[1]), uvw[0]);
                }
                RealT a;
                if constexpr (sizeof(RealT) == 4) a = expf(-sigma * rho * dt);
                else a = std::exp(-double(sigma) * rho * double(dt));
                L += T * (RealT(1) - a) * rho;
                T *= a;
                ++nsteps;
                const float kn = kf + 1.0f;
                if (kn == kf) break; // step index beyond 2^24: float cannot advance it
                kf = kn;
            }
        }
        if (!valid) break;
        const Coord nijk = math::RoundDown<Coord>(R(hdda.time() + Delta));
        const uint32_t ndim = acc.getDim(nijk, R);
        hdda.update(R, int(math::Max(ndim, 8u)));
        on = (ndim == 1u) || acc.isActive(nijk);
    }
}

} // namespace

extern "C" {

unsigned ref_hw_threads() { return std::thread::hardware_concurrency(); }

void ref_free(void* p) { std::free(p); }

// ---------------------------------------------------------------- grid producers
void* ref_level_set_sphere(int kind, double radius, double cx, double cy, double cz, double voxel, double hw,
                           float tol, uint64_t* bytes)
{
    const Vec3d c(cx, cy, cz);
    switch (kind) {
    case kFloat: return copyOut(tools::createLevelSetSphere<float>(radius, c, voxel, hw), bytes);
    case kFp4: return copyOut(tools::createLevelSetSphere<Fp4>(radius, c, voxel, hw), bytes);
    case kFp8: return copyOut(tools::createLevelSetSphere<Fp8>(radius, c, voxel, hw), bytes);
    case kFp16: return copyOut(tools::createLevelSetSphere<Fp16>(radius, c, voxel, hw), bytes);
    case kFpN:
        return copyOut(tools::createLevelSetSphere<FpN>(radius, c, voxel, hw, Vec3d(0), "sphere_ls_FpN",
                                                        tools::StatsMode::Default, CheckMode::Default, tol), bytes);
    }
    return nullptr;
}

void* ref_level_set_torus(int kind, double R, double r, double cx, double cy, double cz, double voxel, double hw,
                          float tol, uint64_t* bytes)
{
    const Vec3d c(cx, cy, cz);
    switch (kind) {
    case kFloat: return copyOut(tools::createLevelSetTorus<float>(R, r, c, voxel, hw), bytes);
    case kFp4: return copyOut(tools::createLevelSetTorus<Fp4>(R, r, c, voxel, hw), bytes);
    case kFp8: return copyOut(tools::createLevelSetTorus<Fp8>(R, r, c, voxel, hw), bytes);
    case kFp16: return copyOut(tools::createLevelSetTorus<Fp16>(R, r, c, voxel, hw), bytes);
    case kFpN:
        return copyOut(tools::createLevelSetTorus<FpN>(R, r, c, voxel, hw, Vec3d(0), "torus_ls",
                                                       tools::StatsMode::Default, CheckMode::Default, tol), bytes);
    }
    return nullptr;
}

void* ref_fog_volume_sphere(double radius, double cx, double cy, double cz, double voxel, double hw, uint64_t* bytes)
{
    return copyOut(tools::createFogVolumeSphere<float>(radius, Vec3d(cx, cy, cz), voxel, hw), bytes);
}

// build::Grid<float>(background, name, FogVolume) filled by operator()(func, bbox)
typedef float (*ref_voxel_fn)(int, int, int, void*);
void* ref_grid_from_function(ref_voxel_fn fn, void* ctx, const int32_t* bbox, float background, int gridClass,
                             uint64_t* bytes)
{
    tools::build::Grid<float> grid(background, "func", GridClass(gridClass));
    const CoordBBox b(Coord(bbox[0], bbox[1], bbox[2]), Coord(bbox[3], bbox[4], bbox[5]));
    grid([&](const Coord& ijk) { return fn(ijk[0], ijk[1], ijk[2], ctx); }, b);
    return copyOut(tools::createNanoGrid(grid), bytes);
}

// Same, with the functor reading a dense array over bbox (x-major, z fastest).
void* ref_grid_from_dense(const float* dense, const int32_t* bbox, float background, int gridClass, uint64_t* bytes)
{
    tools::build::Grid<float> grid(background, "func", GridClass(gridClass));
    const CoordBBox b(Coord(bbox[0], bbox[1], bbox[2]), Coord(bbox[3], bbox[4], bbox[5]));
    const int64_t ny = bbox[4] - bbox[1] + 1, nz = bbox[5] - bbox[2] + 1;
    grid([&](const Coord& ijk) {
        return dense[((int64_t(ijk[0]) - bbox[0]) * ny + (ijk[1] - bbox[1])) * nz + (ijk[2] - bbox[2])];
    }, b);
    return copyOut(tools::createNanoGrid(grid), bytes);
}

// Quantise an FP32 NanoGrid buffer into Fp4/Fp8/Fp16/FpN with the reference converter.
void* ref_quantize(const void* buf, int kind, float tol, uint64_t* bytes)
{
    const auto* src = asGrid<float>(buf);
    switch (kind) {
    case kFp4: return copyOut(tools::createNanoGrid<NanoGrid<float>, Fp4>(*src), bytes);
    case kFp8: return copyOut(tools::createNanoGrid<NanoGrid<float>, Fp8>(*src), bytes);
    case kFp16: return copyOut(tools::createNanoGrid<NanoGrid<float>, Fp16>(*src), bytes);
    case kFpN:
        return copyOut(tools::createNanoGrid<NanoGrid<float>, FpN>(*src, tools::StatsMode::Default, CheckMode::Default,
                                                                   false, 0, tools::AbsDiff(tol)), bytes);
    }
    return nullptr;
}

// ---------------------------------------------------------------- per-item queries
// dim[i] = acc.getDim(ijk, ray) (1 voxel in a leaf, 8/128/4096 tile or background)
int ref_get_value(const void* buf, const int32_t* ijk, uint64_t n, float* val, uint8_t* active, uint32_t* dim,
                  int nthreads)
{
    return dispatch(buf, [&](auto* grid) {
        parallelFor(n, nthreads, [&](uint64_t b, uint64_t e) {
            auto acc = makeAcc(grid);
            const math::Ray<float> ray;
            for (uint64_t i = b; i < e; ++i) {
                const Coord c(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
                if (val) val[i] = acc.getValue(c);
                if (active) active[i] = acc.isActive(c) ? 1 : 0;
                if (dim) dim[i] = acc.getDim(c, ray);
            }
        });
        return 0;
    });
}

// world: 0 = xyz is index space, 1 = world space mapped with Grid::worldToIndex (FP64 Map, as the CUDA op).
// nn  = SampleFromVoxels<Acc,0,false> (Round); val = SampleFromVoxels<Acc,1,false>;
// grad_mode 1 = FD6 of trilinear samples * 0.5/voxelSize (PySampleFromVoxels.cu:42-47)
// grad_mode 2 = GradStencil::gradient at RoundDown(xyz); 3 = TrilinearSampler::gradient (index space)
int ref_sample(const void* buf, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
               int grad_mode, int nthreads)
{
    return dispatch(buf, [&](auto* grid) {
        using GridT = std::remove_cv_t<std::remove_pointer_t<decltype(grid)>>;
        parallelFor(n, nthreads, [&](uint64_t b, uint64_t e) {
            auto acc = makeAcc(grid);
            auto s0 = math::createSampler<0, decltype(acc), false>(acc);
            auto s1 = math::createSampler<1, decltype(acc), false>(acc);
            math::GradStencil<GridT> st(*grid);
            const Vec3f inv2Dx = Vec3f(0.5f / grid->voxelSize());
            for (uint64_t i = b; i < e; ++i) {
                Vec3f p(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
                if (world) p = grid->worldToIndex(p);
                if (nn) nn[i] = s0(p);
                if (val) val[i] = s1(p);
                if (grad) {
                    Vec3f g(0.0f);
                    if (grad_mode == 1) {
                        g = Vec3f(s1(p + Vec3f(1, 0, 0)) - s1(p - Vec3f(1, 0, 0)),
                                  s1(p + Vec3f(0, 1, 0)) - s1(p - Vec3f(0, 1, 0)),
                                  s1(p + Vec3f(0, 0, 1)) - s1(p - Vec3f(0, 0, 1))) * inv2Dx;
                    } else if (grad_mode == 2) {
                        st.moveTo(math::RoundDown<Coord>(p));
                        g = st.gradient();
                    } else if (grad_mode == 3) {
                        g = s1.gradient(p);
                    }
                    grad[3 * i] = g[0];
                    grad[3 * i + 1] = g[1];
                    grad[3 * i + 2] = g[2];
                }
            }
        });
        return 0;
    });
}

// GradStencil at integer voxels: gradient, normSqGrad, laplacian, zeroCrossing (math/Stencils.h:486-538)
int ref_grad_stencil(const void* buf, const int32_t* ijk, uint64_t n, float* grad, float* normsq, float* lap,
                     uint8_t* zc)
{
    const auto* grid = asGrid<float>(buf);
    math::GradStencil<NanoGrid<float>> st(*grid);
    for (uint64_t i = 0; i < n; ++i) {
        st.moveTo(Coord(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]));
        const Vec3f g = st.gradient();
        if (grad) {
            grad[3 * i] = g[0];
            grad[3 * i + 1] = g[1];
            grad[3 * i + 2] = g[2];
        }
        if (normsq) normsq[i] = st.normSqGrad();
        if (lap) lap[i] = st.laplacian();
        if (zc) zc[i] = st.zeroCrossing() ? 1 : 0;
    }
    return 0;
}

// World-space rays -> worldToIndexF -> zeroCrossing; normal = world-space trilinear
// gradient at the hit position ray(t), mapped with indexToWorldGradF and normalised.
int ref_zero_crossing_map(const void* buf, const float* eye, const float* dir, uint64_t n, int f64map, uint8_t* hit,
                          int32_t* ijk, float* t, float* v, float* normal, int nthreads);
int ref_zero_crossing(const void* buf, const float* eye, const float* dir, uint64_t n, uint8_t* hit, int32_t* ijk,
                      float* t, float* v, float* normal, int nthreads)
{
    return ref_zero_crossing_map(buf, eye, dir, n, 0, hit, ijk, t, v, normal, nthreads);
}

// The same, with the index-space ray from Ray::applyInverseMap(grid->map()) (the FP64 map,
// math/Ray.h:223-230) when f64map != 0.
int ref_zero_crossing_map(const void* buf, const float* eye, const float* dir, uint64_t n, int f64map, uint8_t* hit,
                          int32_t* ijk, float* t, float* v, float* normal, int nthreads)
{
    return dispatch(buf, [&](auto* grid) {
        parallelFor(n, nthreads, [&](uint64_t b, uint64_t e) {
            auto acc = makeAcc(grid);
            auto s1 = math::createSampler<1, decltype(acc), false>(acc);
            for (uint64_t i = b; i < e; ++i) {
                math::Ray<float> wRay(Vec3f(eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]),
                                      Vec3f(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
                math::Ray<float> iRay = f64map ? wRay.applyInverseMap(grid->map()) : wRay.worldToIndexF(*grid);
                Coord c(0);
                float vv = 0.0f, tt = 0.0f;
                const bool h = math::zeroCrossing(iRay, acc, c, vv, tt);
                if (hit) hit[i] = h ? 1 : 0;
                if (ijk) {
                    ijk[3 * i] = h ? c[0] : 0;
                    ijk[3 * i + 1] = h ? c[1] : 0;
                    ijk[3 * i + 2] = h ? c[2] : 0;
                }
                if (t) t[i] = h ? tt : 0.0f;
                if (v) v[i] = h ? vv : 0.0f;
                if (normal) {
                    Vec3f nrm(0.0f);
                    if (h) {
                        Vec3f g = grid->indexToWorldGradF(s1.gradient(iRay(tt)));
                        const float len = g.length();
                        if (len > 0.0f) nrm = g * (1.0f / len);
                    }
                    normal[3 * i] = nrm[0];
                    normal[3 * i + 1] = nrm[1];
                    normal[3 * i + 2] = nrm[2];
                }
            }
        });
        return 0;
    });
}

// createNanoGrid<NanoGrid<float>, ValueOnIndex>(grid, channels, includeStats, includeTiles)
// (tools/CreateNanoGrid.h:183-207): an index grid of the active values, with `channels` float
// channels of the source values as blind data.
void* ref_index_grid(const void* float_buf, uint32_t channels, int include_stats, int include_tiles, uint64_t* bytes)
{
    try {
        auto h = tools::createNanoGrid<NanoGrid<float>, ValueOnIndex, HostBuffer>(
            *asGrid<float>(float_buf), channels, include_stats != 0, include_tiles != 0, 0);
        return copyOut(h, bytes);
    } catch (...) {
        *bytes = 0;
        return nullptr;
    }
}

// tools::isValid(grid, CheckMode::Full | Half) (tools/GridValidator.h:57-169): 1 valid, 0 not,
// -4 for a grid type the dispatch does not cover.
int ref_validate(const void* buf, int full)
{
    if (reinterpret_cast<const GridData*>(buf)->mGridType == GridType::PointIndex)
        return tools::isValid(asGrid<Point>(buf), full ? CheckMode::Full : CheckMode::Half, false) ? 1 : 0;
    if (reinterpret_cast<const GridData*>(buf)->mGridType == GridType::Index)
        return tools::isValid(asGrid<ValueIndex>(buf), full ? CheckMode::Full : CheckMode::Half, false) ? 1 : 0;
    return dispatch(buf, [&](auto* grid) {
        return tools::isValid(grid, full ? CheckMode::Full : CheckMode::Half, false) ? 1 : 0;
    });
}

// The host path for a voxel list: build::Grid<float>(0) with value 1 at every voxel, then
// createNanoGrid to float (default statistics) or ValueOnIndex (no stats, no tiles), as
// compareNodeOrdering in unittest/TestNanoVDB.cu builds the CPU grid it compares with
// voxelsToGrid.
void* ref_grid_from_voxels(const int32_t* ijk, uint64_t n, int grid_type, double voxel_size, uint64_t* bytes)
{
    try {
        tools::build::Grid<float> src(0.0f);
        src.setTransform(voxel_size);
        auto acc = src.getAccessor();
        for (uint64_t i = 0; i < n; ++i) acc.setValue(Coord(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]), 1.0f);
        if (grid_type == 20) {
            auto h = tools::createNanoGrid<tools::build::Grid<float>, ValueOnIndex, HostBuffer>(src, 0u, false, false, 0);
            return copyOut(h, bytes);
        }
        auto h = tools::createNanoGrid<tools::build::Grid<float>, float, HostBuffer>(src);
        return copyOut(h, bytes);
    } catch (...) {
        *bytes = 0;
        return nullptr;
    }
}

// Host tools::updateGridStats(grid, mode) in place on a float grid (tools/GridStats.h): bboxes and
// statistics in the host's merge order; mode 1 BBox, 2 MinMax, 3 All (= Default).
int ref_update_grid_stats(void* buf, int mode)
{
    if (reinterpret_cast<const GridData*>(buf)->mGridType != GridType::Float) return -4;
    if (mode < 1 || mode > 3) return -1;
    tools::updateGridStats(reinterpret_cast<NanoGrid<float>*>(buf), tools::StatsMode(mode));
    return 0;
}

// PointAccessor<AttT, Point>::voxelPoints (NanoVDB.h:5655-5716): for each voxel, the number of its
// points and the position of the first one in the grid's blind data.  AttT follows the blind
// data's value size (4: uint32 ids, 12: Vec3f, 24: Vec3d, 6: Vec3u16, 3: Vec3u8).
int ref_point_voxels(const void* buf, const int32_t* ijk, uint64_t n, uint64_t* count, uint64_t* first)
{
    const auto* grid = asGrid<Point>(buf);
    if (grid->gridType() != GridType::PointIndex || grid->blindDataCount() != 1) return -4;
    auto run = [&](auto tag) {
        using AttT = decltype(tag);
        PointAccessor<AttT, Point> acc(*grid);
        const AttT *begin = nullptr, *end = nullptr, *base = nullptr;
        acc.gridPoints(base, end);
        for (uint64_t i = 0; i < n; ++i) {
            count[i] = acc.voxelPoints(Coord(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]), begin, end);
            first[i] = count[i] ? uint64_t(begin - base) : 0;
        }
        return 0;
    };
    switch (grid->blindMetaData(0).mValueSize) {
    case 4: return run(uint32_t());
    case 12: return run(Vec3f());
    case 24: return run(Vec3d());
    case 6: return run(Vec3u16());
    case 3: return run(Vec3u8());
    default: return -4;
    }
}

// ReadAccessor<ValueOnIndex | ValueIndex>::getValue: the value index of each voxel of an index
// grid (LeafData<ValueOnIndex>::getValue, NanoVDB.h:4147-4171; LeafIndexBase for ValueIndex);
// `active` (may be NULL) receives isActive.
int ref_index_values(const void* buf, const int32_t* ijk, uint64_t n, uint64_t* out, uint8_t* active)
{
    auto run = [&](auto* grid) {
        auto acc = grid->getAccessor();
        for (uint64_t i = 0; i < n; ++i) {
            const Coord c(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
            out[i] = acc.getValue(c);
            if (active) active[i] = acc.isActive(c) ? 1 : 0;
        }
        return 0;
    };
    const GridType t = reinterpret_cast<const GridData*>(buf)->mGridType;
    if (t == GridType::OnIndex) return run(asGrid<ValueOnIndex>(buf));
    if (t == GridType::Index) return run(asGrid<ValueIndex>(buf));
    return -4;
}

// math::isoCrossing (math/HDDA.h:225-248) on world-space rays mapped by Ray::worldToIndexF.
int ref_iso_crossing(const void* buf, const float* eye, const float* dir, uint64_t n, float iso, uint8_t* hit,
                     int32_t* ijk, float* t, float* v, int nthreads)
{
    return dispatch(buf, [&](auto* grid) {
        parallelFor(n, nthreads, [&](uint64_t b, uint64_t e) {
            auto acc = makeAcc(grid);
            for (uint64_t i = b; i < e; ++i) {
                math::Ray<float> wRay(Vec3f(eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]),
                                      Vec3f(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
                math::Ray<float> iRay = wRay.worldToIndexF(*grid);
                Coord c(0);
                float vv = 0.0f, tt = 0.0f;
                const bool h = math::isoCrossing(iRay, acc, c, vv, tt, iso);
                if (hit) hit[i] = h ? 1 : 0;
                if (ijk) {
                    ijk[3 * i] = h ? c[0] : 0;
                    ijk[3 * i + 1] = h ? c[1] : 0;
                    ijk[3 * i + 2] = h ? c[2] : 0;
                }
                if (t) t[i] = h ? tt : 0.0f;
                if (v) v[i] = h ? vv : 0.0f;
            }
        });
        return 0;
    });
}

// .nvdb files through the reference's io:: functions (codec NONE).  count grid buffers are copied
// into handles; merge != 0 first merges them into one multi-grid handle (mergeGrids,
// GridHandle.h:507-544); raw != 0 writes the handle's raw buffer (GridHandle::write) instead of a
// file with segments (io::writeGrids, io/IO.h:475-486).
int ref_write_nvdb(const char* path, const void* const* bufs, const uint64_t* sizes, int count, int merge, int raw)
{
    try {
        std::vector<GridHandle<HostBuffer>> hs;
        for (int i = 0; i < count; ++i) {
            HostBuffer b = HostBuffer::create(sizes[i]);
            std::memcpy(b.data(), bufs[i], sizes[i]);
            hs.emplace_back(std::move(b));
        }
        if (merge) {
            GridHandle<HostBuffer> m = mergeGrids(hs);
            hs.clear();
            hs.push_back(std::move(m));
        }
        if (raw) {
            if (hs.size() != 1) return -2;
            hs[0].write(std::string(path));
        } else {
            io::writeGrids(std::string(path), hs, io::Codec::NONE);
        }
        return 0;
    } catch (...) {
        return -1;
    }
}

// io::readGrid(path, n) or io::readGrid(path, name) into a malloc'ed copy
void* ref_read_nvdb(const char* path, int index, const char* name, uint64_t* bytes)
{
    try {
        GridHandle<HostBuffer> h = name ? io::readGrid<HostBuffer>(std::string(path), std::string(name))
                                        : io::readGrid<HostBuffer>(std::string(path), index);
        return copyOut(h, bytes);
    } catch (...) {
        *bytes = 0;
        return nullptr;
    }
}

// Emission-absorption fog march (contract described at fogMarch above).  FP32
// outputs L,T use FP32 arithmetic (the CPU baseline); L64,T64 use FP64
// arithmetic over the identical FP32 step structure (the parity oracle).
int ref_fog_march(const void* buf, const float* eye, const float* dir, uint64_t n, float dt, float sigma, float* L,
                  float* T, double* L64, double* T64, uint32_t* steps, int nthreads)
{
    return dispatch(buf, [&](auto* grid) {
        using GridT = std::remove_cv_t<std::remove_pointer_t<decltype(grid)>>;
        parallelFor(n, nthreads, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) {
                math::Ray<float> wRay(Vec3f(eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]),
                                      Vec3f(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
                uint32_t ns = 0;
                if (L || T) {
                    float l, tr;
                    fogMarch<GridT, float>(grid, wRay, dt, sigma, l, tr, ns);
                    if (L) L[i] = l;
                    if (T) T[i] = tr;
                }
                if (L64 || T64) {
                    double l, tr;
                    fogMarch<GridT, double>(grid, wRay, dt, sigma, l, tr, ns);
                    if (L64) L64[i] = l;
                    if (T64) T64[i] = tr;
                }
                if (steps) steps[i] = ns;
            }
        });
        return 0;
    });
}

// RayGenOp of the level-set example: eye = centre + (0,0,wBBoxDimZ), pinhole FOV 45
int ref_ray_gen(int w, int h, float wBBoxDimZ, float cx, float cy, float cz, float* eye, float* dir)
{
    RayGenOp<Vec3f> op(wBBoxDimZ, Vec3f(cx, cy, cz));
    for (int i = 0; i < w * h; ++i) {
        Vec3f o, d;
        op(i, w, h, o, d);
        for (int a = 0; a < 3; ++a) {
            eye[3 * i + a] = o[a];
            dir[3 * i + a] = d[a];
        }
    }
    return 0;
}

// TreeMarcher<NodeT, Ray, Acc>::init/step (math/HDDA.h:450-536) for world rays mapped with
// worldToIndexF: level 0 LeafNode, 1 lower, 2 upper.  count[i] spans; the first max_spans are
// written at [i*max_spans ...) as t0, t1 and the node's byte offset from the buffer start.
int ref_tree_march(const void* buf, const float* eye, const float* dir, uint64_t n, int level, uint32_t max_spans,
                   uint32_t* count, float* t0, float* t1, int64_t* node)
{
    return dispatch(buf, [&](auto* grid) {
        using GridT = std::remove_pointer_t<decltype(grid)>;
        using BuildT = typename GridT::BuildType;
        if constexpr (std::is_same_v<BuildT, ValueOnIndex>) {
            return -4;
        } else {
            auto run = [&](auto tag) {
                using NodeT = typename decltype(tag)::type;
                for (uint64_t i = 0; i < n; ++i) {
                    auto acc = grid->getAccessor();
                    math::Ray<float> wRay(Vec3f(eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]),
                                          Vec3f(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
                    math::TreeMarcher<NodeT, math::Ray<float>, decltype(acc)> marcher(acc);
                    uint32_t c = 0;
                    if (marcher.init(wRay.worldToIndexF(*grid))) {
                        const NodeT* nd = nullptr;
                        float a = 0.0f, b = 0.0f;
                        while (marcher.step(&nd, a, b)) {
                            if (c < max_spans) {
                                t0[i * max_spans + c] = a;
                                t1[i * max_spans + c] = b;
                                node[i * max_spans + c] =
                                    int64_t(reinterpret_cast<const uint8_t*>(nd) - reinterpret_cast<const uint8_t*>(buf));
                            }
                            ++c;
                        }
                    }
                    count[i] = c;
                }
                return 0;
            };
            struct Leaf { using type = typename GridT::LeafNodeType; };
            struct Lower { using type = NanoLower<BuildT>; };
            struct Upper { using type = NanoUpper<BuildT>; };
            if (level == 0) return run(Leaf{});
            if (level == 1) return run(Lower{});
            return run(Upper{});
        }
    });
}

// math::firstActive (math/HDDA.h:429-444) for world rays mapped with worldToIndexF; zeros on a miss
int ref_first_active(const void* buf, const float* eye, const float* dir, uint64_t n, uint8_t* hit, int32_t* ijk,
                     float* t)
{
    return dispatch(buf, [&](auto* grid) {
        auto acc = makeAcc(grid);
        for (uint64_t i = 0; i < n; ++i) {
            math::Ray<float> wRay(Vec3f(eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]),
                                  Vec3f(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
            math::Ray<float> iRay = wRay.worldToIndexF(*grid);
            Coord c(0);
            float tt = 0.0f;
            const bool h = math::firstActive(iRay, acc, c, tt);
            hit[i] = h ? 1 : 0;
            for (int a = 0; a < 3; ++a) ijk[3 * i + a] = h ? c[a] : 0;
            t[i] = h ? tt : 0.0f;
        }
        return 0;
    });
}

// DDA<Ray, Coord, Dim> known-answer traces (math/HDDA.h:258-379), format of ref_hdda_trace; dim 1, 8, 128
int ref_dda_trace(const float* eye, const float* dir, float t0, float t1, int dim, int nsteps, float* times,
                  float* nexts, int32_t* voxels, uint8_t* ok)
{
    math::Ray<float> ray(Vec3f(eye[0], eye[1], eye[2]), Vec3f(dir[0], dir[1], dir[2]), t0, t1);
    auto run = [&](auto dda) {
        for (int s = 0; s <= nsteps; ++s) {
            if (s > 0) ok[s] = dda.step() ? 1 : 0;
            else ok[0] = 1;
            times[s] = dda.time();
            nexts[s] = dda.next();
            for (int a = 0; a < 3; ++a) voxels[3 * s + a] = dda.voxel()[a];
        }
        return 0;
    };
    if (dim == 1) return run(math::DDA<math::Ray<float>, Coord, 1>(ray));
    if (dim == 8) return run(math::DDA<math::Ray<float>, Coord, 8>(ray));
    if (dim == 128) return run(math::DDA<math::Ray<float>, Coord, 128>(ray));
    return -1;
}

// BoxStencil<NanoGrid<float>> (math/Stencils.h:306-432): moveTo(RoundDown(xyz)), interpolation(xyz),
// gradient(xyz).  BaseStencil::moveTo(const Vec3<RealT>&) (Stencils.h:146-152) does not compile
// (it calls RoundDown without its CoordT argument), so the cell is moved to explicitly.
int ref_box_stencil(const void* buf, const float* xyz, uint64_t n, float* interp, float* grad)
{
    const auto* grid = asGrid<float>(buf);
    math::BoxStencil<NanoGrid<float>> st(*grid);
    for (uint64_t i = 0; i < n; ++i) {
        const Vec3f p(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        st.moveTo(math::RoundDown<Coord>(p));
        if (interp) interp[i] = st.interpolation(p);
        if (grad) {
            const Vec3f g = st.gradient(p);
            for (int a = 0; a < 3; ++a) grad[3 * i + a] = g[a];
        }
    }
    return 0;
}

// GradStencil::gradient(V) (upwind) and cpt() at integer voxels (math/Stencils.h:512-518, 544-554)
int ref_grad_stencil_ext(const void* buf, const int32_t* ijk, uint64_t n, const float* V, float* upwind, float* cpt)
{
    const auto* grid = asGrid<float>(buf);
    math::GradStencil<NanoGrid<float>> st(*grid);
    for (uint64_t i = 0; i < n; ++i) {
        st.moveTo(Coord(ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]));
        if (upwind) {
            const Vec3f g = st.gradient(Vec3f(V[3 * i], V[3 * i + 1], V[3 * i + 2]));
            for (int a = 0; a < 3; ++a) upwind[3 * i + a] = g[a];
        }
        if (cpt) {
            const Vec3f c = st.cpt();
            for (int a = 0; a < 3; ++a) cpt[3 * i + a] = c[a];
        }
    }
    return 0;
}

// SampleFromVoxels<Acc,1,true>::zeroCrossing(xyz) (math/SampleFromVoxels.h:394-399)
int ref_sample_zero_crossing(const void* buf, const float* xyz, uint64_t n, uint8_t* zc)
{
    return dispatch(buf, [&](auto* grid) {
        auto acc = makeAcc(grid);
        auto s = math::createSampler<1>(acc);
        for (uint64_t i = 0; i < n; ++i) zc[i] = s.zeroCrossing(Vec3f(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2])) ? 1 : 0;
        return 0;
    });
}

// HDDA known-answer traces: init(ray(eye,dir,t0,t1), dim) then nsteps of step(),
// recording time(), next() and voxel() after each call (times[0] is after init).
int ref_hdda_trace(const float* eye, const float* dir, float t0, float t1, int dim, int nsteps, float* times,
                   float* nexts, int32_t* voxels, uint8_t* ok)
{
    math::Ray<float> ray(Vec3f(eye[0], eye[1], eye[2]), Vec3f(dir[0], dir[1], dir[2]), t0, t1);
    math::HDDA<math::Ray<float>, Coord> hdda(ray, dim);
    for (int s = 0; s <= nsteps; ++s) {
        if (s > 0) ok[s] = hdda.step() ? 1 : 0;
        else ok[0] = 1;
        times[s] = hdda.time();
        nexts[s] = hdda.next();
        for (int a = 0; a < 3; ++a) voxels[3 * s + a] = hdda.voxel()[a];
    }
    return 0;
}

// Ray::clip against a CoordBBox (max padded by one voxel): returns hit, writes t0,t1
int ref_ray_clip(const float* eye, const float* dir, float t0, float t1, const int32_t* bbox, float* out)
{
    math::Ray<float> ray(Vec3f(eye[0], eye[1], eye[2]), Vec3f(dir[0], dir[1], dir[2]), t0, t1);
    const bool hit = ray.clip(CoordBBox(Coord(bbox[0], bbox[1], bbox[2]), Coord(bbox[3], bbox[4], bbox[5])));
    out[0] = ray.t0();
    out[1] = ray.t1();
    return hit ? 1 : 0;
}

} // extern "C"
