// nvq device API: the in-kernel call shapes of NanoVDB's read path for client HIP kernels.
//
// NanoVDB is consumed mainly as header-only __hostdev__ templates inside the client's own
// kernels.  This header keeps those call shapes over an uploaded grid (nvq_grid of nvq.h), so a
// kernel written against NanoVDB ports by renaming the namespace:
//
//   reference (nanovdb)                                     here (nvq)
//   const NanoGrid<float>* grid  (kernel argument)          nvq::FloatGrid grid   (by value)
//   auto acc = grid->getAccessor();        NanoVDB.h:2210    auto acc = grid.getAccessor();
//   acc.getValue(ijk) / isActive / probeValue  :5370-5377    same
//   acc.getDim(ijk, ray)                       :5431-5447    same
//   acc.getNode<NodeT>()                       :5317-5322    acc.getNode<LEVEL>()  (byte pointer)
//   math::createSampler<0|1>(acc)   math/SampleFromVoxels.h:49-53   nvq::math::createSampler<0|1>(acc)
//     s(xyz), s(ijk), s.gradient(xyz), s.zeroCrossing()    :325-476       same
//   math::GradStencil<GridT> st(grid); st.moveTo(ijk);  math/Stencils.h:455-576   same
//     st.gradient(), normSqGrad(), laplacian(), zeroCrossing(), cpt()
//   math::BoxStencil<GridT>: interpolation(xyz), gradient(xyz)  math/Stencils.h:306-432   same
//   math::Ray<float> ray(eye, dir); ray.worldToIndexF(*grid)    math/Ray.h:45-56, 258-268   same
//   ray.applyInverseMap(grid->map())  (FP64 map)           math/Ray.h:223-230   ray.applyInverseMap(grid)
//   math::zeroCrossing(ray, acc, ijk, v, t)        math/HDDA.h:189-214   same
//   math::isoCrossing(ray, acc, ijk, v, t, iso)    math/HDDA.h:225-248   same
//   math::firstActive(ray, acc, ijk, t)            math/HDDA.h:429-444   same
//   math::HDDA / math::DDA<RayT, Coord, Dim>       math/HDDA.h:28-179, 258-379   same
//   math::TreeMarcher<NodeT, RayT, AccT>::init/step  math/HDDA.h:450-536   same, NodeT = nvq::LeafNode /
//                                                   nvq::LowerNode / nvq::UpperNode, step gives byte pointers
//
// The host obtains the kernel argument from an uploaded grid with nvq::getDeviceGrid(handle, &grid),
// which, like GridHandle::deviceGrid<T>() (GridHandle.h:159-170), fails on a type mismatch.
//
// Values, active states, getDim results, samples, stencils, hit voxels and hit times are
// bit-identical to the reference host code: the float operation order is the reference's and the
// including translation unit must be compiled with -ffp-contract=off (fmaf is written where the
// reference writes fmaf).  The accessor keeps one Coord key and three node pointers (the
// NANOVDB_USE_SINGLE_ACCESSOR_KEY form, NanoVDB.h:5329-5348); returned values do not depend on
// the cache (NanoVDB.h:5379-5403).  Like the reference's, an accessor belongs to one thread.
#pragma once
#include "nvq.h"
#include "nvq_device_impl.h"

namespace nvq {

// ------------------------------------------------------------------ Coord / Vec3f
// nanovdb::math::Coord (math/Math.h) and Vec3<float>, the parts the read path uses.
struct Coord {
    int32_t mVec[3];
    __host__ __device__ Coord() : mVec{0, 0, 0} {}
    __host__ __device__ Coord(int32_t i, int32_t j, int32_t k) : mVec{i, j, k} {}
    __host__ __device__ static Coord max() { return Coord(0x7fffffff, 0x7fffffff, 0x7fffffff); }
    __host__ __device__ int32_t& operator[](int i) { return mVec[i]; }
    __host__ __device__ const int32_t& operator[](int i) const { return mVec[i]; }
    __host__ __device__ int32_t x() const { return mVec[0]; }
    __host__ __device__ int32_t y() const { return mVec[1]; }
    __host__ __device__ int32_t z() const { return mVec[2]; }
    __host__ __device__ Coord offsetBy(int32_t dx, int32_t dy, int32_t dz) const
    {
        return Coord(mVec[0] + dx, mVec[1] + dy, mVec[2] + dz);
    }
    __host__ __device__ bool operator==(const Coord& o) const
    {
        return mVec[0] == o.mVec[0] && mVec[1] == o.mVec[1] && mVec[2] == o.mVec[2];
    }
    __host__ __device__ bool operator!=(const Coord& o) const { return !(*this == o); }
};

struct Vec3f {
    float mVec[3];
    __host__ __device__ Vec3f() : mVec{0.0f, 0.0f, 0.0f} {}
    __host__ __device__ explicit Vec3f(float v) : mVec{v, v, v} {}
    __host__ __device__ Vec3f(float x, float y, float z) : mVec{x, y, z} {}
    __host__ __device__ float& operator[](int i) { return mVec[i]; }
    __host__ __device__ const float& operator[](int i) const { return mVec[i]; }
    __host__ __device__ Vec3f operator*(float s) const { return Vec3f(mVec[0] * s, mVec[1] * s, mVec[2] * s); }
    __host__ __device__ float lengthSqr() const { return mVec[0] * mVec[0] + mVec[1] * mVec[1] + mVec[2] * mVec[2]; }
    __host__ __device__ float length() const { return sqrtf(lengthSqr()); }
};

// Node levels for ReadAccessor::getNode and TreeMarcher (LeafNode, InternalNode<Leaf,4>,
// InternalNode<Lower,5> of NanoVDB.h): LEVEL, the node's edge DIM, and the edge of its children.
struct LeafNode {
    static constexpr int LEVEL = 0;
    static constexpr uint32_t DIM = 8, CHILD_DIM = 1;
};
struct LowerNode {
    static constexpr int LEVEL = 1;
    static constexpr uint32_t DIM = 128, CHILD_DIM = 8;
};
struct UpperNode {
    static constexpr int LEVEL = 2;
    static constexpr uint32_t DIM = 4096, CHILD_DIM = 128;
};

namespace math {
class Ray;
} // namespace math

// ------------------------------------------------------------------ ReadAccessor
// ReadAccessor<BuildT, 0, 1, 2> (NanoVDB.h:5252-5469) over the grid type GT (NVQ_FLOAT, NVQ_FP4,
// NVQ_FP8, NVQ_FP16, NVQ_FPN or NVQ_ONINDEX; index grids return channel values).
template<uint32_t GT>
class ReadAccessor {
    mutable Accessor<GT> mAcc;
    const GridView* mGrid;

public:
    using ValueType = float;
    using CoordType = Coord;
    static constexpr uint32_t kType = GT;

    __device__ explicit ReadAccessor(const GridView& grid) : mAcc(grid), mGrid(&grid) {}

    __device__ const GridView& grid() const { return *mGrid; }
    // the root's index bbox, RootData::mBBox (acc.root().bbox() of the reference)
    __device__ const int32_t* rootBBox() const { return mGrid->bbox; }

    __device__ float getValue(const Coord& ijk) const { return mAcc.getValue(ijk[0], ijk[1], ijk[2]); }
    __device__ float getValue(int i, int j, int k) const { return mAcc.getValue(i, j, k); }
    __device__ float operator()(const Coord& ijk) const { return getValue(ijk); }
    __device__ float operator()(int i, int j, int k) const { return mAcc.getValue(i, j, k); }
    __device__ bool isActive(const Coord& ijk) const { return mAcc.probe(ijk[0], ijk[1], ijk[2]).active; }
    // ProbeValue (NanoVDB.h:6314-6345): the value and the active state
    __device__ bool probeValue(const Coord& ijk, float& v) const
    {
        const Probe p = mAcc.probe(ijk[0], ijk[1], ijk[2]);
        v = p.value;
        return p.active;
    }
    // GetDim (NanoVDB.h:5431-5447): 1 in a leaf, else the edge of the tile or 4096 outside the tree;
    // the ray is not used by the node types of these grids
    template<class RayT>
    __device__ uint32_t getDim(const Coord& ijk, const RayT&) const
    {
        return mAcc.probe(ijk[0], ijk[1], ijk[2]).dim;
    }
    // probeLeaf (GetLeaf, NanoVDB.h:6282-6295): the leaf holding ijk, or nullptr
    __device__ const uint8_t* probeLeaf(const Coord& ijk) const { return mAcc.probe(ijk[0], ijk[1], ijk[2]).leaf; }
    // the node of LEVEL (0 leaf, 1 lower, 2 upper) cached by the last lookup, or nullptr
    template<int LEVEL>
    __device__ const uint8_t* getNode() const
    {
        static_assert(LEVEL >= 0 && LEVEL <= 2, "ReadAccessor::getNode: invalid level");
        if constexpr (LEVEL == 0) return mAcc.leaf;
        else if constexpr (LEVEL == 1) return mAcc.lower;
        else return mAcc.upper;
    }
    template<class NodeT>
    __device__ const uint8_t* getNode() const { return getNode<NodeT::LEVEL>(); }
    // the full lookup record (value, active, dim, leaf, voxel offset, channel index)
    __device__ Probe probe(const Coord& ijk) const { return mAcc.probe(ijk[0], ijk[1], ijk[2]); }
    __device__ void clear() const
    {
        mAcc.kx = mAcc.ky = mAcc.kz = 0x7fffffff;
        mAcc.upper = mAcc.lower = mAcc.leaf = nullptr;
    }
};

// ------------------------------------------------------------------ Grid
// The kernel argument standing in for `const NanoGrid<BuildT>*`: a by-value view of an uploaded
// grid (device pointers into its buffer plus the Map and root bbox).
template<uint32_t GT>
struct Grid {
    GridView view;
    static constexpr uint32_t kType = GT;
    using AccessorType = ReadAccessor<GT>;
    using ValueType = float;

    __device__ ReadAccessor<GT> getAccessor() const { return ReadAccessor<GT>(view); }
    __host__ __device__ Vec3f voxelSize() const
    {
        return Vec3f(float(view.voxelSize[0]), float(view.voxelSize[1]), float(view.voxelSize[2]));
    }
    __host__ __device__ const double* voxelSizeD() const { return view.voxelSize; }
    __host__ __device__ const int32_t* indexBBox() const { return view.bbox; }
    // Grid::worldToIndexF / worldToIndexDirF (Map::applyInverseMapF / applyInverseJacobianF,
    // NanoVDB.h:1518-1521, :1540; math/Math.h:1704-1709)
    __device__ Vec3f worldToIndexF(const Vec3f& p) const
    {
        const float t[3] = {p[0] - view.vecF[0], p[1] - view.vecF[1], p[2] - view.vecF[2]};
        return matMultF(t);
    }
    __device__ Vec3f worldToIndexDirF(const Vec3f& d) const
    {
        const float t[3] = {d[0], d[1], d[2]};
        return matMultF(t);
    }
    // Grid::worldToIndex / worldToIndexDir in FP64 (Map::applyInverseMap / applyInverseJacobian,
    // NanoVDB.h:1507-1510, :1530; math/Math.h:1718-1723): float(xyz - vecD) then a double
    // fma chain rounded to float, as the reference evaluates them on a Vec3f
    __device__ Vec3f worldToIndex(const Vec3f& p) const
    {
        const float t[3] = {float(double(p[0]) - view.vecD[0]), float(double(p[1]) - view.vecD[1]),
                            float(double(p[2]) - view.vecD[2])};
        return matMultD(t);
    }
    __device__ Vec3f worldToIndexDir(const Vec3f& d) const
    {
        const float t[3] = {d[0], d[1], d[2]};
        return matMultD(t);
    }
    // Grid::indexToWorldGradF / indexToWorldGrad (Map::applyIJTF / applyIJT, NanoVDB.h:1548-1550;
    // matMultT, math/Math.h:1758-1776)
    __device__ Vec3f indexToWorldGradF(const Vec3f& g) const
    {
        const float* m = view.invMatF;
        return Vec3f(__builtin_fmaf(g[0], m[0], __builtin_fmaf(g[1], m[3], g[2] * m[6])),
                     __builtin_fmaf(g[0], m[1], __builtin_fmaf(g[1], m[4], g[2] * m[7])),
                     __builtin_fmaf(g[0], m[2], __builtin_fmaf(g[1], m[5], g[2] * m[8])));
    }
    __device__ Vec3f indexToWorldGrad(const Vec3f& g) const
    {
        const double* m = view.invMatD;
        const double x = g[0], y = g[1], z = g[2];
        return Vec3f(float(__builtin_fma(x, m[0], __builtin_fma(y, m[3], z * m[6]))),
                     float(__builtin_fma(x, m[1], __builtin_fma(y, m[4], z * m[7]))),
                     float(__builtin_fma(x, m[2], __builtin_fma(y, m[5], z * m[8]))));
    }

private:
    __device__ Vec3f matMultF(const float t[3]) const
    {
        const float* m = view.invMatF;
        return Vec3f(__builtin_fmaf(t[0], m[0], __builtin_fmaf(t[1], m[1], t[2] * m[2])),
                     __builtin_fmaf(t[0], m[3], __builtin_fmaf(t[1], m[4], t[2] * m[5])),
                     __builtin_fmaf(t[0], m[6], __builtin_fmaf(t[1], m[7], t[2] * m[8])));
    }
    __device__ Vec3f matMultD(const float t[3]) const
    {
        const double* m = view.invMatD;
        const double x = t[0], y = t[1], z = t[2];
        return Vec3f(float(__builtin_fma(x, m[0], __builtin_fma(y, m[1], z * m[2]))),
                     float(__builtin_fma(x, m[3], __builtin_fma(y, m[4], z * m[5]))),
                     float(__builtin_fma(x, m[6], __builtin_fma(y, m[7], z * m[8]))));
    }
};
using FloatGrid = Grid<GT_FLOAT>;
using Fp4Grid = Grid<GT_FP4>;
using Fp8Grid = Grid<GT_FP8>;
using Fp16Grid = Grid<GT_FP16>;
using FpNGrid = Grid<GT_FPN>;
using OnIndexGrid = Grid<GT_ONINDEX>;

// Host: the kernel argument for an uploaded grid of type GT; NVQ_ERR_UNSUPPORTED_TYPE on a type
// mismatch (GridHandle::deviceGrid<T>() returns nullptr then, GridHandle.h:159-170).
template<uint32_t GT>
inline int getDeviceGrid(const nvq_grid* handle, Grid<GT>* out)
{
    if (!handle || !out) return NVQ_ERR_INVALID_ARG;
    GridView v;
    const int st = nvq_grid_get_view(handle, &v, sizeof(GridView));
    if (st != NVQ_OK) return st;
    if (v.type != GT) return NVQ_ERR_UNSUPPORTED_TYPE;
    out->view = v;
    return NVQ_OK;
}

namespace math {

// RoundDown<Coord> (math/Math.h:269-273), Round<Coord> for Vec3<float> (rintf, :255-261),
// Floor<Coord>(Vec3f&) returning the cell and leaving the fraction in xyz (SampleFromVoxels.h:61-69)
__device__ inline Coord RoundDown(const Vec3f& xyz)
{
    return Coord(int32_t(floorf(xyz[0])), int32_t(floorf(xyz[1])), int32_t(floorf(xyz[2])));
}
__device__ inline Coord Round(const Vec3f& xyz)
{
    return Coord(int32_t(rintf(xyz[0])), int32_t(rintf(xyz[1])), int32_t(rintf(xyz[2])));
}
__device__ inline Coord Floor(Vec3f& xyz)
{
    const float f[3] = {floorf(xyz[0]), floorf(xyz[1]), floorf(xyz[2])};
    xyz[0] -= f[0];
    xyz[1] -= f[1];
    xyz[2] -= f[2];
    return Coord(int32_t(f[0]), int32_t(f[1]), int32_t(f[2]));
}

// ------------------------------------------------------------------ Ray
// math::Ray<float> (math/Ray.h:20-520): eye, direction, its reciprocal and the time span
// [t0, t1] (defaults 1e-5 and the host FLT_MAX, math/Math.h:96-124).
class Ray {
public:
    nvq::Ray r;

    __device__ Ray() : Ray(Vec3f(0.0f), Vec3f(1.0f, 0.0f, 0.0f)) {}
    __device__ Ray(const Vec3f& eye, const Vec3f& dir, float t0 = 1.0e-5f, float t1 = kFltMax)
    {
        const float e[3] = {eye[0], eye[1], eye[2]}, d[3] = {dir[0], dir[1], dir[2]};
        r = makeRay(e, d, t0, t1);
    }
    __device__ explicit Ray(const nvq::Ray& raw) : r(raw) {}

    __device__ Vec3f eye() const { return Vec3f(r.eye[0], r.eye[1], r.eye[2]); }
    __device__ Vec3f dir() const { return Vec3f(r.dir[0], r.dir[1], r.dir[2]); }
    __device__ Vec3f invDir() const { return Vec3f(r.inv[0], r.inv[1], r.inv[2]); }
    __device__ float t0() const { return r.t0; }
    __device__ float t1() const { return r.t1; }
    // operator()(t) = eye + t*dir with fmaf (math/Ray.h:154-163)
    __device__ Vec3f operator()(float t) const
    {
        float p[3];
        rayAt(r, t, p);
        return Vec3f(p[0], p[1], p[2]);
    }
    __device__ Vec3f start() const { return (*this)(r.t0); }
    __device__ Vec3f end() const { return (*this)(r.t1); }
    __device__ Ray& setTimes(float t0 = 1.0e-5f, float t1 = kFltMax)
    {
        r.t0 = t0;
        r.t1 = t1;
        return *this;
    }
    // intersects / clip against an index bbox [min, max + 1] (math/Ray.h:398-424, 506-515)
    __device__ bool clip(const int32_t* bbox) { return rayClip(r, bbox); }
    // Ray::worldToIndexF (math/Ray.h:258-268): FP32 map of eye and direction, direction
    // normalised, times scaled by the length (t1 left at FLT_MAX)
    template<uint32_t GT>
    __device__ Ray worldToIndexF(const Grid<GT>& grid) const
    {
        return Ray(mapped(grid.worldToIndexF(eye()), grid.worldToIndexDirF(dir())));
    }
    // Ray::applyInverseMap(map) (math/Ray.h:223-230): the FP64 map (Map::applyInverseMap /
    // applyInverseJacobian, NanoVDB.h:1507-1510, :1530) of eye and direction, for eyes far from
    // the origin of a grid with a large translation, where xyz - vecF in FP32 loses the voxel.
    // As in the reference, both times are scaled by the length (t1 = FLT_MAX may become inf).
    template<uint32_t GT>
    __device__ Ray applyInverseMap(const Grid<GT>& grid) const
    {
        return Ray(mapped(grid.worldToIndex(eye()), grid.worldToIndexDir(dir()), false));
    }

private:
    __device__ nvq::Ray mapped(const Vec3f& e, const Vec3f& d, bool keepMaxT1 = true) const
    {
        const float length = d.length(), invLength = 1.0f / length;
        const float ee[3] = {e[0], e[1], e[2]}, dd[3] = {invLength * d[0], invLength * d[1], invLength * d[2]};
        const float t1 = (!keepMaxT1 || r.t1 < kFltMax) ? length * r.t1 : r.t1;
        return makeRay(ee, dd, length * r.t0, t1);
    }
};

// ------------------------------------------------------------------ samplers
// SampleFromVoxels<AccT, Order, UseCache> (math/SampleFromVoxels.h:92-476); AccT is a ReadAccessor.
template<class AccT, int Order, bool UseCache = true>
class SampleFromVoxels;

// Order 0: the value at Round(xyz) (SampleFromVoxels.h:92-165)
template<class AccT, bool UseCache>
class SampleFromVoxels<AccT, 0, UseCache> {
    const AccT& mAcc;
    mutable Coord mPos;
    mutable float mVal;

public:
    __device__ explicit SampleFromVoxels(const AccT& acc) : mAcc(acc), mPos(Coord::max()), mVal(0.0f) {}
    __device__ const AccT& accessor() const { return mAcc; }
    __device__ float operator()(const Vec3f& xyz) const { return (*this)(Round(xyz)); }
    __device__ float operator()(const Coord& ijk) const
    {
        if constexpr (!UseCache) return mAcc.getValue(ijk);
        if (ijk != mPos) {
            mPos = ijk;
            mVal = mAcc.getValue(ijk);
        }
        return mVal;
    }
};

// Order 1: trilinear (TrilinearSampler, SampleFromVoxels.h:178-288; uncached :295-321, 423-476;
// cached :325-407)
template<class AccT, bool UseCache>
class SampleFromVoxels<AccT, 1, UseCache> {
    const AccT& mAcc;
    mutable Coord mPos;
    mutable Cell mVal;

    // TrilinearSampler::stencil: the eight corners of the cell at ijk
    __device__ void stencil(const Coord& ijk, Cell& c) const
    {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) c.v[a][b][e] = mAcc.getValue(ijk[0] + a, ijk[1] + b, ijk[2] + e);
    }
    __device__ const Cell& cache(Vec3f& xyz, Cell& local) const
    {
        const Coord ijk = Floor(xyz);
        if constexpr (UseCache) {
            if (ijk != mPos) {
                mPos = ijk;
                stencil(ijk, mVal);
            }
            return mVal;
        } else {
            stencil(ijk, local);
            return local;
        }
    }
    __device__ static bool crossing(const Cell& v)
    {
        const bool less = v.v[0][0][0] < 0.0f;
        return (less ^ (v.v[0][0][1] < 0.0f)) || (less ^ (v.v[0][1][1] < 0.0f)) || (less ^ (v.v[0][1][0] < 0.0f)) ||
               (less ^ (v.v[1][0][0] < 0.0f)) || (less ^ (v.v[1][0][1] < 0.0f)) || (less ^ (v.v[1][1][1] < 0.0f)) ||
               (less ^ (v.v[1][1][0] < 0.0f));
    }

public:
    __device__ explicit SampleFromVoxels(const AccT& acc) : mAcc(acc), mPos(Coord::max()) {}
    __device__ const AccT& accessor() const { return mAcc; }

    // the trilinear value at index-space xyz (lerp a + w*(b - a), z then y then x)
    __device__ float operator()(Vec3f xyz) const
    {
        Cell local;
        const Cell& c = cache(xyz, local);
        return trilerp(c, xyz[0], xyz[1], xyz[2]);
    }
    __device__ float operator()(const Coord& ijk) const
    {
        if constexpr (UseCache) return ijk == mPos ? mVal.v[0][0][0] : mAcc.getValue(ijk);
        else return mAcc.getValue(ijk);
    }
    // TrilinearSampler::gradient (index space, SampleFromVoxels.h:232-271)
    __device__ Vec3f gradient(Vec3f xyz) const
    {
        Cell local;
        const Cell& c = cache(xyz, local);
        float g[3];
        trilinearGradient(c, xyz[0], xyz[1], xyz[2], g);
        return Vec3f(g[0], g[1], g[2]);
    }
    // true if the corners of the cell at xyz change sign (SampleFromVoxels.h:274-288)
    __device__ bool zeroCrossing(Vec3f xyz) const
    {
        Cell local;
        return crossing(cache(xyz, local));
    }
    // the cached cell's sign test (cached sampler only)
    __device__ bool zeroCrossing() const
    {
        static_assert(UseCache, "zeroCrossing() needs the cached sampler");
        return crossing(mVal);
    }
};

template<int Order, class AccT, bool UseCache = true>
__device__ inline SampleFromVoxels<AccT, Order, UseCache> createSampler(const AccT& acc)
{
    return SampleFromVoxels<AccT, Order, UseCache>(acc);
}

// ------------------------------------------------------------------ stencils
// BaseStencil (math/Stencils.h:97-297): a value buffer around a centre voxel, filled through the
// stencil's own accessor.
template<class Derived, int SIZE, class GridT>
class BaseStencil {
public:
    using GridType = GridT;
    using ValueType = float;

    __device__ void moveTo(const Coord& ijk)
    {
        mCenter = ijk;
        mValues[0] = mAcc.getValue(ijk);
        static_cast<Derived&>(*this).init(mCenter);
    }
    __device__ void moveTo(const Coord& ijk, float centerValue)
    {
        mCenter = ijk;
        mValues[0] = centerValue;
        static_cast<Derived&>(*this).init(mCenter);
    }
    // RoundDown(xyz); the buffer is refilled only when the centre voxel changes
    __device__ void moveTo(const Vec3f& xyz)
    {
        const Coord ijk = RoundDown(xyz);
        if (ijk != mCenter) moveTo(ijk);
    }
    __device__ float getValue(unsigned int pos = 0) const { return mValues[pos]; }
    template<int i, int j, int k>
    __device__ float getValue() const
    {
        return mValues[Derived::template pos<i, j, k>()];
    }
    template<int i, int j, int k>
    __device__ void setValue(float v)
    {
        mValues[Derived::template pos<i, j, k>()] = v;
    }
    __device__ static int size() { return SIZE; }
    __device__ float mean() const
    {
        float sum = 0.0f;
        for (int i = 0; i < SIZE; ++i) sum += mValues[i];
        return sum / float(SIZE);
    }
    __device__ float min() const
    {
        float v = mValues[0];
        for (int i = 1; i < SIZE; ++i)
            if (mValues[i] < v) v = mValues[i];
        return v;
    }
    __device__ float max() const
    {
        float v = mValues[0];
        for (int i = 1; i < SIZE; ++i)
            if (mValues[i] > v) v = mValues[i];
        return v;
    }
    __device__ const Coord& getCenterCoord() const { return mCenter; }
    __device__ float getCenterValue() const { return mValues[0]; }
    __device__ const GridT& grid() const { return *mGrid; }
    __device__ const typename GridT::AccessorType& accessor() const { return mAcc; }

protected:
    __device__ explicit BaseStencil(const GridT& grid) : mGrid(&grid), mAcc(grid.view), mCenter(Coord::max()) {}

    const GridT* mGrid;
    typename GridT::AccessorType mAcc;
    float mValues[SIZE];
    Coord mCenter;
};

// BoxStencil (math/Stencils.h:306-432): the eight corners of the cell at the centre voxel, in the
// order (000, 001, 011, 010, 100, 101, 111, 110).
template<class GridT>
class BoxStencil : public BaseStencil<BoxStencil<GridT>, 8, GridT> {
    using Base = BaseStencil<BoxStencil<GridT>, 8, GridT>;
    friend Base;
    using Base::mAcc;
    using Base::mCenter;
    using Base::mValues;

    __device__ void init(const Coord& ijk)
    {
        mValues[1] = mAcc.getValue(ijk.offsetBy(0, 0, 1));
        mValues[2] = mAcc.getValue(ijk.offsetBy(0, 1, 1));
        mValues[3] = mAcc.getValue(ijk.offsetBy(0, 1, 0));
        mValues[4] = mAcc.getValue(ijk.offsetBy(1, 0, 0));
        mValues[5] = mAcc.getValue(ijk.offsetBy(1, 0, 1));
        mValues[6] = mAcc.getValue(ijk.offsetBy(1, 1, 1));
        mValues[7] = mAcc.getValue(ijk.offsetBy(1, 1, 0));
    }

public:
    static constexpr int SIZE = 8;
    __device__ explicit BoxStencil(const GridT& grid) : Base(grid) {}

    template<int i, int j, int k>
    __device__ static constexpr unsigned int pos()
    {
        static_assert(i >= 0 && i <= 1 && j >= 0 && j <= 1 && k >= 0 && k <= 1, "BoxStencil: corner out of range");
        return i == 0 ? (j == 0 ? unsigned(k) : 3u - unsigned(k)) : (j == 0 ? 4u + unsigned(k) : 7u - unsigned(k));
    }
    __device__ bool intersects(float iso = 0.0f) const
    {
        const bool less = mValues[0] < iso;
        return (less ^ (mValues[1] < iso)) || (less ^ (mValues[2] < iso)) || (less ^ (mValues[3] < iso)) ||
               (less ^ (mValues[4] < iso)) || (less ^ (mValues[5] < iso)) || (less ^ (mValues[6] < iso)) ||
               (less ^ (mValues[7] < iso));
    }
    // trilinear interpolation at index-space xyz inside the cell (u = x - centre x, ...)
    __device__ float interpolation(const Vec3f& xyz) const
    {
        const float u = xyz[0] - float(mCenter[0]), v = xyz[1] - float(mCenter[1]), w = xyz[2] - float(mCenter[2]);
        float V = mValues[0];
        float A = V + (mValues[1] - V) * w;
        V = mValues[3];
        float B = V + (mValues[2] - V) * w;
        const float C = A + (B - A) * v;
        V = mValues[4];
        A = V + (mValues[5] - V) * w;
        V = mValues[7];
        B = V + (mValues[6] - V) * w;
        const float D = A + (B - A) * v;
        return C + (D - C) * u;
    }
    // the world-space gradient of the interpolation: index-space partials through the FP64
    // Map::applyIJT (math/Stencils.h:397-431)
    __device__ Vec3f gradient(const Vec3f& xyz) const
    {
        const float u = xyz[0] - float(mCenter[0]), v = xyz[1] - float(mCenter[1]), w = xyz[2] - float(mCenter[2]);
        float D[4] = {mValues[1] - mValues[0], mValues[2] - mValues[3], mValues[5] - mValues[4], mValues[6] - mValues[7]};
        float A = D[0] + (D[1] - D[0]) * v;
        float B = D[2] + (D[3] - D[2]) * v;
        Vec3f grad(0.0f, 0.0f, A + (B - A) * u);
        D[0] = mValues[0] + D[0] * w;
        D[1] = mValues[3] + D[1] * w;
        D[2] = mValues[4] + D[2] * w;
        D[3] = mValues[7] + D[3] * w;
        A = D[0] + (D[1] - D[0]) * v;
        B = D[2] + (D[3] - D[2]) * v;
        grad[0] = B - A;
        A = D[1] - D[0];
        B = D[3] - D[2];
        grad[1] = A + (B - A) * u;
        return this->mGrid->indexToWorldGrad(grad);
    }
};

// GradStencil (math/Stencils.h:455-576): the centre and its six face neighbours, in the order
// (0, -x, +x, -y, +y, -z, +z).
template<class GridT>
class GradStencil : public BaseStencil<GradStencil<GridT>, 7, GridT> {
    using Base = BaseStencil<GradStencil<GridT>, 7, GridT>;
    friend Base;
    using Base::mAcc;
    using Base::mValues;
    float mInv2Dx, mInvDx2;

    __device__ void init(const Coord& ijk)
    {
        mValues[1] = mAcc.getValue(ijk.offsetBy(-1, 0, 0));
        mValues[2] = mAcc.getValue(ijk.offsetBy(1, 0, 0));
        mValues[3] = mAcc.getValue(ijk.offsetBy(0, -1, 0));
        mValues[4] = mAcc.getValue(ijk.offsetBy(0, 1, 0));
        mValues[5] = mAcc.getValue(ijk.offsetBy(0, 0, -1));
        mValues[6] = mAcc.getValue(ijk.offsetBy(0, 0, 1));
    }
    __device__ static float sq(float x) { return x * x; }

public:
    static constexpr int SIZE = 7;
    __device__ explicit GradStencil(const GridT& grid) : GradStencil(grid, grid.view.voxelSize[0]) {}
    __device__ GradStencil(const GridT& grid, double dx)
        : Base(grid), mInv2Dx(float(0.5 / dx)), mInvDx2(float(4.0 * double(mInv2Dx) * double(mInv2Dx)))
    {
    }

    template<int i, int j, int k>
    __device__ static constexpr unsigned int pos()
    {
        static_assert((i == 0 && j == 0) || (i == 0 && k == 0) || (j == 0 && k == 0), "GradStencil: not a face neighbour");
        return i != 0 ? (i < 0 ? 1u : 2u) : j != 0 ? (j < 0 ? 3u : 4u) : k != 0 ? (k < 0 ? 5u : 6u) : 0u;
    }
    // second-order central differences (Stencils.h:502-507)
    __device__ Vec3f gradient() const
    {
        return Vec3f(mValues[2] - mValues[1], mValues[4] - mValues[3], mValues[6] - mValues[5]) * mInv2Dx;
    }
    // first-order upwind gradient along V (Stencils.h:512-518)
    __device__ Vec3f gradient(const Vec3f& V) const
    {
        return Vec3f(V[0] > 0 ? mValues[0] - mValues[1] : mValues[2] - mValues[0],
                     V[1] > 0 ? mValues[0] - mValues[3] : mValues[4] - mValues[0],
                     V[2] > 0 ? mValues[0] - mValues[5] : mValues[6] - mValues[0]) *
               (2 * mInv2Dx);
    }
    // Godunov's norm square of the upwind gradient (Stencils.h:486-495, 62-80)
    __device__ float normSqGrad() const
    {
        const float v0 = mValues[0];
        const float xm = v0 - mValues[1], xp = mValues[2] - v0, ym = v0 - mValues[3], yp = mValues[4] - v0,
                    zm = v0 - mValues[5], zp = mValues[6] - v0;
        float d;
        if (v0 > 0.0f) {
            d = fmaxf(sq(fmaxf(xm, 0.0f)), sq(fminf(xp, 0.0f)));
            d += fmaxf(sq(fmaxf(ym, 0.0f)), sq(fminf(yp, 0.0f)));
            d += fmaxf(sq(fmaxf(zm, 0.0f)), sq(fminf(zp, 0.0f)));
        } else {
            d = fmaxf(sq(fminf(xm, 0.0f)), sq(fmaxf(xp, 0.0f)));
            d += fmaxf(sq(fminf(ym, 0.0f)), sq(fmaxf(yp, 0.0f)));
            d += fmaxf(sq(fminf(zm, 0.0f)), sq(fmaxf(zp, 0.0f)));
        }
        return mInvDx2 * d;
    }
    __device__ float laplacian() const
    {
        return mInvDx2 * (mValues[1] + mValues[2] + mValues[3] + mValues[4] + mValues[5] + mValues[6] - 6 * mValues[0]);
    }
    __device__ bool zeroCrossing() const
    {
        return mValues[0] > 0 ? (mValues[1] < 0 || mValues[2] < 0 || mValues[3] < 0 || mValues[4] < 0 ||
                                 mValues[5] < 0 || mValues[6] < 0)
                              : (mValues[1] > 0 || mValues[2] > 0 || mValues[3] > 0 || mValues[4] > 0 ||
                                 mValues[5] > 0 || mValues[6] > 0);
    }
    // closest-point transform in index space (Stencils.h:544-554)
    __device__ Vec3f cpt() const
    {
        const Coord& ijk = this->mCenter;
        const float d = float(mValues[0] * 0.5 * mInvDx2);
        return Vec3f(ijk[0] - d * (mValues[2] - mValues[1]), ijk[1] - d * (mValues[4] - mValues[3]),
                     ijk[2] - d * (mValues[6] - mValues[5]));
    }
    __device__ bool intersects(float iso = 0.0f) const
    {
        const bool less = mValues[0] < iso;
        return (less ^ (mValues[1] < iso)) || (less ^ (mValues[2] < iso)) || (less ^ (mValues[3] < iso)) ||
               (less ^ (mValues[4] < iso)) || (less ^ (mValues[5] < iso)) || (less ^ (mValues[6] < iso));
    }
};

// ------------------------------------------------------------------ HDDA and DDA
// HDDA<RayT, Coord> (math/HDDA.h:28-179): hierarchical DDA over nodes of a changeable edge dim.
class HDDA {
    nvq::HDDA h;

public:
    __device__ HDDA() {}
    __device__ HDDA(const Ray& ray, int dim) { init(ray, ray.t0(), ray.t1(), dim); }
    __device__ void init(const Ray& ray, float startTime, float maxTime, int dim) { h.init(ray.r, startTime, maxTime, dim); }
    __device__ void init(const Ray& ray, int dim) { init(ray, ray.t0(), ray.t1(), dim); }
    __device__ bool update(const Ray& ray, int dim) { return h.update(ray.r, dim); }
    __device__ bool step() { return h.stepOnce(); }
    __device__ int dim() const { return h.dim; }
    __device__ Coord voxel() const { return Coord(h.vx, h.vy, h.vz); }
    __device__ float time() const { return h.t0; }
    __device__ float maxTime() const { return h.t1; }
    __device__ float next() const { return fminf(h.t1, fminf(h.nx, fminf(h.ny, h.nz))); }
    __device__ nvq::HDDA& raw() { return h; }
};

// DDA<RayT, Coord, Dim> (math/HDDA.h:258-379): a DDA of fixed step Dim.  As in the reference, a
// step moves the voxel by Dim but advances the axis' next time by one voxel's time (|1/dir|).
template<int Dim = 1>
class DDA {
    static_assert(Dim >= 1, "Dim must be >= 1");
    float mT0, mT1;
    int32_t mVoxel[3], mStep[3];
    float mDelta[3], mNext[3];

    template<int axis>
    __device__ bool stepAxis()
    {
        if (mNext[axis] <= mT0) mNext[axis] += mT0 - 0.999999f * mNext[axis] + 1.0e-6f; // ENFORCE_FORWARD_STEPPING
        mT0 = mNext[axis];
        mNext[axis] += mDelta[axis];
        mVoxel[axis] += mStep[axis];
        return mT0 <= mT1;
    }

public:
    __device__ DDA() {}
    __device__ explicit DDA(const Ray& ray) { init(ray); }
    __device__ void init(const Ray& ray, float startTime, float maxTime)
    {
        mT0 = startTime;
        mT1 = maxTime;
        const Vec3f pos = ray(mT0);
        const nvq::Ray& r = ray.r;
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            mVoxel[axis] = int32_t(floorf(pos[axis])) & ~(Dim - 1);
            mDelta[axis] = 0.0f;
            if (r.dir[axis] == 0.0f) {
                mNext[axis] = kFltMax;
                mStep[axis] = 0;
            } else if (r.inv[axis] > 0.0f) {
                mStep[axis] = Dim;
                mNext[axis] = mT0 + (float(mVoxel[axis] + Dim) - pos[axis]) * r.inv[axis];
                mDelta[axis] = r.inv[axis];
            } else {
                mStep[axis] = -Dim;
                mNext[axis] = mT0 + (float(mVoxel[axis]) - pos[axis]) * r.inv[axis];
                mDelta[axis] = -r.inv[axis];
            }
        }
    }
    __device__ void init(const Ray& ray) { init(ray, ray.t0(), ray.t1()); }
    // MinIndex (math/Math.h:294-309), then a step on that axis
    __device__ bool step()
    {
        if (mNext[0] < mNext[1] && mNext[0] < mNext[2]) return stepAxis<0>();
        if (mNext[1] < mNext[2]) return stepAxis<1>();
        return stepAxis<2>();
    }
    __device__ int nextAxis() const
    {
        if (mNext[0] < mNext[1] && mNext[0] < mNext[2]) return 0;
        return mNext[1] < mNext[2] ? 1 : 2;
    }
    __device__ Coord voxel() const { return Coord(mVoxel[0], mVoxel[1], mVoxel[2]); }
    __device__ float time() const { return mT0; }
    __device__ float maxTime() const { return mT1; }
    __device__ float next() const { return fminf(mT1, fminf(mNext[0], fminf(mNext[1], mNext[2]))); }
};

// ------------------------------------------------------------------ ray queries
// zeroCrossing(ray, acc, ijk, v, t) (math/HDDA.h:189-214) and isoCrossing (:225-248) on an
// index-space ray; the ray is clipped to the root bbox in place.
template<class AccT>
__device__ inline bool isoCrossing(Ray& ray, AccT& acc, Coord& ijk, float& v, float& t, float iso = 0.0f)
{
    int32_t c[3] = {0, 0, 0};
    Accessor<AccT::kType> impl(acc.grid());
    const bool hit = isoCrossingCore<AccT::kType>(impl, acc.grid(), ray.r, iso, c, v, t);
    ijk = Coord(c[0], c[1], c[2]);
    return hit;
}
template<class AccT>
__device__ inline bool zeroCrossing(Ray& ray, AccT& acc, Coord& ijk, float& v, float& t)
{
    return isoCrossing(ray, acc, ijk, v, t, 0.0f);
}

// firstActive(ray, acc, ijk, t) (math/HDDA.h:429-444): the first active voxel or tile along an
// index-space ray, HDDA-stepped with a 1.0001 nudge; ijk and t give where it was found.
template<class AccT>
__device__ inline bool firstActive(Ray& ray, AccT& acc, Coord& ijk, float& t)
{
    if (!ray.clip(acc.rootBBox()) || double(ray.t1()) > 1e20) return false;
    const float Delta = 1.0001f;
    t = ray.t0();
    ijk = RoundDown(ray.start());
    Probe p = acc.probe(ijk);
    HDDA hdda(ray, int(p.dim));
    while (!p.active) {
        if (!hdda.step()) return false;
        t = hdda.time() + Delta;
        ijk = RoundDown(ray(t));
        p = acc.probe(ijk);
        hdda.update(ray, int(p.dim));
    }
    return true;
}

// TreeMarcher<NodeT, RayT, AccT> (math/HDDA.h:450-536): the spans [t0, t1] of an index-space ray
// inside nodes of type NodeT (LeafNode, LowerNode or UpperNode), with the HDDA stepping at
// NodeT's edge or coarser.  A span begins where the lookup at the span's start lands in a child
// of NodeT (getDim <= NodeT's child edge) and `node` is the NodeT the accessor cached there.
template<class NodeT, class AccT>
class TreeMarcher {
    AccT& mAcc;
    Ray mRay;
    HDDA mHdda;
    float mT0, mTmax;

public:
    __device__ explicit TreeMarcher(AccT& acc) : mAcc(acc), mT0(-1.0f), mTmax(0.0f) {}

    __device__ bool init(const Ray& indexRay)
    {
        mRay = indexRay;
        if (!mRay.clip(mAcc.rootBBox())) return false;
        const float Eps = 0.000001f;
        const float t0 = mRay.t0() + Eps, t1 = mRay.t1() - Eps;
        if (t0 > t1) return false;
        const Coord ijk = RoundDown(mRay(t0));
        const uint32_t dim = mAcc.getDim(ijk, mRay);
        mHdda.init(mRay, t0, t1, int(dim > NodeT::DIM ? dim : NodeT::DIM));
        mT0 = (dim <= NodeT::CHILD_DIM) ? mHdda.time() : -1.0f;
        mTmax = t1;
        return true;
    }

    __device__ bool step(const uint8_t** node, float& t0, float& t1)
    {
        const float Delta = 0.01f;
        bool valid;
        do {
            t0 = mT0;
            const uint8_t* current = mAcc.template getNode<NodeT>();
            valid = mHdda.step();
            const Coord next = RoundDown(mRay(mHdda.time() + Delta));
            const uint32_t nextDim = mAcc.getDim(next, mRay);
            mHdda.update(mRay, int(nextDim > NodeT::DIM ? nextDim : NodeT::DIM));
            mT0 = (nextDim <= NodeT::CHILD_DIM) ? mHdda.time() : -1.0f;
            if (t0 >= 0) {
                t1 = fminf(mTmax, mHdda.time());
                if (t0 >= t1 || current == nullptr) continue;
                *node = current;
                return true;
            }
        } while (valid);
        return false;
    }

    __device__ const Ray& ray() const { return mRay; }
    __device__ Ray& ray() { return mRay; }
};

} // namespace math
} // namespace nvq
