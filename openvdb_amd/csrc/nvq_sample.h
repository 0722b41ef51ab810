// Per-query point sampling through the lane accessor: SampleFromVoxels<Acc,0> (nearest),
// <Acc,1> (trilinear) and the three gradient modes, shared by the direct kernels and the
// fallback of the binned path.
#pragma once
#include "nvq_device.h"

namespace nvq {

// Grid::worldToIndex for a Vec3f (FP64 Map, NanoVDB.h:1507-1510, math/Math.h:1718-1723)
__device__ __forceinline__ void worldToIndexF64(const GridView& g, float p[3])
{
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) t[a] = float(double(p[a]) - g.vecD[a]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
        p[r] = float(__builtin_fma(double(t[0]), g.invMatD[3 * r],
                                   __builtin_fma(double(t[1]), g.invMatD[3 * r + 1], double(t[2]) * g.invMatD[3 * r + 2])));
}

struct SampleConsts {
    float fdx, fdy, fdz; // FD6 scale (BuildT).5/voxelSize, PySampleFromVoxels.cu:42
    float gsInv2dx;      // GradStencil mInv2Dx, math/Stencils.h:477
};

// One trilinear sample of the FD6 stencil at q, served from the 32-voxel plus-shaped block
// around the query's cell when its cell is the expected one, else from the accessor.
template<uint32_t GT>
__device__ __forceinline__ float fd6Sample(Accessor<GT>& acc, float qx, float qy, float qz, int32_t ei, int32_t ej,
                                           int32_t ek, const Cell& blk)
{
    const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
    const int32_t bi = int32_t(fx), bj = int32_t(fy), bk = int32_t(fz);
    if (bi == ei && bj == ej && bk == ek) return trilerp(blk, qx - fx, qy - fy, qz - fz);
    Cell c;
    stencil8(acc, bi, bj, bk, c);
    return trilerp(c, qx - fx, qy - fy, qz - fz);
}

// Result of one point query; fields that were not asked for hold 0.
struct SampleOut {
    float nn, val, g[3];
};

// Evaluate one query at index-space p: the nearest value when wantNN, the trilinear value when
// wantVal, and the gradient of the given mode (0 = none).
template<uint32_t GT>
__device__ __forceinline__ SampleOut sampleEval(Accessor<GT>& acc, const float p[3], bool wantNN, bool wantVal, int mode,
                                                const SampleConsts& k)
{
    SampleOut r{0.0f, 0.0f, {0.0f, 0.0f, 0.0f}};
    const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
    const int32_t i0 = int32_t(fx), j0 = int32_t(fy), k0 = int32_t(fz);
    const float u = p[0] - fx, v = p[1] - fy, w = p[2] - fz;
    Cell C;
    stencil8(acc, i0, j0, k0, C);
    if (wantNN) {
        // SampleFromVoxels<Acc,0>: getValue(Round(xyz)); Round uses rintf, which lands in this cell
        const int32_t ri = int32_t(rintf(p[0])) - i0, rj = int32_t(rintf(p[1])) - j0, rk = int32_t(rintf(p[2])) - k0;
        r.nn = ((ri | rj | rk) & ~1) ? acc.getValue(i0 + ri, j0 + rj, k0 + rk) : pick(C, ri, rj, rk);
    }
    if (wantVal) r.val = trilerp(C, u, v, w);
    float* o = r.g;
    if (mode == 1) {
        // FD6 of trilinear samples, python/cuda/PySampleFromVoxels.cu:42-47
        Cell B;
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                B.v[0][b][e] = C.v[1][b][e];
                B.v[1][b][e] = acc.getValue(i0 + 2, j0 + b, k0 + e);
            }
        const float xp = fd6Sample(acc, p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f, i0 + 1, j0, k0, B);
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                B.v[1][b][e] = C.v[0][b][e];
                B.v[0][b][e] = acc.getValue(i0 - 1, j0 + b, k0 + e);
            }
        const float xm = fd6Sample(acc, p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f, i0 - 1, j0, k0, B);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                B.v[a][0][e] = C.v[a][1][e];
                B.v[a][1][e] = acc.getValue(i0 + a, j0 + 2, k0 + e);
            }
        const float yp = fd6Sample(acc, p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f, i0, j0 + 1, k0, B);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                B.v[a][1][e] = C.v[a][0][e];
                B.v[a][0][e] = acc.getValue(i0 + a, j0 - 1, k0 + e);
            }
        const float ym = fd6Sample(acc, p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f, i0, j0 - 1, k0, B);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                B.v[a][b][0] = C.v[a][b][1];
                B.v[a][b][1] = acc.getValue(i0 + a, j0 + b, k0 + 2);
            }
        const float zp = fd6Sample(acc, p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f, i0, j0, k0 + 1, B);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                B.v[a][b][1] = C.v[a][b][0];
                B.v[a][b][0] = acc.getValue(i0 + a, j0 + b, k0 - 1);
            }
        const float zm = fd6Sample(acc, p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f, i0, j0, k0 - 1, B);
        o[0] = (xp - xm) * k.fdx;
        o[1] = (yp - ym) * k.fdy;
        o[2] = (zp - zm) * k.fdz;
    } else if (mode == 2) {
        // GradStencil::moveTo(RoundDown(xyz)).gradient(), math/Stencils.h:148-152, 502-507
        o[0] = k.gsInv2dx * (acc.getValue(i0 + 1, j0, k0) - acc.getValue(i0 - 1, j0, k0));
        o[1] = k.gsInv2dx * (acc.getValue(i0, j0 + 1, k0) - acc.getValue(i0, j0 - 1, k0));
        o[2] = k.gsInv2dx * (acc.getValue(i0, j0, k0 + 1) - acc.getValue(i0, j0, k0 - 1));
    } else if (mode == 3) {
        trilinearGradient(C, u, v, w, o);
    }
    return r;
}

// Evaluate one query at index-space p; writes the requested outputs at index i.
template<uint32_t GT>
__device__ __forceinline__ void sampleOne(Accessor<GT>& acc, const float p[3], uint64_t i, float* __restrict__ nn,
                                          float* __restrict__ val, float* __restrict__ grad, int mode,
                                          const SampleConsts& k)
{
    const SampleOut r = sampleEval(acc, p, nn != nullptr, val != nullptr, grad ? mode : 0, k);
    if (nn) nn[i] = r.nn;
    if (val) val[i] = r.val;
    if (grad) {
        grad[3 * i] = r.g[0];
        grad[3 * i + 1] = r.g[1];
        grad[3 * i + 2] = r.g[2];
    }
}

inline SampleConsts sampleConsts(const GridView& g)
{
    SampleConsts k;
    k.fdx = float(0.5 / g.voxelSize[0]);
    k.fdy = float(0.5 / g.voxelSize[1]);
    k.fdz = float(0.5 / g.voxelSize[2]);
    k.gsInv2dx = float(0.5 / g.voxelSize[0]);
    return k;
}

} // namespace nvq
