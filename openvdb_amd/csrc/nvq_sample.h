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

// ------------------------------------------------------------------ adjoint of the point queries
// (no counterpart in the reference, whose sampler has no backward)
//
// Every output of a query is linear in the voxel values it reads, so dLoss/dvalue of one query is
// a sum of terms, one per output: the trilinear value (math/SampleFromVoxels.h:423-476), the
// gradient (FD6 of six trilinear samples, python/cuda/PySampleFromVoxels.cu:42-47; GradStencil,
// math/Stencils.h:502-507; TrilinearSampler::gradient, math/SampleFromVoxels.h:232-271) and the
// nearest value.  A term reaches the eight corners (o + (a,b,e)) of one trilinear cell with weight
// ((s * A[a]) * B[b]) * C[e], where s is the upstream gradient times the output's constant and
// A, B, C the per-axis weights of d output / d corner: (1-u, u) for an interpolation along the
// axis, (-1, +1) for a difference along it.  A single term reaches one voxel with weight s.
//
// Contract (what makes every schedule and both launch paths bit-identical): for each distinct
// voxel, the query's contribution is the FP32 sum, in term order, of that voxel's term weights,
// starting from +0; terms whose s is zero are left out.  Term order: value, then gradient (FD6:
// +x, -x, +y, -y, +z, -z samples; GradStencil: +x, -x, +y, -y, +z, -z voxels; trilinear: d/dx,
// d/dy, d/dz), then nearest.  The contribution is converted once, to int64 fixed point
// (round to nearest even, saturating), and integer-added into the accumulator.
struct SampleGrad {
    float gn, gv, gg[3]; // dLoss / d(nearest, value, gradient xyz) of one query
};

// float -> int64 fixed point, round to nearest, saturating at +-2^62 (float->int conversion is
// undefined beyond the int64 range); a non-finite or saturated value clears `ok`.
__device__ __forceinline__ long long toFixedSat(float x, bool& ok)
{
    constexpr float kLim = 4611686018427387904.0f; // 2^62
    if (fabsf(x) < kLim) return __float2ll_rn(x);
    ok = false;
    if (x != x) return 0;
    return x > 0.0f ? (1ll << 62) : -(1ll << 62);
}

struct AdjTerm {
    int32_t o[3];
    float s;
    float A[2], B[2], C[2];
    bool single;
};

// The six FD6 sample points of the forward (sampleEval): p + (+-1 along one axis, +-0 on the others)
__host__ __device__ __forceinline__ void fd6Points(const float p[3], float q[6][3])
{
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int ax = t >> 1;
        const bool plus = (t & 1) == 0;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float dlt = b == ax ? 1.0f : 0.0f;
            q[t][b] = plus ? p[b] + dlt : p[b] - dlt;
        }
    }
}

// General form: the query's terms as a list, each voxel's sum formed in term order.  Used when an
// FD6 sample cell is not the one next to the query's cell (float rounding of p +- 1) and in
// kernels that must not assume the plus-shaped block.  emit(x, y, z, contribution) once per voxel.
template<class Emit>
__host__ __device__ __noinline__ void sampleAdjointGeneral(const float p[3], const SampleGrad& d, int mode,
                                                   const SampleConsts& k, Emit& emit)
{
    AdjTerm T[8];
    int nt = 0;
    const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
    const int32_t i0 = int32_t(fx), j0 = int32_t(fy), k0 = int32_t(fz);
    const float u = p[0] - fx, v = p[1] - fy, w = p[2] - fz;
    auto cellTerm = [&](int32_t x, int32_t y, int32_t z, float s, float a0, float a1, float b0, float b1, float c0,
                        float c1) {
        AdjTerm& t = T[nt++];
        t.o[0] = x; t.o[1] = y; t.o[2] = z;
        t.s = s;
        t.A[0] = a0; t.A[1] = a1; t.B[0] = b0; t.B[1] = b1; t.C[0] = c0; t.C[1] = c1;
        t.single = false;
    };
    auto singleTerm = [&](int32_t x, int32_t y, int32_t z, float s) {
        AdjTerm& t = T[nt++];
        t.o[0] = x; t.o[1] = y; t.o[2] = z;
        t.s = s;
        t.single = true;
    };
    if (d.gv != 0.0f) cellTerm(i0, j0, k0, d.gv, 1.0f - u, u, 1.0f - v, v, 1.0f - w, w);
    if (mode == 1) {
        float q[6][3];
        fd6Points(p, q);
        const float fd[3] = {k.fdx, k.fdy, k.fdz};
        for (int t = 0; t < 6; ++t) {
            const float s0 = d.gg[t >> 1] * fd[t >> 1];
            const float s = (t & 1) ? -s0 : s0;
            if (s == 0.0f) continue;
            const float gx = floorf(q[t][0]), gy = floorf(q[t][1]), gz = floorf(q[t][2]);
            const float a = q[t][0] - gx, b = q[t][1] - gy, c = q[t][2] - gz;
            cellTerm(int32_t(gx), int32_t(gy), int32_t(gz), s, 1.0f - a, a, 1.0f - b, b, 1.0f - c, c);
        }
    } else if (mode == 2) {
        for (int ax = 0; ax < 3; ++ax) {
            const float s = k.gsInv2dx * d.gg[ax];
            if (s == 0.0f) continue;
            const int32_t dx = ax == 0, dy = ax == 1, dz = ax == 2;
            singleTerm(i0 + dx, j0 + dy, k0 + dz, s);
            singleTerm(i0 - dx, j0 - dy, k0 - dz, -s);
        }
    } else if (mode == 3) {
        if (d.gg[0] != 0.0f) cellTerm(i0, j0, k0, d.gg[0], -1.0f, 1.0f, 1.0f - v, v, 1.0f - w, w);
        if (d.gg[1] != 0.0f) cellTerm(i0, j0, k0, d.gg[1], 1.0f - u, u, -1.0f, 1.0f, 1.0f - w, w);
        if (d.gg[2] != 0.0f) cellTerm(i0, j0, k0, d.gg[2], 1.0f - u, u, 1.0f - v, v, -1.0f, 1.0f);
    }
    if (d.gn != 0.0f) singleTerm(int32_t(rintf(p[0])), int32_t(rintf(p[1])), int32_t(rintf(p[2])), d.gn);
    auto corners = [&](int t) { return T[t].single ? 1 : 8; };
    for (int t = 0; t < nt; ++t) {
        for (int c = 0; c < corners(t); ++c) {
            const int32_t X = T[t].o[0] + (c >> 2), Y = T[t].o[1] + ((c >> 1) & 1), Z = T[t].o[2] + (c & 1);
            bool seen = false;
            for (int t2 = 0; t2 < t && !seen; ++t2)
                for (int c2 = 0; c2 < corners(t2); ++c2)
                    if (T[t2].o[0] + (c2 >> 2) == X && T[t2].o[1] + ((c2 >> 1) & 1) == Y && T[t2].o[2] + (c2 & 1) == Z) {
                        seen = true;
                        break;
                    }
            if (seen) continue;
            float sum = 0.0f;
            for (int t2 = t; t2 < nt; ++t2)
                for (int c2 = 0; c2 < corners(t2); ++c2)
                    if (T[t2].o[0] + (c2 >> 2) == X && T[t2].o[1] + ((c2 >> 1) & 1) == Y && T[t2].o[2] + (c2 & 1) == Z)
                        sum += T[t2].single ? T[t2].s
                                            : ((T[t2].s * T[t2].A[c2 >> 2]) * T[t2].B[(c2 >> 1) & 1]) * T[t2].C[c2 & 1];
            emit(X, Y, Z, sum);
        }
    }
}

// True when every voxel the query's terms reach lies in the plus-shaped block around its cell
// (cell + [-1, 2] on one axis, cell + [0, 1] on the others): always, except for FD6 sample points
// whose p +- 1 rounds into a cell other than the neighbouring one.
__host__ __device__ __forceinline__ bool adjointInBlock(const float p[3], int mode)
{
    if (mode != 1) return true;
    float q[6][3];
    fd6Points(p, q);
    bool ok = true;
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int ax = t >> 1;
        const float expect = floorf(p[ax]) + ((t & 1) ? -1.0f : 1.0f);
        ok = ok && floorf(q[t][ax]) == expect;
    }
    return ok;
}

// Block form of the same sums (requires adjointInBlock): 32 register accumulators at constant
// offsets; emit(x, y, z, contribution) for each of the 32 voxels of the plus-shaped block.
template<class Emit>
__host__ __device__ __forceinline__ void sampleAdjointBlock(const float p[3], const SampleGrad& d, int mode,
                                                   const SampleConsts& k, Emit& emit)
{
    const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
    const int32_t i0 = int32_t(fx), j0 = int32_t(fy), k0 = int32_t(fz);
    const float u = p[0] - fx, v = p[1] - fy, w = p[2] - fz;
    float bx[4][4][4]; // voxel (i0-1+x, j0-1+y, k0-1+z); only the plus-shaped 32 are used
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int z = 0; z < 4; ++z) bx[x][y][z] = 0.0f;
    auto cell = [&](int da, int db, int de, float s, const float A[2], const float B[2], const float C[2]) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) bx[da + a + 1][db + b + 1][de + e + 1] += ((s * A[a]) * B[b]) * C[e];
    };
    const float Wu[2] = {1.0f - u, u}, Wv[2] = {1.0f - v, v}, Ww[2] = {1.0f - w, w}, Dm[2] = {-1.0f, 1.0f};
    if (d.gv != 0.0f) cell(0, 0, 0, d.gv, Wu, Wv, Ww);
    if (mode == 1) {
        float q[6][3];
        fd6Points(p, q);
        const float fd[3] = {k.fdx, k.fdy, k.fdz};
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            const int ax = t >> 1;
            const float s0 = d.gg[ax] * fd[ax];
            const float s = (t & 1) ? -s0 : s0;
            if (s == 0.0f) continue;
            float A[3][2];
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const float fr = q[t][b] - floorf(q[t][b]);
                A[b][0] = 1.0f - fr;
                A[b][1] = fr;
            }
            const int off = (t & 1) ? -1 : 1;
            cell(ax == 0 ? off : 0, ax == 1 ? off : 0, ax == 2 ? off : 0, s, A[0], A[1], A[2]);
        }
    } else if (mode == 2) {
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float s = k.gsInv2dx * d.gg[ax];
            if (s == 0.0f) continue;
            bx[1 + (ax == 0)][1 + (ax == 1)][1 + (ax == 2)] += s;
            bx[1 - (ax == 0)][1 - (ax == 1)][1 - (ax == 2)] += -s;
        }
    } else if (mode == 3) {
        if (d.gg[0] != 0.0f) cell(0, 0, 0, d.gg[0], Dm, Wv, Ww);
        if (d.gg[1] != 0.0f) cell(0, 0, 0, d.gg[1], Wu, Dm, Ww);
        if (d.gg[2] != 0.0f) cell(0, 0, 0, d.gg[2], Wu, Wv, Dm);
    }
    if (d.gn != 0.0f) {
        // Round lands in the cell: offsets in {0, 1}; adding +0 to the other corners leaves their
        // fixed-point value unchanged
        const int32_t ri = int32_t(rintf(p[0])) - i0, rj = int32_t(rintf(p[1])) - j0, rk = int32_t(rintf(p[2])) - k0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 2; ++e) bx[a + 1][b + 1][e + 1] += (ri == a && rj == b && rk == e) ? d.gn : 0.0f;
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y)
#pragma unroll
            for (int z = 0; z < 4; ++z) {
                const int out = (x == 0 || x == 3) + (y == 0 || y == 3) + (z == 0 || z == 3);
                if (out <= 1) emit(i0 - 1 + x, j0 - 1 + y, k0 - 1 + z, bx[x][y][z]);
            }
}

// Leaf-value key of a probed voxel: Float grids leaf index * 512 + voxel offset (-1 outside
// leaves), index grids the channel index.  Float leaves are 2144 B = 32 * 67.
template<uint32_t GT>
__device__ __forceinline__ int64_t adjointKey(const GridView& g, const Probe& pr)
{
    if constexpr (GT == GT_ONINDEX) return int64_t(pr.index);
    else return pr.leaf ? int64_t((uint64_t(pr.leaf - g.leaf0) >> 5) / 67u) * 512 + pr.n : -1;
}

// One query's adjoint through the lane accessor, added into accum with int64 atomics.
template<uint32_t GT>
__device__ __forceinline__ void sampleAdjointOne(Accessor<GT>& acc, const GridView& g, const float p[3],
                                                 const SampleGrad& d, int mode, const SampleConsts& k, float scale,
                                                 unsigned long long* __restrict__ accum, bool& ok)
{
    auto emit = [&](int32_t x, int32_t y, int32_t z, float sum) {
        const long long q = toFixedSat(sum * scale, ok);
        if (q == 0) return;
        const int64_t key = adjointKey<GT>(g, acc.probe(x, y, z));
        if (key >= 0) atomicAdd(accum + key, static_cast<unsigned long long>(q));
    };
    if (adjointInBlock(p, mode)) sampleAdjointBlock(p, d, mode, k, emit);
    else sampleAdjointGeneral(p, d, mode, k, emit);
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
