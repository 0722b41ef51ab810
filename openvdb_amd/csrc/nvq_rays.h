// Per-ray query bodies shared by the single-grid kernels and the persistent multi-grid kernel.
#pragma once
#include "nvq_device.h"

namespace nvq {

// Batched ray query: Ray::worldToIndexF (math/Ray.h:258-268), isoCrossing, and the normal from
// the trilinear gradient at ray(t), Grid::indexToWorldGradF (Map::applyIJTF,
// math/Math.h:1758-1763), normalised.  v is the value at the hit voxel minus iso.
template<uint32_t GT>
__device__ __forceinline__ void zeroCrossingRay(const GridView& g, const float* __restrict__ eye,
                                                const float* __restrict__ dir, uint64_t i, float iso,
                                                uint8_t* __restrict__ hit, int32_t* __restrict__ ijk,
                                                float* __restrict__ tout, float* __restrict__ vout,
                                                float* __restrict__ normal)
{
    Accessor<GT> acc(g);
    const float e[3] = {eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]};
    const float d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    Ray ray = worldToIndexRay(g, e, d);
    int32_t c[3] = {0, 0, 0};
    float v = 0.0f, t = 0.0f;
    const bool found = isoCrossingCore<GT>(acc, g, ray, iso, c, v, t);
    if (hit) hit[i] = found ? 1 : 0;
    if (ijk) {
        ijk[3 * i] = found ? c[0] : 0;
        ijk[3 * i + 1] = found ? c[1] : 0;
        ijk[3 * i + 2] = found ? c[2] : 0;
    }
    if (tout) tout[i] = found ? t : 0.0f;
    if (vout) vout[i] = found ? v : 0.0f;
    if (normal) {
        float nr[3] = {0.0f, 0.0f, 0.0f};
        if (found) {
            float p[3], gi[3], gw[3];
            rayAt(ray, t, p);
            const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
            Cell C;
            stencil8(acc, int32_t(fx), int32_t(fy), int32_t(fz), C);
            trilinearGradient(C, p[0] - fx, p[1] - fy, p[2] - fz, gi);
#pragma unroll
            for (int a = 0; a < 3; ++a)
                gw[a] = __builtin_fmaf(gi[0], g.invMatF[a], __builtin_fmaf(gi[1], g.invMatF[3 + a], gi[2] * g.invMatF[6 + a]));
            const float len = sqrtf(gw[0] * gw[0] + gw[1] * gw[1] + gw[2] * gw[2]);
            if (len > 0.0f) {
                const float sc = 1.0f / len;
                nr[0] = sc * gw[0];
                nr[1] = sc * gw[1];
                nr[2] = sc * gw[2];
            }
        }
        normal[3 * i] = nr[0];
        normal[3 * i + 1] = nr[1];
        normal[3 * i + 2] = nr[2];
    }
}

// ------------------------------------------------------------------ fog march
// Emission-absorption over the shared step structure (marchSteps): per step
// a = exp(-sigma*rho*dt), L += T*(1-a)*rho, T *= a.
template<uint32_t GT>
__device__ __forceinline__ void fogMarchRay(const GridView& g, const float* __restrict__ eye,
                                            const float* __restrict__ dir, uint64_t i, float dt, float sigma,
                                            float* __restrict__ Lout, float* __restrict__ Tout)
{
    Accessor<GT> acc(g);
    const float e[3] = {eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]};
    const float d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    float L = 0.0f, T = 1.0f;
    CellWindow<false> W;
    NoSink none;
    marchSteps(acc, g, e, d, dt, [&](float x0, float x1, float x2) {
        const float fx = floorf(x0), fy = floorf(x1), fz = floorf(x2);
        W.moveTo(acc, g, int32_t(fx), int32_t(fy), int32_t(fz), none);
        const float rho = trilerp(W.c, x0 - fx, x1 - fy, x2 - fz);
        const float a = stepAbsorption(sigma, rho, dt);
        L += T * (1.0f - a) * rho;
        T *= a;
    });
    if (Lout) Lout[i] = L;
    if (Tout) Tout[i] = T;
}

} // namespace nvq
