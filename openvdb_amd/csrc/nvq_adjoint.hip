// Adjoint of the emission-absorption fog march with respect to the leaf values of a Float grid,
// or the channel values of an index grid (no counterpart in the reference, which has no
// differentiable path).
//
// Forward (per ray, steps k = 0..N-1 from the shared marchSteps structure):
//   rho_k = trilinear(x_k), a_k = exp(-s*rho_k) with s = sigma*dt,
//   L = sum_k T_k (1-a_k) rho_k,  T_{k+1} = T_k a_k,  T_0 = 1.
// Adjoint:
//   dLoss/drho_k = gL * (T_k [(1-a_k) + rho_k s a_k] - s R_k) - gT * s * T_N,
//   R_k = sum_{j>k} T_j (1-a_j) rho_j = L - L_{<=k}.
// The backward kernel re-marches each ray with the identical FP32 step structure and
// arithmetic, so T_k and the running radiance L_{<=k} are recomputed bit for bit instead of
// stored per step; L and T_N come from the forward outputs.  dLoss/drho_k is spread over the
// eight trilinear corners of step k with weights (1-u|u)(1-v|v)(1-w|w).
//
// Deterministic reduction: contributions are first summed per ray and per corner voxel over
// the run of consecutive steps whose trilinear cells contain that voxel (a segmented reduction
// in the sliding corner window), then added into an int64 fixed-point accumulator per leaf
// value (key = leaf index * 512 + voxel offset) with integer atomics.  Integer addition is associative, so the result does not
// depend on thread scheduling; the accumulator is converted through FP64.
#include "nvq_device.h"
#include "nvq_internal.h"
#include "nvq_profile.h"
#include "../../include/nvq.h"

#include <hip/hip_runtime.h>

#include <cstring>

namespace nvq {
namespace {

// float -> int64 fixed point, round to nearest, saturating at +-2^62 (float->int conversion is
// undefined beyond the int64 range); a non-finite or saturated value clears `ok`.
__device__ __forceinline__ long long toFixed(float x, bool& ok)
{
    constexpr float kLim = 4611686018427387904.0f; // 2^62
    if (fabsf(x) < kLim) return __float2ll_rn(x);
    ok = false;
    if (x != x) return 0;
    return x > 0.0f ? (1ll << 62) : -(1ll << 62);
}

__global__ void k_clear_flag(uint32_t* flag) { *flag = 0u; }

template<uint32_t GT>
__global__ __launch_bounds__(kBlock) void k_fog_march_bwd(GridView g, const float* __restrict__ eye,
                                                          const float* __restrict__ dir, Sched s,
                                                          const uint32_t* __restrict__ order, float dt, float sigma,
                                                          const float* __restrict__ Lfwd, const float* __restrict__ Tfwd,
                                                          const float* __restrict__ gLin, const float* __restrict__ gTin,
                                                          unsigned long long* __restrict__ accum, float scale,
                                                          uint32_t* __restrict__ overflow)
{
    uint64_t i;
    if (!scheduleOrdered(s, order, i)) return;
    const float gL = gLin ? gLin[i] : 0.0f, gT = gTin ? gTin[i] : 0.0f;
    if (gL == 0.0f && gT == 0.0f) return;
    Accessor<GT> acc(g);
    const float e[3] = {eye[3 * i], eye[3 * i + 1], eye[3 * i + 2]};
    const float d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    const float Ltot = Lfwd[i];
    const float sdt = sigma * dt;
    const float tTerm = gTin ? gT * sdt * Tfwd[i] : 0.0f; // Tfwd may be NULL when no dL/dT is given
    float L = 0.0f, T = 1.0f;
    CellWindow<true> W;
    bool range = true; // every contribution converted without saturation
    auto sink = [&](int32_t key, float sum) {
        const long long q = toFixed(sum * scale, range);
        if (q != 0) atomicAdd(accum + key, static_cast<unsigned long long>(q));
    };
    marchSteps(acc, g, e, d, dt, [&](float x0, float x1, float x2) {
        const float fx = floorf(x0), fy = floorf(x1), fz = floorf(x2);
        W.moveTo(acc, g, int32_t(fx), int32_t(fy), int32_t(fz), sink);
        const float u = x0 - fx, v = x1 - fy, w = x2 - fz;
        const float rho = trilerp(W.c, u, v, w);
        const float a = stepAbsorption(sigma, rho, dt);
        L += T * (1.0f - a) * rho;        // same operations as the forward march
        const float R = Ltot - L;         // radiance of the later steps
        const float gk = gL * (T * ((1.0f - a) + rho * sdt * a) - sdt * R) - tTerm;
        T *= a;
        // corner (a,b,c) receives gk * wx[a] * wy[b] * wz[c], factored as (gk wx wy) wz and fused
        // into the running sum: 6 products and 8 fmas per step
        const float gx[2] = {gk * (1.0f - u), gk * u};
        const float wy[2] = {1.0f - v, v}, wz[2] = {1.0f - w, w};
#pragma unroll
        for (int aa = 0; aa < 2; ++aa)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const float gxy = gx[aa] * wy[b];
#pragma unroll
                for (int c = 0; c < 2; ++c) W.sum[aa * 4 + b * 2 + c] = fmaf(gxy, wz[c], W.sum[aa * 4 + b * 2 + c]);
            }
    });
    W.flush(sink);
    if (!range) atomicOr(overflow, 1u);
}

__global__ __launch_bounds__(kBlock) void k_fixed_to_float(const long long* __restrict__ in, uint64_t n, double inv,
                                                           float* __restrict__ out)
{
    const uint64_t stride = uint64_t(gridDim.x) * kBlock;
    for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) out[i] = float(double(in[i]) * inv);
}

} // namespace

int fogMarchBackward(const GridView& g, const float* eye, const float* dir, uint64_t n, uint32_t width, float dt,
                     float sigma, const float* L, const float* T, const float* gL, const float* gT, int64_t* accum,
                     int shift, hipStream_t s)
{
    const Sched sc = makeSched(n, width);
    const float scale = ldexpf(1.0f, shift);
    TileOrder to;
    if (tileOrder(g, eye, dir, sc, to, s, sizeof(uint32_t)) != 0) return -1;
    uint32_t* flag = reinterpret_cast<uint32_t*>(to.extra);
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, flag);
    auto kernel = g.type == GT_ONINDEX ? k_fog_march_bwd<GT_ONINDEX> : k_fog_march_bwd<GT_FLOAT>;
    {
        ProfScope prof("k_fog_march_bwd", s);
        hipLaunchKernelGGL(kernel, dim3(orderedBlocks(sc)), dim3(kBlock), 0, s, g, eye, dir, sc, to.order, dt, sigma,
                           L, T, gL, gT, reinterpret_cast<unsigned long long*>(accum), scale, flag);
    }
    bool ok = hipGetLastError() == hipSuccess;
    const void* st = stageCopy(flag, sizeof(uint32_t), s);
    ok = st && ok;
    ok = hipStreamSynchronize(s) == hipSuccess && ok;
    if (!ok) return -1;
    uint32_t over = 0;
    std::memcpy(&over, st, sizeof(over));
    return over ? -2 : 0;
}

int fixedToFloat(const int64_t* accum, uint64_t count, int shift, float* out, hipStream_t s)
{
    const uint64_t b = (count + kBlock - 1) / kBlock;
    const uint32_t blocks = uint32_t(b < 1 ? 1 : (b > (1u << 20) ? (1u << 20) : b));
    hipLaunchKernelGGL(k_fixed_to_float, dim3(blocks), dim3(kBlock), 0, s, reinterpret_cast<const long long*>(accum),
                       count, ldexp(1.0, -shift), out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace nvq
