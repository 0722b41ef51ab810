// Large-batch point sampling: spatial binning + LDS-staged voxel bricks.
//
// Random query points touch a dozen distinct cache lines each over a grid that
// does not fit in L2, so a lane-per-query kernel is bound by line traffic from
// the last-level cache.  For large batches the queries are first counting-sorted
// into bins of BIN^3 voxels (one bin per leaf-sized cell of the grid's active
// bounding box, points outside clamped to the border bins):
//   k_bin_count   : key per point + global histogram
//   k_scan_*      : exclusive scan of the histogram (bin offsets)
//   k_bin_scatter : index-space position + original index written to the bin's slot
//   k_bin_sample  : one workgroup per bin stages the (BIN+4)^3 brick of voxel
//                   values around the bin in LDS (same getValue semantics, so
//                   results stay bit-identical) and evaluates the bin's queries
//                   from LDS; outputs go back to each query's original index.
// A query whose stencil leaves the brick falls back to the per-lane accessor.
#include "nvq_device.h"
#include "nvq_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace nvq {
namespace {

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t m, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = v;
}

// Keep freed stream-ordered allocations in the device's default pool instead of returning them
// to the driver at every synchronisation, so repeated large queries do not re-map memory.
void keepPoolMemory(hipStream_t)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
    uint64_t thr = ~uint64_t(0);
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
}

constexpr int kThreads = 256;
constexpr int kBin = 8;                 // bin edge in voxels (one leaf)
constexpr int kApronLo = 1;             // stencil reaches 1 voxel below the bin
constexpr int kBrick = kBin + 4;        // covers [bin-1, bin+10]: floor(p+1) may round up by one more voxel
constexpr int kBrickVox = kBrick * kBrick * kBrick;

struct BinLayout {
    int32_t lo[3];   // index-space origin of bin (0,0,0)
    int32_t nb[3];   // bins per axis
    uint32_t log2;   // bin edge = 1 << log2 voxels
    uint32_t nbins;
};

struct Record {
    float x, y, z;
    uint32_t idx;
};

__device__ __forceinline__ void loadPoint(const GridView& g, const float* xyz, uint64_t i, int world, float p[3])
{
    p[0] = xyz[3 * i];
    p[1] = xyz[3 * i + 1];
    p[2] = xyz[3 * i + 2];
    if (world) {
        float t[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) t[a] = float(double(p[a]) - g.vecD[a]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            p[r] = float(__builtin_fma(double(t[0]), g.invMatD[3 * r],
                                       __builtin_fma(double(t[1]), g.invMatD[3 * r + 1], double(t[2]) * g.invMatD[3 * r + 2])));
    }
}

__device__ __forceinline__ uint32_t binOf(const BinLayout& L, const float p[3])
{
    int32_t b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        // floorf of NaN / huge values is clamped into the border bins; the key is only a locality hint
        const float f = floorf(p[a]);
        const float rel = f - float(L.lo[a]);
        int32_t v = (rel > 0.0f) ? (rel < float(L.nb[a] << L.log2) ? int32_t(rel) >> L.log2 : L.nb[a] - 1) : 0;
        b[a] = v;
    }
    return (uint32_t(b[0]) * uint32_t(L.nb[1]) + uint32_t(b[1])) * uint32_t(L.nb[2]) + uint32_t(b[2]);
}

__global__ __launch_bounds__(kThreads) void k_bin_count(GridView g, BinLayout L, const float* __restrict__ xyz,
                                                        uint64_t n, int world, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ counts)
{
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
        float p[3];
        loadPoint(g, xyz, i, world, p);
        const uint32_t k = binOf(L, p);
        keys[i] = k;
        atomicAdd(&counts[k], 1u);
    }
}

// exclusive scan of counts[0..m) into offs[0..m]; three phases over blocks of 1024 entries
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kThreads) void k_scan_reduce(const uint32_t* __restrict__ counts, uint32_t m,
                                                          uint32_t* __restrict__ blockSums)
{
    __shared__ uint32_t s[kThreads];
    const uint32_t base = blockIdx.x * kScanBlock;
    uint32_t v = 0;
    for (int j = 0; j < kScanBlock / kThreads; ++j) {
        const uint32_t i = base + threadIdx.x * (kScanBlock / kThreads) + j;
        if (i < m) v += counts[i];
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) blockSums[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kThreads) void k_scan_blocks(uint32_t* __restrict__ blockSums, uint32_t nb)
{
    // single workgroup: in-place exclusive scan of nb block sums
    __shared__ uint32_t s[kThreads];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? blockSums[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kThreads; o <<= 1) {
            const uint32_t t = threadIdx.x >= uint32_t(o) ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) blockSums[i] = carry + s[threadIdx.x] - v;
        const uint32_t total = s[kThreads - 1];
        __syncthreads();
        carry += total;
    }
}

__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint32_t* __restrict__ counts, uint32_t m,
                                                         const uint32_t* __restrict__ blockSums,
                                                         uint32_t* __restrict__ offs, uint32_t* __restrict__ cursor)
{
    __shared__ uint32_t s[kThreads];
    const uint32_t base = blockIdx.x * kScanBlock;
    constexpr int per = kScanBlock / kThreads;
    uint32_t v[per];
    uint32_t sum = 0;
    for (int j = 0; j < per; ++j) {
        const uint32_t i = base + threadIdx.x * per + j;
        v[j] = i < m ? counts[i] : 0u;
        sum += v[j];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
        const uint32_t t = threadIdx.x >= uint32_t(o) ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = blockSums[blockIdx.x] + s[threadIdx.x] - sum;
    for (int j = 0; j < per; ++j) {
        const uint32_t i = base + threadIdx.x * per + j;
        if (i < m) {
            offs[i] = run;
            cursor[i] = run;
        }
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) offs[m] = run;
}

__global__ __launch_bounds__(kThreads) void k_bin_scatter(GridView g, const float* __restrict__ xyz, uint64_t n,
                                                          int world, const uint32_t* __restrict__ keys,
                                                          uint32_t* __restrict__ cursor, Record* __restrict__ recs)
{
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
        float p[3];
        loadPoint(g, xyz, i, world, p);
        const uint32_t pos = atomicAdd(&cursor[keys[i]], 1u);
        Record r;
        r.x = p[0];
        r.y = p[1];
        r.z = p[2];
        r.idx = uint32_t(i);
        recs[pos] = r;
    }
}

// Trilinear sample with the cell values read from the LDS brick.
__device__ __forceinline__ float brickTrilinear(const float* __restrict__ brick, const int32_t o[3], float qx, float qy,
                                                float qz, bool& ok)
{
    const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
    const int32_t bx = int32_t(fx) - o[0], by = int32_t(fy) - o[1], bz = int32_t(fz) - o[2];
    ok = ok && uint32_t(bx) < uint32_t(kBrick - 1) && uint32_t(by) < uint32_t(kBrick - 1) &&
         uint32_t(bz) < uint32_t(kBrick - 1);
    const int32_t base = ok ? (bx * kBrick + by) * kBrick + bz : 0;
    Cell c;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) c.v[a][b][e] = brick[base + (a * kBrick + b) * kBrick + e];
    return trilerp(c, qx - fx, qy - fy, qz - fz);
}

template<uint32_t GT>
__global__ __launch_bounds__(kThreads) void k_bin_sample(GridView g, BinLayout L, const Record* __restrict__ recs,
                                                         const uint32_t* __restrict__ offs, float* __restrict__ nn,
                                                         float* __restrict__ val, float* __restrict__ grad, int mode,
                                                         float fdx, float fdy, float fdz, float gsInv2dx)
{
    __shared__ float brick[kBrickVox];
    const uint32_t b = blockIdx.x;
    const uint32_t begin = offs[b], end = offs[b + 1];
    if (begin == end) return;
    const uint32_t bz = b % uint32_t(L.nb[2]), by = (b / uint32_t(L.nb[2])) % uint32_t(L.nb[1]),
                   bx = b / (uint32_t(L.nb[2]) * uint32_t(L.nb[1]));
    const int32_t o[3] = {L.lo[0] + (int32_t(bx) << L.log2) - kApronLo, L.lo[1] + (int32_t(by) << L.log2) - kApronLo,
                          L.lo[2] + (int32_t(bz) << L.log2) - kApronLo};
    Accessor<GT> acc(g);
    for (int v = threadIdx.x; v < kBrickVox; v += kThreads) {
        const int32_t z = v % kBrick, y = (v / kBrick) % kBrick, x = v / (kBrick * kBrick);
        brick[v] = acc.getValue(o[0] + x, o[1] + y, o[2] + z);
    }
    __syncthreads();
    for (uint32_t q = begin + threadIdx.x; q < end; q += kThreads) {
        const Record r = recs[q];
        const float p[3] = {r.x, r.y, r.z};
        const float fx = floorf(p[0]), fy = floorf(p[1]), fz = floorf(p[2]);
        const int32_t i0 = int32_t(fx), j0 = int32_t(fy), k0 = int32_t(fz);
        bool ok = true;
        const float sv = brickTrilinear(brick, o, p[0], p[1], p[2], ok);
        float o3[3] = {0.0f, 0.0f, 0.0f};
        float nnv = 0.0f;
        if (ok && nn) {
            const int32_t ri = int32_t(rintf(p[0])) - o[0], rj = int32_t(rintf(p[1])) - o[1], rk = int32_t(rintf(p[2])) - o[2];
            nnv = brick[(ri * kBrick + rj) * kBrick + rk];
        }
        if (ok && grad) {
            if (mode == 1) {
                const float xp = brickTrilinear(brick, o, p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f, ok);
                const float xm = brickTrilinear(brick, o, p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f, ok);
                const float yp = brickTrilinear(brick, o, p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f, ok);
                const float ym = brickTrilinear(brick, o, p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f, ok);
                const float zp = brickTrilinear(brick, o, p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f, ok);
                const float zm = brickTrilinear(brick, o, p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f, ok);
                o3[0] = (xp - xm) * fdx;
                o3[1] = (yp - ym) * fdy;
                o3[2] = (zp - zm) * fdz;
            } else if (mode == 2) {
                const int32_t c = ((i0 - o[0]) * kBrick + (j0 - o[1])) * kBrick + (k0 - o[2]);
                ok = ok && i0 - o[0] >= 1 && j0 - o[1] >= 1 && k0 - o[2] >= 1;
                if (ok) {
                    o3[0] = gsInv2dx * (brick[c + kBrick * kBrick] - brick[c - kBrick * kBrick]);
                    o3[1] = gsInv2dx * (brick[c + kBrick] - brick[c - kBrick]);
                    o3[2] = gsInv2dx * (brick[c + 1] - brick[c - 1]);
                }
            } else if (mode == 3) {
                const int32_t c = ((i0 - o[0]) * kBrick + (j0 - o[1])) * kBrick + (k0 - o[2]);
                Cell C;
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                        for (int e = 0; e < 2; ++e) C.v[a][bb][e] = brick[c + (a * kBrick + bb) * kBrick + e];
                trilinearGradient(C, p[0] - fx, p[1] - fy, p[2] - fz, o3);
            }
        }
        if (!ok) { // stencil leaves the brick (border bins, huge coordinates): per-lane accessor
            Cell C;
            stencil8(acc, i0, j0, k0, C);
            nnv = acc.getValue(int32_t(rintf(p[0])), int32_t(rintf(p[1])), int32_t(rintf(p[2])));
            const float s2 = trilerp(C, p[0] - fx, p[1] - fy, p[2] - fz);
            if (val) val[r.idx] = s2;
            if (grad) {
                if (mode == 1) {
                    o3[0] = (trilinear(acc, p[0] + 1.0f, p[1] + 0.0f, p[2] + 0.0f) -
                             trilinear(acc, p[0] - 1.0f, p[1] - 0.0f, p[2] - 0.0f)) * fdx;
                    o3[1] = (trilinear(acc, p[0] + 0.0f, p[1] + 1.0f, p[2] + 0.0f) -
                             trilinear(acc, p[0] - 0.0f, p[1] - 1.0f, p[2] - 0.0f)) * fdy;
                    o3[2] = (trilinear(acc, p[0] + 0.0f, p[1] + 0.0f, p[2] + 1.0f) -
                             trilinear(acc, p[0] - 0.0f, p[1] - 0.0f, p[2] - 1.0f)) * fdz;
                } else if (mode == 2) {
                    o3[0] = gsInv2dx * (acc.getValue(i0 + 1, j0, k0) - acc.getValue(i0 - 1, j0, k0));
                    o3[1] = gsInv2dx * (acc.getValue(i0, j0 + 1, k0) - acc.getValue(i0, j0 - 1, k0));
                    o3[2] = gsInv2dx * (acc.getValue(i0, j0, k0 + 1) - acc.getValue(i0, j0, k0 - 1));
                } else if (mode == 3) {
                    trilinearGradient(C, p[0] - fx, p[1] - fy, p[2] - fz, o3);
                }
            }
        } else if (val) {
            val[r.idx] = sv;
        }
        if (nn) nn[r.idx] = nnv;
        if (grad) {
            grad[3 * uint64_t(r.idx)] = o3[0];
            grad[3 * uint64_t(r.idx) + 1] = o3[1];
            grad[3 * uint64_t(r.idx) + 2] = o3[2];
        }
    }
}

template<uint32_t GT> struct RunBinSample {
    static void run(const GridView& g, const BinLayout& L, const Record* recs, const uint32_t* offs, float* nn, float* val,
                    float* grad, int mode, float fdx, float fdy, float fdz, float gs, hipStream_t s)
    {
        hipLaunchKernelGGL(k_bin_sample<GT>, dim3(L.nbins), dim3(kThreads), 0, s, g, L, recs, offs, nn, val, grad, mode,
                           fdx, fdy, fdz, gs);
    }
};

bool makeLayout(const GridView& g, BinLayout& L)
{
    if (g.bbox[0] > g.bbox[3] || g.bbox[1] > g.bbox[4] || g.bbox[2] > g.bbox[5]) return false; // empty grid
    for (uint32_t lg = 3; lg <= 12; ++lg) {
        uint64_t total = 1;
        for (int a = 0; a < 3; ++a) {
            const int64_t lo = int64_t(g.bbox[a]) - 2, hi = int64_t(g.bbox[3 + a]) + 2;
            L.lo[a] = int32_t(lo);
            const int64_t nb = ((hi - lo) >> lg) + 1;
            L.nb[a] = int32_t(nb);
            total *= uint64_t(nb);
        }
        if (total <= (uint64_t(1) << 22)) {
            if (lg != 3) return false; // bricks assume leaf-sized bins: coarse layouts use the direct kernel
            L.log2 = lg;
            L.nbins = uint32_t(total);
            return true;
        }
    }
    return false;
}

} // namespace

bool sampleBinnedSupported(const GridView& g, uint64_t n)
{
    BinLayout L;
    return n >= (uint64_t(1) << 20) && n < (uint64_t(1) << 32) && makeLayout(g, L);
}

int sampleBinned(const GridView& g, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
                 int mode, hipStream_t s)
{
    BinLayout L;
    if (!makeLayout(g, L)) return -1;
    const uint32_t m = L.nbins;
    const uint32_t nScanBlocks = (m + kScanBlock - 1) / kScanBlock;
    // scratch: keys[n] | counts[m] | offs[m+1] | cursor[m] | blockSums[nScanBlocks] | records[n]
    const size_t bytesKeys = size_t(n) * 4, bytesBins = (size_t(m) * 3 + 1 + nScanBlocks) * 4;
    const size_t bytesRecs = size_t(n) * sizeof(Record);
    keepPoolMemory(s);
    char* scratch = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&scratch), bytesKeys + bytesBins + bytesRecs + 256, s) != hipSuccess)
        return -2;
    uint32_t* keys = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* counts = keys + n;
    uint32_t* offs = counts + m;
    uint32_t* cursor = offs + m + 1;
    uint32_t* blockSums = cursor + m;
    Record* recs = reinterpret_cast<Record*>(scratch + ((bytesKeys + bytesBins + 255) & ~size_t(255)));
    int st = 0;
    const uint32_t gridN = uint32_t(std::min<uint64_t>((n + kThreads - 1) / kThreads, 1u << 20));
    hipLaunchKernelGGL(k_fill_u32, dim3((m + kThreads - 1) / kThreads), dim3(kThreads), 0, s, counts, m, 0u);
    hipLaunchKernelGGL(k_bin_count, dim3(gridN), dim3(kThreads), 0, s, g, L, xyz, n, world, keys, counts);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nScanBlocks), dim3(kThreads), 0, s, counts, m, blockSums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kThreads), 0, s, blockSums, nScanBlocks);
    hipLaunchKernelGGL(k_scan_apply, dim3(nScanBlocks), dim3(kThreads), 0, s, counts, m, blockSums, offs, cursor);
    hipLaunchKernelGGL(k_bin_scatter, dim3(gridN), dim3(kThreads), 0, s, g, xyz, n, world, keys, cursor, recs);
    const float fdx = float(0.5 / g.voxelSize[0]), fdy = float(0.5 / g.voxelSize[1]), fdz = float(0.5 / g.voxelSize[2]);
    const float gs = float(0.5 / g.voxelSize[0]);
    switch (g.type) {
    case GT_FLOAT: RunBinSample<GT_FLOAT>::run(g, L, recs, offs, nn, val, grad, mode, fdx, fdy, fdz, gs, s); break;
    case GT_FP4: RunBinSample<GT_FP4>::run(g, L, recs, offs, nn, val, grad, mode, fdx, fdy, fdz, gs, s); break;
    case GT_FP8: RunBinSample<GT_FP8>::run(g, L, recs, offs, nn, val, grad, mode, fdx, fdy, fdz, gs, s); break;
    case GT_FP16: RunBinSample<GT_FP16>::run(g, L, recs, offs, nn, val, grad, mode, fdx, fdy, fdz, gs, s); break;
    case GT_FPN: RunBinSample<GT_FPN>::run(g, L, recs, offs, nn, val, grad, mode, fdx, fdy, fdz, gs, s); break;
    default: st = -3;
    }
    if (hipGetLastError() != hipSuccess) st = -4;
    if (hipFreeAsync(scratch, s) != hipSuccess) st = -5;
    return st;
}

} // namespace nvq
