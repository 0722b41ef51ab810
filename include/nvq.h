/*
 * nvq — MI355X-native (gfx950) read-only query path for NanoVDB grids.
 *
 * C ABI of libnvq.so.  Plain pointers and sizes only.  Unless a name says
 * "host_", every pointer argument is a DEVICE pointer.  A `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  Every call is
 * stream-ordered and asynchronous unless its comment says it waits.  Every call returns NVQ_OK (0)
 * or a negative NVQ_ERR_* status.  Nothing throws across the ABI.  Query scratch (ray-tile orders,
 * queue counters, the sort buffers of binned point queries) lives in one grow-only block per
 * (device, stream) that the library keeps until the process ends; besides it, the library holds
 * the opt-in kernel profiler (nvq_profile_*), the work-queue array of nvq_multi_query_noalloc and
 * a small pinned host buffer per calling thread for values it reads back.  Calls on distinct
 * streams are thread-safe; calls on one stream from several threads are serialised.  The caller
 * owns every buffer and keeps the grid alive until work queued on it has finished.
 *
 * The grid is an unmodified NanoVDB buffer (magic "NanoVDB1", ABI major 32),
 * as written by NanoVDB's own tools (createLevelSetSphere, createNanoGrid, …)
 * or by this library's host builder (nvq_build.h).  The kernels read it by
 * byte offset; NanoVDB.h is never compiled for the device.
 *
 * Reference citations are relative to the reference tree (AcademySoftwareFoundation/openvdb).
 */
#ifndef NVQ_H
#define NVQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NVQ_ABI_VERSION 1

enum nvq_status {
    NVQ_OK = 0,
    NVQ_ERR_INVALID_ARG = -1,     /* NULL pointer, n too large, bad mode */
    NVQ_ERR_BAD_MAGIC = -2,       /* GridData::isValid magic check, NanoVDB.h:2002-2013 */
    NVQ_ERR_BAD_VERSION = -3,     /* major version != 32, NanoVDB.h:729-769 */
    NVQ_ERR_UNSUPPORTED_TYPE = -4, /* grid type not Float/Fp4/Fp8/Fp16/FpN/OnIndex, or op needs Float */
    NVQ_ERR_MISALIGNED = -5,      /* buffer not 32-byte aligned, NanoVDB.h:69-74 */
    NVQ_ERR_TRUNCATED = -6,       /* buffer shorter than GridData::mGridSize / node offsets out of range */
    NVQ_ERR_HIP = -7,             /* a HIP runtime call failed */
    NVQ_ERR_NO_DEVICE = -8,       /* no usable gfx950 device */
    NVQ_ERR_OVERFLOW = -9         /* a value left the fixed-point range of an accumulator (adjoint) */
};

/* NanoVDB GridType values the kernels decode (NanoVDB.h:219-247).  NVQ_ONINDEX grids (ValueOnIndex,
 * NanoVDB.h:4147-4171) hold indices; their values come from a float channel, the grid's own blind
 * data #0 or an array given with nvq_grid_set_channel (ChannelAccessor, NanoVDB.h:5724-5800). */
enum nvq_grid_type { NVQ_FLOAT = 1, NVQ_FP4 = 13, NVQ_FP8 = 14, NVQ_FP16 = 15, NVQ_FPN = 16, NVQ_ONINDEX = 20 };

typedef struct nvq_grid nvq_grid; /* opaque device-resident grid */

typedef struct nvq_grid_info {
    uint32_t grid_type;   /* enum nvq_grid_type */
    uint32_t grid_class;  /* NanoVDB GridClass: 1 LevelSet, 2 FogVolume, ... */
    uint64_t bytes;       /* buffer size */
    uint32_t node_count[3]; /* leaf, lower, upper */
    uint32_t root_tiles;
    int32_t index_bbox[6];  /* RootData::mBBox (min xyz, max xyz) */
    double voxel_size[3];
    float background;
    uint32_t pad;
    const void* device_buffer; /* the flat buffer in device memory */
} nvq_grid_info;

/* Human-readable text for a status code (static storage). */
const char* nvq_status_string(int status);

/* Validate a host NanoVDB buffer and upload it with one hipMalloc + one
 * hipMemcpyAsync on `stream`.  Replaces GridHandle<BufferT>::deviceUpload +
 * deviceGrid<float>() (GridHandle.h:159-194, cuda/DeviceBuffer.h:422-437); where
 * deviceGrid<T>() returns nullptr on a type mismatch, this returns
 * NVQ_ERR_UNSUPPORTED_TYPE / NVQ_ERR_BAD_MAGIC / ... and leaves *out NULL. */
int nvq_grid_upload(const void* host_buf, uint64_t bytes, void* stream, nvq_grid** out);
/* Same as nvq_grid_upload, but copies into caller-owned device memory `device_dst` (at least
 * `bytes`, 32-byte aligned) instead of allocating; nvq_grid_free then leaves that memory alone.
 * Lets a framework own the buffer, e.g. to expose the leaf values as a trainable tensor. */
int nvq_grid_upload_into(const void* host_buf, uint64_t bytes, void* device_dst, void* stream, nvq_grid** out);
/* Index grids: read values from `channel` (a DEVICE array of at least `count` >= the grid's value
 * count floats, e.g. a trainable tensor) instead of the grid's own channel; NULL restores the
 * grid's own channel.  ChannelAccessor(grid, channelPtr), NanoVDB.h:5750-5757.  Waits for the
 * device before switching; every query on an index grid without a channel returns
 * NVQ_ERR_INVALID_ARG. */
int nvq_grid_set_channel(nvq_grid* grid, const float* channel, uint64_t count);
/* How the ray queries (nvq_zero_crossing, nvq_iso_crossing, nvq_render_iso_surface, nvq_fog_march and
 * its adjoint, nvq_multi_query) map world rays into index space.  NVQ_RAY_MAP_F32 (the default):
 * Ray::worldToIndexF (math/Ray.h:258-268), the FP32 Map.  NVQ_RAY_MAP_F64: Ray::applyInverseMap
 * (math/Ray.h:223-230), the FP64 Map (NanoVDB.h:1507-1510, :1530), for eyes far from the origin of a
 * grid with a large translation or a non-power-of-two voxel size, where the FP32 map loses the
 * voxel.  The traversal stays FP32.  Waits for the device before switching. */
enum nvq_ray_map { NVQ_RAY_MAP_F32 = 0, NVQ_RAY_MAP_F64 = 1 };
int nvq_grid_set_ray_map(nvq_grid* grid, int mode);
/* Free the grid handle and, for nvq_grid_upload, the device copy (waits for the device). */
int nvq_grid_free(nvq_grid* grid);
int nvq_grid_get_info(const nvq_grid* grid, nvq_grid_info* info);
/* Copy the grid's kernel-argument view (nvq::GridView of nvq_device_impl.h: device pointers to
 * the root, the first leaf and, for index grids, the channel set at the time of the call, plus
 * the Map and root bbox) into host memory `host_view` of exactly `view_bytes` = sizeof(GridView),
 * else NVQ_ERR_INVALID_ARG.  Client kernels take it through nvq::getDeviceGrid (nvq_device.h), the
 * counterpart of GridHandle::deviceGrid<T>() (GridHandle.h:159-170); it stays valid until the grid
 * is freed. */
int nvq_grid_get_view(const nvq_grid* grid, void* host_view, uint64_t view_bytes);

/* Batched ReadAccessor::getValue / isActive / getDim (NanoVDB.h:5370-5377, 5431-5447)
 * for n integer voxels ijk[n][3].  Any of val / active / dim may be NULL.
 * dim follows ReadAccessor::getDim: 1 inside a leaf, 8 / 128 / 4096 for tiles and background. */
int nvq_get_value(const nvq_grid* grid, const int32_t* ijk, uint64_t n, float* val, uint8_t* active,
                  uint32_t* dim, void* stream);

/* Batched point queries at xyz[n][3] (index space if world == 0, else world space mapped with
 * Grid::worldToIndex in FP64 as the reference CUDA op does, python/cuda/PySampleFromVoxels.cu:23,38).
 *   nn   : SampleFromVoxels<Acc,0>   (getValue at Round(xyz)),  math/SampleFromVoxels.h:149-156
 *   val  : SampleFromVoxels<Acc,1>   (trilinear),               math/SampleFromVoxels.h:423-464
 *   grad : grad_mode 1 = FD6 of trilinear * 0.5/voxelSize        python/cuda/PySampleFromVoxels.cu:42-47
 *          grad_mode 2 = GradStencil::gradient at RoundDown(xyz)  math/Stencils.h:502-507
 *          grad_mode 3 = TrilinearSampler::gradient (index space) math/SampleFromVoxels.h:232-271
 * nn / val / grad may be NULL.  Replaces the nanobind op sampleFromVoxels(points, d_grid, values[, gradients]).
 * Batches of 2^20 points or more are binned through the stream's scratch block (about 110 bytes
 * per point) and wait for `stream` once (to size the sort of the points that touch a leaf).
 * Smaller batches do not wait and take no scratch. */
int nvq_sample(const nvq_grid* grid, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
               int grad_mode, void* stream);

/* GradStencil at integer voxels (math/Stencils.h:455-576): gradient, normSqGrad, laplacian,
 * zeroCrossing.  Float grids only.  Outputs may be NULL. */
int nvq_grad_stencil(const nvq_grid* grid, const int32_t* ijk, uint64_t n, float* grad, float* norm_sq, float* lap,
                     uint8_t* zero_crossing, void* stream);

/* World-space rays eye[n][3], dir[n][3] -> Ray::worldToIndexF -> math::zeroCrossing
 * (math/Ray.h:258-268, math/HDDA.h:189-214).  Outputs per ray: hit (0/1), ijk[3], t (index
 * space), v (value at ijk), normal[3] (normalised world-space trilinear gradient at ray(t)).
 * Misses write zeros.  Any output may be NULL.  If image_width > 0 the rays are a row-major
 * image of that width (n a multiple of it) and are scheduled in 8x8 pixel tiles per wavefront;
 * results are written at the input index either way. */
int nvq_zero_crossing(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                      uint8_t* hit, int32_t* ijk, float* t, float* v, float* normal, void* stream);

/* math::isoCrossing (math/HDDA.h:225-248) for world-space rays: as nvq_zero_crossing, with the
 * crossing taken where value - iso changes sign; v is the value at the hit voxel minus iso.
 * nvq_zero_crossing is this call with iso = 0 (bit for bit). */
int nvq_iso_crossing(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                     float iso, uint8_t* hit, int32_t* ijk, float* t, float* v, float* normal, void* stream);

/* Camera of the iso-surface renderer (RenderOp, examples/ex_raytrace_iso_surface/common.h:18-37).
 * nvqb_iso_camera (nvq_build.h) fills it from a grid the way RenderOp's constructor does. */
typedef struct nvq_iso_camera {
    int32_t width, height; /* image size in pixels */
    float iso;             /* iso value (RenderOp: the voxel size) */
    float voxel_size;      /* mDx */
    float wbbox_dim_z;     /* mWBBoxDimZ: twice the z extent of the world bbox */
    float tan_half_fov;    /* tanf(45 / 2 degrees), evaluated on the host */
    float origin[3];       /* mWBBoxCenter + (0, 0, mWBBoxDimZ) */
} nvq_iso_camera;

/* Render a width x height image (RenderOp::renderPixel, common.h:70-78): per pixel one ray
 * (getIndexRay, common.h:130-145), isoCrossing, and composite (common.h:147-155): depth
 * (t * voxel_size) / (2 * wbbox_dim_z) where hit, else a checkerboard of 1.0 / 0.5 in 128-pixel
 * squares.  persistent != 0 uses a persistent launch whose wavefronts pull 8x8 pixel tiles from a
 * device counter (renderIsoSurfacePersistentKernel, common.h:166-189); otherwise one thread per
 * pixel in 8x8 tiles.  `camera` is a HOST pointer; image[width*height] (row-major), optional
 * hit[width*height] and t[width*height] (index-space hit time, 0 on miss) are device pointers. */
int nvq_render_iso_surface(const nvq_grid* grid, const nvq_iso_camera* camera, int persistent, float* image,
                           uint8_t* hit, float* t, void* stream);

/* Emission-absorption march of world-space rays through a Float fog volume.  HDDA over the
 * tree at leaf-node granularity skips inactive tiles and background (math/HDDA.h:28-179,
 * 450-536); inside leaves and active tiles the ray is sampled trilinearly at globally aligned
 * steps t = k*dt; per step a = exp(-sigma_t*rho*dt), L += T*(1-a)*rho, T *= a (the EA integrator
 * of openvdb/openvdb/tools/RayTracer.h:993-1064 with density emission).  radiance / transmittance
 * may be NULL.  image_width as for nvq_zero_crossing. */
int nvq_fog_march(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                  float dt, float sigma_t, float* radiance, float* transmittance, void* stream);

/* Per-grid query of nvq_multi_query */
enum nvq_multi_op { NVQ_OP_ZERO_CROSSING = 0, NVQ_OP_FOG_MARCH = 1 };

/* One persistent launch over k <= 32 grids of any leaf codecs and classes.  host_grids[k] and
 * host_ray_offsets[k+1] / host_image_width[k] / host_ops[k] are HOST arrays; the rays of grid j are
 * eye/dir[host_ray_offsets[j] .. host_ray_offsets[j+1]), a row-major image of width
 * host_image_width[j] (its ray count must be a multiple of the width; width 0 = rows of 64).
 * host_ops[j] chooses the query: NVQ_OP_ZERO_CROSSING (hit, ijk, t -> out_a, v -> out_b, normal) or
 * NVQ_OP_FOG_MARCH (radiance -> out_a, transmittance -> out_b).  With host_ops NULL the class
 * decides (level set: zeroCrossing, fog volume: march) and other classes, e.g. index grids, are
 * rejected.  Wavefronts pull 8x8 ray tiles from a device work queue (the persistent-warp launch of
 * examples/ex_raytrace_iso_surface/common.h:166-189).  hit / ijk / normal may be NULL. */
int nvq_multi_query(const nvq_grid* const* host_grids, uint32_t k, const uint64_t* host_ray_offsets,
                    const uint32_t* host_image_width, const uint32_t* host_ops, const float* eye, const float* dir,
                    float dt, float sigma_t, uint8_t* hit, int32_t* ijk, float* out_a, float* out_b, float* normal,
                    void* stream);

/* nvq_multi_query for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments, limits and results (bit for bit) as nvq_multi_query; the only scratch, the work
 * queue's counters, lives in a fixed 2 KiB array that the library holds on each device until the
 * process ends, so the call allocates and frees nothing (no hipMalloc, no stream-ordered pool).
 * The scratch does not grow with the batch: any launch nvq_multi_query accepts fits (fewer than
 * 2^31 ray tiles, else NVQ_ERR_INVALID_ARG).  The counters are cleared by a kernel at the start of
 * each call, so they keep the last call's contents in between; calls that overlap on one device
 * share them and race: serialise them (one stream, or events).  Does not wait. */
int nvq_multi_query_noalloc(const nvq_grid* const* host_grids, uint32_t k, const uint64_t* host_ray_offsets,
                            const uint32_t* host_image_width, const uint32_t* host_ops, const float* eye,
                            const float* dir, float dt, float sigma_t, uint8_t* hit, int32_t* ijk, float* out_a,
                            float* out_b, float* normal, void* stream);

/* Adjoint of nvq_fog_march (no counterpart in the reference): dLoss/dvalue for every leaf value
 * of a Float grid, or every channel value of an index grid (NVQ_ONINDEX), given the forward
 * outputs radiance[n] (and transmittance[n] when dL_dtransmittance is given) and the loss
 * gradients per ray.  Contributions are ADDED to grad_accum — Float grids: [leafCount*512] (leaf
 * order of the buffer, voxel offset n = x<<6|y<<3|z within the leaf); index grids: [value count]
 * (the channel's own order, a dense scatter) — as int64 fixed point with 2^-fixed_shift resolution; the
 * sum is deterministic for any schedule.  The caller zeroes grad_accum before the first call
 * and may accumulate several calls (e.g. cameras).  Transmittance is recomputed by re-marching
 * each ray; no per-step state is stored.  Range: each ray's contribution to one value is converted
 * with saturation at +-2^62 (a non-finite one counts as 0); when any contribution saturates or is
 * not finite the call returns NVQ_ERR_OVERFLOW and grad_accum holds the saturated sums.  The int64
 * additions wrap, so a value's total is exact whenever the true total lies within
 * +-2^(63-fixed_shift), however its partial sums ran.  Waits for `stream` (to read the overflow flag). */
int nvq_fog_march_bwd(const nvq_grid* grid, const float* eye, const float* dir, uint64_t n, uint32_t image_width,
                      float dt, float sigma_t, const float* radiance, const float* transmittance,
                      const float* dL_dradiance, const float* dL_dtransmittance, int64_t* grad_accum, int fixed_shift,
                      void* stream);
/* Adjoint of nvq_sample (no counterpart in the reference; it differentiates the forward of
 * nanovdb.tools.cuda.sampleFromVoxels, python/cuda/PySampleFromVoxels.cu:16-52, and
 * SampleFromVoxels<Acc,0|1>, math/SampleFromVoxels.h:149-156, 423-476): given the same points
 * and dLoss/d(nearest)[n], dLoss/d(value)[n] and dLoss/d(gradient)[n][3] of grad_mode (each may
 * be NULL), ADDS dLoss/dvalue into grad_accum with the layout, fixed point and range rules of
 * nvq_fog_march_bwd.  Float grids: only leaf values receive gradients (tile and background values
 * are constants); index grids: every channel entry a query reads.  Each query's contribution to
 * one value is the FP32 sum of its terms (value, gradient, nearest, in that order) converted once
 * to fixed point, so the result is bit-identical for every schedule and batch size.  Float and
 * NVQ_ONINDEX grids only.  Waits for `stream`. */
int nvq_sample_bwd(const nvq_grid* grid, const float* xyz, uint64_t n, int world, const float* dL_dnn,
                   const float* dL_dval, const float* dL_dgrad, int grad_mode, int64_t* grad_accum, int fixed_shift,
                   void* stream);
/* out[i] = float(double(accum[i]) * 2^-fixed_shift) */
int nvq_grad_fixed_to_float(const int64_t* accum, uint64_t count, int fixed_shift, float* out, void* stream);

/* Device-side construction: tools::cuda::voxelsToGrid<float | ValueOnIndex | ValueIndex>(d_ijk, n, voxelSize)
 * (tools/cuda/PointsToGrid.cuh:119-125, :1417-1424).  ijk is a DEVICE array of n integer voxels
 * (nanovdb::Coord, int32 x3; duplicates allowed; 1 <= n < 2^31).  grid_type NVQ_FLOAT gives a
 * grid whose active voxels hold 1 and inactive 0; NVQ_ONINDEX an IndexGrid whose active voxels
 * hold indices 1..count in breadth-first order; grid_type 19 (GridType::Index, ValueIndex) an
 * IndexGrid that indexes all 512 values of every leaf, 1 + 512 * leaf + offset (tiles and the
 * background are index 0).  The buffer equals the reference's byte for
 * byte.  `name` may be NULL.  `index` (index grids only; may be NULL) is a DEVICE array of n that
 * receives each input voxel's value index, so that channel[index[i]] = value[i] fills a float
 * channel for nvq_grid_set_channel.  The new grid owns its device buffer.  The input may span at
 * most 2^28 upper-node tiles (4096^3 voxels each) in its bbox, else NVQ_ERR_INVALID_ARG.
 * Waits for `stream` (twice to size the buffer, and before returning).  Scratch of about 16 B
 * (inputs within one upper node) or 24 B per voxel, plus the sort's temporaries, is kept per device
 * between calls and grows to the largest build; builds on one device run one at a time. */
int nvq_voxels_to_grid(const int32_t* ijk, uint64_t n, uint32_t grid_type, double voxel_size, const char* name,
                       uint64_t* index, void* stream, nvq_grid** out);
/* The same from index-space voxel coordinates given as floating point (PtrT = Vec3f* or Vec3d*,
 * xyz_f64 selects double): each is rounded to the nearest voxel, Vec3<T>::round (math/Math.h
 * :1515-1524, TileKeyFunctor tools/cuda/PointsToGrid.cuh:576-577). */
int nvq_voxels_to_grid_fp(const void* xyz, int xyz_f64, uint64_t n, uint32_t grid_type, double voxel_size,
                          const char* name, uint64_t* index, void* stream, nvq_grid** out);
/* Device-side construction of a point grid: tools::cuda::pointsToGrid(d_points, n, voxelSize, type)
 * (tools/cuda/PointsToGrid.cuh:57-64) for n float world-space points xyz[n][3] (a DEVICE array).
 * Each point lies in voxel Round(Map(voxel_size).applyInverseMapF(p)).  The grid (GridType
 * PointIndex, class PointData, or PointIndex for point_type 1) has per-voxel point counts in its
 * leaves and the points, sorted by voxel, as blind data encoded by point_type: 1 PointID (uint32),
 * 2 World64, 3 World32, 4 Grid64, 5 Grid32, 6 Voxel32, 7 Voxel16, 8 Voxel8, 9 Default (= World32
 * for float points) (PointType, NanoVDB.h:395-405).  The buffer equals the reference's byte for
 * byte.  A leaf may hold at most 65535 points, else NVQ_ERR_INVALID_ARG (the reference throws).
 * Point grids are built and downloaded here; the queries of this library do not read them. */
int nvq_points_to_grid(const float* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                       void* stream, nvq_grid** out);
/* The same for double world points (PtrT = Vec3d*, tools/cuda/PointsToGrid.cuh:571-572): voxels
 * are rounded in FP64, and point_type 9 (Default) stores the points as Vec3<double>. */
int nvq_points_to_grid_f64(const double* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                           void* stream, nvq_grid** out);
/* Point grid with the voxel size chosen for a target point density: PointsToGrid(maxPointsPerVoxel,
 * tolerance, maxIterations) (tools/cuda/PointsToGrid.cuh:335-342, the search in countNodes
 * :630-774).  Starting from voxel size 1, the largest number of points in one voxel is counted and
 * the voxel size updated until that count is within `tolerance` of max_points_per_voxel (clamped
 * to n), for at most max_iterations counts; the grid is then built as nvq_points_to_grid(_f64) with
 * the last voxel size.  xyz holds float (xyz_f64 = 0) or double world points.  voxel_size and
 * points_per_voxel (either may be NULL) receive the chosen voxel size and its largest count
 * (PointsToGrid::maxPointsPerVoxel). */
int nvq_points_to_grid_density(const void* xyz, int xyz_f64, uint64_t n, uint32_t max_points_per_voxel, int tolerance,
                               int max_iterations, uint32_t point_type, const char* name, void* stream, nvq_grid** out,
                               double* voxel_size, uint32_t* points_per_voxel);
/* Copy the first `bytes` bytes (at most nvq_grid_info.bytes; pass that for the whole grid) of a
 * grid's device buffer to host memory and wait for it: GridHandle::deviceDownload
 * (GridHandle.h:186-194). */
int nvq_grid_download(const nvq_grid* grid, void* host_dst, uint64_t bytes, void* stream);

/* tools::cuda::updateGridStats(d_grid, StatsMode::Default) (tools/cuda/GridStats.cuh:343-362) for
 * NVQ_FLOAT grids, in place on the device buffer: leaf bboxes (LeafNode::updateBBox), node and
 * root bboxes, min / max / average / standard deviation of every node, render flags, the world
 * bbox and the grid's statistics flags.  The merge order is the reference device code's (32
 * lanes per leaf, 128 threads per internal node), so the bits are its bits; they can differ in
 * the last place of an average from the host tools::updateGridStats.  Leaves without active
 * voxels keep their statistics and add nothing to their parent.  Waits for `stream` (the
 * grid's cached bbox is refreshed).  Other grid types: NVQ_ERR_UNSUPPORTED_TYPE. */
int nvq_update_grid_stats(nvq_grid* grid, void* stream);
/* The same for StatsMode `mode` (tools/cuda/GridStats.cuh:344-362): 1 BBox (bboxes and render
 * flags only; NoopStats), 2 MinMax (plus the minimum and maximum of every node; a parent merges
 * the extrema its children store, Extrema), 3 All (= nvq_update_grid_stats).  The grid's
 * HasMinMax / HasAverage / HasStdDeviation flags follow the mode.  0 (Disable) and other values:
 * NVQ_ERR_INVALID_ARG, as the reference throws for them. */
int nvq_update_grid_stats_mode(nvq_grid* grid, int mode, void* stream);

/* tools::cuda::signedFloodFill(d_grid) (tools/cuda/SignedFloodFill.cuh:240-253) for NVQ_FLOAT level
 * sets, in place on the device buffer: every inactive leaf voxel and every non-child node entry
 * becomes +-background with the sign of its nearest active neighbour in index order (next, else
 * previous), and inactive root tiles between two inside children on a z-scanline become
 * -background.  This is the reference device algorithm, which differs from the host
 * build::Grid fill on leaves and nodes that straddle the surface.  NVQ_ERR_INVALID_ARG when a root
 * tile the fill must write is missing (the reference throws); NVQ_ERR_UNSUPPORTED_TYPE for other
 * grid types.  A full checksum is not recomputed.  Waits for `stream`. */
int nvq_signed_flood_fill(nvq_grid* grid, void* stream);

/* Opt-in kernel timing for measurement tools (no counterpart in the reference).  While enabled,
 * each call records HIP events on its stream around its dominant kernel: "k_sample" (direct
 * point queries), "k_bin_sample" (binned point queries, n >= 2^20), "k_zero_crossing",
 * "k_fog_march", "k_fog_march_bwd", "k_multi_query".  nvq_profile_read waits for the recorded
 * launches and returns the summed milliseconds and launch count of one kernel name since the
 * last reset.  The iso-surface renderer records "k_iso_render" / "k_iso_render_persistent"; the
 * device builder "v2g_sort", "v2g_scan" (library sort and scan), "k_v2g_heads", "k_v2g_leaves".
 * Process-wide; thread-safe. */
int nvq_profile_enable(int enable);
int nvq_profile_read(const char* kernel, double* total_ms, uint64_t* launches);
int nvq_profile_reset(void);

#ifdef __cplusplus
}
#endif
#endif /* NVQ_H */
