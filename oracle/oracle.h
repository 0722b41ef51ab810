/* Oracle: CPU restatement of NanoVDB's read-only query path, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
 * It is never linked into, loaded by or called from the product library.
 *
 * Everything reads the flat NanoVDB buffer by byte offsets (the layout pinned
 * by nanovdb/nanovdb/NanoVDB.h and its C99 port nanovdb/nanovdb/PNanoVDB.h).
 * FP32 functions reproduce the reference's float operation order exactly
 * (built with -ffp-contract=off, fmaf only where the reference writes fmaf);
 * the *_f64 functions keep the FP32 structural decisions (cells, steps,
 * stencils) and evaluate the arithmetic in FP64.
 */
#ifndef NVQ_ORACLE_H
#define NVQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 0 ok, <0 invalid (magic / version / unsupported type / truncated) */
int or_validate(const void* buf, uint64_t bytes);
uint32_t or_grid_type(const void* buf);
uint32_t or_leaf_count(const void* buf);

/* ReadAccessor::getValue / isActive / getDim(ijk, ray) per voxel */
void or_get_value(const void* buf, const int32_t* ijk, uint64_t n, float* val, uint8_t* active, uint32_t* dim);

/* nn = NN sampler (rintf), val = trilinear, grad per grad_mode
 * (1 FD6 of trilinear * 0.5/voxelSize, 2 GradStencil central difference, 3 trilinear analytic, index space).
 * world=1 maps xyz with the FP64 inverse map (Grid::worldToIndex). */
void or_sample(const void* buf, const float* xyz, uint64_t n, int world, float* nn, float* val, float* grad,
               int grad_mode);
/* FP64 evaluation: corner values decoded in FP64 (code*quantum+min) and FP64 lerps over
 * the same FP32 positions; used to split arithmetic error from codec error. */
void or_sample_f64(const void* buf, const float* xyz, uint64_t n, double* val, double* grad_fd6);

/* Adjoint of or_sample (no reference implementation exists).  or_sample_bwd_fixed adds the
 * FP32 / int64 fixed-point adjoint (2^-shift resolution) into accum and sets *overflow when a
 * contribution saturated or was not finite; or_sample_bwd_f64 adds the exact FP64 derivative
 * into grad; or_sample_loss_f64 is the FP64 forward loss sum(gnn*nn + gval*val + ggrad.grad)
 * with the value whose key is `key` raised by delta (index-space points).  Keys: Float grids
 * leafIndex*512 + n (leaf voxels only), index grids the channel index. */
void or_sample_bwd_fixed(const void* buf, const float* xyz, uint64_t n, int world, const float* gnn,
                         const float* gval, const float* ggrad, int grad_mode, int shift, int64_t* accum,
                         int* overflow);
void or_sample_bwd_f64(const void* buf, const float* xyz, uint64_t n, const double* gnn, const double* gval,
                       const double* ggrad, int grad_mode, double* grad);
double or_sample_loss_f64(const void* buf, const float* xyz, uint64_t n, const double* gnn, const double* gval,
                          const double* ggrad, int grad_mode, int64_t key, double delta);

/* GradStencil at integer voxels: gradient, normSqGrad (Godunov), laplacian, zeroCrossing */
void or_grad_stencil(const void* buf, const int32_t* ijk, uint64_t n, float* grad, float* normsq, float* lap,
                     uint8_t* zc);

/* world rays -> worldToIndexF -> zeroCrossing (+ normalised world-space trilinear-gradient normal) */
uint64_t or_value_count(const void* buf);
void or_iso_crossing(const void* buf, const float* eye, const float* dir, uint64_t n, float iso, uint8_t* hit,
                     int32_t* ijk, float* t, float* v, float* normal);
void or_iso_render(const void* buf, int32_t width, int32_t height, float iso, float dx, float wbbox_dim_z,
                   float tan_half_fov, const float origin[3], float* image, uint8_t* hit, float* t);
/* the same with the ray mapped by the FP64 Ray::applyInverseMap when f64map != 0 */
void or_iso_crossing_map(const void* buf, const float* eye, const float* dir, uint64_t n, float iso, int f64map,
                         uint8_t* hit, int32_t* ijk, float* t, float* v, float* normal);
void or_zero_crossing(const void* buf, const float* eye, const float* dir, uint64_t n, uint8_t* hit, int32_t* ijk,
                      float* t, float* v, float* normal);

/* Emission-absorption fog march: FP32 HDDA/step structure, FP64 (L,T); steps per ray. */
void or_fog_march(const void* buf, const float* eye, const float* dir, uint64_t n, float dt, float sigma, double* L,
                  double* T, uint32_t* steps);
/* FP64 adjoint of or_fog_march: accumulates dLoss/dvalue into grad (leafCount*512 doubles,
 * leaf order of the buffer) given dLoss/dL and dLoss/dT per ray (dT may be NULL). */
void or_fog_march_adjoint(const void* buf, const float* eye, const float* dir, uint64_t n, float dt, float sigma,
                          const double* dL, const double* dT, double* grad);

/* tools/cuda/PointsToGrid.cuh voxelsToGrid<float | ValueOnIndex> restated on the CPU: the
 * buffer (malloc'd, free with or_free) for n integer voxels ijk[n][3], duplicates allowed.
 * grid_type 1 (float: active voxels 1, inactive 0) or 20 (ValueOnIndex).  out_index (index
 * grids only, may be NULL) receives each input voxel's value index.  0 ok, -1 bad argument. */
int or_voxels_to_grid(const int32_t* ijk, uint64_t n, uint32_t grid_type, double voxel_size, const char* name,
                      void** out, uint64_t* bytes_out, uint64_t* out_index);
void or_free(void* p);
/* tools/cuda/PointsToGrid.cuh pointsToGrid restated for n float world points xyz[n][3]: a Point
 * grid (GridType PointIndex) with the points as blind data encoded by point_type (PointType
 * 1..9).  0 ok, -1 bad argument, -3 more than 65535 points in one leaf. */
int or_points_to_grid(const float* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                      void** out, uint64_t* bytes_out);
/* The same for double world points (Vec3d input). */
int or_points_to_grid_f64(const double* xyz, uint64_t n, double voxel_size, uint32_t point_type, const char* name,
                          void** out, uint64_t* bytes_out);
/* PointsToGrid(maxPointsPerVoxel, tolerance, maxIterations): the voxel-size search, then the grid
 * at the chosen voxel size (dx_out) and its largest points-per-voxel count (density_out). */
int or_points_to_grid_density(const void* xyz, int f64, uint64_t n, uint32_t target, int tolerance, int max_iterations,
                              uint32_t point_type, const char* name, void** out, uint64_t* bytes_out,
                              double* dx_out, uint32_t* density_out);
/* tools/cuda/GridStats.cuh updateGridStats<float>(StatsMode::Default) restated, in place on a host
 * float grid buffer (bboxes, node statistics, flags, world bbox).  0 ok, -4 not a float grid. */
int or_update_grid_stats(void* buf);
/* The same for StatsMode 1 (BBox), 2 (MinMax) or 3 (All = Default); -1 for any other mode. */
int or_update_grid_stats_mode(void* buf, int mode);
/* tools/cuda/SignedFloodFill.cuh signedFloodFill<float> restated, in place on a host float grid
 * buffer.  0 ok, -4 not a float grid, -3 a root tile the fill must write is missing. */
int or_signed_flood_fill(void* buf);

/* TreeMarcher<NodeT>::init/step for world rays (worldToIndexF): level 0 leaf, 1 lower, 2 upper
 * nodes.  count[i] = the ray's number of spans; the first max_spans are written at
 * [i*max_spans ...): t0, t1 and the node's byte offset from the start of the buffer. */
void or_tree_march(const void* buf, const float* eye, const float* dir, uint64_t n, int level, uint32_t max_spans,
                   uint32_t* count, float* t0, float* t1, int64_t* node);
/* firstActive for world rays (worldToIndexF): hit, ijk, t (zeros on a miss) */
void or_first_active(const void* buf, const float* eye, const float* dir, uint64_t n, uint8_t* hit, int32_t* ijk,
                     float* t);
/* DDA<Ray, Coord, dim> trace in the format of or_hdda_trace */
void or_dda_trace(const float* eye, const float* dir, float t0, float t1, int dim, int nsteps, float* times,
                  float* nexts, int32_t* voxels, uint8_t* ok);
/* BoxStencil at index-space xyz: interpolation and the world-space (FP64 applyIJT) gradient */
void or_box_stencil(const void* buf, const float* xyz, uint64_t n, float* interp, float* grad);
/* GradStencil upwind gradient along V[n][3] and closest-point transform at integer voxels */
void or_grad_stencil_ext(const void* buf, const int32_t* ijk, uint64_t n, const float* V, float* upwind, float* cpt);
/* TrilinearSampler::zeroCrossing of the cell at Floor(xyz) */
void or_sample_zero_crossing(const void* buf, const float* xyz, uint64_t n, uint8_t* zc);

/* Known-answer helpers */
void or_hdda_trace(const float* eye, const float* dir, float t0, float t1, int dim, int nsteps, float* times,
                   float* nexts, int32_t* voxels, uint8_t* ok);
int or_ray_clip(const float* eye, const float* dir, float t0, float t1, const int32_t* bbox, float* out);

#ifdef __cplusplus
}
#endif
#endif
