/*
 * nvq host builder: writes NanoVDB grid buffers (ABI 32.9, unmodified layout)
 * on the host for upload with nvq_grid_upload.  Part of libnvq.so; plain C ABI.
 *
 * Each producer returns a malloc'ed, 32-byte aligned buffer in *out (size in
 * *bytes) that the caller releases with nvqb_free.  Status codes are those of
 * nvq.h.  grid_type selects the leaf codec: NVQ_FLOAT, NVQ_FP4, NVQ_FP8,
 * NVQ_FP16 or NVQ_FPN (FpN uses the absolute-error tolerance `tol`).
 *
 * The buffers match, byte for byte, the ones the reference host tools write
 * for the same inputs, except the 8-byte checksum field, which is written in
 * its "disabled" state (all bits set, GridData::init in NanoVDB.h:1979).
 */
#ifndef NVQ_BUILD_H
#define NVQ_BUILD_H

#include <stdint.h>

#include "nvq.h"

#ifdef __cplusplus
extern "C" {
#endif

/* createLevelSetSphere<BuildT>(radius, center, voxelSize, halfWidth, origin=0, "sphere_ls")
 * (tools/CreatePrimitives.h:1016-1100): narrow-band SDF, signed-flood-filled. */
int nvqb_level_set_sphere(double radius, const double center[3], double voxel_size, double half_width,
                          uint32_t grid_type, float tol, void** out, uint64_t* bytes);

/* createLevelSetTorus<BuildT>(R, r, center, voxelSize, halfWidth, origin=0, "torus_ls")
 * (tools/CreatePrimitives.h:1186-1243). */
int nvqb_level_set_torus(double major_radius, double minor_radius, const double center[3], double voxel_size,
                         double half_width, uint32_t grid_type, float tol, void** out, uint64_t* bytes);

/* createFogVolumeSphere<float>(radius, center, voxelSize, halfWidth, origin=0, "sphere_fog")
 * (tools/CreatePrimitives.h:1103-1160): level set mapped by levelSetToFog. */
int nvqb_fog_volume_sphere(double radius, const double center[3], double voxel_size, double half_width, void** out,
                           uint64_t* bytes);

/* Fog volume over the index box bbox[6] = (min xyz, max xyz) filled the way
 * build::Grid<float>::operator()(func, bbox) fills it (tools/GridBuilder.h:1927-2019):
 * density = clamp(amplitude * fBm(p), 0, 1) with fBm = sum_{o<octaves} 0.5^o * noise(p * base_freq * 2^o),
 * noise = improved Perlin noise (Perlin 2002) over a permutation shuffled from `seed`;
 * densities below `threshold` become the background (0, inactive); constant full leaves
 * are pruned into active tiles.  Voxel size 1, origin 0, class FogVolume. */
int nvqb_fog_fbm(const int32_t bbox[6], uint32_t octaves, double base_freq, uint64_t seed, float amplitude,
                 float threshold, uint32_t grid_type, float tol, void** out, uint64_t* bytes);

/* Evaluate the density used by nvqb_fog_fbm at n voxels ijk[n][3] (host). */
int nvqb_fbm_density(const int32_t* ijk, uint64_t n, uint32_t octaves, double base_freq, uint64_t seed,
                     float amplitude, float threshold, float* out);

/* Re-encode an FP32 NanoVDB buffer with another leaf codec (createNanoGrid<NanoGrid<float>, BuildT>,
 * tools/CreateNanoGrid.h:1331-1474). */
int nvqb_quantize(const void* float_buf, uint64_t bytes_in, uint32_t grid_type, float tol, void** out,
                  uint64_t* bytes);

/* Pinhole camera rays, the RayGenOp convention of examples/ex_raytrace_level_set/common.h:91-125
 * (pixel i -> x = i % w, y = i / w; P = (2u-1)*tan(fov/2)*aspect, (2v-1)*tan(fov/2); dir =
 * normalize(P.x*right + P.y*up - forward)) for an eye looking at `look_at`.  With eye on the
 * +z axis looking at the origin and up = +y the rays equal RayGenOp's bit for bit. */
int nvqb_camera_rays(uint32_t width, uint32_t height, const float eye[3], const float look_at[3], const float up[3],
                     float fov_deg, float* out_eye, float* out_dir);

void nvqb_free(void* p);

/* The camera of the iso-surface example's RenderOp (examples/ex_raytrace_iso_surface/common.h:26-37)
 * for a width x height image of a NanoVDB grid buffer: voxel_size = float(voxelSize[0]), iso =
 * voxel_size, wbbox_dim_z = float(worldBBox.dim()[2]) * 2, origin = float(worldBBox.min() +
 * worldBBox.dim() * 0.5) + (0, 0, wbbox_dim_z), tan_half_fov = tanf(45 / 2 degrees). */
int nvqb_iso_camera(const void* grid_buf, uint64_t bytes, int32_t width, int32_t height, nvq_iso_camera* out);

/* Read grid number `index` (name == NULL) or the first grid called `name` from a file: either a
 * NanoVDB file of segments (io::readGrid, nanovdb/nanovdb/io/IO.h:489-600; codec NONE only, a
 * ZIP or BLOSC segment gives NVQ_ERR_UNSUPPORTED_TYPE) or a raw file of grid buffers
 * (GridHandle::write).  As io::readGrid does, the returned grid is numbered 0 of 1
 * (tools::updateGridCount, tools/GridChecksum.h:412-420, header CRC refreshed when the grid
 * carries a checksum).  Missing grid: NVQ_ERR_INVALID_ARG. */
int nvqb_read_nvdb(const char* path, int32_t index, const char* name, void** out, uint64_t* bytes);

/* Write one grid buffer as a single-segment NanoVDB file, byte for byte what
 * io::writeGrid(path, handle, io::Codec::NONE) writes for a one-grid handle (io/IO.h:444-467). */
int nvqb_write_nvdb(const char* path, const void* grid_buf, uint64_t bytes);

/* CRC-32 (reflected, polynomial 0xEDB88320) of size bytes, as NanoVDB's checksums use it. */
uint32_t nvqb_crc32(const void* data, uint64_t size);

#ifdef __cplusplus
}
#endif
#endif /* NVQ_BUILD_H */
