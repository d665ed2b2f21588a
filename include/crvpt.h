/*
 * crvpt — C ABI of the MI355X (gfx950) volumetric path tracer.
 *
 * The entry points mirror the reference's PyTorch module `vpt` (src/PyTorch/Module.cpp, stub vpt.pyi) one to one:
 * same names with a crvpt_ prefix, same argument meaning and order, plain pointers and sizes instead of torch
 * tensors and std::vector. Every function returns 0 on success and a non-zero CRVPT_ERR_* code on failure;
 * crvpt_last_error() then returns the message the reference would have raised through sgl::Logfile::throwError.
 * One global, reference-counted renderer per process, as in the reference (Module.cpp:283-284, 701, 732-813);
 * calls are not thread-safe (the reference has no locks either).
 */
#ifndef CRVPT_H
#define CRVPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CRVPT_OK = 0,
    CRVPT_ERR_INVALID_ARGUMENT = 1,
    CRVPT_ERR_NOT_INITIALIZED = 2,
    CRVPT_ERR_HIP = 3,
    CRVPT_ERR_IO = 4,
    CRVPT_ERR_UNSUPPORTED = 5,
};

/* Last error message of the calling process ("" when none). */
const char* crvpt_last_error(void);

/* Module.cpp:732-802 initialize(): ref-counted; the first call creates the renderer with the reference defaults
 * (fog-sphere scene of radius 0.25 and voxel size 0.01, DELTA_TRACKING, linear RGB). */
int crvpt_initialize(void);
/* Module.cpp:804-813 cleanup(): drops one reference; the last one frees all device memory. */
int crvpt_cleanup(void);

/* Module.cpp:286-317 render_frame(input_tensor[C,H,W] f32, frame_count): renders frame_count samples per pixel
 * into the running mean kept since the last reset and writes the image as C planes of H*W floats (C = 3 or 4,
 * alpha = 1). out_on_device != 0: `out` is a device pointer written on `stream` (a hipStream_t, 0 = null stream)
 * without host round trips, as the CUDA branch (Module.cpp:307-309); otherwise `out` is host memory and the
 * call returns after the copy, as renderFrameCpu. */
int crvpt_render_frame(float* out, int channels, int height, int width, int frame_count, int out_on_device,
                       void* stream);

/* Module.cpp:815-824 load_cloud_file / load_volume_file (.xyz dense float grids, CloudData.cpp:236-316;
 * .nvdb NanoVDB float grids, CloudData.cpp:1360-1365). */
int crvpt_load_cloud_file(const char* filename);
/* No reference counterpart, host only (no GPU needed): the .xyz parse that load_cloud_file applies
 * (CloudData::loadFromXyzFile, CloudData.cpp:236-316): grid sizes, voxel sizes, and with out != NULL the field x
 * fastest (transposed from the file's z-fastest order) and min-max normalised when max - min > 1e-6. */
int crvpt_read_xyz_file(const char* filename, int32_t dims[3], double voxel_size[3], float* out, uint64_t capacity);
/* No reference counterpart, host only (no GPU needed): the .nvdb read that load_cloud_file applies
 * (CloudData::loadFromNvdbFile, CloudData.cpp:1360-1365: io::readGrid of the file's first grid, uncompressed): the
 * grid's bytes, validated as a NanoVDB 32.3 FP32 or Fp16 grid. With out == NULL only *size is written. */
int crvpt_read_nvdb_file(const char* filename, uint8_t* out, uint64_t capacity, uint64_t* size);
/* CloudData::setDensityField (CloudData.cpp:147-167): dense grid, x fastest, size nx*ny*nz (copied). */
int crvpt_set_density_field(int nx, int ny, int nz, const float* values);
/* CloudData::setNanoVdbGridHandle (CloudData.cpp:169-172): a NanoVDB 32.3 FP32 grid buffer (copied). */
int crvpt_set_nanovdb_buffer(const uint8_t* data, uint64_t size);
/* No reference counterpart: seeded fBm procedural cloud generated on the GPU as the dense density field
 * (octaves, lacunarity, gain, base frequency over the unit cube, smoothstep(edge0, edge1) remap, optional anvil
 * height profile). */
int crvpt_generate_fbm_cloud(int nx, int ny, int nz, uint32_t seed, int octaves, float lacunarity, float gain,
                             float base_frequency, float edge0, float edge1, int height_falloff);
/* No reference counterpart, host only (no GPU needed): the dense -> NanoVDB conversion of
 * CloudData::getSparseDensityField (CloudData.cpp:1367-1438) on caller data: FP32 fog volume, background 0,
 * voxel size dx, index (0,0,0) at world origin[3], grid name "tmp". With out == NULL only *size is written. */
int crvpt_build_nanovdb_from_dense(const float* values, int nx, int ny, int nz, double dx, const double origin[3],
                                   uint8_t* out, uint64_t capacity, uint64_t* size);
/* The same with a leaf value type: 0 FP32 (as above), 1 Fp16 (NanoVDB GridType::Fp16, 16-bit codes per leaf with
 * the leaf's minimum and quantum, as nanovdb::GridBuilder<float, Fp16> writes them without dithering). No reference
 * counterpart: the reference renders FP32 grids only. */
int crvpt_build_nanovdb_from_dense_typed(const float* values, int nx, int ny, int nz, double dx,
                                         const double origin[3], int value_type, uint8_t* out, uint64_t capacity,
                                         uint64_t* size);
/* Host only (no GPU needed): nanovdb::createFogVolumeSphere<float>(radius, center, voxel_size, half_width)
 * (src/nanovdb/util/Primitives.h:84-95, 893-914), the grid of the module's default scene
 * (Module.cpp:793-795: radius 0.25, centre 0, voxel 0.01, half width 3). With out == NULL only *size is written. */
int crvpt_build_fog_volume_sphere(float radius, const float center[3], double voxel_size, double half_width,
                                  uint8_t* out, uint64_t capacity, uint64_t* size);
/* No reference counterpart: leaf value type of the NanoVDB grid that set_use_sparse_grid builds from a dense field
 * (0 FP32, the reference's; 1 Fp16: half the leaf memory, values quantised per leaf). Loaded .nvdb files and
 * crvpt_set_nanovdb_buffer grids may be either type. */
int crvpt_set_sparse_value_type(int value_type);
/* Copies the current dense field (normalised or raw) or the current NanoVDB buffer to host memory. */
int crvpt_get_density_field(float* out, uint64_t count, int normalized);
int crvpt_get_nanovdb_buffer(uint8_t* out, uint64_t capacity, uint64_t* size);

/* VolumetricPathTracingPass::setUseSparseGrid / setSparseGridInterpolationType (VolumetricPathTracingPass.cpp:
 * 863-875); interpolation: 0 nearest, 1 stochastic, 2 trilinear (RenderSettings.hpp:52-56). */
int crvpt_set_use_sparse_grid(int use_sparse);
/* Module.cpp:1074-1076 set_output_foreground_map (VolumetricPathTracingPass::setOutputForegroundMap,
 * VolumetricPathTracingPass.cpp:1111-1126; the OUTPUT_FOREGROUND_MAP store of Clouds.glsl:181-191): render_frame
 * returns the running mean of (radiance, 1) over samples whose first scattering or absorption event has a value and
 * (0, 0, 0, 0) over the others, so alpha is the foreground coverage. Residual ratio and decomposition tracking record
 * no first event, so their foreground map is 0. Changing it resets the accumulation. */
int crvpt_set_output_foreground_map(int output_foreground_map);
int crvpt_set_grid_interpolation_type(int type);

/* Module.cpp set_vpt_mode / set_vpt_mode_from_name (VptMode order and names of RenderSettings.hpp:32-45). */
int crvpt_set_vpt_mode(int64_t mode);
int crvpt_set_vpt_mode_from_name(const char* mode_name);

/* Module.cpp medium setters: set_scattering_albedo(vec3), set_extinction_scale, set_extinction_base(vec3),
 * set_phase_g, set_seed_offset. */
int crvpt_set_scattering_albedo(const double albedo[3]);
int crvpt_set_extinction_scale(double extinction_scale);
int crvpt_set_extinction_base(const double extinction_base[3]);
int crvpt_set_phase_g(double phase_g);
int crvpt_set_seed_offset(int64_t offset);
/* Pass setting superVoxelSize (VolumetricPathTracingPass.hpp:254, GUI slider VolumetricPathTracingPass.cpp:2569,
 * range 1-64): edge of the supervoxel tables of residual ratio and decomposition tracking (SuperVoxelGrid.cpp:40-60;
 * halved until it divides grids smaller than the edge). */
int crvpt_set_super_voxel_size(int32_t super_voxel_size);
/* Scattering events per path before the path ends with zero further radiance (0 = unbounded). No counterpart in
 * the reference, whose trackers never cut a path (DeltaTracking.glsl:167-345 loops until the path leaves or is
 * absorbed); it bounds path length for benchmark workloads that state a maximum bounce count. */
int crvpt_set_max_bounces(int32_t max_bounces);

/* Module.cpp camera API: get_camera_position, get_camera_view_matrix (16, column-major like glm),
 * get_camera_fovy (radians), set_camera_position, set_camera_target, overwrite_camera_view_matrix,
 * set_camera_fovy (degrees, Module.cpp:1386). */
int crvpt_get_camera_position(double out[3]);
int crvpt_get_camera_view_matrix(double out[16]);
int crvpt_get_camera_fovy(double* out);
int crvpt_set_camera_position(const double position[3]);
int crvpt_set_camera_target(const double target[3]);
int crvpt_overwrite_camera_view_matrix(const double view[16]);
int crvpt_set_camera_fovy(double fovy_degrees);

/* Module.cpp environment API: set_use_builtin_environment_map("Default" | "Black" | "Single Color"),
 * load_environment_map_from_tensor (float32 [C=3|4, H, W] rows top to bottom, linear RGB),
 * set_environment_map_intensity, set_environment_map_intensity_rgb. */
int crvpt_set_use_builtin_environment_map(const char* name);
int crvpt_load_environment_map_from_buffer(const float* pixels_chw, int channels, int height, int width);
int crvpt_set_environment_map_intensity(double factor);
int crvpt_set_environment_map_intensity_rgb(const double factor[3]);

/* Module.cpp light API (LightEditorWidget.cpp:280-320): set_use_lights, clear_lights, add_light, remove_light,
 * set_light_property(idx, key, value) with keys type, color, intensity, use_distance, position, direction, space,
 * spot_total_width, spot_falloff_start, spot_direction. */
int crvpt_set_use_lights(int use_lights);
int crvpt_clear_lights(void);
int crvpt_add_light(void);
int crvpt_remove_light(int64_t light_idx);
int crvpt_set_light_property(int64_t light_idx, const char* key, const char* value);
int crvpt_get_num_lights(int64_t* out);

/* Module.cpp volume placement: set_max_grid_extent, get_render_bounding_box (xmin, xmax, ymin, ymax, zmin, zmax),
 * get_volume_voxel_size (z, y, x). */
int crvpt_set_max_grid_extent(double max_grid_extent);
int crvpt_get_render_bounding_box(double out[6]);
int crvpt_get_volume_voxel_size(int64_t out[3]);

/* Module.cpp denoiser API: set_denoiser(name) with "None", "Default" (EAW), "EAW", "SVGF"; set_denoiser_property
 * (EAW: maxNumIterations, phiColor, phiPosition, phiNormal, useColorDistance; SVGF: numIterations). */
int crvpt_set_denoiser(const char* name);
int crvpt_set_denoiser_property(const char* key, const char* value);

/* No reference counterpart: batched transmittance along n_rays segments o + t d, t in [0, t_max] (t_max may be
 * NULL: to the box exit), n_samples estimates per ray with estimator 0 delta, 1 ratio, 2 residual ratio,
 * 3 decomposition. All arrays are device pointers when on_device != 0, host otherwise; mean and m2 receive the
 * per-ray mean and second moment (m2 may be NULL). */
int crvpt_transmittance_batch(int estimator, int n_rays, const float* origins, const float* directions,
                              const float* t_max, int n_samples, uint32_t seed, float* mean, float* m2,
                              int on_device, void* stream);

/* No reference counterpart: exactness probes of the device grid accessors on the current volume.
 * grid_lookup_batch: voxel values at integer coordinates (NanoVDB index space for the sparse grid, normalised texels
 * for the dense grid); grid_lookup_world_batch (sparse grid only): pnanovdb_grid_world_to_indexf + floor + lookup of
 * NanoVDB world coordinates, also returning the selected voxels. Host arrays. */
int crvpt_grid_lookup_batch(int n, const int32_t* ijk, float* out);
int crvpt_grid_lookup_world_batch(int n, const float* xyz, float* out, int32_t* ijk_out);
/* No reference counterpart: the supervoxel tables the device built for the residual-ratio and decomposition trackers
 * (SuperVoxelGrid.cpp:40-405): edge, grid size, (min, max) and (mu_c, mu_r_bar) pairs (capacity in floats). */
int crvpt_get_supervoxel_tables(int* supervoxel_size, int grid_size[3], float* decomposition_min_max,
                                float* rrt_mu_c_mu_r, uint64_t capacity);

/* No reference counterpart: performance controls of the gfx950 path pool. use_majorant_grid: local majorants
 * from the supervoxel grid (1, default) or the reference's global majorant (0); supervoxel_size: edge in voxels of
 * the majorant grid (default 8 dense, 16 sparse when 0); blocks: persistent workgroups (0 = occupancy based);
 * steps_per_refill: walker events between refill ballots (0 = the estimator's default: 24 for residual ratio and
 * next-event tracking, 16 for the others). */
int crvpt_set_use_majorant_grid(int use_majorant_grid);
int crvpt_set_launch_config(int supervoxel_size, int blocks, int steps_per_refill);

/* Module.cpp:319-330 set_use_feature_maps(names) with the names mapped to FeatureMapTypeVpt values
 * (VolumetricPathTracingPass.hpp:67-80): the listed maps are accumulated by the following renders. Produced here:
 * 0 Result, 1 First X, 2 First W, 3 Normal, 4 Normal (Length 1), 5 Cloud Only, 6 Depth, 7 Depth Far, 8 Flow,
 * 9 Flow Reverse, 10 Depth (nabla), 11 Depth (fwidth), 12 Density, 13 Background, 14 Reprojected UV, 17 Albedo;
 * the others report CRVPT_ERR_UNSUPPORTED. */
int crvpt_set_use_feature_maps(const int32_t* maps, int count);
/* Module.cpp:1078-1081 set_feature_map_type(type): the map render_frame returns instead of the result. */
int crvpt_set_feature_map_type(int32_t type);
/* Module.cpp:343-400 get_feature_map(input_tensor, map): the map of the last render as C planes of H*W floats
 * (the reference's RGBA blit: RG and R maps have 0 in the missing channels and alpha 1). Device / host output as in
 * crvpt_render_frame. */
int crvpt_get_feature_map(int32_t map, float* out, int channels, int height, int width, int out_on_device,
                          void* stream);
/* Module.cpp:1339-1341 set_view_projection_matrix_as_previous(): the current view-projection becomes the
 * previous-frame matrix of the Flow / Flow Reverse / Reprojected UV maps (until then the current one is used). */
int crvpt_set_view_projection_matrix_as_previous(void);

/* modules/kpn_module per_pixel_kernel::forward (perPixelKernel.cpp:31-40, Kernel.cu:38-131) on device buffers:
 * image [n, c, img_h, img_w], weights [n, k*k, weight_h, weight_w] (already normalised) -> out
 * [n, c, min(img_h, weight_h), min(img_w, weight_w)]; c <= 4; stream is a hipStream_t. Device tensors only, as in
 * the reference (which rejects CPU tensors). */
int crvpt_kpn_per_pixel_kernel(const float* image, const float* weights, float* out, int n, int c, int img_h,
                               int img_w, int weight_h, int weight_w, int kernel_size, void* stream);
/* PyTorchDenoiser.cpp:174-179 inv_iter_kernel: getInvIterBase (Kernel.cu:137-186; sigmoid-gated blend of the
 * low-resolution prediction, the previous frame and the detail, with the reference's weight batch stride) followed by
 * the per-pixel kernel with softmax over the first k*k channels of fused_weights [n, weight_c >= k*k + 3, weight_h,
 * weight_w]. x, prev: [n, c, img_h, img_w]; low_res_pred, low_res_x: [n, c, lo_h, lo_w]. */
int crvpt_kpn_inv_iter_kernel(const float* x, const float* low_res_pred, const float* low_res_x, const float* prev,
                              const float* fused_weights, float* out, int n, int c, int img_h, int img_w, int lo_h,
                              int lo_w, int weight_c, int weight_h, int weight_w, int kernel_size, int use_previous,
                              void* stream);
/* No reference counterpart: crvpt_kpn_inv_iter_kernel without device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops). The blended base image and the normalised weights live in
 * a fixed 128 MiB device array per GPU (32 Mi floats; n*c*min(img_h, weight_h)*min(img_w, weight_w) +
 * n*k*k*weight_h*weight_w must fit, else CRVPT_ERR_INVALID_ARGUMENT); the array is shared by every call on that GPU,
 * so calls that overlap on one GPU race. Same arguments and results as crvpt_kpn_inv_iter_kernel. */
int crvpt_kpn_inv_iter_kernel_noalloc(const float* x, const float* low_res_pred, const float* low_res_x,
                                      const float* prev, const float* fused_weights, float* out, int n, int c,
                                      int img_h, int img_w, int lo_h, int lo_w, int weight_c, int weight_h,
                                      int weight_w, int kernel_size, int use_previous, void* stream);

/* No reference counterpart: statistics of the last crvpt_render_frame: kernel milliseconds of the path kernels
 * (device events on the render stream), number of paths, number of path-kernel launches. */
int crvpt_get_last_render_stats(double* path_kernel_ms, double* total_ms, uint64_t* num_paths, int* launches);

/* No reference counterpart: work counters of the path kernels for the algorithmic-byte roofline. When enabled,
 * each crvpt_render_frame runs the instrumented variant of the path kernel and accumulates [density taps,
 * supervoxel majorant / table reads, paths, NanoVDB tree descents (lookups that missed the per-lane 8^3-block
 * memo)] (wave-reduced atomics at kernel exit); crvpt_get_last_render_counters reads the totals of the last call. */
int crvpt_set_collect_stats(int collect);
int crvpt_get_last_render_counters(uint64_t out[4]);

/* Kernel-prediction denoiser network. No reference counterpart in C++ (the reference loads a TorchScript model
 * trained elsewhere, PyTorchDenoiser.cpp:644, and applies its kernels with modules/kpn_module Kernel.cu:38-83): six 3x3
 * convolutions (15 -> 32 -> 32 -> 32 -> 32 -> 32 -> 25 channels, ReLU between them) in bf16 with FP32 accumulation on
 * the matrix cores, whose 25 logits per pixel are turned by a softmax into a 5x5 filter of the noisy radiance.
 * Parameters: CRVPT_KPN_PARAMS floats, weights [layer][co][ci][ky * 3 + kx] (32 x 32 x 9 per layer, padding channels
 * zero), then biases [layer][co]. All tensors are device buffers on the caller's stream (a hipStream_t):
 *   features [n, 15, h, w] f32: radiance mean (3), albedo (3), depth, depth standard deviation, transmittance
 *   (1 - coverage), screen-space gradient (2), radiance variance (3), density; radiance and variance enter the
 *   network as log(1 + v);
 *   noisy, target, out [n, 3, h, w] f32 (linear radiance).
 * Loss: mean over pixels and channels of (x - r)^2 / (r^2 + 0.01). Optimiser: Adam (0.9, 0.999, 1e-8). */
#define CRVPT_KPN_PARAMS 55488
int crvpt_kpn_init(uint32_t seed);
int crvpt_kpn_set_parameters(const float* params, uint64_t count);
int crvpt_kpn_get_parameters(float* out, uint64_t count);
/* Forward and backward pass; gradients replace the previous ones; *loss (host, nullable) gets the batch loss. */
int crvpt_kpn_compute_gradients(const float* features, const float* noisy, const float* target, int n, int h, int w,
                                float* loss, void* stream);
int crvpt_kpn_get_gradients(float* out, uint64_t count);
int crvpt_kpn_adam_step(float lr, void* stream);
int crvpt_kpn_train_step(const float* features, const float* noisy, const float* target, int n, int h, int w,
                         float lr, float* loss, void* stream);
int crvpt_kpn_denoise(const float* features, const float* noisy, float* out, int n, int h, int w, void* stream);
/* Logits [n, h, w, 32] f32 of the last forward pass (channels 25..31 are padding). */
int crvpt_kpn_get_logits(float* out, int n, int h, int w, void* stream);
int crvpt_kpn_release(void);

/* Batched render queue. No reference counterpart (the reference renders one frame per render_frame call,
 * Module.cpp:286-317); for callers that render many scenes at once, e.g. training-pair generation.
 * crvpt_batch_add_request snapshots the current scene exactly as crvpt_render_frame would render it (volume in its
 * active representation with the supervoxel tables of the current estimator, camera, medium, phase, lights,
 * environment, seed offset, bounce limit) and queues spp fresh samples per pixel whose mean goes to out, a device
 * buffer of [channels, height, width] floats (channels 3 or 4, alpha 1) that must stay valid while the batch is
 * rendered. Consecutive requests of an unchanged volume share one snapshot. All requests of a batch use the estimator
 * of the first. Request i draws the random numbers of a single-frame render of the same scene with frames
 * frame_offset .. frame_offset + spp - 1, so a batch with one sample per unit reproduces those images up to the
 * last bits of floating-point rounding. */
int crvpt_batch_clear(void);
int crvpt_batch_add_request(float* out, int channels, int height, int width, int spp, int32_t* request_id);
/* Renders every request with one persistent launch on stream (a hipStream_t; 0 = the null stream): the counter
 * reset, the path pool and the image resolve are captured once in a HIP graph and replayed while the batch and its
 * launch configuration are unchanged. Asynchronous: synchronise the stream before reading the outputs. */
int crvpt_batch_render(uint32_t frame_offset, void* stream);
/* Launch configuration {workgroups (0 = occupancy limit x CUs), threads per workgroup (64..256, multiple of 64),
 * walker steps between refills (compaction interval), samples per claimed unit (paths per lane)}. */
int crvpt_batch_set_launch_config(int blocks, int block_size, int steps_per_refill, int samples_per_unit);
int crvpt_batch_get_launch_config(int32_t config[4]);
/* Times candidate launch configurations on the batch with spp / spp_divisor samples per request (coordinate descent
 * over the compaction interval, samples per unit, block size and workgroup count), keeps the fastest and returns it
 * as in crvpt_batch_get_launch_config. The tuning renders overwrite the request outputs. */
int crvpt_batch_autotune(int spp_divisor, int32_t config[4]);
/* Device time of the last crvpt_batch_render (events around the graph launch; waits for it), total paths and
 * requests of the batch. */
int crvpt_batch_get_stats(double* kernel_ms, uint64_t* num_paths, int32_t* num_requests);

#ifdef __cplusplus
}
#endif

#endif /* CRVPT_H */
