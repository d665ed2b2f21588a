/*
 * CPU oracle for the volumetric path tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
 * this library, and only as the checker or the timed CPU baseline. The product (libcrvpt) never links or calls it.
 *
 * Two parts:
 *   1. An FP32 restatement of the reference's GLSL megakernel (Clouds.glsl + VptUtils.glsl + the tracker files)
 *      with the reference's own hybrid-Taus RNG, seeding and running-mean accumulation, so per-path arithmetic
 *      follows the reference shaders statement by statement (citations at each function in vpt_oracle.cpp).
 *   2. An FP64 exact optical-depth integrator (3D-DDA over voxels, Gauss-Legendre on trilinear cells) used as the
 *      closed-form check of every transmittance estimator.
 */
#ifndef CRVPT_ORACLE_H
#define CRVPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* VptMode (RenderSettings.hpp:32-36) subset on the hot path. */
enum { OR_DELTA = 0, OR_DELTA_SPECTRAL = 1, OR_RATIO = 2, OR_DECOMPOSITION = 3, OR_RESIDUAL_RATIO = 4, OR_NEXT_EVENT = 5,
       OR_NEXT_EVENT_SPECTRAL = 6 };
/* GridInterpolationType (RenderSettings.hpp:52-56). */
enum { OR_NEAREST = 0, OR_STOCHASTIC = 1, OR_TRILINEAR = 2 };
/* Environment kinds: BuiltinEnvMap DEFAULT/BLACK/SINGLE_COLOR (RenderSettings.hpp:118-121) + equirect image. */
enum { OR_ENV_DEFAULT = 0, OR_ENV_BLACK = 1, OR_ENV_SINGLE_COLOR = 2, OR_ENV_IMAGE = 3 };
/* Transmittance estimators for ray queries. */
enum { OR_T_DELTA = 0, OR_T_RATIO = 1, OR_T_RESIDUAL_RATIO = 2, OR_T_DECOMPOSITION = 3 };

/* Light (VptHeader.glsl:233-248 / LightEditorWidget.hpp:69-85), std140 layout, 64 bytes. */
typedef struct OracleLight {
    uint32_t light_type;   /* 0 point, 1 spot, 2 directional */
    uint32_t light_space;  /* 0 world, 1 view, 2 view orientation */
    float spot_total_width, spot_falloff_start;
    float color[3]; float intensity;
    float position[3]; uint32_t use_distance;
    float spot_direction[3]; float padding;
} OracleLight;

typedef struct OracleScene {
    /* Dense density grid, x fastest: index = x + (y + z*ny)*nx. Raw values; normalised with vmin/vmax. */
    const float* density;
    int nx, ny, nz;
    float voxel_value_min, voxel_value_max;
    float box_min[3], box_max[3];
    int interpolation;
    /* Medium. */
    float extinction[3];       /* extinction base * scale */
    float scattering_albedo[3];
    float phase_g;
    /* Sky. */
    int env_kind;
    float sun_direction[3];
    float sun_intensity[3];    /* sunlightIntensity * sunlightColor */
    const float* env_image;    /* RGBA32F, env_h rows of env_w texels, row 0 = top (v = 0) */
    int env_w, env_h;
    float env_intensity_factor[3];
    int is_env_map_black;
    int use_linear_rgb;
    /* Lights for next-event tracking. */
    int num_lights;
    const OracleLight* lights;
    /* Camera (column-major 4x4 like glm). */
    float inverse_view_proj[16];
    float inverse_view[16];
    /* Estimator. */
    int mode;
    uint32_t custom_seed_offset;
    int super_voxel_size;      /* requested supervoxel edge for RRT / decomposition (default 8) */
    int max_bounces;           /* scattering events before a path ends with zero further radiance; 0 = unbounded
                                  (a benchmark-workload bound with no counterpart in the reference) */
} OracleScene;

/* Camera matrices of the build's camera model (pos, target-or-front, up, fovy in radians, near, far). */
void oracle_camera_matrices(const double pos[3], const double target[3], const double up[3], double fovy,
                            double aspect, double znear, double zfar, float inv_view_proj[16], float inv_view[16],
                            float view_proj[16]);

/*
 * Renders frames [frame_start, frame_start + num_frames) into a running mean exactly as Clouds.glsl:172-179 does
 * (FP32 mix with 1/(frame+1)); acc_rgb is [h][w][3] and is read as the previous mean when frame_start > 0.
 * sum_sq (optional, [h][w][3] doubles) accumulates per-pixel squared path radiance for variance estimates.
 * Returns 0 on success.
 */
int oracle_render(const OracleScene* scene, int width, int height, int frame_start, int num_frames,
                  float* acc_rgb, double* sum_sq, int num_threads);

/*
 * The same samples as oracle_render, summed in FP64 into sum_rgb and sum_sq ([h][w][3] doubles each, added to):
 * the exact sample mean, free of the FP32 running mean's rounding drift (up to 5e-4 relative after 65536 frames),
 * for comparisons of converged images. Returns 0 on success.
 */
int oracle_render_sums(const OracleScene* scene, int width, int height, int frame_start, int num_frames,
                       double* sum_rgb, double* sum_sq, int num_threads);

/*
 * Feature maps of Clouds.glsl:146-383 (WRITE_*_MAP): every pointer that is not null receives its map as height x
 * width pixels of the reference image format's channels (RGBA: result, first_x, first_w, normal, cloud_only,
 * background, albedo; RG: depth (mean, std), density (mean, std), reproj_uv, flow, flow_reverse, depth_nabla;
 * R: depth_far, depth_fwidth) and is read back as the running state when frame_start > 0. prev_view_proj is the
 * previous frame's view-projection matrix (column-major). Returns 0 on success.
 */
typedef struct OracleFeatureMaps {
    float* result;
    float* first_x;
    float* first_w;
    float* normal;
    float* cloud_only;
    float* background;
    float* depth;
    float* depth_far;
    float* density;
    float* reproj_uv;
    float* albedo;
    float* flow;
    float* flow_reverse;
    float* depth_nabla;
    float* depth_fwidth;
} OracleFeatureMaps;
int oracle_render_features(const OracleScene* scene, int width, int height, int frame_start, int num_frames,
                           const float prev_view_proj[16], float far_distance, OracleFeatureMaps* maps,
                           int num_threads);

/*
 * Transmittance queries: for ray i, origin o[3i..], unit direction d[3i..], segment [0, t_max[i]] clipped to the box.
 * Each estimator is evaluated n_samples times with hybrid-Taus streams seeded from seed + i*n_samples + s; the
 * mean and the second moment per ray are written to mean[i], m2[i].
 */
int oracle_transmittance(const OracleScene* scene, int estimator, int n_rays, const float* o, const float* d,
                         const float* t_max, int n_samples, uint32_t seed, double* mean, double* m2,
                         int num_threads);

/* FP64 optical depth along o + t d, t in [0, t_max] (clipped to the box): exact for NEAREST (piecewise constant)
 * and for STOCHASTIC/TRILINEAR (trilinear field, per-cell 2-point Gauss-Legendre, exact for cubics). */
double oracle_optical_depth(const OracleScene* scene, const double o[3], const double d[3], double t_max);
int oracle_optical_depth_batch(const OracleScene* scene, int n_rays, const float* o, const float* d,
                               const float* t_max, double* tau, int num_threads);

/* Supervoxel tables of SuperVoxelGrid.cpp:40-405 (normalised densities). Output sizes: gx*gy*gz entries.
 * rrt: (mu_c, mu_r_bar) pairs; decomp: (min, max) pairs; occupancy bytes. Returns the effective supervoxel size. */
int oracle_supervoxel_tables(const OracleScene* scene, float extinction_x, int* grid_size3, float* rrt_pairs,
                             float* decomp_pairs, uint8_t* occupancy);

/* One path of the scene's tracker per ray, started at o[3i..] in unit direction d[3i..] instead of at the camera
 * (hybrid-Taus stream seeded with seed + i), radiance to out_rgb[3i..]. rrt_max_iterations >= 0 replaces the
 * residual-ratio tracker's fixed limit of 10 in-scattering iterations (0: the first traversal alone); out_iterations
 * (nullable) receives the iterations made. Lets tests pin the radiance loops against closed forms per ray. */
int oracle_trace_rays(const OracleScene* scene, int n_rays, const float* o, const float* d, uint32_t seed,
                      int rrt_max_iterations, float* out_rgb, int* out_iterations, int num_threads);

/* sampleSkybox(dir) + sampleLight(dir) (VptUtils.glsl:301-391) for n unit directions. */
int oracle_sky(const OracleScene* scene, int n, const float* dirs, float* out_rgb);

/* sampleHeadlight(pos, light_idx) (VptUtils.glsl:394-453): the radiance light light_idx sends towards each
 * position. */
int oracle_light_radiance(const OracleScene* scene, int light_idx, int n, const float* positions, float* out_rgb);

/* Hybrid Taus stream (VptUtils.glsl:46-65): n values after initializeRandom(seed). */
void oracle_hybrid_taus(uint32_t seed, int n, float* out);

#ifdef __cplusplus
}
#endif

#endif
