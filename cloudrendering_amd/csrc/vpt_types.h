// Plain-old-data types shared by the host library and the gfx950 kernels of the volumetric path tracer.
// Every struct is fixed-size and passed by value to a launch; device pointers inside point at HBM allocations
// owned by the host context.
#pragma once

#include <stdint.h>

namespace crvpt {

// VptMode of the reference (src/PathTracer/RenderSettings.hpp:32-36); values follow the reference's enum order.
enum VptMode : int32_t {
    VPT_DELTA_TRACKING = 0,
    VPT_SPECTRAL_DELTA_TRACKING = 1,
    VPT_RATIO_TRACKING = 2,
    VPT_DECOMPOSITION_TRACKING = 3,
    VPT_RESIDUAL_RATIO_TRACKING = 4,
    VPT_NEXT_EVENT_TRACKING = 5,
    VPT_NEXT_EVENT_TRACKING_SPECTRAL = 6,
    VPT_ISOSURFACE_RENDERING = 7,
    VPT_RAY_MARCHING_EMISSION_ABSORPTION = 8,
};

// GridInterpolationType (RenderSettings.hpp:52-56).
enum GridInterpolation : int32_t { INTERP_NEAREST = 0, INTERP_STOCHASTIC = 1, INTERP_TRILINEAR = 2 };

// Environment: builtin DEFAULT / BLACK / SINGLE_COLOR (RenderSettings.hpp:118-121) or an equirectangular image.
enum EnvKind : int32_t { ENV_DEFAULT = 0, ENV_BLACK = 1, ENV_SINGLE_COLOR = 2, ENV_IMAGE = 3 };

// Transmittance estimators for batched ray queries.
enum TransmittanceEstimator : int32_t {
    T_DELTA = 0,           // binary: any real collision ends the ray
    T_RATIO = 1,           // T *= 1 - sigma/majorant per tentative collision
    T_RESIDUAL_RATIO = 2,  // exp(-mu_c d) times the residual ratio product, per supervoxel
    T_DECOMPOSITION = 3,   // binary, control and residual competing exponentials per supervoxel
};

constexpr int kMaxLights = 8;

// Exact unsigned division of numerators below 2^31 by a divisor d fixed per launch: n / d = (n * m) >> s with
// s = 31 + ceil(log2 d) and m = floor(2^s / d) + 1 < 2^32 (Granlund and Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, theorem 4.2). Set on the host by makeDiv31.
struct Div31 {
    uint32_t m;
    uint32_t s;
};

inline Div31 makeDiv31(uint32_t d) {
    if (d == 0u) d = 1u;  // callers pass sizes of at least 1; a zero divisor is treated as 1
    uint32_t l = 0;
    while (l < 32u && (uint64_t(1) << l) < d) l++;
    return Div31{uint32_t((uint64_t(1) << (31u + l)) / d + 1u), 31u + l};
}

// std140 light record of the reference (VptHeader.glsl:233-248), 64 bytes.
struct Light {
    uint32_t lightType;   // 0 point, 1 spot, 2 directional
    uint32_t lightSpace;  // 0 world, 1 view, 2 view orientation
    float spotTotalWidth;
    float spotFalloffStart;
    float color[3];
    float intensity;
    float position[3];    // direction (its length is the shadow-ray distance) for directional lights
    uint32_t useDistance;
    float spotDirection[3];
    float padding;
};
static_assert(sizeof(Light) == 64, "Light must match the 64-byte std140 record");

// Dense grid: normalised FP32 values, x fastest, resident in HBM.
struct DenseGrid {
    const float* values;
    int32_t dim[3];
    float borderValue;  // normalised value of the black border texel: (0 - vmin) / (vmax - vmin)
};

// Sparse grid: a NanoVDB 32.3 buffer with FP32 or Fp16 leaves, resident in HBM, read with 32-bit byte offsets.
struct NanoGrid {
    const uint8_t* buf;
    uint64_t rootOffset;  // byte offset of the root node
    uint32_t tileCount;
    float invMat[9];      // world -> index map, FP32 copy (Map::mInvMatF, row-major)
    float vec[3];         // Map::mVecF
    float background;
    int32_t indexMin[3];  // index bounding box minimum: texel (0,0,0) of the dense-equivalent grid
    int32_t fp16;         // 1: Fp16 leaves (codes at +96, minimum at +80, quantum at +84); 0: FP32 leaves
    int32_t wide;         // 1: the grid spans 4 GiB or more (64-bit byte offsets)
    // Optional flat index of the 8^3 blocks in [blockMin, blockMin + blockDim) (block coordinates, x fastest): per
    // block the leaf's byte offset / 32 (nodes are 32-byte aligned), or 0 with the block's tile value in blockTile,
    // exactly what the tree descent returns. Without one (blockDim 0, null pointers) every lookup descends the tree.
    const uint32_t* blockLeaf;
    const float* blockTile;
    int32_t blockMin[3];
    int32_t blockDim[3];
    // Render-box world position -> index coordinate in one affine map, tapMat * p + tapOff: the box -> grid map of
    // sampleCloud composed with world_to_indexf (composed in double, rounded once). tapDiag: the map has no
    // off-diagonal terms (grids whose voxels are axis-aligned boxes), so a tap needs three FMAs.
    float tapMat[9];
    float tapOff[3];
    int32_t tapDiag;
};

// Supervoxel tables. The majorant table drives the build's local-majorant free-flight sampling; the RRT and
// decomposition tables are the reference's (SuperVoxelGrid.cpp) and feed those estimators.
struct SupervoxelGrid {
    const float* majorant;  // per supervoxel: max normalised density reachable by a lookup (halo for jitter/filter)
    const float* rrt;       // (mu_c, mu_r_bar) pairs, reference RRT table
    const float* decomp;    // (min, max) pairs, reference decomposition table
    int32_t size;           // supervoxel edge in voxels (majorant grid)
    int32_t grid[3];        // majorant grid dimensions
    int32_t refSize;        // supervoxel edge of the reference tables
    int32_t refGrid[3];
    // Walker constants of the two cell grids, computed on the host from the scene box (uniform values the kernels
    // would otherwise derive at launch and hold in vector registers): cell edge in world units and cells per world
    // unit per axis, and the upper corner of the reference grid (it overhangs the box when the grid size is not a
    // multiple of the edge).
    float cellWorld[3];
    float cellsPerWorld[3];
    float refCellWorld[3];
    float refCellsPerWorld[3];
    float refHi[3];
};

// Constants of the Henyey-Greenstein phase function derived from g once per launch (SceneParams::phaseC).
enum PhaseConstant : int {
    PHASE_ONE_MINUS_G2 = 0,  // 1 - g^2
    PHASE_ONE_PLUS_G2 = 1,   // 1 + g^2
    PHASE_TWO_G = 2,         // 2 g
    PHASE_HALF_INV_G = 3,    // 0.5 / g (0 for |g| < 0.001, where sampling is isotropic)
    PHASE_ONE_MINUS_G = 4,   // 1 - g
    PHASE_PDF_SCALE = 5,     // (1 - g^2) / (4 pi)
};

// Per-launch scene constants (the build's counterpart of the reference's UniformData, VptHeader.glsl:45-112).
struct SceneParams {
    float invViewProj[16];  // column-major
    float invView[16];      // column-major
    float prevViewProj[16]; // previous frame's view-projection (feature maps: flow, reprojected UV), column-major
    float farDistance;      // camera far plane (the "Depth Far" feature map of rays without an event)
    float boxMin[3];
    float boxMax[3];
    float worldToTex[3];    // grid dim / (boxMax - boxMin): texel coordinates of a world position
    float gridMin[3];       // sparse grids: NanoVDB world bounding box the box maps onto (sampleCloud)
    float gridMax[3];
    float extinction;       // extinction.x = majorant of the reference trackers
    float albedo;           // scatteringAlbedo.x
    float extinctionRGB[3]; // extinction = extinctionBase * extinctionScale (spectral trackers)
    float albedoRGB[3];     // scatteringAlbedo (spectral trackers)
    float extinctionMax;    // maxComponent(extinction): majorant of the spectral trackers
    float phaseG;
    float phaseC[6];        // Henyey-Greenstein constants of phaseG, indexed by PhaseConstant (set on the host)
    int32_t mode;
    int32_t interpolation;
    int32_t envKind;
    int32_t useLinearRGB;
    uint32_t isEnvMapBlack;
    int32_t useMajorantGrid;
    float sunDirection[3];
    float sunIntensity[3];  // sunlightIntensity * sunlightColor
    float envFactor[3];
    const float* envImage;  // RGBA32F rows of envW texels
    int32_t envW, envH;
    int32_t numLights;
    int32_t sparse;         // 1: NanoVDB grid
    Light lights[kMaxLights];
    uint32_t seedOffset;
    int32_t width, height;
    uint32_t frameBase;     // global frame number of sample 0 of this launch
    uint32_t numFrames;     // samples per pixel in this launch
    uint32_t maxBounces;    // scattering events per path before it ends with zero radiance; 0 = unbounded
    uint64_t numUnits;      // width * height * numFrames
    // Host-computed helpers of the per-path index arithmetic (uniform values the kernels would otherwise derive at
    // launch in vector registers): divisions by the image width, the pixel count, the claims of one screen tile
    // (tile pixels x numFrames) and the tiles per image row; the image size as floats; 1 - scatteringAlbedo.x.
    Div31 divWidth;
    Div31 divPixels;
    Div31 divTileClaims;
    Div31 divTilesX;
    float widthF, heightF;
    float absorption;
    // Device copy of these constants, for branches the path kernels take rarely (claims and camera rays at path
    // start, the sky at segment ends): read from memory there, those constants do not occupy scalar registers across
    // the persistent loop. The single-frame launcher stores the copy before each launch; the batched queue points it
    // at the request's own record in the request table.
    const SceneParams* mem;
};

// One request of a batched render (the batched queue's device table; no reference counterpart). The scene
// constants and the volume are snapshots taken when the request was added; the request's paths are the batch units
// unitBegin + g * (W*H) + pixel for groups g < groups, each unit tracing samples [g * G, min(spp, (g + 1) * G)) of
// its pixel (G = samples per unit of the launch).
struct BatchRequest {
    SceneParams P;          // P.frameBase: global frame number of sample 0; P.sparse selects the volume below
    SupervoxelGrid S;
    DenseGrid dense;
    NanoGrid nano;
    uint64_t unitBegin;     // first batch unit of the request
    uint64_t pixBegin;      // first pixel of the request in the batch's flattened pixel range
    uint32_t spp;           // samples per pixel
    uint32_t groups;        // units per pixel: ceil(spp / G)
    float* out;             // [channels, height, width] float32 image on the device
    int32_t channels;
    int32_t pad;
};

}  // namespace crvpt
