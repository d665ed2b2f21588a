// Host side of the gfx950 volumetric path tracer: the process-global renderer behind the crvpt C ABI.
//
// The renderer keeps the reference's settings semantics (src/PathTracer/VolumetricPathTracingPass.cpp,
// src/PyTorch/Module.cpp, src/PyTorch/VolumetricPathTracingModuleRenderer.cpp): every setting that changes the
// image resets the progressive accumulation; render_frame adds frame_count samples per pixel to the running mean.
// Device work runs on one HIP stream; scene data (grid, supervoxel tables, environment image) stays resident in
// HBM until the scene or a setting that shapes it changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/crvpt.h"
#include "host_common.hpp"
#include "nanovdb_build.hpp"
#include "parallel_for.hpp"
#include "render_api.hpp"
#include "vpt_types.h"

namespace crvpt {

thread_local std::string g_lastError;

int fail(int code, const std::string& msg) {
    g_lastError = msg;
    return code;
}

namespace {

// ------------------------------------------------------------------------------------------------------------
// Small double-precision matrix helpers (column-major 4x4 like glm).
// ------------------------------------------------------------------------------------------------------------
struct Mat4 {
    double m[16];
};

Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a.m[k * 4 + row] * b.m[c * 4 + k];
            r.m[c * 4 + row] = s;
        }
    return r;
}

Mat4 inverse(const Mat4& in) {
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = in.m[c * 4 + r];
            a[r][c + 4] = r == c ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; col++) {
        int piv = col;
        for (int r = col + 1; r < 4; r++)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        for (int c = 0; c < 8; c++) std::swap(a[col][c], a[piv][c]);
        const double p = a[col][col];
        for (int c = 0; c < 8; c++) a[col][c] /= p;
        for (int r = 0; r < 4; r++) {
            if (r == col) continue;
            const double f = a[r][col];
            for (int c = 0; c < 8; c++) a[r][c] -= f * a[col][c];
        }
    }
    Mat4 o;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) o.m[c * 4 + r] = a[r][c + 4];
    return o;
}

// Right-handed look-at view matrix.
Mat4 lookAt(const double eye[3], const double center[3], const double up[3]) {
    double f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f) v /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s) v /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    Mat4 V{};
    V.m[0] = s[0]; V.m[4] = s[1]; V.m[8] = s[2];
    V.m[1] = u[0]; V.m[5] = u[1]; V.m[9] = u[2];
    V.m[2] = -f[0]; V.m[6] = -f[1]; V.m[10] = -f[2];
    V.m[12] = -(s[0] * eye[0] + s[1] * eye[1] + s[2] * eye[2]);
    V.m[13] = -(u[0] * eye[0] + u[1] * eye[1] + u[2] * eye[2]);
    V.m[14] = f[0] * eye[0] + f[1] * eye[1] + f[2] * eye[2];
    V.m[15] = 1.0;
    return V;
}

// Right-handed perspective with depth in [0, 1] and y flipped so that image row 0 is the top of the view.
Mat4 perspective(double fovy, double aspect, double zn, double zf) {
    const double th = std::tan(fovy / 2.0);
    Mat4 P{};
    P.m[0] = 1.0 / (aspect * th);
    P.m[5] = -1.0 / th;
    P.m[10] = zf / (zn - zf);
    P.m[11] = -1.0;
    P.m[14] = -(zf * zn) / (zf - zn);
    return P;
}

const char* const kModeNames[] = {"Delta Tracking", "Delta Tracking (Spectral)", "Ratio Tracking",
                                  "Decomposition Tracking", "Residual Ratio Tracking", "Next Event Tracking",
                                  "Next Event Tracking (Spectral)", "Isosurfaces",
                                  "Ray Marching (Emission/Absorption)"};

struct Renderer {
    // ---- settings (VolumetricPathTracingPass.hpp defaults; Module.cpp:797-798 sets delta tracking) ----
    int mode = VPT_DELTA_TRACKING;
    int interpolation = INTERP_STOCHASTIC;
    double extinctionBase[3] = {1.0, 1.0, 1.0};
    double extinctionScale = 1024.0;
    double albedo[3] = {0.9, 0.9, 0.9};
    double phaseG = 0.5;
    double sunColor[3] = {1.0, 0.961538462, 0.884615385};
    double sunIntensity = 2.6;
    double sunDirection[3] = {0.5826, 0.7660, 0.2717};
    int envKind = ENV_DEFAULT;
    bool useEnvImage = false;
    std::vector<float> envImage;  // RGBA32F
    int envW = 0, envH = 0;
    bool envImageBlack = false;
    bool envFactorRgbMode = false;
    double envFactor = 1.0;
    double envFactorRgb[3] = {1.0, 1.0, 1.0};
    int useLinearRGB = 1;
    uint32_t seedOffset = 0;
    uint32_t maxBounces = 0;  // 0: unbounded like the reference trackers
    bool useLights = false;
    std::vector<Light> lights;
    // camera (VolumetricPathTracingModuleRenderer.cpp:124-131)
    double camPos[3] = {0.0, 0.0, 0.8};
    double memberPos[3] = {0.0, 0.0, 0.0};
    double memberTarget[3] = {0.0, 0.0, 0.0};
    Mat4 view{};
    double fovy = 2.0 * std::atan(0.5);
    double znear = 1.0 / 512.0, zfar = 80.0;
    double maxGridExtent = 0.25;
    bool useSparse = false;
    bool outputForeground = false;  // OUTPUT_FOREGROUND_MAP (VolumetricPathTracingPass.cpp:1669-1671)
    int sparseValueType = 0;  // leaf type of grids built from the dense field: 0 FP32, 1 Fp16
    int denoiser = 0;  // 0 none, 1 EAW, 2 SVGF
    // feature maps (VolumetricPathTracingPass::setUseFeatureMaps / setFeatureMapType)
    std::vector<int> featureMaps;
    int featureMapType = 0;          // map render_frame returns (0: Result)
    bool prevViewProjSet = false;    // setViewProjectionMatrixAsPrevious was called
    float prevViewProj[16] = {};
    int featW = 0, featH = 0;        // resolution of the feature state (0: none)
    int svgfW = 0, svgfH = 0;        // resolution of the SVGF history (0: none yet)
    int svgfIterations = 5;
    int eawIterations = 0;
    float eawPhiColor = 5.0f, eawPhiPosition = 0.1f, eawPhiNormal = 0.1f;
    int eawUseColorDistance = 0;
    bool useMajorantGrid = true;
    int svSizeSetting = 0;
    int blocksSetting = 0;
    int stepsPerRefill = 0;  // walker steps between refills; 0: the estimator's default (refillInterval)
    bool collectStats = false;

    // ---- volume ----
    int dims[3] = {0, 0, 0};        // dense grid (or, for loaded NanoVDB grids, the index bounding box) size
    int sparseDims[3] = {0, 0, 0};  // NanoVDB index bounding box size
    double voxelSize[3] = {1.0, 1.0, 1.0};
    bool hasDense = false;
    bool hasSparse = false;
    bool denseOnHostValid = false;
    std::vector<float> denseHost;
    float vmin = 0.0f, vmax = 1.0f;
    std::vector<uint8_t> nvdbHost;
    NanoInfo nano{};
    double boxMinDense[3], boxMaxDense[3], boxMinSparse[3], boxMaxSparse[3], gridMinSparse[3], gridMaxSparse[3];

    // ---- device state ----
    hipStream_t stream = nullptr;
    DevBuf dRaw, dNorm, dNvdb, dMajorant, dRrt, dDecomp, dEnv, dAcc, dPaths, dCounter, dStage, dRays, dParams;
    DevBuf dSvMin, dSvMax, dSvAvg, dDen0, dDen1, dStats, dFeat, dFeatAcc, dSvgf, dKpn, dBlockLeaf, dBlockTile, dBlockBad;
    // flat 8^3-block index of the NanoVDB grid (NanoGrid::blockLeaf), valid for volume version blockTableVersion
    bool blockTableOk = false;
    uint64_t blockTableVersion = ~0ull;
    int blockMin[3] = {0, 0, 0}, blockDim[3] = {0, 0, 0};
    bool sceneDirty = true;     // grid contents changed
    bool tablesDirty = true;    // supervoxel tables must be rebuilt
    bool envDirty = true;
    int svSize = 8;
    int svGrid[3] = {1, 1, 1};
    int superVoxelSize = 8;  // requested edge of the reference tables (VolumetricPathTracingPass.hpp:254)
    int refSvSize = 8;
    int refGrid[3] = {1, 1, 1};
    bool refTablesValid = false;
    int refTablesInterp = -1;
    double refTablesExtinction = -1.0;
    int cuCount = 256;
    uint64_t volumeVersion = 0;  // bumped whenever the grid contents change (batched requests snapshot the volume)
    uint64_t envVersion = 0;     // bumped whenever an environment image is uploaded

    // ---- accumulation ----
    int accW = 0, accH = 0, accC = 0;
    uint32_t frameCount = 0;
    uint32_t globalFrameNumber = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evT0 = nullptr, evT1 = nullptr;
    double lastPathMs = 0.0, lastTotalMs = 0.0;
    uint64_t lastPaths = 0;
    int lastLaunches = 0;

    void resetAccumulation() {
        frameCount = 0;
        globalFrameNumber = 0;
    }
};

Renderer* g_r = nullptr;
int g_refCount = 0;

int requireInit() {
    if (!g_r) return fail(CRVPT_ERR_NOT_INITIALIZED, "crvpt: renderer not initialised (call initialize first)");
    return CRVPT_OK;
}

void computeDenseBounds(Renderer& r) {
    // CloudData::computeGridBounds (CloudData.cpp:63-78)
    const double sx = r.dims[0] * r.voxelSize[0], sy = r.dims[1] * r.voxelSize[1], sz = r.dims[2] * r.voxelSize[2];
    const double maxSize = std::max(sx, std::max(sy, sz));
    const float ext = float(r.maxGridExtent);
    const float bx = float(sx) * ext / float(maxSize), by = float(sy) * ext / float(maxSize),
                bz = float(sz) * ext / float(maxSize);
    r.boxMaxDense[0] = bx; r.boxMaxDense[1] = by; r.boxMaxDense[2] = bz;
    r.boxMinDense[0] = -bx; r.boxMinDense[1] = -by; r.boxMinDense[2] = -bz;
}

void computeSparseBounds(Renderer& r, bool fromDense) {
    // CloudData::computeSparseGridMetadata (CloudData.cpp:1289-1358): world bbox of the grid, normalised into
    // the +-maxGridExtent box unless the grid was built from the dense field (then world coordinates are kept).
    for (int a = 0; a < 3; a++) {
        r.gridMinSparse[a] = float(r.nano.worldBBoxMin[a]);
        r.gridMaxSparse[a] = float(r.nano.worldBBoxMax[a]);
    }
    if (fromDense) {
        for (int a = 0; a < 3; a++) {
            r.boxMinSparse[a] = r.gridMinSparse[a];
            r.boxMaxSparse[a] = r.gridMaxSparse[a];
        }
        return;
    }
    float diff[3], maxDim = 0.0f;
    for (int a = 0; a < 3; a++) {
        diff[a] = float(r.gridMaxSparse[a]) - float(r.gridMinSparse[a]);
        maxDim = std::max(maxDim, diff[a]);
    }
    const float ext = float(r.maxGridExtent);
    const float A = 2.0f * ext / maxDim;
    for (int a = 0; a < 3; a++) {
        const float L = diff[a] / maxDim;
        const float B = ext * L - float(r.gridMaxSparse[a]) * A;
        r.boxMinSparse[a] = A * float(r.gridMinSparse[a]) + B;
        r.boxMaxSparse[a] = A * float(r.gridMaxSparse[a]) + B;
    }
}

int setDenseHost(Renderer& r, int nx, int ny, int nz, std::vector<float>&& values) {
    HIP_TRY(hipDeviceSynchronize());  // in-flight renders on any stream may read the grid
    r.dims[0] = nx; r.dims[1] = ny; r.dims[2] = nz;
    r.voxelSize[0] = r.voxelSize[1] = r.voxelSize[2] = 1.0;
    r.denseHost = std::move(values);
    r.denseOnHostValid = true;
    float mn = 3.4e38f, mx = -3.4e38f;
    {
        std::vector<float> pmn(16, 3.4e38f), pmx(16, -3.4e38f);
        parallelFor(r.denseHost.size(), [&](size_t b, size_t e, size_t t) {
            float a = 3.4e38f, c = -3.4e38f;
            for (size_t i = b; i < e; i++) {
                a = std::min(a, r.denseHost[i]);
                c = std::max(c, r.denseHost[i]);
            }
            pmn[t] = a;
            pmx[t] = c;
        }, 1 << 20);
        for (size_t t = 0; t < 16; t++) {
            mn = std::min(mn, pmn[t]);
            mx = std::max(mx, pmx[t]);
        }
    }
    if (mn == mx) mn = 0.0f;  // DensityField::computeMinMax (DensityField.cpp:149-152)
    r.vmin = mn;
    r.vmax = mx;
    r.hasDense = true;
    r.hasSparse = false;
    r.nvdbHost.clear();
    computeDenseBounds(r);
    HIP_TRY(r.dRaw.ensure(r.denseHost.size() * sizeof(float)));
    HIP_TRY(hipMemcpy(r.dRaw.ptr, r.denseHost.data(), r.denseHost.size() * sizeof(float), hipMemcpyHostToDevice));
    r.volumeVersion++;
    r.sceneDirty = r.tablesDirty = true;
    r.refTablesValid = false;
    r.resetAccumulation();
    return CRVPT_OK;
}

int setSparseHost(Renderer& r, std::vector<uint8_t>&& buf, bool fromDense) {
    NanoInfo info;
    std::string err;
    if (!parseNanoVdb(buf.data(), buf.size(), info, err)) return fail(CRVPT_ERR_INVALID_ARGUMENT, err);
    HIP_TRY(hipDeviceSynchronize());
    r.nvdbHost = std::move(buf);
    r.nano = info;
    r.hasSparse = true;
    for (int a = 0; a < 3; a++) r.sparseDims[a] = info.indexBBoxMax[a] - info.indexBBoxMin[a] + 1;
    if (!fromDense) {
        r.hasDense = false;
        r.denseOnHostValid = false;
        for (int a = 0; a < 3; a++) {
            r.dims[a] = info.indexBBoxMax[a] - info.indexBBoxMin[a] + 1;
            r.voxelSize[a] = info.voxelSize[a];
        }
    }
    computeSparseBounds(r, fromDense);
    if (!fromDense) {
        // the dense field a dense render decodes from this grid is placed in the grid's box
        // (CloudData::getDenseDensityField, CloudData.cpp:1190-1191)
        for (int a = 0; a < 3; a++) {
            r.boxMinDense[a] = r.boxMinSparse[a];
            r.boxMaxDense[a] = r.boxMaxSparse[a];
        }
    }
    HIP_TRY(r.dNvdb.ensure(r.nvdbHost.size()));
    HIP_TRY(hipMemcpy(r.dNvdb.ptr, r.nvdbHost.data(), r.nvdbHost.size(), hipMemcpyHostToDevice));
    r.volumeVersion++;
    r.sceneDirty = r.tablesDirty = true;
    r.refTablesValid = false;
    r.resetAccumulation();
    return CRVPT_OK;
}

bool sparseActive(const Renderer& r) { return r.useSparse && r.hasSparse; }

// Texel grid of the active representation: dense voxels, or the NanoVDB active index box.
const int* activeDims(const Renderer& r) { return sparseActive(r) ? r.sparseDims : r.dims; }

// Builds the NanoVDB buffer from the dense field when the sparse grid is requested and missing
// (CloudData::getSparseDensityField, CloudData.cpp:1367-1438).
int ensureSparse(Renderer& r) {
    if (!r.useSparse || r.hasSparse) return CRVPT_OK;
    if (!r.hasDense) return fail(CRVPT_ERR_INVALID_ARGUMENT, "crvpt: no density field loaded");
    if (!r.denseOnHostValid) {
        r.denseHost.resize(size_t(r.dims[0]) * r.dims[1] * r.dims[2]);
        HIP_TRY(hipMemcpy(r.denseHost.data(), r.dRaw.ptr, r.denseHost.size() * sizeof(float), hipMemcpyDeviceToHost));
        r.denseOnHostValid = true;
    }
    std::vector<float> norm(r.denseHost.size());
    const float mn = r.vmin, mx = r.vmax;
    parallelFor(norm.size(), [&](size_t b, size_t e, size_t) {
        for (size_t i = b; i < e; i++) norm[i] = (r.denseHost[i] - mn) / (mx - mn);
    }, 1 << 20);
    const double dx = double(float(r.boxMaxDense[0]) - float(r.boxMinDense[0])) / double(r.dims[0]);
    const double origin[3] = {float(r.boxMinDense[0]), float(r.boxMinDense[1]), float(r.boxMinDense[2])};
    std::vector<uint8_t> buf = buildNanoVdbFromDense(norm.data(), r.dims, dx, origin, "tmp", r.sparseValueType);
    const bool keepDense = r.hasDense;
    int rc = setSparseHost(r, std::move(buf), true);
    r.hasDense = keepDense;
    return rc;
}

int uploadEnv(Renderer& r) {
    if (!r.envDirty) return CRVPT_OK;
    if (r.useEnvImage && !r.envImage.empty()) {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(r.dEnv.ensure(r.envImage.size() * sizeof(float)));
        HIP_TRY(hipMemcpy(r.dEnv.ptr, r.envImage.data(), r.envImage.size() * sizeof(float), hipMemcpyHostToDevice));
        r.envVersion++;
    }
    r.envDirty = false;
    return CRVPT_OK;
}

float extinctionX(const Renderer& r) { return float(r.extinctionBase[0]) * float(r.extinctionScale); }

int chooseSvSize(const Renderer& r) {
    if (r.svSizeSetting > 0) return r.svSizeSetting;
    return sparseActive(r) ? 16 : 8;
}

// Grid views of the current volume.
void gridViews(Renderer& r, DenseGrid& dg, NanoGrid& ng, bool& sparse) {
    sparse = sparseActive(r);
    dg.values = r.dNorm.as<float>();
    for (int a = 0; a < 3; a++) dg.dim[a] = r.dims[a];
    dg.borderValue = (0.0f - r.vmin) / (r.vmax - r.vmin);
    ng.buf = r.dNvdb.as<uint8_t>();
    ng.rootOffset = r.nano.rootOffset;
    ng.tileCount = r.nano.tileCount;
    for (int i = 0; i < 9; i++) ng.invMat[i] = r.nano.invMatF[i];
    for (int i = 0; i < 3; i++) {
        ng.vec[i] = r.nano.vecF[i];
        ng.indexMin[i] = r.nano.indexBBoxMin[i];
    }
    ng.background = r.nano.background;
    ng.fp16 = r.nano.valueType == 1 ? 1 : 0;
    ng.wide = r.nvdbHost.size() >= (size_t(1) << 32) ? 1 : 0;
    // composed tap map: idx = invMat * ((p - boxMin) * scale + gridMin - vec), scale = (gridMax - gridMin) /
    // (boxMax - boxMin), with the float values the path kernels see for the box and the grid bounds
    double scale[3], shift[3];
    for (int a = 0; a < 3; a++) {
        const float bMin = float(r.boxMinSparse[a]), bMax = float(r.boxMaxSparse[a]);
        const float gMin = float(r.gridMinSparse[a]), gMax = float(r.gridMaxSparse[a]);
        scale[a] = double(gMax - gMin) / double(bMax - bMin);
        shift[a] = double(gMin) - double(r.nano.vecF[a]) - double(bMin) * scale[a];
    }
    ng.tapDiag = 1;
    for (int row = 0; row < 3; row++) {
        double off = 0.0;
        for (int c = 0; c < 3; c++) {
            const double m = double(r.nano.invMatF[3 * row + c]);
            ng.tapMat[3 * row + c] = float(m * scale[c]);
            off += m * shift[c];
            if (row != c && m != 0.0) ng.tapDiag = 0;
        }
        ng.tapOff[row] = float(off);
    }
    const bool table = r.blockTableOk && r.blockTableVersion == r.volumeVersion;
    ng.blockLeaf = table ? r.dBlockLeaf.as<uint32_t>() : nullptr;
    ng.blockTile = table ? r.dBlockTile.as<float>() : nullptr;
    for (int a = 0; a < 3; a++) {
        ng.blockMin[a] = table ? r.blockMin[a] : 0;
        ng.blockDim[a] = table ? r.blockDim[a] : 0;
    }
}

// CloudData::getDenseDensityField (CloudData.cpp:1173-1253): a volume loaded only as NanoVDB renders through a dense
// field decoded from the grid over its index bounding box and placed in the sparse grid's box. (The reference stores
// that field as half floats from 1024^3 voxels on; this build keeps FP32.)
int decodeDenseFromSparse(Renderer& r) {
    if (r.hasDense || !r.hasSparse) return CRVPT_OK;
    int dims[3];
    size_t n = 1;
    for (int a = 0; a < 3; a++) {
        dims[a] = r.sparseDims[a];
        n *= size_t(dims[a]);
    }
    HIP_TRY(hipDeviceSynchronize());
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    HIP_TRY(r.dRaw.ensure(n * sizeof(float)));
    HIP_TRY(launchDecodeDense(ng, dims, r.dRaw.as<float>(), r.stream));
    r.denseHost.resize(n);
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipMemcpy(r.denseHost.data(), r.dRaw.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
    float mn = 3.4e38f, mx = -3.4e38f;
    for (float v : r.denseHost) {
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    if (mn == mx) mn = 0.0f;  // DensityField::computeMinMax (DensityField.cpp:149-152)
    r.vmin = mn;
    r.vmax = mx;
    for (int a = 0; a < 3; a++) {
        r.dims[a] = dims[a];
        r.boxMinDense[a] = r.boxMinSparse[a];
        r.boxMaxDense[a] = r.boxMaxSparse[a];
    }
    r.denseOnHostValid = true;
    r.hasDense = true;
    r.volumeVersion++;
    r.sceneDirty = r.tablesDirty = true;
    r.refTablesValid = false;
    return CRVPT_OK;
}

int ensureDense(Renderer& r) { return r.useSparse ? CRVPT_OK : decodeDenseFromSparse(r); }

// Builds the flat 8^3-block index over the grid's index bounding box (at most 2^27 blocks: 1 GiB of index; larger
// sparse domains, and grids whose leaves lie beyond 128 GiB, keep descending the tree).
int buildBlockTable(Renderer& r) {
    r.blockTableOk = false;
    r.blockTableVersion = r.volumeVersion;
    auto floorDiv8 = [](int v) { return v >= 0 ? v / 8 : -((-v + 7) / 8); };
    int b0[3], nb[3];
    size_t n = 1;
    for (int a = 0; a < 3; a++) {
        b0[a] = floorDiv8(r.nano.indexBBoxMin[a]);
        nb[a] = floorDiv8(r.nano.indexBBoxMax[a]) - b0[a] + 1;
        if (nb[a] <= 0) return CRVPT_OK;
        n *= size_t(nb[a]);
    }
    if (n > (size_t(1) << 27)) return CRVPT_OK;
    HIP_TRY(r.dBlockLeaf.ensure(n * sizeof(uint32_t)));
    HIP_TRY(r.dBlockTile.ensure(n * sizeof(float)));
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    HIP_TRY(r.dBlockBad.ensure(sizeof(uint32_t)));
    HIP_TRY(hipMemsetAsync(r.dBlockBad.ptr, 0, sizeof(uint32_t), r.stream));
    HIP_TRY(launchBlockTable(ng, b0, nb, r.dBlockLeaf.as<uint32_t>(), r.dBlockTile.as<float>(),
                             r.dBlockBad.as<uint32_t>(), r.stream));
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, r.dBlockBad.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, r.stream));
    HIP_TRY(hipStreamSynchronize(r.stream));
    if (bad) return CRVPT_OK;  // offsets the index cannot encode: lookups descend the tree
    for (int a = 0; a < 3; a++) {
        r.blockMin[a] = b0[a];
        r.blockDim[a] = nb[a];
    }
    r.blockTableOk = true;
    return CRVPT_OK;
}

// Uploads the grid and rebuilds the supervoxel tables the current mode needs.
int prepareScene(Renderer& r) {
    if (!r.hasDense && !r.hasSparse) return fail(CRVPT_ERR_INVALID_ARGUMENT, "crvpt: no volume loaded");
    int rc = ensureSparse(r);
    if (rc) return rc;
    rc = ensureDense(r);
    if (rc) return rc;
    rc = uploadEnv(r);
    if (rc) return rc;
    const bool sparse = sparseActive(r);
    if (r.sceneDirty) {
        if (!sparse) {
            const size_t n = size_t(r.dims[0]) * r.dims[1] * r.dims[2];
            HIP_TRY(r.dNorm.ensure(n * sizeof(float)));
            HIP_TRY(launchNormalize(r.dRaw.as<float>(), r.dNorm.as<float>(), n, r.vmin, r.vmax, r.stream));
        }
        r.sceneDirty = false;
        r.tablesDirty = true;
    }
    if (sparse && r.blockTableVersion != r.volumeVersion) {
        rc = buildBlockTable(r);
        if (rc) return rc;
    }
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    if (r.tablesDirty) {
        r.svSize = chooseSvSize(r);
        const int* ad = activeDims(r);
        for (int a = 0; a < 3; a++) r.svGrid[a] = (ad[a] + r.svSize - 1) / r.svSize;
        const size_t n = size_t(r.svGrid[0]) * r.svGrid[1] * r.svGrid[2];
        // the path kernels index cells with 32-bit arithmetic
        if (n >= (size_t(1) << 32)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "majorant grid of 2^32 cells or more");
        HIP_TRY(r.dMajorant.ensure(n * sizeof(float)));
        HIP_TRY(launchMajorant(sp ? nullptr : &dg, sp ? &ng : nullptr, ad, r.svSize, r.svGrid,
                               r.dMajorant.as<float>(), r.stream));
        r.tablesDirty = false;
        r.refTablesValid = false;
    }
    const bool needRef = r.mode == VPT_RESIDUAL_RATIO_TRACKING || r.mode == VPT_DECOMPOSITION_TRACKING;
    if (needRef && (!r.refTablesValid || r.refTablesInterp != r.interpolation ||
                    r.refTablesExtinction != double(extinctionX(r)))) {
        // SuperVoxelGridResidualRatioTracking / ...DecompositionTracking constructors (SuperVoxelGrid.cpp:40-91,
        // 269-405): the requested edge, halved until it divides grids smaller than the edge.
        const int* ad = activeDims(r);
        int sv = std::max(r.superVoxelSize, 1);
        if (ad[0] < sv || ad[1] < sv || ad[2] < sv) {
            while (!(ad[0] % sv == 0 && ad[1] % sv == 0 && ad[2] % sv == 0)) sv /= 2;
        }
        r.refSvSize = sv;
        for (int a = 0; a < 3; a++) r.refGrid[a] = (ad[a] + sv - 1) / sv;
        const size_t n = size_t(r.refGrid[0]) * r.refGrid[1] * r.refGrid[2];
        // the path kernels index cells with 32-bit arithmetic
        if (n >= (size_t(1) << 32)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "supervoxel grid of 2^32 cells or more");
        HIP_TRY(r.dSvMin.ensure(n * sizeof(float)));
        HIP_TRY(r.dSvMax.ensure(n * sizeof(float)));
        HIP_TRY(r.dSvAvg.ensure(n * sizeof(float)));
        HIP_TRY(launchRefSupervoxelStats(sp ? nullptr : &dg, sp ? &ng : nullptr, ad, sv, r.refGrid,
                                         r.interpolation == INTERP_NEAREST ? 0 : 1, r.dSvMin.as<float>(),
                                         r.dSvMax.as<float>(), r.dSvAvg.as<float>(), r.stream));
        std::vector<float> mn(n), mx(n), avg(n);
        HIP_TRY(hipStreamSynchronize(r.stream));
        HIP_TRY(hipMemcpy(mn.data(), r.dSvMin.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(mx.data(), r.dSvMax.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(avg.data(), r.dSvAvg.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
        // Host part of SuperVoxelGrid.cpp:180-256 (std::pow on the host as the reference does).
        std::vector<float> rrt(2 * n), dec(2 * n);
        const float ext = extinctionX(r);
        const float gamma = 2.0f;
        const float D = std::sqrt(3.0f) * float(sv);
        for (size_t i = 0; i < n; i++) {
            const int sx = int(i % r.refGrid[0]), sy = int((i / r.refGrid[0]) % r.refGrid[1]),
                      sz = int(i / (size_t(r.refGrid[0]) * r.refGrid[1]));
            dec[2 * i] = mn[i];
            dec[2 * i + 1] = mx[i];
            float dmin = mn[i];
            if ((ad[0] % sv != 0 && sx == r.refGrid[0] - 1) || (ad[1] % sv != 0 && sy == r.refGrid[1] - 1) ||
                (ad[2] % sv != 0 && sz == r.refGrid[2] - 1))
                dmin = 0.0f;
            const float muMin = dmin * ext, muMax = mx[i] * ext, muAvg = avg[i] * ext;
            const float muR = std::max(muMax - muMin, 0.1f);
            const float muC = muMin + muR * std::pow(gamma, (1.0f / (D * muR)) - 1.0f);
            rrt[2 * i] = std::min(std::max(muC, muMin), muAvg);
            rrt[2 * i + 1] = muR;
        }
        HIP_TRY(r.dRrt.ensure(2 * n * sizeof(float)));
        HIP_TRY(r.dDecomp.ensure(2 * n * sizeof(float)));
        HIP_TRY(hipMemcpy(r.dRrt.ptr, rrt.data(), 2 * n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(r.dDecomp.ptr, dec.data(), 2 * n * sizeof(float), hipMemcpyHostToDevice));
        r.refTablesValid = true;
        r.refTablesInterp = r.interpolation;
        r.refTablesExtinction = double(ext);
    }
    return CRVPT_OK;
}

Mat4 projection(const Renderer& r, int w, int h) { return perspective(r.fovy, double(w) / double(h), r.znear, r.zfar); }

// Feature maps are traced whenever a map is requested, returned by render_frame, or read by the EAW denoiser
// (COLOR + POSITION + NORMAL, EAWDenoiser.cpp:67-74).
bool featuresActive(const Renderer& r) {
    return !r.featureMaps.empty() || r.featureMapType != 0 || (r.denoiser == 1 && r.eawIterations > 0) ||
           r.denoiser == 2;
}

FeatureAcc featureAcc(Renderer& r, uint32_t npix) {
    float4* b = r.dFeatAcc.as<float4>();
    FeatureAcc a;
    float4** slots[kFeatureAccArrays] = {&a.firstX, &a.firstW,  &a.normal, &a.cloudOnly, &a.background, &a.albedo,
                                         &a.depth,  &a.dens,    &a.flow,   &a.flowRev,   &a.radianceSq};
    for (int i = 0; i < kFeatureAccArrays; i++) *slots[i] = b + size_t(i) * npix;
    return a;
}

// The walker's per-axis constants of the majorant and reference cell grids (cell edge in world units, cells per
// world unit, the reference grid's upper corner), from the box and texel scale of P.
void setWalkerConstants(const SceneParams& P, SupervoxelGrid& S) {
    for (int a = 0; a < 3; a++) {
        S.cellWorld[a] = float(S.size) / P.worldToTex[a];
        S.cellsPerWorld[a] = P.worldToTex[a] * (1.0f / float(S.size));
        S.refCellWorld[a] = float(S.refSize) / P.worldToTex[a];
        S.refCellsPerWorld[a] = P.worldToTex[a] * (1.0f / float(S.refSize));
        S.refHi[a] = P.boxMin[a] + float(S.refGrid[a] * S.refSize) / P.worldToTex[a];
    }
}

void fillSceneParams(Renderer& r, int w, int h, SceneParams& P, SupervoxelGrid& S) {
    std::memset(&P, 0, sizeof(P));
    const Mat4 vp = mul(projection(r, w, h), r.view);
    const Mat4 ivp = inverse(vp);
    const Mat4 iv = inverse(r.view);
    for (int i = 0; i < 16; i++) {
        P.invViewProj[i] = float(ivp.m[i]);
        P.invView[i] = float(iv.m[i]);
        // no previous matrix set: the current one (VolumetricPathTracingPass.cpp:1985-1989)
        P.prevViewProj[i] = r.prevViewProjSet ? r.prevViewProj[i] : float(vp.m[i]);
    }
    P.farDistance = float(r.zfar);
    const bool sparse = sparseActive(r);
    for (int a = 0; a < 3; a++) {
        P.boxMin[a] = float(sparse ? r.boxMinSparse[a] : r.boxMinDense[a]);
        P.boxMax[a] = float(sparse ? r.boxMaxSparse[a] : r.boxMaxDense[a]);
        P.gridMin[a] = sparse ? float(r.gridMinSparse[a]) : 0.0f;
        P.gridMax[a] = sparse ? float(r.gridMaxSparse[a]) : 1.0f;
        P.worldToTex[a] = float(activeDims(r)[a]) / (P.boxMax[a] - P.boxMin[a]);
    }
    P.extinction = extinctionX(r);
    P.albedo = float(r.albedo[0]);
    for (int a = 0; a < 3; a++) {
        P.extinctionRGB[a] = float(r.extinctionBase[a]) * float(r.extinctionScale);
        P.albedoRGB[a] = float(r.albedo[a]);
    }
    P.extinctionMax = std::max(std::max(P.extinctionRGB[0], P.extinctionRGB[1]), P.extinctionRGB[2]);
    P.phaseG = float(r.phaseG);
    {
        // the device evaluated these per launch with fused multiply-adds; fmaf keeps the same roundings
        const float g = P.phaseG;
        P.phaseC[PHASE_ONE_MINUS_G2] = std::fmaf(-g, g, 1.0f);
        P.phaseC[PHASE_ONE_PLUS_G2] = std::fmaf(g, g, 1.0f);
        P.phaseC[PHASE_TWO_G] = 2.0f * g;
        P.phaseC[PHASE_HALF_INV_G] = std::fabs(g) < 0.001f ? 0.0f : 0.5f / g;
        P.phaseC[PHASE_ONE_MINUS_G] = 1.0f - g;
        P.phaseC[PHASE_PDF_SCALE] = (0.25f / 3.14159265359f) * std::fmaf(-g, g, 1.0f);
    }
    P.mode = r.mode;
    P.interpolation = r.interpolation;
    P.useLinearRGB = r.useLinearRGB;
    P.useMajorantGrid = r.useMajorantGrid ? 1 : 0;
    // sun = sunlightIntensity * sunlightColor; direction normalised like glm::normalize of the default
    double dl = std::sqrt(r.sunDirection[0] * r.sunDirection[0] + r.sunDirection[1] * r.sunDirection[1] +
                          r.sunDirection[2] * r.sunDirection[2]);
    for (int a = 0; a < 3; a++) {
        P.sunDirection[a] = float(r.sunDirection[a] / dl);
        P.sunIntensity[a] = float(r.sunIntensity * r.sunColor[a]);
        P.envFactor[a] = float(r.envFactorRgbMode ? r.envFactorRgb[a] : r.envFactor);
    }
    if (r.useEnvImage && !r.envImage.empty()) {
        P.envKind = ENV_IMAGE;
        P.envImage = r.dEnv.as<float>();
        P.envW = r.envW;
        P.envH = r.envH;
        // VolumetricPathTracingPass.cpp:2114-2131
        bool black = r.envImageBlack;
        if (r.envFactorRgbMode) {
            if (r.envFactorRgb[0] < 1e-5 && r.envFactorRgb[1] < 1e-5 && r.envFactorRgb[2] < 1e-5) black = true;
        } else if (r.envFactor < 1e-5) {
            black = true;
        }
        P.isEnvMapBlack = black ? 1u : 0u;
    } else {
        P.envKind = r.envKind;
        P.isEnvMapBlack = r.envKind == ENV_BLACK ? 1u : 0u;
    }
    P.numLights = r.useLights ? int(std::min<size_t>(r.lights.size(), kMaxLights)) : 0;
    for (int i = 0; i < P.numLights; i++) P.lights[i] = r.lights[i];
    P.sparse = sparse ? 1 : 0;
    P.seedOffset = r.seedOffset;
    P.maxBounces = r.maxBounces;
    P.width = w;
    P.height = h;
    P.widthF = float(w);
    P.heightF = float(h);
    P.divWidth = makeDiv31(uint32_t(std::max(w, 1)));
    P.divPixels = makeDiv31(uint32_t(std::max(w, 1)) * uint32_t(std::max(h, 1)));
    P.absorption = 1.0f - P.albedo;
    S.majorant = r.dMajorant.as<float>();
    S.rrt = r.dRrt.as<float>();
    S.decomp = r.dDecomp.as<float>();
    S.size = r.svSize;
    S.refSize = r.refSvSize;
    for (int a = 0; a < 3; a++) {
        S.grid[a] = r.svGrid[a];
        S.refGrid[a] = r.refGrid[a];
    }
    setWalkerConstants(P, S);
}

// Walker steps between refill ballots: the setting, else 24 for residual ratio and next-event tracking (their
// segment ends and scatter events are rarer per step) and 16 for the other estimators (measured on MI355X, 1024^2
// renders of the 512^3 sparse grid).
int refillInterval(const Renderer& r) {
    if (r.stepsPerRefill > 0) return r.stepsPerRefill;
    return (r.mode == VPT_RESIDUAL_RATIO_TRACKING || r.mode == VPT_NEXT_EVENT_TRACKING ||
            r.mode == VPT_NEXT_EVENT_TRACKING_SPECTRAL) ? 24 : 16;
}

int launchBlocks(Renderer& r) {
    if (r.blocksSetting > 0) return r.blocksSetting;
    return r.cuCount * occupancyBlocksPerCU(r.mode, sparseActive(r));
}

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

bool parseVec3(const std::string& s, float out[3]) {
    std::vector<float> v;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ' ' || s[i] == ',' || s[i] == '\t')) i++;
        if (i >= s.size()) break;
        size_t j = i;
        while (j < s.size() && s[j] != ' ' && s[j] != ',' && s[j] != '\t') j++;
        try {
            v.push_back(std::stof(s.substr(i, j - i)));
        } catch (...) {
            return false;
        }
        i = j;
    }
    if (v.size() != 3) return false;
    for (int a = 0; a < 3; a++) out[a] = v[a];
    return true;
}

Light defaultLight() {
    // Light defaults (LightEditorWidget.hpp:69-85)
    Light L{};
    L.lightType = 0;
    L.lightSpace = 1;
    L.spotTotalWidth = 0.0981747704f;
    L.spotFalloffStart = 0.0245436926f;
    L.color[0] = 1.0f; L.color[1] = 0.961538462f; L.color[2] = 0.884615385f;
    L.intensity = 0.5f;
    L.useDistance = 1u;
    L.spotDirection[2] = -1.0f;
    return L;
}

void releaseBatch();  // batched queue state (defined with the batch entry points)

}  // namespace
}  // namespace crvpt

using namespace crvpt;

extern "C" {

const char* crvpt_last_error(void) { return g_lastError.c_str(); }

int crvpt_initialize(void) {
    if (g_refCount == 0) {
        int ndev = 0;
        hipError_t e = hipGetDeviceCount(&ndev);
        if (e != hipSuccess || ndev == 0) return fail(CRVPT_ERR_HIP, "crvpt: no HIP device available");
        Renderer* r = new Renderer();
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) r->cuCount = prop.multiProcessorCount;
        HIP_TRY(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&r->ev0));
        HIP_TRY(hipEventCreate(&r->ev1));
        HIP_TRY(hipEventCreate(&r->evT0));
        HIP_TRY(hipEventCreate(&r->evT1));
        const double eye[3] = {0.0, 0.0, 0.8}, ctr[3] = {0.0, 0.0, -0.2}, up[3] = {0.0, 1.0, 0.0};
        r->view = lookAt(eye, ctr, up);
        // LightEditorWidget constructor (LightEditorWidget.cpp:51-63): one white point light, view orientation.
        Light l0 = defaultLight();
        l0.color[0] = l0.color[1] = l0.color[2] = 1.0f;
        l0.intensity = 1.0f;
        l0.position[0] = 0.68f; l0.position[1] = 0.22f; l0.position[2] = 0.70f;
        l0.lightSpace = 2;
        r->lights.push_back(l0);
        g_r = r;
        HIP_TRY(kpnScratchSetup(nullptr));  // address of the no-allocation entry point's scratch on this device
        // Default scene (Module.cpp:793-795): createFogVolumeSphere<float>(0.25, 0, 0.01) as the only volume; the
        // pass keeps rendering a dense grid (useSparseGrid defaults to false, VolumetricPathTracingPass.hpp:260),
        // which is decoded from the sphere's NanoVDB grid.
        const float centre[3] = {0.0f, 0.0f, 0.0f};
        int rc = setSparseHost(*r, buildFogVolumeSphere(0.25f, centre, double(0.01f), 3.0, "sphere_fog"), false);
        if (rc) return rc;
    }
    g_refCount++;
    return CRVPT_OK;
}

int crvpt_cleanup(void) {
    if (g_refCount <= 0) return fail(CRVPT_ERR_NOT_INITIALIZED, "crvpt: cleanup without initialize");
    g_refCount--;
    if (g_refCount == 0 && g_r) {
        Renderer* r = g_r;
        (void)hipStreamSynchronize(r->stream);
        releaseBatch();
        for (DevBuf* b : {&r->dRaw, &r->dNorm, &r->dNvdb, &r->dMajorant, &r->dRrt, &r->dDecomp, &r->dEnv, &r->dAcc,
                          &r->dPaths, &r->dCounter, &r->dStage, &r->dRays, &r->dSvMin, &r->dSvMax, &r->dSvAvg,
                          &r->dDen0, &r->dDen1, &r->dStats, &r->dFeat, &r->dFeatAcc, &r->dSvgf, &r->dKpn,
                          &r->dBlockLeaf, &r->dBlockTile, &r->dBlockBad, &r->dParams})
            b->release();
        (void)hipEventDestroy(r->ev0);
        (void)hipEventDestroy(r->ev1);
        (void)hipEventDestroy(r->evT0);
        (void)hipEventDestroy(r->evT1);
        (void)hipStreamDestroy(r->stream);
        delete r;
        g_r = nullptr;
    }
    return CRVPT_OK;
}

int crvpt_render_frame(float* out, int channels, int height, int width, int frame_count, int out_on_device,
                       void* stream) {
    if (int rc = requireInit()) return rc;
    // Module.cpp:287-300
    if (channels != 3 && channels != 4)
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "Error in renderFrame: The number of image channels is not equal to 3 or 4.");
    if (height <= 0 || width <= 0) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in renderFrame: empty image.");
    if (frame_count < 0) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in renderFrame: negative frame count.");
    if (!out) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in renderFrame: null output.");
    Renderer& r = *g_r;
    hipStream_t s = out_on_device ? reinterpret_cast<hipStream_t>(stream) : r.stream;
    if (int rc = prepareScene(r)) return rc;
    const uint32_t npix = uint32_t(width) * uint32_t(height);
    if (width != r.accW || height != r.accH || channels != r.accC) {
        HIP_TRY(r.dAcc.ensure(size_t(npix) * sizeof(float4)));
        r.accW = width;
        r.accH = height;
        r.accC = channels;
        r.resetAccumulation();
    }
    // Scene preparation ran on the renderer stream; order it before work on the caller's stream.
    if (s != r.stream) HIP_TRY(hipStreamSynchronize(r.stream));
    SceneParams P;
    SupervoxelGrid S;
    fillSceneParams(r, width, height, P, S);
    // the path kernels read their rarely used constants (camera, sky, claim arithmetic) from a device copy
    // (SceneParams::mem), which launchRender stores before each launch
    HIP_TRY(r.dParams.ensure(sizeof(SceneParams)));
    P.mem = r.dParams.as<SceneParams>();
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    HIP_TRY(r.dCounter.ensure(64 * 1024));  // claim counters of the path pool (one per claim range)
    unsigned long long* stats = nullptr;
    if (r.collectStats) {
        HIP_TRY(r.dStats.ensure(4 * sizeof(unsigned long long)));
        HIP_TRY(hipMemsetAsync(r.dStats.ptr, 0, 4 * sizeof(unsigned long long), s));
        stats = r.dStats.as<unsigned long long>();
    }
    // SVGF wants non-accumulated input and the global frame number (SVGF.hpp:81-82, DISABLE_ACCUMULATION): every
    // frame overwrites the previous one, so only the call's last frame is traced.
    const bool svgf = r.denoiser == 2;
    const int traceFrames = svgf ? std::min(frame_count, 1) : frame_count;
    const uint32_t skipFrames = uint32_t(frame_count - traceFrames);
    if (svgf) r.frameCount = 0;
    const bool denoise = r.denoiser == 1 && r.eawIterations > 0;
    const bool features = featuresActive(r);
    const int foreground = !r.outputForeground ? 0
                           : (r.mode == VPT_RESIDUAL_RATIO_TRACKING || r.mode == VPT_DECOMPOSITION_TRACKING) ? 2 : 1;
    // Paths per launch: every launch ends with a drain in which the pool's lanes run out of work, so fewer, larger
    // launches trace faster (16 M -> 256 M paths: +13 % on a 1024^2 x 256-spp sparse render). The records of one
    // launch take 16 bytes per path (4 GiB at 2^28), plus 48 bytes per path with feature maps (2^26 paths then).
    const uint32_t maxPaths = features ? (1u << 26) : (1u << 28);
    const uint32_t framesPerLaunch = std::max(
            1u, std::min({uint32_t(std::max(traceFrames, 1)), maxPaths / npix, maxFramesPerLaunch(width, height)}));
    HIP_TRY(r.dPaths.ensure(size_t(npix) * framesPerLaunch * sizeof(float4)));
    LaunchConfig cfg;
    cfg.numBlocks = uint32_t(launchBlocks(r));
    cfg.stepsPerRefill = refillInterval(r);
    float4* feat = nullptr;
    FeatureAcc facc{};
    if (features) {
        HIP_TRY(r.dFeat.ensure(size_t(npix) * framesPerLaunch * 3 * sizeof(float4)));
        HIP_TRY(r.dFeatAcc.ensure(size_t(npix) * kFeatureAccArrays * sizeof(float4)));
        feat = r.dFeat.as<float4>();
        facc = featureAcc(r, npix);
        r.featW = width;
        r.featH = height;
    }
    const bool lateOut = denoise || svgf || r.featureMapType != 0;  // the output is written after the sample loop
    float* directOut = (out_on_device && !lateOut) ? out : nullptr;
    if (!out_on_device) {
        HIP_TRY(r.dStage.ensure(size_t(npix) * channels * sizeof(float)));
        if (!lateOut) directOut = r.dStage.as<float>();
    }
    int launches = 0;
    HIP_TRY(hipEventRecord(r.evT0, s));
    HIP_TRY(hipEventRecord(r.ev0, s));
    uint32_t done = 0;
    while (done < uint32_t(traceFrames)) {
        const uint32_t f = std::min(framesPerLaunch, uint32_t(traceFrames) - done);
        P.frameBase = r.globalFrameNumber + skipFrames + done;
        P.numFrames = f;
        P.numUnits = uint64_t(npix) * f;
        HIP_TRY(launchRender(P, sp ? nullptr : &dg, sp ? &ng : nullptr, S, r.dPaths.as<float4>(),
                             r.dCounter.as<uint32_t>(), cfg, stats, feat, s));
        const bool last = done + f == uint32_t(traceFrames);
        if (last) HIP_TRY(hipEventRecord(r.ev1, s));
        HIP_TRY(launchAccumulate(r.dPaths.as<float4>(), npix, f, r.dAcc.as<float4>(), r.frameCount + done,
                                 last ? directOut : nullptr, channels, foreground, s));
        if (features)
            HIP_TRY(launchAccumulateFeatures(P, r.dPaths.as<float4>(), feat, npix, f, r.frameCount + done, svgf,
                                             facc, s));
        done += f;
        launches++;
    }
    if (traceFrames == 0) HIP_TRY(hipEventRecord(r.ev1, s));
    if (frame_count == 0 || lateOut) {
        float* target = out_on_device ? out : r.dStage.as<float>();
        if (frame_count == 0 && r.frameCount == 0) {
            HIP_TRY(hipMemsetAsync(r.dAcc.ptr, 0, size_t(npix) * sizeof(float4), s));
        }
        if (denoise) {
            HIP_TRY(r.dDen0.ensure(size_t(npix) * sizeof(float4)));
            HIP_TRY(r.dDen1.ensure(size_t(npix) * sizeof(float4)));
            EawParams ep{width, height, r.eawIterations, r.eawPhiColor, r.eawPhiPosition, r.eawPhiNormal,
                         r.eawUseColorDistance};
            float4* res = nullptr;
            HIP_TRY(launchEaw(r.dAcc.as<float4>(), facc.firstX, facc.normal, r.dDen0.as<float4>(),
                              r.dDen1.as<float4>(), ep, &res, s));
            HIP_TRY(launchWriteCHW(res, npix, target, channels, s));
        } else if (r.featureMapType != 0) {
            HIP_TRY(launchFeatureOut(r.featureMapType, facc, r.dAcc.as<float4>(), npix, target, channels, s));
        } else if (svgf && frame_count > 0) {
            HIP_TRY(r.dSvgf.ensure(size_t(npix) * kSvgfBuffers * sizeof(float4)));
            float4* b = r.dSvgf.as<float4>();
            if (r.svgfW != width || r.svgfH != height) {
                // new history: all zero (SVGF.cpp:244-268)
                HIP_TRY(hipMemsetAsync(b, 0, size_t(npix) * 4 * sizeof(float4), s));
                r.svgfW = width;
                r.svgfH = height;
            }
            const SvgfState st{b, b + npix, b + 2 * size_t(npix), b + 3 * size_t(npix), b + 4 * size_t(npix),
                               b + 5 * size_t(npix), b + 6 * size_t(npix), b + 7 * size_t(npix)};
            const SvgfInputs in{r.dAcc.as<float4>(), facc.flow, facc.normal, facc.depth, width, height};
            float4* res = nullptr;
            HIP_TRY(launchSvgf(in, st, r.svgfIterations, &res, s));
            HIP_TRY(launchWriteCHW(res, npix, target, channels, s));
        } else {
            HIP_TRY(launchWriteCHW(r.dAcc.as<float4>(), npix, target, channels, s));
        }
    }
    HIP_TRY(hipEventRecord(r.evT1, s));
    r.frameCount = svgf ? uint32_t(traceFrames) : r.frameCount + uint32_t(frame_count);
    r.globalFrameNumber += uint32_t(frame_count);
    r.lastPaths = uint64_t(npix) * uint64_t(frame_count);
    r.lastLaunches = launches;
    if (!out_on_device) {
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(out, r.dStage.ptr, size_t(npix) * channels * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CRVPT_OK;
}

namespace {
// FeatureMapTypeVpt values the gfx950 path produces (VolumetricPathTracingPass.hpp:67-73); Depth Blended, Depth
// Nearest Opaque, the transmittance volume and the absorption moments belong to renderers outside this path.
bool featureMapSupported(int m) { return (m >= 0 && m <= 14) || m == 17 || m == kFeatureMapRadianceVariance; }
}  // namespace

int crvpt_set_use_feature_maps(const int32_t* maps, int count) {
    if (int rc = requireInit()) return rc;
    if (count < 0 || (count > 0 && !maps)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_use_feature_maps: bad list");
    std::vector<int> v;
    for (int i = 0; i < count; i++) {
        if (!featureMapSupported(maps[i]))
            return fail(CRVPT_ERR_UNSUPPORTED, "set_use_feature_maps: feature map " + std::to_string(maps[i]) +
                                                   " is not produced by the gfx950 path tracer");
        v.push_back(maps[i]);
    }
    g_r->featureMaps = v;
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_feature_map_type(int32_t type) {
    if (int rc = requireInit()) return rc;
    if (!featureMapSupported(type))
        return fail(CRVPT_ERR_UNSUPPORTED, "set_feature_map_type: feature map " + std::to_string(type) +
                                               " is not produced by the gfx950 path tracer");
    if (type != g_r->featureMapType) g_r->resetAccumulation();
    g_r->featureMapType = type;
    return CRVPT_OK;
}

int crvpt_get_feature_map(int32_t map, float* out, int channels, int height, int width, int out_on_device,
                          void* stream) {
    if (int rc = requireInit()) return rc;
    if (channels != 3 && channels != 4)
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "Error in renderFrame: The number of image channels is not equal to 3 or 4.");
    if (!out || height <= 0 || width <= 0) return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_feature_map: bad output");
    if (!featureMapSupported(map))
        return fail(CRVPT_ERR_UNSUPPORTED, "get_feature_map: feature map " + std::to_string(map) +
                                               " is not produced by the gfx950 path tracer");
    Renderer& r = *g_r;
    if (width != r.accW || height != r.accH)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_feature_map: no frame of this size was rendered");
    if (map != 0 && (r.featW != width || r.featH != height || !r.dFeatAcc.ptr))
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "get_feature_map: enable the map with set_use_feature_maps before rendering");
    const uint32_t npix = uint32_t(width) * uint32_t(height);
    hipStream_t s = out_on_device ? reinterpret_cast<hipStream_t>(stream) : r.stream;
    float* target = out;
    if (!out_on_device) {
        HIP_TRY(r.dStage.ensure(size_t(npix) * channels * sizeof(float)));
        target = r.dStage.as<float>();
    }
    const FeatureAcc acc = map != 0 ? featureAcc(r, npix) : FeatureAcc{};
    HIP_TRY(launchFeatureOut(map, acc, r.dAcc.as<float4>(), npix, target, channels, s));
    if (!out_on_device) {
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(out, r.dStage.ptr, size_t(npix) * channels * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CRVPT_OK;
}

int crvpt_set_view_projection_matrix_as_previous(void) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const int w = r.accW > 0 ? r.accW : 1, h = r.accH > 0 ? r.accH : 1;
    const Mat4 vp = mul(projection(r, w, h), r.view);
    for (int i = 0; i < 16; i++) r.prevViewProj[i] = float(vp.m[i]);
    r.prevViewProjSet = true;
    return CRVPT_OK;
}

int crvpt_kpn_per_pixel_kernel(const float* image, const float* weights, float* out, int n, int c, int img_h,
                                int img_w, int weight_h, int weight_w, int kernel_size, void* stream) {
    if (int rc = requireInit()) return rc;
    if (!image || !weights || !out || n <= 0 || c <= 0 || img_h <= 0 || img_w <= 0 || weight_h <= 0 ||
        weight_w <= 0 || kernel_size <= 0 || kernel_size % 2 == 0)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_per_pixel_kernel: invalid arguments");
    if (c > 4) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error, must have <= 4 channels");
    HIP_TRY(launchPerPixelKernel(image, weights, out, n, c, img_h, img_w, weight_h, weight_w, kernel_size,
                                 reinterpret_cast<hipStream_t>(stream)));
    return CRVPT_OK;
}

int crvpt_kpn_inv_iter_kernel(const float* x, const float* low_res_pred, const float* low_res_x, const float* prev,
                              const float* fused_weights, float* out, int n, int c, int img_h, int img_w, int lo_h,
                              int lo_w, int weight_c, int weight_h, int weight_w, int kernel_size, int use_previous,
                              void* stream) {
    if (int rc = requireInit()) return rc;
    const int kk = kernel_size * kernel_size;
    if (!x || !low_res_pred || !low_res_x || !fused_weights || !out || (use_previous && !prev) || n <= 0 || c <= 0 ||
        kernel_size <= 0 || kernel_size % 2 == 0 || weight_c < kk + 3)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_inv_iter_kernel: invalid arguments");
    if (c > 4) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error, must have <= 4 channels");
    Renderer& r = *g_r;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int h = std::min(img_h, weight_h), w = std::min(img_w, weight_w);
    HIP_TRY(r.dKpn.ensure((size_t(n) * c * h * w + size_t(n) * kk * weight_h * weight_w) * sizeof(float)));
    float* base = r.dKpn.as<float>();
    float* soft = base + size_t(n) * c * h * w;
    HIP_TRY(launchInvIterBase(x, low_res_pred, low_res_x, prev, fused_weights, base, use_previous != 0, n, c, img_h,
                              img_w, lo_h, lo_w, weight_h, weight_w, kernel_size, s));
    HIP_TRY(launchSoftmaxChannels(fused_weights, soft, n, weight_c, kk, weight_h, weight_w, s));
    HIP_TRY(launchPerPixelKernel(base, soft, out, n, c, h, w, weight_h, weight_w, kernel_size, s));
    return CRVPT_OK;
}

int crvpt_kpn_inv_iter_kernel_noalloc(const float* x, const float* low_res_pred, const float* low_res_x,
                                      const float* prev, const float* fused_weights, float* out, int n, int c,
                                      int img_h, int img_w, int lo_h, int lo_w, int weight_c, int weight_h,
                                      int weight_w, int kernel_size, int use_previous, void* stream) {
    if (int rc = requireInit()) return rc;
    const int kk = kernel_size * kernel_size;
    if (!x || !low_res_pred || !low_res_x || !fused_weights || !out || (use_previous && !prev) || n <= 0 || c <= 0 ||
        img_h <= 0 || img_w <= 0 || weight_h <= 0 || weight_w <= 0 || kernel_size <= 0 || kernel_size % 2 == 0 ||
        weight_c < kk + 3)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "kpn_inv_iter_kernel_noalloc: invalid arguments");
    if (c > 4) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error, must have <= 4 channels");
    const int h = std::min(img_h, weight_h), w = std::min(img_w, weight_w);
    const uint64_t need = uint64_t(n) * c * h * w + uint64_t(n) * kk * weight_h * weight_w;
    if (need > kKpnNoAllocScratchFloats)
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "kpn_inv_iter_kernel_noalloc: the call needs " + std::to_string(need) +
                        " scratch floats (n*c*min(img_h, weight_h)*min(img_w, weight_w) + n*k*k*weight_h*weight_w), "
                        "more than the fixed " + std::to_string(kKpnNoAllocScratchFloats) +
                        "; use kpn_inv_iter_kernel");
    HIP_TRY(launchInvIterNoAlloc(x, low_res_pred, low_res_x, prev, fused_weights, out, use_previous != 0, n, c, img_h,
                                 img_w, lo_h, lo_w, weight_c, weight_h, weight_w, kernel_size,
                                 reinterpret_cast<hipStream_t>(stream)));
    return CRVPT_OK;
}

int crvpt_get_last_render_stats(double* path_kernel_ms, double* total_ms, uint64_t* num_paths, int* launches) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    float a = 0.0f, b = 0.0f;
    HIP_TRY(hipEventSynchronize(r.evT1));
    HIP_TRY(hipEventElapsedTime(&a, r.ev0, r.ev1));
    HIP_TRY(hipEventElapsedTime(&b, r.evT0, r.evT1));
    if (path_kernel_ms) *path_kernel_ms = a;
    if (total_ms) *total_ms = b;
    if (num_paths) *num_paths = r.lastPaths;
    if (launches) *launches = r.lastLaunches;
    return CRVPT_OK;
}

int crvpt_set_collect_stats(int collect) {
    if (int rc = requireInit()) return rc;
    g_r->collectStats = collect != 0;
    return CRVPT_OK;
}

int crvpt_get_last_render_counters(uint64_t out[4]) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    if (!r.collectStats || !r.dStats.ptr) return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_last_render_counters: "
                                                      "enable crvpt_set_collect_stats before rendering");
    unsigned long long v[4];
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(v, r.dStats.ptr, sizeof(v), hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; i++) out[i] = v[i];
    return CRVPT_OK;
}

int crvpt_load_cloud_file(const char* filename) {
    if (int rc = requireInit()) return rc;
    if (!filename) return fail(CRVPT_ERR_INVALID_ARGUMENT, "load_cloud_file: null file name");
    std::string fn(filename);
    std::ifstream f(fn, std::ios::binary);
    if (!f) return fail(CRVPT_ERR_IO, "Error in CloudData::loadFromFile: File \"" + fn + "\" does not exist.");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const std::string ext = fn.size() >= 4 ? fn.substr(fn.size() - 4) : "";
    Renderer& r = *g_r;
    if (ext == ".xyz") {
        int dims[3];
        double vs[3];
        std::vector<float> v;
        std::string err;
        if (!parseXyzFile(bytes, dims, vs, v, err)) return fail(CRVPT_ERR_IO, err);
        int rc = setDenseHost(r, dims[0], dims[1], dims[2], std::move(v));
        if (rc) return rc;
        r.voxelSize[0] = float(vs[0]);
        r.voxelSize[1] = float(vs[1]);
        r.voxelSize[2] = float(vs[2]);
        computeDenseBounds(r);
        return CRVPT_OK;
    }
    if (ext == "nvdb") {
        std::vector<uint8_t> grid;
        std::string err;
        if (!extractNanoVdbGridFromFile(bytes, grid, err)) return fail(CRVPT_ERR_IO, err);
        const int rc = setSparseHost(r, std::move(grid), false);
        // VolumetricPathTracingPass::setCloudData (VolumetricPathTracingPass.cpp:780-795): grids of 1024^3 index
        // entries and more switch the pass to the sparse grid, smaller ones keep the current choice (by default the
        // dense field decoded from the grid). The reference also requires the grid to stay below 2 GiB, a limit of
        // its GLSL accessor that this renderer does not have.
        if (rc == CRVPT_OK && uint64_t(r.dims[0]) * uint64_t(r.dims[1]) * uint64_t(r.dims[2]) >= (1ull << 30))
            r.useSparse = true;
        return rc;
    }
    return fail(CRVPT_ERR_UNSUPPORTED, "load_cloud_file: unsupported file format \"" + fn + "\"");
}

int crvpt_read_xyz_file(const char* filename, int32_t dims[3], double voxel_size[3], float* out,
                        uint64_t capacity) {
    if (!filename || !dims || !voxel_size) return fail(CRVPT_ERR_INVALID_ARGUMENT, "read_xyz_file: null argument");
    std::ifstream f(filename, std::ios::binary);
    if (!f) return fail(CRVPT_ERR_IO, std::string("read_xyz_file: cannot open \"") + filename + "\"");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    int d[3];
    std::vector<float> v;
    std::string err;
    if (!parseXyzFile(bytes, d, voxel_size, v, err)) return fail(CRVPT_ERR_IO, err);
    for (int a = 0; a < 3; a++) dims[a] = d[a];
    if (out) {
        if (capacity < v.size()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "read_xyz_file: output too small");
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    }
    return CRVPT_OK;
}

int crvpt_read_nvdb_file(const char* filename, uint8_t* out, uint64_t capacity, uint64_t* size) {
    if (!filename || !size) return fail(CRVPT_ERR_INVALID_ARGUMENT, "read_nvdb_file: null argument");
    std::ifstream f(filename, std::ios::binary);
    if (!f) return fail(CRVPT_ERR_IO, std::string("read_nvdb_file: cannot open \"") + filename + "\"");
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<uint8_t> grid;
    std::string err;
    if (!extractNanoVdbGridFromFile(bytes, grid, err)) return fail(CRVPT_ERR_IO, err);
    NanoInfo info;
    if (!parseNanoVdb(grid.data(), grid.size(), info, err)) return fail(CRVPT_ERR_INVALID_ARGUMENT, err);
    *size = grid.size();
    if (out) {
        if (capacity < grid.size()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "read_nvdb_file: output too small");
        std::memcpy(out, grid.data(), grid.size());
    }
    return CRVPT_OK;
}

int crvpt_set_density_field(int nx, int ny, int nz, const float* values) {
    if (int rc = requireInit()) return rc;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !values)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_density_field: invalid grid size or null data");
    std::vector<float> v(values, values + size_t(nx) * ny * nz);
    return setDenseHost(*g_r, nx, ny, nz, std::move(v));
}

int crvpt_set_nanovdb_buffer(const uint8_t* data, uint64_t size) {
    if (int rc = requireInit()) return rc;
    if (!data || size == 0) return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_nanovdb_buffer: empty buffer");
    std::vector<uint8_t> b(data, data + size);
    const int rc = setSparseHost(*g_r, std::move(b), false);
    if (rc == CRVPT_OK) g_r->useSparse = true;  // a rejected buffer leaves the renderer unchanged
    return rc;
}

int crvpt_generate_fbm_cloud(int nx, int ny, int nz, uint32_t seed, int octaves, float lacunarity, float gain,
                             float base_frequency, float edge0, float edge1, int height_falloff) {
    if (int rc = requireInit()) return rc;
    if (nx <= 0 || ny <= 0 || nz <= 0 || octaves <= 0 || !(edge1 > edge0))
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "generate_fbm_cloud: invalid parameters");
    Renderer& r = *g_r;
    const size_t n = size_t(nx) * ny * nz;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(r.dRaw.ensure(n * sizeof(float)));
    FbmParams F{seed, octaves, lacunarity, gain, base_frequency, edge0, edge1, height_falloff};
    HIP_TRY(launchFbm(r.dRaw.as<float>(), nx, ny, nz, F, r.stream));
    r.denseHost.resize(n);
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipMemcpy(r.denseHost.data(), r.dRaw.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<float> v = std::move(r.denseHost);
    return setDenseHost(r, nx, ny, nz, std::move(v));
}

int crvpt_get_density_field(float* out, uint64_t count, int normalized) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    if (int rc = decodeDenseFromSparse(r)) return rc;  // NanoVDB-only volumes: CloudData::getDenseDensityField
    if (!r.hasDense) return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_density_field: no dense field");
    const size_t n = size_t(r.dims[0]) * r.dims[1] * r.dims[2];
    if (count < n || !out) return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_density_field: output too small");
    if (!r.denseOnHostValid) {
        r.denseHost.resize(n);
        HIP_TRY(hipMemcpy(r.denseHost.data(), r.dRaw.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
        r.denseOnHostValid = true;
    }
    for (size_t i = 0; i < n; i++) out[i] = normalized ? (r.denseHost[i] - r.vmin) / (r.vmax - r.vmin) : r.denseHost[i];
    return CRVPT_OK;
}

int crvpt_build_nanovdb_from_dense(const float* values, int nx, int ny, int nz, double dx, const double origin[3],
                                   uint8_t* out, uint64_t capacity, uint64_t* size) {
    return crvpt_build_nanovdb_from_dense_typed(values, nx, ny, nz, dx, origin, 0, out, capacity, size);
}

int crvpt_build_nanovdb_from_dense_typed(const float* values, int nx, int ny, int nz, double dx,
                                         const double origin[3], int value_type, uint8_t* out, uint64_t capacity,
                                         uint64_t* size) {
    if (!values || nx <= 0 || ny <= 0 || nz <= 0 || !(dx > 0.0) || !origin || !size ||
        (value_type != 0 && value_type != 1))
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "build_nanovdb_from_dense: invalid arguments");
    const int dims[3] = {nx, ny, nz};
    std::vector<uint8_t> b = buildNanoVdbFromDense(values, dims, dx, origin, "tmp", value_type);
    *size = b.size();
    if (out) {
        if (capacity < b.size()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "build_nanovdb_from_dense: too small");
        std::memcpy(out, b.data(), b.size());
    }
    return CRVPT_OK;
}

int crvpt_build_fog_volume_sphere(float radius, const float center[3], double voxel_size, double half_width,
                                  uint8_t* out, uint64_t capacity, uint64_t* size) {
    if (!center || !size || !(radius > 0.0f) || !(voxel_size > 0.0) || !(half_width > 0.0))
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "build_fog_volume_sphere: invalid arguments");
    std::vector<uint8_t> b = buildFogVolumeSphere(radius, center, voxel_size, half_width, "sphere_fog");
    *size = b.size();
    if (out) {
        if (capacity < b.size()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "build_fog_volume_sphere: too small");
        std::memcpy(out, b.data(), b.size());
    }
    return CRVPT_OK;
}

int crvpt_set_sparse_value_type(int value_type) {
    if (int rc = requireInit()) return rc;
    if (value_type != 0 && value_type != 1)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_sparse_value_type: 0 (FP32) or 1 (Fp16)");
    Renderer& r = *g_r;
    if (value_type == r.sparseValueType) return CRVPT_OK;
    r.sparseValueType = value_type;
    // a grid built from the dense field is rebuilt with the new leaf type when next needed
    if (r.hasDense && r.hasSparse) {
        HIP_TRY(hipDeviceSynchronize());
        r.hasSparse = false;
        r.nvdbHost.clear();
        r.sceneDirty = r.tablesDirty = true;
        r.refTablesValid = false;
        r.resetAccumulation();
    }
    return CRVPT_OK;
}

int crvpt_get_nanovdb_buffer(uint8_t* out, uint64_t capacity, uint64_t* size) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const bool prev = r.useSparse;
    r.useSparse = true;
    int rc = ensureSparse(r);
    r.useSparse = prev;
    if (rc) return rc;
    if (size) *size = r.nvdbHost.size();
    if (out) {
        if (capacity < r.nvdbHost.size()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_nanovdb_buffer: too small");
        std::memcpy(out, r.nvdbHost.data(), r.nvdbHost.size());
    }
    return CRVPT_OK;
}

int crvpt_set_output_foreground_map(int output_foreground_map) {
    if (int rc = requireInit()) return rc;
    const bool v = output_foreground_map != 0;
    if (v != g_r->outputForeground) {
        g_r->outputForeground = v;
        g_r->resetAccumulation();
    }
    return CRVPT_OK;
}

int crvpt_set_use_sparse_grid(int use_sparse) {
    if (int rc = requireInit()) return rc;
    // A volume loaded only as NanoVDB gets its dense field decoded from the grid at the next render (ensureDense).
    g_r->useSparse = use_sparse != 0;
    g_r->sceneDirty = g_r->tablesDirty = true;
    g_r->refTablesValid = false;
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_grid_interpolation_type(int type) {
    if (int rc = requireInit()) return rc;
    if (type < 0 || type > 2) return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_grid_interpolation_type: invalid type");
    g_r->interpolation = type;
    g_r->refTablesValid = false;
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_vpt_mode(int64_t mode) {
    if (int rc = requireInit()) return rc;
    if (mode < VPT_DELTA_TRACKING || mode > VPT_NEXT_EVENT_TRACKING_SPECTRAL)
        return fail(CRVPT_ERR_UNSUPPORTED, "set_vpt_mode: mode " + std::to_string(mode) +
                                               " is not on the gfx950 hot path (supported: 0-6)");
    g_r->mode = int(mode);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_vpt_mode_from_name(const char* mode_name) {
    if (int rc = requireInit()) return rc;
    if (!mode_name) return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_vpt_mode_from_name: null name");
    for (int m = 0; m < 9; m++)
        if (std::string(mode_name) == kModeNames[m]) return crvpt_set_vpt_mode(m);
    // Module.cpp:1046-1055 ignores unknown names; the C ABI reports them.
    return fail(CRVPT_ERR_INVALID_ARGUMENT, std::string("set_vpt_mode_from_name: unknown mode \"") + mode_name + "\"");
}

int crvpt_set_scattering_albedo(const double albedo[3]) {
    if (int rc = requireInit()) return rc;
    for (int a = 0; a < 3; a++) g_r->albedo[a] = float(albedo[a]);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_extinction_scale(double s) {
    if (int rc = requireInit()) return rc;
    g_r->extinctionScale = float(s);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_extinction_base(const double b[3]) {
    if (int rc = requireInit()) return rc;
    for (int a = 0; a < 3; a++) g_r->extinctionBase[a] = float(b[a]);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_phase_g(double g) {
    if (int rc = requireInit()) return rc;
    g_r->phaseG = float(g);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_seed_offset(int64_t offset) {
    if (int rc = requireInit()) return rc;
    g_r->seedOffset = uint32_t(offset);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_super_voxel_size(int32_t super_voxel_size) {
    if (int rc = requireInit()) return rc;
    if (super_voxel_size < 1 || super_voxel_size > 64)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_super_voxel_size: the edge must lie in [1, 64]");
    g_r->superVoxelSize = super_voxel_size;
    g_r->refTablesValid = false;
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_max_bounces(int32_t max_bounces) {
    if (int rc = requireInit()) return rc;
    if (max_bounces < 0) return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_max_bounces: negative bounce limit");
    g_r->maxBounces = uint32_t(max_bounces);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_get_camera_position(double out[3]) {
    if (int rc = requireInit()) return rc;
    for (int a = 0; a < 3; a++) out[a] = g_r->camPos[a];
    return CRVPT_OK;
}

int crvpt_get_camera_view_matrix(double out[16]) {
    if (int rc = requireInit()) return rc;
    for (int i = 0; i < 16; i++) out[i] = float(g_r->view.m[i]);
    return CRVPT_OK;
}

int crvpt_get_camera_fovy(double* out) {
    if (int rc = requireInit()) return rc;
    *out = g_r->fovy;
    return CRVPT_OK;
}

int crvpt_set_camera_position(const double p[3]) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const double up[3] = {0.0, 1.0, 0.0};
    for (int a = 0; a < 3; a++) r.memberPos[a] = r.camPos[a] = float(p[a]);
    r.view = lookAt(r.memberPos, r.memberTarget, up);
    r.resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_camera_target(const double t[3]) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const double up[3] = {0.0, 1.0, 0.0};
    for (int a = 0; a < 3; a++) r.memberTarget[a] = float(t[a]);
    r.view = lookAt(r.memberPos, r.memberTarget, up);
    r.resetAccumulation();
    return CRVPT_OK;
}

int crvpt_overwrite_camera_view_matrix(const double v[16]) {
    if (int rc = requireInit()) return rc;
    for (int i = 0; i < 16; i++) g_r->view.m[i] = float(v[i]);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_camera_fovy(double fovy_degrees) {
    if (int rc = requireInit()) return rc;
    g_r->fovy = fovy_degrees * 3.14159265358979323846 / 180.0;
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_use_builtin_environment_map(const char* name) {
    if (int rc = requireInit()) return rc;
    const std::string n = name ? name : "";
    Renderer& r = *g_r;
    r.useEnvImage = false;
    if (n == "Default") r.envKind = ENV_DEFAULT;
    else if (n == "Black") r.envKind = ENV_BLACK;
    else if (n == "Single Color") r.envKind = ENV_SINGLE_COLOR;
    else return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_use_builtin_environment_map: unknown map \"" + n + "\"");
    r.resetAccumulation();
    return CRVPT_OK;
}

int crvpt_load_environment_map_from_buffer(const float* px, int channels, int height, int width) {
    if (int rc = requireInit()) return rc;
    if (!px || (channels != 3 && channels != 4) || height <= 0 || width <= 0)
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "Error in loadEnvironmentMapFromTensor: Number of channels must be 3 or 4.");
    Renderer& r = *g_r;
    const size_t npix = size_t(width) * height;
    r.envImage.assign(npix * 4, 1.0f);
    bool black = true;
    for (size_t i = 0; i < npix; i++) {
        for (int c = 0; c < 3; c++) {
            const float v = px[size_t(c) * npix + i];
            r.envImage[i * 4 + c] = v;
            if (v != 0.0f) black = false;
        }
        if (channels == 4) r.envImage[i * 4 + 3] = px[3 * npix + i];
    }
    r.envW = width;
    r.envH = height;
    r.envImageBlack = black;
    r.useEnvImage = true;
    r.envDirty = true;
    r.resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_environment_map_intensity(double f) {
    if (int rc = requireInit()) return rc;
    g_r->envFactorRgbMode = false;
    g_r->envFactor = float(f);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_environment_map_intensity_rgb(const double f[3]) {
    if (int rc = requireInit()) return rc;
    g_r->envFactorRgbMode = true;
    for (int a = 0; a < 3; a++) g_r->envFactorRgb[a] = float(f[a]);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_use_lights(int use_lights) {
    if (int rc = requireInit()) return rc;
    g_r->useLights = use_lights != 0;
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_clear_lights(void) {
    if (int rc = requireInit()) return rc;
    g_r->lights.clear();
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_add_light(void) {
    if (int rc = requireInit()) return rc;
    if (g_r->lights.size() >= size_t(kMaxLights))
        return fail(CRVPT_ERR_UNSUPPORTED, "add_light: at most " + std::to_string(kMaxLights) + " lights");
    g_r->lights.push_back(defaultLight());
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_remove_light(int64_t idx) {
    if (int rc = requireInit()) return rc;
    if (idx >= 0 && size_t(idx) < g_r->lights.size()) g_r->lights.erase(g_r->lights.begin() + idx);
    g_r->resetAccumulation();
    return CRVPT_OK;
}

int crvpt_get_num_lights(int64_t* out) {
    if (int rc = requireInit()) return rc;
    *out = int64_t(g_r->lights.size());
    return CRVPT_OK;
}

int crvpt_set_light_property(int64_t idx, const char* key_c, const char* value_c) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    // LightEditorWidget::setLightProperty (LightEditorWidget.cpp:286-320)
    if (idx < 0 || size_t(idx) >= r.lights.size())
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in LightEditorWidget::setLightProperty: Invalid light index.");
    Light& L = r.lights[size_t(idx)];
    const std::string key = key_c ? key_c : "", value = trim(value_c ? value_c : "");
    float v3[3];
    try {
        if (key == "type") {
            const char* names[] = {"Point Light", "Spotlight", "Directional Light"};
            for (int i = 0; i < 3; i++)
                if (value == names[i]) L.lightType = uint32_t(i);
        } else if (key == "color") {
            if (!parseVec3(value, v3)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in stringToVec3: Invalid number of entries.");
            for (int a = 0; a < 3; a++) L.color[a] = v3[a];
        } else if (key == "intensity") {
            L.intensity = std::stof(value);
        } else if (key == "use_distance") {
            // The reference writes this key into the intensity field (LightEditorWidget.cpp:302-303).
            L.intensity = (value == "1" || value == "true") ? 1.0f : 0.0f;
        } else if (key == "position" || key == "direction") {
            if (!parseVec3(value, v3)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in stringToVec3: Invalid number of entries.");
            for (int a = 0; a < 3; a++) L.position[a] = v3[a];
        } else if (key == "space") {
            const char* names[] = {"World", "View", "View Orientation"};
            for (int i = 0; i < 3; i++)
                if (value == names[i]) L.lightSpace = uint32_t(i);
        } else if (key == "spot_total_width") {
            L.spotTotalWidth = std::stof(value);
        } else if (key == "spot_falloff_start") {
            L.spotFalloffStart = std::stof(value);
        } else if (key == "spot_direction") {
            if (!parseVec3(value, v3)) return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in stringToVec3: Invalid number of entries.");
            for (int a = 0; a < 3; a++) L.spotDirection[a] = v3[a];
        }
    } catch (...) {
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_light_property: cannot parse \"" + value + "\"");
    }
    r.resetAccumulation();
    return CRVPT_OK;
}

int crvpt_set_max_grid_extent(double e) {
    if (int rc = requireInit()) return rc;
    g_r->maxGridExtent = float(e);
    // Like CloudData::setMaxGridExtent, this takes effect for volumes loaded afterwards.
    return CRVPT_OK;
}

int crvpt_get_render_bounding_box(double out[6]) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const bool sp = sparseActive(r);
    for (int a = 0; a < 3; a++) {
        out[2 * a] = sp ? r.boxMinSparse[a] : r.boxMinDense[a];
        out[2 * a + 1] = sp ? r.boxMaxSparse[a] : r.boxMaxDense[a];
    }
    return CRVPT_OK;
}

int crvpt_get_volume_voxel_size(int64_t out[3]) {
    if (int rc = requireInit()) return rc;
    out[0] = g_r->dims[2];
    out[1] = g_r->dims[1];
    out[2] = g_r->dims[0];
    return CRVPT_OK;
}

int crvpt_set_denoiser(const char* name) {
    if (int rc = requireInit()) return rc;
    const std::string n = name ? name : "";
    Renderer& r = *g_r;
    // DENOISER_NAMES (Denoiser.hpp:63-79) plus the short names "Default", "EAW"
    if (n == "None") r.denoiser = 0;
    else if (n == "Default" || n == "EAW" || n == "Edge-Avoiding À-Trous Wavelet Transform" ||
             n == "Edge-Avoiding À-Trous Wavelet")
        r.denoiser = 1;
    else if (n == "SVGF") r.denoiser = 2;
    else return fail(CRVPT_ERR_INVALID_ARGUMENT, "Error in setDenoiser: Invalid denoiser name \"" + n + "\".");
    r.svgfW = r.svgfH = 0;  // a new denoiser starts without history (SVGF.cpp:244-268)
    r.svgfIterations = 5;   // SVGF_ATrous_Pass::maxNumIterations (SVGF.hpp:160)
    r.resetAccumulation();
    // VolumetricPathTracingModuleRenderer::setDenoiserType clears the settings (EAW defaults, EAWDenoiser.hpp:107-119)
    r.eawIterations = 0;
    r.eawPhiColor = 5.0f;
    r.eawPhiPosition = 0.1f;
    r.eawPhiNormal = 0.1f;
    return CRVPT_OK;
}

int crvpt_set_denoiser_property(const char* key_c, const char* value_c) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const std::string key = key_c ? key_c : "", value = value_c ? value_c : "";
    try {
        if (key == "maxNumIterations" || key == "numIterations") {
            r.eawIterations = std::stoi(value);
            r.svgfIterations = r.eawIterations;
        }
        else if (key == "phiColor") r.eawPhiColor = std::stof(value);
        else if (key == "phiPosition") r.eawPhiPosition = std::stof(value);
        else if (key == "phiNormal") r.eawPhiNormal = std::stof(value);
        else if (key == "useColorDistance") r.eawUseColorDistance = std::stoi(value);
        else return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_denoiser_property: unknown key \"" + key + "\"");
    } catch (...) {
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_denoiser_property: cannot parse \"" + value + "\"");
    }
    return CRVPT_OK;
}

int crvpt_transmittance_batch(int estimator, int n_rays, const float* origins, const float* directions,
                              const float* t_max, int n_samples, uint32_t seed, float* mean, float* m2,
                              int on_device, void* stream) {
    if (int rc = requireInit()) return rc;
    if (estimator < 0 || estimator > 3 || n_rays < 0 || n_samples <= 0 || !origins || !directions || !mean)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "transmittance_batch: invalid arguments");
    Renderer& r = *g_r;
    const int prevMode = r.mode;
    if (estimator == T_RESIDUAL_RATIO) r.mode = VPT_RESIDUAL_RATIO_TRACKING;
    if (estimator == T_DECOMPOSITION) r.mode = VPT_DECOMPOSITION_TRACKING;
    int rc = prepareScene(r);
    r.mode = prevMode;
    if (rc) return rc;
    if (n_rays == 0) return CRVPT_OK;
    hipStream_t s = on_device ? reinterpret_cast<hipStream_t>(stream) : r.stream;
    if (s != r.stream) HIP_TRY(hipStreamSynchronize(r.stream));
    SceneParams P;
    SupervoxelGrid S;
    fillSceneParams(r, 1, 1, P, S);
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    const float *o = origins, *d = directions, *tm = t_max;
    float *mo = mean, *m2o = m2;
    if (!on_device) {
        const size_t nb = size_t(n_rays) * (3 + 3 + 1 + 1 + 1) * sizeof(float);
        HIP_TRY(r.dRays.ensure(nb));
        float* base = r.dRays.as<float>();
        float* dO = base;
        float* dD = dO + 3 * size_t(n_rays);
        float* dT = dD + 3 * size_t(n_rays);
        float* dM = dT + n_rays;
        float* dM2 = dM + n_rays;
        HIP_TRY(hipMemcpy(dO, origins, 3 * size_t(n_rays) * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dD, directions, 3 * size_t(n_rays) * sizeof(float), hipMemcpyHostToDevice));
        if (t_max) HIP_TRY(hipMemcpy(dT, t_max, size_t(n_rays) * sizeof(float), hipMemcpyHostToDevice));
        o = dO;
        d = dD;
        tm = t_max ? dT : nullptr;
        mo = dM;
        m2o = m2 ? dM2 : nullptr;
    }
    HIP_TRY(hipEventRecord(r.ev0, s));
    HIP_TRY(launchTransmittance(P, sp ? nullptr : &dg, sp ? &ng : nullptr, S, estimator, o, d, tm, n_rays, n_samples,
                                seed, mo, m2o, s));
    HIP_TRY(hipEventRecord(r.ev1, s));
    HIP_TRY(hipEventRecord(r.evT1, s));
    if (!on_device) {
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipMemcpy(mean, mo, size_t(n_rays) * sizeof(float), hipMemcpyDeviceToHost));
        if (m2) HIP_TRY(hipMemcpy(m2, m2o, size_t(n_rays) * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CRVPT_OK;
}

int crvpt_grid_lookup_batch(int n, const int32_t* ijk, float* out) {
    if (int rc = requireInit()) return rc;
    if (n < 0 || (n > 0 && (!ijk || !out))) return fail(CRVPT_ERR_INVALID_ARGUMENT, "grid_lookup_batch: bad arguments");
    Renderer& r = *g_r;
    if (int rc = prepareScene(r)) return rc;
    if (n == 0) return CRVPT_OK;
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    HIP_TRY(r.dRays.ensure(size_t(n) * 16));
    int32_t* dIjk = r.dRays.as<int32_t>();
    float* dOut = reinterpret_cast<float*>(dIjk + 3 * size_t(n));
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipMemcpy(dIjk, ijk, 12 * size_t(n), hipMemcpyHostToDevice));
    HIP_TRY(launchGridLookup(sp ? nullptr : &dg, sp ? &ng : nullptr, n, dIjk, dOut, r.stream));
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipMemcpy(out, dOut, 4 * size_t(n), hipMemcpyDeviceToHost));
    return CRVPT_OK;
}

int crvpt_grid_lookup_world_batch(int n, const float* xyz, float* out, int32_t* ijk_out) {
    if (int rc = requireInit()) return rc;
    if (n < 0 || (n > 0 && (!xyz || !out || !ijk_out)))
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "grid_lookup_world_batch: bad arguments");
    Renderer& r = *g_r;
    if (int rc = prepareScene(r)) return rc;
    if (!sparseActive(r)) return fail(CRVPT_ERR_UNSUPPORTED, "grid_lookup_world_batch: needs the sparse grid");
    if (n == 0) return CRVPT_OK;
    DenseGrid dg;
    NanoGrid ng;
    bool sp;
    gridViews(r, dg, ng, sp);
    HIP_TRY(r.dRays.ensure(size_t(n) * 28));
    float* dXyz = r.dRays.as<float>();
    float* dOut = dXyz + 3 * size_t(n);
    int32_t* dIjk = reinterpret_cast<int32_t*>(dOut + n);
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipMemcpy(dXyz, xyz, 12 * size_t(n), hipMemcpyHostToDevice));
    HIP_TRY(launchNanoLookupWorld(ng, n, dXyz, dOut, dIjk, r.stream));
    HIP_TRY(hipStreamSynchronize(r.stream));
    HIP_TRY(hipMemcpy(out, dOut, 4 * size_t(n), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(ijk_out, dIjk, 12 * size_t(n), hipMemcpyDeviceToHost));
    return CRVPT_OK;
}

int crvpt_get_supervoxel_tables(int* supervoxel_size, int grid_size[3], float* decomposition_min_max,
                                float* rrt_mu_c_mu_r, uint64_t capacity) {
    if (int rc = requireInit()) return rc;
    Renderer& r = *g_r;
    const int prevMode = r.mode;
    r.mode = VPT_DECOMPOSITION_TRACKING;
    int rc = prepareScene(r);
    r.mode = prevMode;
    if (rc) return rc;
    const size_t n = size_t(r.refGrid[0]) * r.refGrid[1] * r.refGrid[2];
    if (supervoxel_size) *supervoxel_size = r.refSvSize;
    if (grid_size)
        for (int a = 0; a < 3; a++) grid_size[a] = r.refGrid[a];
    if (decomposition_min_max || rrt_mu_c_mu_r) {
        if (capacity < 2 * n) return fail(CRVPT_ERR_INVALID_ARGUMENT, "get_supervoxel_tables: output too small");
        HIP_TRY(hipStreamSynchronize(r.stream));
        if (decomposition_min_max)
            HIP_TRY(hipMemcpy(decomposition_min_max, r.dDecomp.ptr, 2 * n * sizeof(float), hipMemcpyDeviceToHost));
        if (rrt_mu_c_mu_r) HIP_TRY(hipMemcpy(rrt_mu_c_mu_r, r.dRrt.ptr, 2 * n * sizeof(float), hipMemcpyDeviceToHost));
    }
    return CRVPT_OK;
}

int crvpt_set_use_majorant_grid(int use) {
    if (int rc = requireInit()) return rc;
    g_r->useMajorantGrid = use != 0;
    return CRVPT_OK;
}

int crvpt_set_launch_config(int supervoxel_size, int blocks, int steps_per_refill) {
    if (int rc = requireInit()) return rc;
    if (supervoxel_size < 0 || blocks < 0 || steps_per_refill < 0)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "set_launch_config: negative value");
    Renderer& r = *g_r;
    if (supervoxel_size != r.svSizeSetting) {
        r.svSizeSetting = supervoxel_size;
        r.tablesDirty = true;
    }
    r.blocksSetting = blocks;
    r.stepsPerRefill = steps_per_refill;
    return CRVPT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------------------
// Batched render queue (no reference counterpart): requests snapshot the renderer's current scene (volume in its
// active representation with its supervoxel tables, camera, medium, phase, lights, environment, seed offset and
// bounce limit) and are traced together by one persistent launch captured in a HIP graph.
// ------------------------------------------------------------------------------------------------------------
namespace crvpt {
namespace {

struct BatchVolume {
    DevBuf grid, majorant, rrt, decomp, blockLeaf, blockTile;
    DenseGrid dense{};
    NanoGrid nano{};
    SupervoxelGrid S{};
    ~BatchVolume() {
        grid.release();
        majorant.release();
        rrt.release();
        decomp.release();
        blockLeaf.release();
        blockTile.release();
    }
};

struct BatchEnvImage {
    DevBuf img;
    ~BatchEnvImage() { img.release(); }
};

// Identity of the renderer's active volume and tables: a request reuses the previous request's snapshot when the
// identity is unchanged.
struct VolumeKey {
    uint64_t version = ~0ull;
    int sparse = -1, svSize = 0, refSvSize = 0, interp = -1, needRef = 0;
    double extinction = -1.0;
    bool operator==(const VolumeKey& o) const {
        return version == o.version && sparse == o.sparse && svSize == o.svSize && refSvSize == o.refSvSize &&
               interp == o.interp && needRef == o.needRef && extinction == o.extinction;
    }
};

struct Batch {
    std::vector<std::unique_ptr<BatchVolume>> volumes;
    std::vector<std::unique_ptr<BatchEnvImage>> envs;
    std::vector<BatchRequest> reqs;
    int mode = -1;
    VolumeKey lastVolume;
    uint64_t lastEnvVersion = ~0ull;
    BatchLaunch cfg{0, 256, 16, 16};
    bool blocksAuto = true;  // numBlocks follows the occupancy of the block size
    DevBuf dReqs, dUnitBegins, dPixBegins, dRecords, dCounter, dFrame;
    uint32_t totalUnits = 0;
    uint64_t totalPix = 0, totalPaths = 0;
    bool tableDirty = true;
    hipStream_t captureStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;

    void dropGraph() {
        if (exec) (void)hipDeviceSynchronize();  // a launch of the graph may still run
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~Batch() {
        dropGraph();
        for (DevBuf* b : {&dReqs, &dUnitBegins, &dPixBegins, &dRecords, &dCounter, &dFrame}) b->release();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (captureStream) (void)hipStreamDestroy(captureStream);
    }
};

Batch* g_batch = nullptr;

void releaseBatch() {
    if (!g_batch) return;
    (void)hipDeviceSynchronize();  // kernels on any stream may still read the snapshots
    delete g_batch;
    g_batch = nullptr;
}

int requireBatch() {
    if (int rc = requireInit()) return rc;
    if (!g_batch) {
        Batch* b = new Batch();
        if (hipStreamCreateWithFlags(&b->captureStream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
            delete b;
            return fail(CRVPT_ERR_HIP, "crvpt: cannot create the batch stream or events");
        }
        g_batch = b;
    }
    return CRVPT_OK;
}

hipError_t copyDevice(DevBuf& dst, const void* src, size_t bytes, hipStream_t s) {
    hipError_t e = dst.ensure(bytes);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(dst.ptr, src, bytes, hipMemcpyDeviceToDevice, s);
}

// Snapshot of the renderer's active volume and the tables of its mode (prepareScene has run).
int snapshotVolume(Renderer& r, BatchVolume& v) {
    const bool sp = sparseActive(r);
    DenseGrid dg;
    NanoGrid ng;
    bool spv;
    gridViews(r, dg, ng, spv);
    if (sp) {
        HIP_TRY(copyDevice(v.grid, r.dNvdb.ptr, r.nvdbHost.size(), r.stream));
        ng.buf = v.grid.as<uint8_t>();
        if (ng.blockLeaf) {
            // leaf offsets are relative to the grid buffer, so the index stays valid for the copy
            const size_t nb = size_t(ng.blockDim[0]) * ng.blockDim[1] * ng.blockDim[2];
            HIP_TRY(copyDevice(v.blockLeaf, ng.blockLeaf, nb * sizeof(uint32_t), r.stream));
            HIP_TRY(copyDevice(v.blockTile, ng.blockTile, nb * sizeof(float), r.stream));
            ng.blockLeaf = v.blockLeaf.as<uint32_t>();
            ng.blockTile = v.blockTile.as<float>();
        }
        v.nano = ng;
    } else {
        const size_t n = size_t(r.dims[0]) * r.dims[1] * r.dims[2];
        HIP_TRY(copyDevice(v.grid, r.dNorm.ptr, n * sizeof(float), r.stream));
        dg.values = v.grid.as<float>();
        v.dense = dg;
    }
    const size_t nm = size_t(r.svGrid[0]) * r.svGrid[1] * r.svGrid[2];
    HIP_TRY(copyDevice(v.majorant, r.dMajorant.ptr, nm * sizeof(float), r.stream));
    v.S.majorant = v.majorant.as<float>();
    v.S.size = r.svSize;
    v.S.refSize = r.refSvSize;
    for (int a = 0; a < 3; a++) {
        v.S.grid[a] = r.svGrid[a];
        v.S.refGrid[a] = r.refGrid[a];
    }
    if (r.mode == VPT_RESIDUAL_RATIO_TRACKING || r.mode == VPT_DECOMPOSITION_TRACKING) {
        const size_t nr = size_t(r.refGrid[0]) * r.refGrid[1] * r.refGrid[2];
        HIP_TRY(copyDevice(v.rrt, r.dRrt.ptr, 2 * nr * sizeof(float), r.stream));
        HIP_TRY(copyDevice(v.decomp, r.dDecomp.ptr, 2 * nr * sizeof(float), r.stream));
        v.S.rrt = v.rrt.as<float>();
        v.S.decomp = v.decomp.as<float>();
    }
    HIP_TRY(hipStreamSynchronize(r.stream));
    return CRVPT_OK;
}

// Unit and pixel ranges of the requests for G samples per unit, uploaded with the request table.
int uploadBatchTable(Batch& b, uint32_t samplesPerUnit) {
    const size_t n = b.reqs.size();
    std::vector<uint64_t> unitBegins(n), pixBegins(n);
    uint64_t units = 0, pix = 0, paths = 0;
    for (size_t i = 0; i < n; i++) {
        BatchRequest& q = b.reqs[i];
        const uint64_t npix = uint64_t(q.P.width) * uint64_t(q.P.height);
        q.groups = (q.spp + samplesPerUnit - 1) / samplesPerUnit;
        q.unitBegin = units;
        q.pixBegin = pix;
        unitBegins[i] = units;
        pixBegins[i] = pix;
        units += npix * q.groups;
        pix += npix;
        paths += npix * q.spp;
    }
    if (units >= (1ull << 32))
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "batch: more than 2^32 units; raise the samples per unit or split the batch");
    HIP_TRY(b.dReqs.ensure(n * sizeof(BatchRequest)));
    // each request's constants point at their own record in the device table (SceneParams::mem)
    for (size_t i = 0; i < n; i++) b.reqs[i].P.mem = &b.dReqs.as<BatchRequest>()[i].P;
    HIP_TRY(b.dUnitBegins.ensure(n * sizeof(uint64_t)));
    HIP_TRY(b.dPixBegins.ensure(n * sizeof(uint64_t)));
    HIP_TRY(b.dRecords.ensure(std::max<uint64_t>(units, 1) * sizeof(float4)));
    HIP_TRY(b.dCounter.ensure(64));
    HIP_TRY(b.dFrame.ensure(64));
    HIP_TRY(hipMemcpy(b.dReqs.ptr, b.reqs.data(), n * sizeof(BatchRequest), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b.dUnitBegins.ptr, unitBegins.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b.dPixBegins.ptr, pixBegins.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
    b.totalUnits = uint32_t(units);
    b.totalPix = pix;
    b.totalPaths = paths;
    return CRVPT_OK;
}

BatchLaunch resolvedLaunch(const Batch& b, const Renderer& r, BatchLaunch c) {
    if (c.numBlocks == 0) c.numBlocks = uint32_t(r.cuCount * batchOccupancyBlocksPerCU(b.mode, c.blockSize));
    return c;
}

hipError_t enqueueBatch(Batch& b, const BatchLaunch& c, hipStream_t s) {
    return launchBatch(b.mode, b.dReqs.as<BatchRequest>(), b.dUnitBegins.as<uint64_t>(), b.dPixBegins.as<uint64_t>(),
                       int(b.reqs.size()), b.totalUnits, b.totalPix, c, b.dFrame.as<uint32_t>(),
                       b.dRecords.as<float4>(), b.dCounter.as<uint32_t>(), s);
}

bool validBatchLaunch(const BatchLaunch& c) {
    return c.blockSize >= 64 && c.blockSize <= 256 && c.blockSize % 64 == 0 && c.stepsPerRefill > 0 &&
           c.samplesPerUnit > 0;
}

}  // namespace
}  // namespace crvpt

extern "C" {

int crvpt_batch_clear(void) {
    if (int rc = requireInit()) return rc;
    releaseBatch();
    return CRVPT_OK;
}

int crvpt_batch_add_request(float* out, int channels, int height, int width, int spp, int32_t* request_id) {
    if (int rc = requireBatch()) return rc;
    if (channels != 3 && channels != 4)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_add_request: the number of channels must be 3 or 4");
    if (!out || height <= 0 || width <= 0 || spp <= 0)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_add_request: empty image, no samples or null output");
    Renderer& r = *g_r;
    Batch& b = *g_batch;
    if (b.mode >= 0 && b.mode != r.mode)
        return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_add_request: all requests of a batch use one estimator");
    if (r.outputForeground)
        return fail(CRVPT_ERR_UNSUPPORTED, "batch_add_request: the foreground map is written by render_frame only");
    if (int rc = prepareScene(r)) return rc;
    HIP_TRY(hipStreamSynchronize(r.stream));
    VolumeKey key;
    key.version = r.volumeVersion;
    key.sparse = sparseActive(r) ? 1 : 0;
    key.svSize = r.svSize;
    key.refSvSize = r.refSvSize;
    key.needRef = (r.mode == VPT_RESIDUAL_RATIO_TRACKING || r.mode == VPT_DECOMPOSITION_TRACKING) ? 1 : 0;
    key.interp = key.needRef ? r.refTablesInterp : -1;
    key.extinction = key.needRef ? r.refTablesExtinction : -1.0;
    if (b.volumes.empty() || !(key == b.lastVolume)) {
        std::unique_ptr<BatchVolume> v(new BatchVolume());
        if (int rc = snapshotVolume(r, *v)) return rc;
        b.volumes.push_back(std::move(v));
        b.lastVolume = key;
    }
    const BatchVolume& vol = *b.volumes.back();
    BatchRequest q;
    std::memset(&q, 0, sizeof(q));
    SupervoxelGrid S;
    fillSceneParams(r, width, height, q.P, S);
    if (q.P.envKind == ENV_IMAGE) {
        if (b.envs.empty() || b.lastEnvVersion != r.envVersion) {
            std::unique_ptr<BatchEnvImage> e(new BatchEnvImage());
            HIP_TRY(copyDevice(e->img, r.dEnv.ptr, r.envImage.size() * sizeof(float), r.stream));
            HIP_TRY(hipStreamSynchronize(r.stream));
            b.envs.push_back(std::move(e));
            b.lastEnvVersion = r.envVersion;
        }
        q.P.envImage = b.envs.back()->img.as<float>();
    }
    q.P.frameBase = 0;
    q.P.numFrames = uint32_t(spp);
    q.P.numUnits = uint64_t(width) * uint64_t(height) * uint64_t(spp);
    q.S = vol.S;
    setWalkerConstants(q.P, q.S);
    q.dense = vol.dense;
    q.nano = vol.nano;
    q.spp = uint32_t(spp);
    q.out = out;
    q.channels = channels;
    b.reqs.push_back(q);
    b.mode = r.mode;
    b.tableDirty = true;
    b.dropGraph();
    if (request_id) *request_id = int32_t(b.reqs.size() - 1);
    return CRVPT_OK;
}

int crvpt_batch_set_launch_config(int blocks, int block_size, int steps_per_refill, int samples_per_unit) {
    if (int rc = requireBatch()) return rc;
    BatchLaunch c{uint32_t(std::max(blocks, 0)), block_size, steps_per_refill, uint32_t(std::max(samples_per_unit, 0))};
    if (blocks < 0 || !validBatchLaunch(c))
        return fail(CRVPT_ERR_INVALID_ARGUMENT,
                    "batch_set_launch_config: block size 64..256 in steps of 64, positive steps and samples per unit");
    Batch& b = *g_batch;
    b.cfg = c;
    b.blocksAuto = blocks == 0;
    b.tableDirty = true;
    b.dropGraph();
    return CRVPT_OK;
}

int crvpt_batch_get_launch_config(int32_t config[4]) {
    if (int rc = requireBatch()) return rc;
    if (!config) return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_get_launch_config: null output");
    Batch& b = *g_batch;
    const BatchLaunch c = b.mode >= 0 ? resolvedLaunch(b, *g_r, b.blocksAuto ? BatchLaunch{0, b.cfg.blockSize,
                                                                                          b.cfg.stepsPerRefill,
                                                                                          b.cfg.samplesPerUnit}
                                                                         : b.cfg)
                                      : b.cfg;
    config[0] = int32_t(c.numBlocks);
    config[1] = c.blockSize;
    config[2] = c.stepsPerRefill;
    config[3] = int32_t(c.samplesPerUnit);
    return CRVPT_OK;
}

int crvpt_batch_render(uint32_t frame_offset, void* stream) {
    if (int rc = requireBatch()) return rc;
    Batch& b = *g_batch;
    if (b.reqs.empty()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_render: the batch holds no request");
    Renderer& r = *g_r;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    BatchLaunch c = b.cfg;
    if (b.blocksAuto) c.numBlocks = 0;
    c = resolvedLaunch(b, r, c);
    if (b.tableDirty) {
        HIP_TRY(hipDeviceSynchronize());  // a previous launch may still read the table
        if (int rc = uploadBatchTable(b, c.samplesPerUnit)) return rc;
        b.tableDirty = false;
        b.dropGraph();
    }
    if (!b.exec) {
        // zero the counter, the persistent pool, the resolve: one graph, replayed while the batch is unchanged
        HIP_TRY(hipStreamBeginCapture(b.captureStream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = enqueueBatch(b, c, b.captureStream);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b.captureStream, &g);
        if (le != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(CRVPT_ERR_HIP, std::string("batch_render: launch: ") + hipGetErrorString(le));
        }
        HIP_TRY(ce);
        b.graph = g;
        HIP_TRY(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    }
    HIP_TRY(launchSetU32(b.dFrame.as<uint32_t>(), frame_offset, s));
    HIP_TRY(hipEventRecord(b.ev0, s));
    HIP_TRY(hipGraphLaunch(b.exec, s));
    HIP_TRY(hipEventRecord(b.ev1, s));
    b.timed = true;
    return CRVPT_OK;
}

int crvpt_batch_get_stats(double* kernel_ms, uint64_t* num_paths, int32_t* num_requests) {
    if (int rc = requireBatch()) return rc;
    Batch& b = *g_batch;
    if (kernel_ms) {
        float ms = 0.0f;
        if (b.timed) {
            HIP_TRY(hipEventSynchronize(b.ev1));
            HIP_TRY(hipEventElapsedTime(&ms, b.ev0, b.ev1));
        }
        *kernel_ms = ms;
    }
    if (num_paths) {
        uint64_t p = 0;
        for (const BatchRequest& q : b.reqs) p += uint64_t(q.P.width) * uint64_t(q.P.height) * q.spp;
        *num_paths = p;
    }
    if (num_requests) *num_requests = int32_t(b.reqs.size());
    return CRVPT_OK;
}

int crvpt_batch_autotune(int spp_divisor, int32_t config[4]) {
    if (int rc = requireBatch()) return rc;
    Batch& b = *g_batch;
    if (b.reqs.empty()) return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_autotune: the batch holds no request");
    if (spp_divisor < 1) return fail(CRVPT_ERR_INVALID_ARGUMENT, "batch_autotune: spp_divisor must be >= 1");
    Renderer& r = *g_r;
    HIP_TRY(hipDeviceSynchronize());
    b.dropGraph();
    // Tuning batch: every request with spp / spp_divisor samples (at least 1) into its own output.
    std::vector<uint32_t> spp(b.reqs.size());
    for (size_t i = 0; i < b.reqs.size(); i++) {
        spp[i] = b.reqs[i].spp;
        b.reqs[i].spp = std::max<uint32_t>(1u, spp[i] / uint32_t(spp_divisor));
    }
    hipStream_t s = b.captureStream;
    auto restore = [&]() {
        for (size_t i = 0; i < b.reqs.size(); i++) b.reqs[i].spp = spp[i];
        b.tableDirty = true;
    };
    auto timeConfig = [&](BatchLaunch c, double& best) -> int {
        c = resolvedLaunch(b, r, c);
        if (int rc = uploadBatchTable(b, c.samplesPerUnit)) return rc;
        HIP_TRY(launchSetU32(b.dFrame.as<uint32_t>(), 0u, s));
        float msMin = 3.0e38f;
        for (int rep = 0; rep < 2; rep++) {
            HIP_TRY(hipEventRecord(b.ev0, s));
            HIP_TRY(enqueueBatch(b, c, s));
            HIP_TRY(hipEventRecord(b.ev1, s));
            HIP_TRY(hipEventSynchronize(b.ev1));
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, b.ev0, b.ev1));
            msMin = std::min(msMin, ms);
        }
        best = msMin;
        return CRVPT_OK;
    };
    // Coordinate descent from the current configuration: compaction interval, paths per unit, block size, then the
    // number of persistent workgroups as a fraction of the occupancy limit.
    BatchLaunch cur = b.cfg;
    cur.numBlocks = 0;
    double curMs = 0.0;
    if (int rc = timeConfig(cur, curMs)) {
        restore();
        return rc;
    }
    auto tryConfig = [&](BatchLaunch c) -> int {
        double ms = 0.0;
        if (int rc = timeConfig(c, ms)) return rc;
        if (ms < curMs) {
            curMs = ms;
            cur = c;
        }
        return CRVPT_OK;
    };
    int rc = CRVPT_OK;
    for (int steps : {2, 4, 8, 16, 24, 32})
        if (!rc && steps != cur.stepsPerRefill) rc = tryConfig(BatchLaunch{cur.numBlocks, cur.blockSize, steps,
                                                                           cur.samplesPerUnit});
    // Samples per unit whose full-size batch keeps at most 2^28 unit records (4 GiB).
    auto fullUnits = [&](uint32_t g) {
        uint64_t u = 0;
        for (size_t i = 0; i < b.reqs.size(); i++)
            u += uint64_t(b.reqs[i].P.width) * uint64_t(b.reqs[i].P.height) * ((spp[i] + g - 1) / g);
        return u;
    };
    for (uint32_t g : {1u, 4u, 16u, 64u})
        if (!rc && g != cur.samplesPerUnit && fullUnits(g) <= (1ull << 28))
            rc = tryConfig(BatchLaunch{cur.numBlocks, cur.blockSize, cur.stepsPerRefill, g});
    for (int bs : {64, 128, 256})
        if (!rc && bs != cur.blockSize) rc = tryConfig(BatchLaunch{0, bs, cur.stepsPerRefill, cur.samplesPerUnit});
    if (!rc) {
        const uint32_t full = resolvedLaunch(b, r, BatchLaunch{0, cur.blockSize, 1, 1}).numBlocks;
        for (int num : {1, 2})
            if (!rc) rc = tryConfig(BatchLaunch{std::max(1u, full * uint32_t(num) / 3u), cur.blockSize,
                                                cur.stepsPerRefill, cur.samplesPerUnit});
    }
    restore();
    if (rc) return rc;
    b.cfg = cur;
    b.blocksAuto = cur.numBlocks == 0;
    if (config) {
        const BatchLaunch c = resolvedLaunch(b, r, cur);
        config[0] = int32_t(c.numBlocks);
        config[1] = c.blockSize;
        config[2] = c.stepsPerRefill;
        config[3] = int32_t(c.samplesPerUnit);
    }
    return CRVPT_OK;
}

}  // extern "C"
