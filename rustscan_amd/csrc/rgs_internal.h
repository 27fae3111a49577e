// Internal declarations shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rgs_math.h"

namespace rgs {

constexpr int kMaxViews = 64;
constexpr int kMaxShDegree = 4;

// Per-view constants.  Lives in device memory so a captured graph can be
// replayed with new cameras after a plain host->device copy.
struct ViewDev {
  float vm[16];  // world->camera, column-major
  float fx, fy, cx, cy;
  uint32_t W, H;
  float campos[3];
  uint32_t tiles_x, tiles_y;
  uint32_t tile_base;   // global id of this view's first tile
  uint32_t pixel_base;  // offset of this view's first pixel in the image arena
  // EWA clip limits of calc_cam_j (helpers.wgsl:144-145)
  float lim_px, lim_py, lim_nx, lim_ny;
  float bg[3];
  float loss_scale;  // loss weight / (3*W*H)
  // Per-view image pointers (device).  img: [H][W][4] output of the forward;
  // target: [H][W][3] (fused L1); vout: [H][W][4] image gradient (unfused).
  float* img;
  const float* target;
  const float* vout;
  // combined loss: [9][H][W] partial derivatives of the SSIM map (channel c,
  // map m at plane 3c + m), written by the loss forward, read by its backward
  float* lgrad;
  float loss_weight;
};

// Camera of the evaluation renderer (TiledRenderer): rotation rows and
// translation of the world->camera transform, intrinsics, the covariance
// blur floor and the tile grid (16x16 pixel tiles).
struct EvalCam {
  float r[9];  // row-major rows of the rotation
  float t[3];
  float pos[3];  // camera centre in world space
  float fx, fy, cx, cy;
  float blur;
  uint32_t W, H, ntx, nty;  // image size; reference tiles across and down
  uint32_t nbx, nby;        // binning tiles (2x2 reference tiles) across and down
};
constexpr uint32_t kEvalTile = 16;
// The evaluation renderer's lists are binned on tiles of 2x2 reference tiles
// (32 pixels); a 16-pixel tile keeps the entries whose box holds it.
constexpr uint32_t kEvalBin = 2 * kEvalTile;
constexpr uint32_t kEvalMaxSide = 16384;  // image side limit (tile ranges pack in 16 bits)
constexpr int kEvalSums = 8;               // double sums per block of the metrics kernel

// Projected splat instance, 64 bytes, the unit every later stage gathers by
// instance.
//   p0 = (x, y, a/2, b)        conic = [[a, b], [b, c]]
//   p1 = (c/2, opacity, r, g)  colour before the max(., 0) clamp
//   p2 = (blue, threshold, bbox_lo, bbox_hi)
//   p3 = (mask_lo, mask_hi, -, -)
// threshold = log(255 * opacity) (the power threshold of the tile test); the
// tile bbox is stored as two packed uint16 pairs (x, y) reinterpreted as floats.
// mask: for bboxes of at most 64 tiles, bit k (row-major within the bbox) is
// set when tile k passes the tile test; the emission pass reuses it.
struct ProjSplat {
  float4 p0, p1, p2, p3;
};

// Counters block in device memory (one per workspace).
enum Counter : int {
  kCntVisible = 0,      // visible instances over all views
  kCntIsect = 1,        // intersections emitted (may exceed capacity)
  kCntOverflow = 2,     // nonzero if any capacity was exceeded this step
  kCntQueueFwd = 3,     // persistent raster work counters
  kCntQueueBwd = 4,
  kCntStepValid = 5,    // 0 when the step must not update parameters
  kCntOverflowSticky = 6,  // set by a step that overflowed; cleared by the host once it has re-sized
  kCntIsectMax = 7,     // max intersections of any forward in this step
  kCntSkipped = 8,      // steps whose update was skipped
  kCntIsectMaxSticky = 9,  // largest kCntIsectMax of the overflowed steps since the last clear
  kCntNum = 16,
};

// Float scalars block
enum FScalar : int {
  kFLossBase = 0,      // [kMaxViews] per-view loss sums
  kFBc1 = kMaxViews,   // Adam bias corrections (1 - b1^t, 1 - b2^t) of each group g at kFBc1 + 2g, + 2g + 1
  kFNum = kMaxViews + 8,
};

// Adam parameter groups of AdamScaled (optimizer.rs:46-51), each with its own
// step counter: 0 = transforms (planes 0-9), 1 = SH coefficients (planes 11
// and up), 2 = opacity logits (plane 10).
constexpr int kAdamGroups = 3;
__host__ __device__ inline int adam_group_of_plane(int plane) { return plane < 10 ? 0 : (plane == 10 ? 2 : 1); }

constexpr int kAcc = 16;

// Per-step parameters of the topology kernels (k_topology.hip), host-computed.
struct TopoParams {
  uint32_t n, npad, nsh;  // Gaussians, plane stride, SH planes (3K)
  uint32_t iteration;
  uint32_t densify, prune, reset_opacity;
  uint32_t split_score_mode, prune_mode, prune_min_age, prune_invisible_epochs;
  float growth_grad_threshold, split_grad_threshold, depth_scale_gamma;
  float prune_opacity_threshold, prune_visibility_threshold, prune_high_opacity_threshold;
  uint64_t seed;
};

// Parameters of the rebuild kernel k_topo_apply.
struct TopoApply {
  uint32_t nn, old_npad, new_npad, nplanes, iteration;
  uint32_t decay, reset_opacity;
  float scale_decay, opacity_decay, train_t;
};

// counters of the topology analysis and sampling kernels
enum TopoCount : int {
  kTopoPrune = 0,
  kTopoGrowth = 1,
  kTopoClone = 2,
  kTopoSplit = 3,
  kTopoReplacementSources = 4,
  kTopoEligible = 5,  // eligible sources of the last sampling pass
  kTopoCountNum = 8,
};


// Coefficients of the combined loss kernels (k_loss.hip), passed by value.
struct LossCoef {
  float l1, ssim, grad;  // term weights
  int residual;          // 0: L1, 1: saturating robust, 2: soft outlier
  int degenerate_nan;    // the loss of an image one pixel wide or high is NaN (the reference's
                         // image-gradient term averages an empty difference there)
  float delta;           // robust delta, or outlier threshold
  float floor_w;         // outlier gradient floor
  float c1, c2;          // SSIM constants (k1 * range)^2, (k2 * range)^2
  float k[11];           // gaussian_kernel_1d
  float kcum[11];        // k[0] + ... + k[m]
  float kdown[11];       // k[m] + ... + k[10]
};

// Arguments of the compositing kernels.  Work items are global tile ids
// [tile_first, tile_first + tile_count).
struct RasterArgs {
  const ProjSplat* proj;
  const uint32_t* isect_vals;  // Gaussian index of each sorted intersection (instance = view * n + index)
  const uint32_t* ranges;      // [2 * tiles] (start, end) into isect_vals
  const ViewDev* views;
  int v0, nv;
  uint32_t tile_first, tile_count;
  uint32_t* queue;  // persistent work counter, or null for one workgroup per tile
  int cull;         // nonzero: skip splats per 8x8 block that provably miss it
  uint32_t n;       // Gaussians per view; instance = view * n + gid
  float* visible;   // [n] OR over views, or null
  uint8_t* vis_inst;  // [instances] 1 where a fragment blended the instance, or null
  float* loss;      // per-view sums of |pred - target|, or null
  float* acc;       // backward: [instances][16]
  // optional fragment counters: forward [0] evaluated, [1] blended;
  // backward [2] evaluated, [3] contributing; per-wave splat iterations:
  // [4] forward, [5] forward with a lane needing exp, [6] backward,
  // [7] backward with a contributing lane
  unsigned long long* frag;
};  // floats per instance in the raster-gradient accumulator

constexpr int kFragCounters = 8;

inline int sh_coeffs(int degree) { return (degree + 1) * (degree + 1); }

}  // namespace rgs
