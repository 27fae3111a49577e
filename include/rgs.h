/* rgs.h - C ABI of the MI355X Gaussian-splatting render/train path.
 *
 * Drop-in boundary for RustGS's rasterizer/trainer seams (reference paths are
 * relative to the RustScan repository).  Every entry point names the
 * reference interface it replaces.  All functions return RGS_OK (0) or a
 * positive status; they never abort.  rgs_last_error() gives the message.
 *
 * Conventions
 *  - One context = one HIP stream = one host thread at a time.  Entry points
 *    are asynchronous on that stream unless stated otherwise.
 *  - Array arguments are plain pointers; unless stated they may be host or
 *    device memory (copies use unified addressing).  Arguments documented as
 *    "device" must be device memory.
 *  - Layouts at the boundary are the reference's
 *    (training/engine/splats.rs:7-18):
 *      transforms [N][10] = mean xyz, quaternion wxyz, log-scale xyz
 *      sh         [N][K][3], K = (degree+1)^2, coefficient-major, RGB interleaved
 *      raw_opac   [N]        opacity logits
 *    Images are [H][W][4] RGBA float (forward/rasterize.rs:84-85), targets
 *    [H][W][3] RGB float (engine/trainer.rs:265-268).
 *  - Internally the splat store is SoA planes in HBM; the reference AoS layout
 *    exists only at import/export.
 */
#ifndef RGS_H
#define RGS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RGS_OK = 0,
  RGS_ERR_INVALID = 1,     /* bad argument */
  RGS_ERR_CAPACITY = 2,    /* a capacity was exceeded; retry after it grows */
  RGS_ERR_HIP = 3,         /* HIP runtime error */
  RGS_ERR_STATE = 4,       /* call out of order (e.g. backward without forward) */
  RGS_ERR_UNSUPPORTED = 5  /* configuration not built (e.g. SH degree > 4) */
};

typedef struct rgs_ctx rgs_ctx;
typedef struct rgs_splats rgs_splats;

/* Camera of one view: core/camera.rs:30-37 (view_matrix, position) plus the
 * intrinsics of rustscan-types/src/camera.rs:15-29, as packed into
 * ProjectUniforms (forward/mod.rs:31-68). */
typedef struct rgs_camera {
  float viewmat[16]; /* world->camera, column-major (glam Mat4::to_cols_array) */
  float fx, fy, cx, cy;
  uint32_t width, height;
  float cam_pos[3]; /* camera centre in world space */
} rgs_camera;

/* AdamScaled (engine/optimizer.rs:7-24, 122-162) with the per-column
 * learning-rate scaling tensors of trainer.rs:127-151 folded in:
 * lr_transforms[10] = lr * transform scaling, lr_sh[25] = lr * SH scaling per
 * coefficient (first K used), lr_opacity = lr * opacity scaling.
 * weight_decay: AdamScaledConfig::weight_decay (gradient += weight_decay *
 * parameter before the moments, optimizer.rs:128-130; 0 in RustGS). */
typedef struct rgs_adam_cfg {
  float beta1, beta2, eps;
  float lr_transforms[10];
  float lr_sh[25];
  float lr_opacity;
  float weight_decay;
} rgs_adam_cfg;

enum { RGS_MODE_PER_VIEW = 0, RGS_MODE_BATCHED = 1 };
enum { RGS_GRAPH_OFF = 0, RGS_GRAPH_ON = 1 };

/* One training step over n_views views (WgpuTrainer::train_step,
 * engine/trainer.rs:253-440, generalised to several views per step):
 * forward, L1 loss with weight loss_weight per view (loss.rs:144-196,
 * mean over H*W*3), backward summed over views, one Adam update. */
/* The training loss (combined_loss_with_kernel, engine/loss.rs:112-142, with
 * TrainingLossConfig, config.rs:566-608).  combined = 0: plain L1 (the
 * configurations' loss; fused into the compositing kernels).  combined = 1:
 * l1_weight * residual (L1; saturating when robust_delta > 0; soft outlier
 * weighting when outlier_threshold > 0 and outlier_weight < 1) + ssim_weight
 * * (1 - mean SSIM, 11-tap Gaussian window, sigma 1.5) + gradient_weight *
 * image-gradient difference.  The dynamic residual mask (loss.rs:177-211) is
 * not supported.  RustGS's defaults: combined, 0.8, 0.2, 0, 0, 0, 1. */
typedef struct rgs_loss_cfg {
  int combined;
  float l1_weight, ssim_weight, gradient_weight;
  float robust_delta, outlier_threshold, outlier_weight;
} rgs_loss_cfg;

typedef struct rgs_step_cfg {
  int mode;          /* RGS_MODE_PER_VIEW: views one after another, grads
                        accumulated; RGS_MODE_BATCHED: all views per launch
                        (split into consecutive groups of views when one
                        launch would exceed 1.6e9 intersections) */
  int graph;         /* RGS_GRAPH_ON: capture the step as a HIP graph on first
                        use and replay it while shapes and capacities hold */
  float cov_blur;    /* ProjectUniforms::cov_blur (clamped at 0) */
  float background[3];
  float loss_weight; /* L1 weight per view */
  int check_loss;    /* skip the update when the loss is not finite
                        (trainer.rs:297-302) */
  rgs_adam_cfg adam;
  int densify;       /* RGS_DENSIFY_* flags: accumulate the densification
                        statistics of each view (accumulate_gradients,
                        trainer.rs:651-728; see rgs_densify_export) */
  rgs_loss_cfg loss; /* loss of each view, times loss_weight */
} rgs_step_cfg;

enum {
  RGS_DENSIFY_ACCUMULATE = 1,        /* update the accumulators each step (per view) */
  RGS_DENSIFY_ACTUAL_VISIBILITY = 2, /* visible_observations += actual visibility
                                        (use_actual_visibility), else += 1 */
  RGS_DENSIFY_COLLECT_VISIBILITY = 4 /* actual_visible_observations += actual visibility
                                        (collect_actual_visibility_diagnostics) */
};

typedef struct rgs_raster_cfg {
  uint32_t tile_width;  /* 8, 16 or 32 (reference: 16, forward/mod.rs:25) */
  uint32_t batch_depth; /* LDS splat batch: 128, 256 or 512 (reference: 256) */
  uint32_t persistent;  /* 1: persistent tile-queue launch, 0: one workgroup per tile */
} rgs_raster_cfg;

typedef struct rgs_counts {
  int64_t num_visible;       /* V: Gaussians passing the projection cull */
  int64_t num_intersections; /* M: (tile, Gaussian) pairs */
  int64_t num_tiles;
} rgs_counts;

/* ---- context (replaces GsBackendBase/GsDevice, engine/backend.rs:9-18) ---- */
int rgs_ctx_create(int device, rgs_ctx** out);
void rgs_ctx_destroy(rgs_ctx* ctx);
const char* rgs_last_error(const rgs_ctx* ctx);
int rgs_sync(rgs_ctx* ctx);              /* blocks until the stream is idle */
void* rgs_get_stream(rgs_ctx* ctx);      /* hipStream_t of the context */
int rgs_set_raster_config(rgs_ctx* ctx, const rgs_raster_cfg* cfg);
int rgs_get_raster_config(rgs_ctx* ctx, rgs_raster_cfg* cfg);
/* Most intersections one batched launch may hold (default 1.6e9, at most
 * 3e9); a batched step whose views exceed it is run as consecutive groups of
 * views.  No reference counterpart: RustGS renders one view per launch. */
int rgs_set_launch_limit(rgs_ctx* ctx, int64_t max_intersections);

/* ---- splat store (replaces DeviceSplats, engine/splats.rs:13-63) ---- */
int rgs_splats_create(rgs_ctx* ctx, int64_t n, int sh_degree, rgs_splats** out);
void rgs_splats_destroy(rgs_ctx* ctx, rgs_splats* s);
int64_t rgs_splats_count(const rgs_splats* s);
int rgs_splats_sh_degree(const rgs_splats* s);
/* host_splats_to_device (splats.rs:26-63); resets the Adam state */
int rgs_splats_import_aos(rgs_ctx* ctx, rgs_splats* s, const float* transforms, const float* sh,
                          const float* raw_opac);
/* device_splats_to_host (splats.rs:65-...); synchronous */
int rgs_splats_export_aos(rgs_ctx* ctx, const rgs_splats* s, float* transforms, float* sh, float* raw_opac);
/* Gradients of the last backward (or train step), reference layouts:
 * v_transforms [N][10], v_sh [N][K][3], v_raw_opac [N]; screen_stats [N][7]
 * (project_backwards.wgsl:33-36,106-108, summed over views); synchronous.
 * Any pointer may be NULL. */
int rgs_splats_export_grads(rgs_ctx* ctx, const rgs_splats* s, float* v_transforms, float* v_sh,
                            float* v_raw_opac, float* screen_stats);
/* visible[N]: 1.0 where some fragment of the last forward blended the
 * Gaussian, OR over views (rasterize.wgsl:79); synchronous */
int rgs_splats_export_visible(rgs_ctx* ctx, const rgs_splats* s, float* visible);
/* AdamScaled::reset (optimizer.rs:62-66): zero moments and step counters */
int rgs_adam_reset(rgs_ctx* ctx, rgs_splats* s);
/* Densification accumulators of the trainer (trainer.rs:66-75), updated by
 * rgs_train_step when cfg->densify has RGS_DENSIFY_ACCUMULATE, once per view
 * (a one-view step is the reference's accumulate_gradients, trainer.rs:651-728;
 * a step over several views accumulates as that many one-view steps at the
 * same parameters); skipped steps add nothing.  out: [10][N], the
 * accumulators in this order: grad_2d (mean |v_mean| over xyz), screen_grad_2d
 * (|screen v_xy|), abs_grad_2d (|sum |v_xy||), abs_pixel_grad_2d,
 * pixel_coverage, camera_depth, grad_color (mean |v_sh|), num_observations,
 * visible_observations, actual_visible_observations.  Synchronous.
 * rgs_densify_reset zeroes them (reset_accumulators, trainer.rs:861-871). */
int rgs_densify_export(rgs_ctx* ctx, const rgs_splats* s, float* out);
/* Sets the accumulators from [10][N] (same order), e.g. to resume a run. */
int rgs_densify_import(rgs_ctx* ctx, rgs_splats* s, const float* in);
int rgs_densify_reset(rgs_ctx* ctx, rgs_splats* s);
/* ---- topology step: densify / prune / opacity reset ---------------------
 * WgpuTrainer::apply_topology_mutations (engine/trainer.rs:792-859) with its
 * planning (topology/bridge.rs:25-172, topology/mod.rs:203-981,
 * topology/schedule.rs:38-133) and rebuild (topology/apply.rs:10-188) on the
 * device: candidate analysis from the densification accumulators, weighted
 * sampling of the sources to refine, survivors compacted, refined pairs
 * appended, refine decay and opacity reset, the per-Gaussian birth iteration
 * and invisible-window state remapped, the accumulators cleared, and the Adam
 * state cleared when the splats changed (trainer.rs:851-853).  The splat
 * handle keeps its identity; its count may change (rgs_splats_count).
 * Differences from the reference: its random draws (rand 0.8.5's StdRng and
 * sample_weighted) are replaced by SplitMix64 draws with the same algorithms
 * (Efraimidis-Spirakis keys, largest first; Box-Muller), and the scene centre
 * is summed in double.  Synchronous. */
typedef struct rgs_topology_cfg {
  uint32_t iterations;                  /* TrainingConfig::iterations (30000) */
  uint32_t max_initial_gaussians;       /* growth budget floor (100000) */
  uint64_t frame_shuffle_seed;          /* (0) */
  uint32_t refine_every;                /* (160) */
  uint32_t opacity_reset_interval;      /* epochs (10) */
  uint32_t opacity_reset_mode;          /* 0 decay (default), 1 reset */
  int32_t topology_freeze_after_epoch;  /* -1: none */
  int32_t growth_freeze_after_epoch;    /* -1: none */
  int32_t prune_until_epoch;            /* -1: none */
  uint32_t split_score_mode;            /* 0 baseline, 1 abs, 2 abs-pixel, 3 abs-pixel-depth */
  uint32_t growth_stop_iter;            /* (15000) */
  float growth_grad_threshold;          /* (0.00014) */
  float split_grad_threshold;           /* (0.00014) */
  float depth_scale_gamma;              /* (0.37) */
  float growth_select_fraction;         /* (0.25) */
  float opacity_decay, scale_decay;     /* refine decay (0, 0) */
  uint32_t prune_mode;                  /* 0 threshold, 1 weight (default), 2 visibility-weight */
  uint32_t prune_min_age;               /* iterations (5) */
  uint32_t prune_invisible_epochs;      /* (10) */
  uint32_t prune_visibility_dry_run;    /* diagnostics only (0) */
  float prune_opacity_threshold;        /* (1/255) */
  float prune_visibility_threshold;     /* (0.05) */
  float prune_high_opacity_threshold;   /* (0.8) */
} rgs_topology_cfg;

typedef struct rgs_topology_schedule {
  int32_t completed_epoch; /* -1: none */
  int32_t densify, prune, reset_opacity, allow_extra_growth;
} rgs_topology_schedule;

typedef struct rgs_topology_result {
  rgs_topology_schedule schedule;
  int64_t before, after, added, pruned;
  int64_t clone_candidates, split_candidates, growth_candidates, prune_candidates;
  int64_t replacement_count, extra_growth_count, selected_count;
  int32_t mutated;    /* the splats were rebuilt (plan.mutates_splats) */
  int32_t adam_reset; /* the optimizer state was cleared */
} rgs_topology_result;

/* should_apply_topology_step (schedule.rs:97-112): 1 when this iteration's
 * schedule densifies, prunes or resets opacities; rgs_topology_get_schedule
 * returns the schedule itself (schedule_topology, schedule.rs:38-95). */
int rgs_topology_should_apply(const rgs_topology_cfg* cfg, uint32_t iteration, uint32_t frame_count);
int rgs_topology_get_schedule(const rgs_topology_cfg* cfg, uint32_t iteration, uint32_t frame_count,
                              rgs_topology_schedule* out);
int rgs_topology_step(rgs_ctx* ctx, rgs_splats* s, const rgs_topology_cfg* cfg, uint32_t iteration,
                      uint32_t frame_count, rgs_topology_result* result);
/* The per-Gaussian topology state (trainer.rs:76-77, 171-172): birth
 * iteration and invisible windows, [N] each; either pointer may be NULL.
 * Synchronous.  Both start at 0 when splats are created or imported. */
int rgs_topology_export_state(rgs_ctx* ctx, const rgs_splats* s, uint32_t* birth, uint32_t* invisible_windows);
int rgs_topology_import_state(rgs_ctx* ctx, rgs_splats* s, const uint32_t* birth, const uint32_t* invisible_windows);

/* AdamState::step of the three groups (optimizer.rs:27-32, 46-51):
 * steps[0] transforms, steps[1] SH coefficients, steps[2] opacities.  Read,
 * or set (to resume from a checkpoint); synchronous. */
int rgs_adam_get_steps(rgs_ctx* ctx, const rgs_splats* s, uint32_t* steps);
int rgs_adam_set_steps(rgs_ctx* ctx, rgs_splats* s, const uint32_t* steps);
/* AdamState::moment1 (m_*) and moment2 (v_*) of the three groups
 * (optimizer.rs:27-32) in the reference layouts: [N][10], [N][K][3], [N].
 * Export: synchronous, any pointer may be NULL.  Import: all six required;
 * the step counters are left as they are. */
int rgs_adam_export_moments(rgs_ctx* ctx, const rgs_splats* s, float* m_transforms, float* m_sh, float* m_raw_opac,
                            float* v_transforms, float* v_sh, float* v_raw_opac);
int rgs_adam_import_moments(rgs_ctx* ctx, rgs_splats* s, const float* m_transforms, const float* m_sh,
                            const float* m_raw_opac, const float* v_transforms, const float* v_sh,
                            const float* v_raw_opac);

/* ---- evaluation render: the reference's CPU renderer on the device -------
 * TiledRenderer::render_camera_splats (render/tiled_renderer.rs:174-492):
 * colour from sh.rs evaluate_sh_rgb (degrees 0-4), projection without EWA
 * clip, 16x16 tile lists from compute_tile_bounds, per-tile stable depth
 * order, alpha accumulation to 0.999, background, clamp to [0, 1].
 * color [H][W][3] floats, depth [H][W] (f32::MAX where no splat reached
 * alpha 0.01); device pointers, either may be NULL.  background: 3 host
 * floats or NULL (black).  cov_blur: raster_cov_blur (negative or NaN -> 0).
 * Image sides 1..16384.  Synchronous.  Independent of the training forward
 * state. */
int rgs_tiled_render(rgs_ctx* ctx, const rgs_splats* s, const rgs_camera* cam, const float* background,
                     float cov_blur, float* color, float* depth);
/* GaussianRenderer::render_splats (render/renderer.rs:72-98): the same render
 * with colour quantised to `(clamp(v, 0, 1) * 255) as u8` ([H][W][3] bytes)
 * and depth f32::MAX mapped to 0. */
int rgs_render_splats(rgs_ctx* ctx, const rgs_splats* s, const rgs_camera* cam, const float* background,
                      float cov_blur, uint8_t* color, float* depth);

/* Per-frame evaluation numbers (evaluation/core.rs:442-449). */
typedef struct rgs_frame_metrics {
  float psnr_db;                                   /* compute_psnr_f32, core.rs:238-256 */
  float render_grad, target_grad;                  /* compute_gradient_sharpness_f32, :258-280 */
  float render_lap, target_lap;                    /* compute_laplacian_sharpness_f32, :282-306 */
  float sharpness_grad_ratio, sharpness_lap_ratio; /* sharpness_ratio, :313-319 */
} rgs_frame_metrics;
/* The metrics of a rendered image against a target, both [H][W][3] floats
 * on the device.  The per-pixel terms are the reference's float32 terms;
 * their sums are float64 (the reference sums in float32, left to right).
 * Synchronous. */
int rgs_image_metrics(rgs_ctx* ctx, const float* rendered, const float* target, uint32_t width, uint32_t height,
                      rgs_frame_metrics* out);
/* One frame of evaluate_splats (evaluation/core.rs:431-449): render with
 * SplatEvaluationRenderer (GaussianRenderer, black background, colour as
 * u8 / 255, core.rs:362-377) and the metrics against target ([H][W][3]
 * floats, device).  color: [H][W][3] bytes (device) receiving the render, or
 * NULL.  Synchronous. */
int rgs_evaluate_frame(rgs_ctx* ctx, const rgs_splats* s, const rgs_camera* cam, float cov_blur,
                       const float* target, uint8_t* color, rgs_frame_metrics* out);

/* ---- scene files: io/scene_io.rs -----------------------------------------
 * SplatMetadata (scene_io.rs:12-19). */
typedef struct rgs_splat_metadata {
  uint64_t iterations;
  float final_loss;
  uint64_t gaussian_count;
  uint32_t sh_degree;
} rgs_splat_metadata;
typedef struct rgs_scene rgs_scene; /* splats read from a file, in host memory */
/* save_splats (scene_io.rs:617-627): by extension, .ply (save_splats_ply,
 * :485-550: ASCII 3DGS layout, f_rest channel-major, floats as Rust prints
 * them) or .splat (save_splats_splat, :552-615).  HostSplats arrays (host):
 * positions [N*3], log_scales [N*3], rotations wxyz [N*4], opacity_logits [N],
 * sh_coeffs [N*K*3] (K = (sh_degree+1)^2, coefficient-major).  meta NULL:
 * zeros, count N, degree sh_degree.  ctx may be NULL (no device is used);
 * then rgs_io_last_error() holds the message. */
int rgs_save_splats(rgs_ctx* ctx, const char* path, int64_t n, int sh_degree, const float* positions,
                    const float* log_scales, const float* rotations, const float* opacity_logits,
                    const float* sh_coeffs, const rgs_splat_metadata* meta);
/* load_splats (scene_io.rs:872-878): .ply (load_splats_ply, :629-757: ASCII or
 * binary little-endian, raw-logit or probability opacity, channel-major or
 * interleaved f_rest, SH degree <= 3) or .splat (load_splats_splat,
 * :759-870).  *out is freed with rgs_scene_free.  ctx may be NULL. */
int rgs_scene_load(rgs_ctx* ctx, const char* path, rgs_scene** out);
int rgs_scene_info(const rgs_scene* scene, int64_t* n, int* sh_degree, rgs_splat_metadata* meta);
/* Copies the HostSplats arrays (layouts as in rgs_save_splats); any pointer may be NULL. */
int rgs_scene_get(const rgs_scene* scene, float* positions, float* log_scales, float* rotations,
                  float* opacity_logits, float* sh_coeffs);
void rgs_scene_free(rgs_scene* scene);
/* A device splat store from a loaded scene (host_splats_to_device,
 * engine/splats.rs:26-63); N >= 1. */
int rgs_scene_to_splats(rgs_ctx* ctx, const rgs_scene* scene, rgs_splats** out);
/* save_splats of a device store (its current parameters). */
int rgs_splats_save(rgs_ctx* ctx, const rgs_splats* s, const char* path, const rgs_splat_metadata* meta);
/* Message of the last failed scene-file call made with ctx == NULL on this thread. */
const char* rgs_io_last_error(void);

/* ---- forward: render_forward (forward/mod.rs:156-294) --------------------
 * Projection + SH (ProjectionBackend, projection.rs:44-50; ProjectVisibleBackend,
 * project_visible.rs:36-45), depth sort (SortingBackend, sorting.rs:8-61), tile
 * binning (TileMappingBackend, tile_mapping.rs:82-95), compositing
 * (RasterizeBackend, rasterize.rs:48-59) for n_views cameras in one batch.
 * out_rgba: n_views pointers (H*W*4 floats each) that receive copies of the
 * images, or NULL.  The library renders into its own buffers and keeps them,
 * with the rest of the forward state, for rgs_render_backward
 * (RenderCheckpoint, autodiff.rs:56-71).  Synchronous. */
int rgs_render_forward(rgs_ctx* ctx, rgs_splats* s, const rgs_camera* cams, int n_views, const float* background,
                       float cov_blur, float* const* out_rgba);

/* ---- backward: RenderBackward::backward (backward/autodiff.rs:85-141) ----
 * Rasterize backward (RasterizeBwdBackend, rasterize_bwd.rs:41-55) and
 * projection backward (ProjectBwdBackend, project_bwd.rs:44-51) of the last
 * rgs_render_forward.  v_out: n_views pointers to [H][W][4] image gradients.
 * Gradients are summed over views into the splat store (see
 * rgs_splats_export_grads). */
int rgs_render_backward(rgs_ctx* ctx, rgs_splats* s, const float* const* v_out);

/* ---- checkpoints: RenderCheckpoint (backward/autodiff.rs:56-71) per call --
 * The context keeps one forward state for rgs_render_backward.  A checkpoint
 * holds one more: rgs_render_forward_checkpoint renders like
 * rgs_render_forward into the checkpoint's own buffers (projected splats,
 * tile lists, images) and copies the store's parameters as the forward read
 * them; rgs_render_backward_checkpoint runs the backward of that forward,
 * with those parameters, into the store's gradients, whatever forwards,
 * backwards, steps or parameter updates came in between (the store must keep
 * its count and SH degree).  Each checkpoint holds a workspace of its own,
 * grown like the context's; destroy checkpoints before their context
 * (ctx may then be NULL).  Synchronous. */
typedef struct rgs_checkpoint rgs_checkpoint;
int rgs_checkpoint_create(rgs_ctx* ctx, rgs_checkpoint** out);
void rgs_checkpoint_destroy(rgs_ctx* ctx, rgs_checkpoint* ck);
int rgs_render_forward_checkpoint(rgs_ctx* ctx, rgs_splats* s, const rgs_camera* cams, int n_views,
                                  const float* background, float cov_blur, float* const* out_rgba,
                                  rgs_checkpoint* ck);
int rgs_render_backward_checkpoint(rgs_ctx* ctx, rgs_splats* s, rgs_checkpoint* ck, const float* const* v_out);

/* ---- L1 loss (engine/loss.rs:144-196, trainer.rs:278) ---------------------
 * pred [H][W][4], target [H][W][3] (device); v_out [H][W][4] (device, may be
 * NULL); *loss_out (host) = weight * mean |pred_rgb - target|.  Synchronous
 * when loss_out is non-NULL. */
int rgs_l1_loss_grad(rgs_ctx* ctx, const float* pred, const float* target, uint32_t width, uint32_t height,
                     float weight, float* v_out, float* loss_out);

/* ---- combined loss (combined_loss_with_kernel, loss.rs:112-142) ----------
 * pred [H][W][4], target [H][W][3], v_out [H][W][4] (device; v_out may be
 * NULL); *loss_out (host, may be NULL) = weight * loss.  v_out = weight *
 * d loss / d pred on RGB, 0 on alpha.  Synchronous. */
int rgs_image_loss_grad(rgs_ctx* ctx, const float* pred, const float* target, uint32_t width, uint32_t height,
                        float weight, const rgs_loss_cfg* cfg, float* v_out, float* loss_out);

/* ---- Adam: AdamScaled::step_device_splats (optimizer.rs:80-120) -----------
 * One update of all parameter groups from the splat store's gradients. */
int rgs_adam_step(rgs_ctx* ctx, rgs_splats* s, const rgs_adam_cfg* cfg);

/* ---- fused training step: WgpuTrainer::train_step (trainer.rs:253-440) ----
 * targets: n_views device pointers to [H][W][3].  loss_out: n_views host
 * floats or NULL (NULL keeps the step free of host synchronisation). */
int rgs_train_step(rgs_ctx* ctx, rgs_splats* s, const rgs_camera* cams, const float* const* targets, int n_views,
                   const rgs_step_cfg* cfg, float* loss_out);

/* Picks tile width and LDS batch depth by timing rgs_train_step over the
 * candidates {8,16,32} x {128,256,512} (persistent launch) on the given views;
 * the choice is cached per (N, view sizes, SH degree) and made current. */
int rgs_autotune(rgs_ctx* ctx, rgs_splats* s, const rgs_camera* cams, const float* const* targets, int n_views,
                 const rgs_step_cfg* cfg, rgs_raster_cfg* chosen);

/* Counts of the last forward per view (n_views entries); synchronous. */
int rgs_get_counts(rgs_ctx* ctx, rgs_counts* out, int n_views);
/* Number of steps whose update was skipped (capacity overflow or non-finite
 * loss) since the context was created; synchronous.  A synchronous
 * rgs_train_step (loss_out non-null) that overflows grows its buffers and
 * redoes the step, which is then not counted. */
int64_t rgs_skipped_steps(rgs_ctx* ctx);

/* ---- GPU primitives (gpu_primitives/, device pointers) ---------------------
 * RadixSortBackend::radix_sort_by_key_u32_primitive (radix_sort.rs:48-53):
 * stable ascending sort of (key, value) pairs.
 * PrefixSumBackend::prefix_sum_u32_primitive (prefix_sum.rs:41-44):
 * inclusive scan. */
int rgs_sort_pairs_u32(rgs_ctx* ctx, const uint32_t* keys, const uint32_t* values, uint32_t* out_keys,
                       uint32_t* out_values, int64_t n);
int rgs_inclusive_scan_u32(rgs_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n);

/* rgs_sort_pairs_u32 without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP stream capture, allocation-free serving loops).  Same results as
 * rgs_sort_pairs_u32 (stable ascending sort of (key, value) pairs, device
 * pointers; the outputs may alias the inputs) for n <= RGS_SORT_NOALLOC_MAX_PAIRS;
 * larger n is rejected with RGS_ERR_INVALID.  Its scratch memory is one
 * fixed device array per GPU (18.9 MB, held until the process ends), shared
 * by every call on that GPU: calls that overlap on one GPU (several
 * contexts or threads) race.  Asynchronous on the context stream; calls no
 * hipMalloc or hipFree. */
#define RGS_SORT_NOALLOC_MAX_PAIRS 2097152
int rgs_sort_pairs_u32_noalloc(rgs_ctx* ctx, const uint32_t* keys, const uint32_t* values, uint32_t* out_keys,
                               uint32_t* out_values, int64_t n);

/* ---- intermediate state of the last forward, for parity checks ------------
 * Copied into host or device memory; synchronous.  `view` selects a view of
 * the last batch; returns the element count through *count (call with
 * dst = NULL to query). */
enum {
  RGS_DEBUG_DEPTH_ORDER = 0, /* uint32 gids of visible Gaussians, front to back */
  RGS_DEBUG_ISECT_TILES = 1, /* uint32 tile ids (view-local) of sorted intersections */
  RGS_DEBUG_ISECT_GIDS = 2,  /* uint32 gids of sorted intersections */
  RGS_DEBUG_TILE_RANGES = 3, /* uint32 [2*T] (start, end) per tile, view-local offsets */
  RGS_DEBUG_PROJECTED = 4,   /* float [N][9] x,y,a,b,c,r,g,b,opacity by gid (0 if culled) */
  RGS_DEBUG_TILE_COUNTS = 5, /* uint32 [N] tiles per Gaussian */
  RGS_DEBUG_IMAGE = 6,       /* float [H][W][4] */
  RGS_DEBUG_RASTER_GRADS = 7 /* float [N][16] rasterizer accumulators of the last backward
                                (v_xy, v_conic, v_rgb, v_opacity, v_refine, |v_xy|, coverage) */
};
int rgs_debug_get(rgs_ctx* ctx, int what, int view, void* dst, int64_t* count);
/* flags bit 0: keep a copy of the rasterizer accumulators of each backward
 * (RGS_DEBUG_RASTER_GRADS); costs one extra copy per backward.
 * flags bit 1: count compositing fragments in eager launches.
 * flags bit 2: plain sorts: the depth sort gathers the tile counts and the
 * tile sort moves (key, value) pairs in every pass, instead of carrying the
 * counts in the sort values and packing the middle pass (same results).
 * flags bit 3: no block culling in the compositing kernels: every wave
 * walks every splat of its tile (same results). */
int rgs_set_debug(rgs_ctx* ctx, int flags);
/* Fragment counts since the last reset (bit 1 of rgs_set_debug), 8 entries:
 * [0] forward (pixel, splat) pairs evaluated, [1] forward pairs blended,
 * [2] backward pairs evaluated, [3] backward pairs contributing a gradient,
 * [4] forward (wave, splat) iterations, [5] of those where some lane
 * evaluates exp, [6] backward (wave, splat) iterations, [7] of those with a
 * contributing lane.  Synchronous. */
int rgs_get_fragment_counts(rgs_ctx* ctx, int64_t* out8, int reset);

/* ---- per-stage timing (HIP events on the context stream) ----------------
 * When enabled, eager (non-graph) launches record events around each stage;
 * rgs_get_timing synchronises and returns the accumulated milliseconds and
 * launch counts per stage (arrays of RGS_STAGE_COUNT). */
enum {
  RGS_STAGE_PROJECT_FWD = 0,
  RGS_STAGE_DEPTH_SORT = 1,
  RGS_STAGE_COUNT_SCAN = 2,
  RGS_STAGE_EMIT = 3,
  RGS_STAGE_TILE_SORT = 4,
  RGS_STAGE_TILE_RANGES = 5,
  RGS_STAGE_RASTER_FWD = 6,
  RGS_STAGE_RASTER_BWD = 7,
  RGS_STAGE_PROJECT_BWD = 8,
  RGS_STAGE_ADAM = 9,
  RGS_STAGE_COUNT = 10
};
int rgs_set_timing(rgs_ctx* ctx, int enable);
int rgs_get_timing(rgs_ctx* ctx, double* ms, int64_t* calls, int reset);

/* Library version string. */
const char* rgs_version(void);

#ifdef __cplusplus
}
#endif
#endif /* RGS_H */
