/* Oracle for the Gaussian-splatting render/train path: a CPU restatement of
 * RustGS's WGSL pipeline (the .wgsl files under RustGS/src/training/shaders/ and the Rust glue
 * in RustGS/src/training/{forward,backward,engine}).
 *
 * THIS IS TEST INFRASTRUCTURE.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it, and only as the checker (or
 * the timed CPU baseline).  The product library never links or calls it.
 *
 * How each function is pinned (golden vectors, closed forms, brute force,
 * finite differences) is stated in the docstrings of tests/test_oracle_*.py.
 */
#ifndef RGS_ORACLE_H
#define RGS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as rgs_camera in include/rgs.h (kept separate so the oracle does
 * not depend on the product header). */
typedef struct orc_camera {
  float viewmat[16]; /* world->camera, column-major (glam Mat4 cols) */
  float fx, fy, cx, cy;
  uint32_t width, height;
  float cam_pos[3]; /* camera centre in world space */
} orc_camera;

typedef struct orc_render orc_render;

typedef struct orc_fwd_stats {
  int64_t num_visible;        /* V: Gaussians passing the projection cull */
  int64_t num_intersections;  /* M: (tile, Gaussian) pairs emitted */
  int64_t k1_count_mismatch;  /* Gaussians whose K1 count != K5 emission */
  int64_t depth_tie_runs;     /* runs of equal depth bits among visible */
  int64_t evaluated_fragments;/* (pixel, splat) pairs evaluated by K7 */
  int64_t blended_fragments;  /* (pixel, splat) pairs accepted by K7 */
  int64_t num_tiles;
} orc_fwd_stats;

/* ---- forward (FP32, reference semantics, canonical tie rule) ---------- */
orc_render* orc_forward(int64_t n, int sh_degree, const float* transforms,
                        const float* sh, const float* raw_opac,
                        const orc_camera* cam, const float* background,
                        float cov_blur, int tile_width);
void orc_free(orc_render* r);
void orc_stats(const orc_render* r, orc_fwd_stats* out);
void orc_get_image(const orc_render* r, float* out_rgba);   /* H*W*4 */
void orc_get_visible(const orc_render* r, float* out);      /* N */
void orc_get_depth_order(const orc_render* r, uint32_t* gids); /* V */
void orc_get_depths(const orc_render* r, float* depth_by_gid); /* N, 0 if culled */
void orc_get_projected(const orc_render* r, float* out);    /* V*9, compact order */
void orc_get_counts(const orc_render* r, uint32_t* counts_by_gid); /* N */
/* sorted intersections: tile ids and GLOBAL gids, M each */
void orc_get_isect(const orc_render* r, uint32_t* tile_ids, uint32_t* gids);
void orc_get_tile_offsets(const orc_render* r, uint32_t* offsets); /* 2*T */

/* ---- backward ------------------------------------------------------------
 * mode 0: reference semantics (rasterize_backwards.wgsl + project_backwards.wgsl)
 * mode 1: exact gradient of orc_forward's image (crossing splat included,
 *         forward sigma rounding, SH view-direction and blur-compensation
 *         derivatives added).
 * Outputs are overwritten (not accumulated); any pointer may be NULL. */
int orc_backward(orc_render* r, const float* v_out_rgba, int mode,
                 float* v_transforms, float* v_sh, float* v_raw_opac,
                 float* screen_stats, float* v_splats_compact,
                 float* screen_splats_compact);

/* FP64 replay of orc_backward over this FP32 forward: the same sorted lists,
 * projected splats (promoted to double), image, and the FP32 backward's skip
 * and stop decisions; the gradient arithmetic of K8 and K9 in double.  With
 * vs_in (V*10 compact splat gradients, FP32) K8 is skipped and only K9 runs
 * in double on those values.  Measures how much of an FP32 backward's error
 * is rounding.  Outputs as orc_backward (doubles); any pointer may be NULL. */
int orc_backward_f64(orc_render* r, const float* v_out_rgba, int mode, const float* vs_in,
                     double* v_transforms, double* v_sh, double* v_raw_opac,
                     double* v_splats_compact);

/* ---- L1 loss (engine/loss.rs reconstruction_residual_loss, plain L1) ---- */
double orc_l1_loss_grad(const float* pred_rgba, const float* target_rgb,
                        int64_t width, int64_t height, float weight,
                        float* v_out_rgba);

/* ---- Adam (engine/optimizer.rs step_tensor) ------------------------------
 * One group: param/grad/m/v of n_rows*row_len floats; lr_row[row_len] is the
 * per-column learning rate (lr * scaling); step is the group's step counter
 * AFTER increment.  beta1/beta2/eps as in AdamScaledConfig. */
void orc_adam_group(int64_t n_rows, int64_t row_len, float* param,
                    const float* grad, float* m, float* v, const float* lr_row,
                    float beta1, float beta2, float eps, int32_t step,
                    float weight_decay);
float orc_powi(float base, int32_t exp);

/* ---- FP64 forward for finite differences ---------------------------------
 * Renders with double arithmetic and returns the L1 loss against target_rgb
 * (weight 1/(3HW)), optionally writing the image.
 * flags bit0: freeze the SH view direction to frozen_viewdir[N*3]
 * flags bit1: freeze the blur compensation to frozen_comp[N]
 * flags bit2: reference-backward termination (stop before the splat that
 *             takes T to <= 1e-4 instead of after it)
 * Tile lists are rebuilt from the double projection. */
double orc_forward_loss_f64(int64_t n, int sh_degree, const double* transforms,
                            const double* sh, const double* raw_opac,
                            const orc_camera* cam, const double* background,
                            double cov_blur, int tile_width, const float* target_rgb,
                            int flags, const double* frozen_viewdir,
                            const double* frozen_comp, double* out_rgba,
                            int64_t frozen_g, const uint8_t* frozen_pix, const double* frozen_proj);
/* The unperturbed FP64 forward's crossing map: pix_gid [H*W] (global id of
 * the splat that takes T below 1e-4, -1 for none), proj9 [N*9] (projection
 * x, y, a, b, c, r, g, b, opacity by global id). */
void orc_crossing_map_f64(int64_t n, int deg, const double* tf, const double* sh, const double* op,
                          const orc_camera* cam, const double* bg, double blur, int tw, int flags, const double* fvd,
                          const double* fcomp, int64_t* pix_gid, double* proj9);
/* The per-Gaussian view direction and compensation of the FP64 projection,
 * for building the frozen surrogate (NaN rows for culled Gaussians). */
void orc_projection_terms_f64(int64_t n, const double* transforms,
                              const double* raw_opac, const orc_camera* cam,
                              double cov_blur, double* viewdir, double* comp);
/* Global ids of Gaussians that are the terminating ("crossing") splat of at
 * least one pixel in the FP32 forward: 1/0 per gid. */
void orc_crossing_splats(const orc_render* r, uint8_t* flags);

/* ---- brute-force renderer: no tiles, every Gaussian tested at every pixel
 * in depth order, same per-pixel arithmetic as K7.  Only valid where the tile
 * test is exact (used on small scenes to check the binning). */
void orc_render_bruteforce(const orc_render* r, float* out_rgba);

/* The gradient of the FP64 pipeline with its own decisions (forward, raster
 * and projection backward in double): v_out [H*W*4] double; outputs
 * v_tf [N*10], v_sh [N*K*3], v_op [N] double. */
int orc_grad_f64(int64_t n, int deg, const double* tf, const double* sh, const double* op, const orc_camera* cam,
                 const double* bg, double blur, int tw, const double* v_out, int mode, double* v_tf, double* v_sh,
                 double* v_op);

/* ---- the reference's CPU renderer (render/tiled_renderer.rs), restated ----
 * positions/log_scales [N*3], rotations wxyz [N*4], opacity_logits [N],
 * sh [N*K*3].  Writes color[H*W*3], depth[H*W]. Single-threaded.
 * gpu_math: 0 = libm expf (the reference's), 1 = rgs_exp. */
void orc_tiled_renderer(int64_t n, int sh_degree, const float* positions,
                        const float* log_scales, const float* rotations,
                        const float* opacity_logits, const float* sh,
                        const orc_camera* cam, const float* background,
                        float cov_blur, int gpu_math, float* color, float* depth);
/* sh.rs evaluate_sh_rgb (degrees 0-4): sh [N*K*3], dirs [N*3] -> rgb [N*3] */
void orc_tr_sh_rgb(int64_t n, int degree, const float* sh, const float* dirs, float* rgb);

/* ---- scalar math exported for pinning tests ---- */
void orc_math_exp(const float* x, float* y, int64_t n);
/* rgs_expf_libm against the C library's expf on all 2^32 inputs: the count of
 * inputs with different bits, the first max_out of them (bits) in out. */
int64_t orc_expf_libm_check(uint32_t* out, int max_out);
void orc_math_log(const float* x, float* y, int64_t n);
void orc_math_exp_core(const float* x, float* y, int64_t n);
void orc_math_exp_core_fast(const float* x, float* y, int64_t n);
/* SH colour (project_visible.wgsl:44-127): coeffs [K*3], viewdir[3] -> rgb[3] */
void orc_sh_color(int degree, const float* coeffs, const float* viewdir, float* rgb);
/* sort contract: stable ascending by key, ties by input position */
void orc_sort_by_key(int64_t n, const uint32_t* keys, const uint32_t* values,
                     uint32_t* out_keys, uint32_t* out_values);
void orc_inclusive_scan(int64_t n, const uint32_t* in, uint32_t* out);

int orc_num_threads(void);


/* ---- topology step (densify / prune / opacity reset); see topology.cpp ---- */
/* Same layout as rgs_topology_cfg in include/rgs.h. */
typedef struct orc_topology_cfg {
  uint32_t iterations;
  uint32_t max_initial_gaussians;
  uint64_t frame_shuffle_seed;
  uint32_t refine_every;
  uint32_t opacity_reset_interval;
  uint32_t opacity_reset_mode;          /* 0 decay, 1 reset */
  int32_t topology_freeze_after_epoch;  /* -1: none */
  int32_t growth_freeze_after_epoch;    /* -1: none */
  int32_t prune_until_epoch;            /* -1: none */
  uint32_t split_score_mode;            /* 0 baseline, 1 abs, 2 abs-pixel, 3 abs-pixel-depth */
  uint32_t growth_stop_iter;
  float growth_grad_threshold;
  float split_grad_threshold;
  float depth_scale_gamma;
  float growth_select_fraction;
  float opacity_decay;
  float scale_decay;
  uint32_t prune_mode;                  /* 0 threshold, 1 weight, 2 visibility-weight */
  uint32_t prune_min_age;
  uint32_t prune_invisible_epochs;
  uint32_t prune_visibility_dry_run;
  float prune_opacity_threshold;
  float prune_visibility_threshold;
  float prune_high_opacity_threshold;
} orc_topology_cfg;

typedef struct orc_topology_schedule_t {
  int32_t completed_epoch; /* -1: none */
  int32_t densify, prune, reset_opacity, allow_extra_growth;
} orc_topology_schedule_t;

#define ORC_TOPOLOGY_MAX_REPORT 4096

typedef struct orc_topology_out {
  orc_topology_schedule_t schedule;
  int64_t before, after, added, pruned;
  int64_t clone_candidates, split_candidates, growth_candidates, prune_candidates;
  int64_t replacement_count, extra_growth_count, selected_count;
  int32_t mutated, adam_reset;
  float center[3];
  float scene_extent, brush_extent;
  int64_t selected[ORC_TOPOLOGY_MAX_REPORT]; /* first selected sources, in order */
  /* arrays of `after` rows (flags: `before` entries; bit 0 clone, 1 split,
     2 growth, 3 prune), owned by the oracle until orc_topology_free */
  float* transforms;
  float* sh;
  float* raw_opac;
  uint32_t* birth;
  uint32_t* invisible_windows;
  int64_t* row_src;
  int* row_kind; /* 0 existing, 1 refined existing, 2 refined new */
  uint8_t* flags;
} orc_topology_out;

void orc_topology_schedule(const orc_topology_cfg* cfg, uint32_t iteration, uint32_t frame_count,
                           orc_topology_schedule_t* out);
uint64_t orc_topology_seed(uint64_t base_seed, uint32_t iteration);
float orc_topology_refine_scalar(uint64_t seed, uint64_t j);
float orc_topology_key(uint64_t seed, int pass, uint64_t i, float w);
/* dens: [10][n] accumulators (order of rgs_densify_export); birth and
 * invisible windows per Gaussian.  Returns 0 on success. */
int orc_topology_step(const orc_topology_cfg* cfg, int64_t n, int deg, const float* transforms, const float* sh,
                      const float* raw_opac, const float* dens, const uint32_t* birth, const uint32_t* invis,
                      uint32_t iteration, uint32_t frame_count, orc_topology_out* out);
void orc_topology_free(orc_topology_out* out);
float orc_cos_2pi(float u); /* the refine offset's cos(2 pi u) */
void orc_brush_refine_offset(const float* rot_wxyz, const float* scale, float sample_scalar, float* out);

#ifdef __cplusplus
}
#endif
#endif
