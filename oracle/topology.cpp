// Oracle of the topology step (densify / prune / opacity reset), test
// infrastructure only: the product library never links or calls it.
//
// Restates, sequentially on the CPU:
//   schedule_topology                 RustGS/src/training/topology/schedule.rs:38-95
//   the trainer's visibility state    engine/trainer.rs:877-944 (update, ages, remap)
//   TopologySplatMetrics, extents     topology/splat_metrics.rs:1-114, core/splats.rs:258-284
//   build_host_snapshot_stats         topology/mod.rs:1124-1235 (with bridge.rs:82-128 sanitising)
//   analyze_topology_candidates       topology/mod.rs:293-436
//   litegs_should_prune_candidate     topology/mod.rs:1079-1122
//   litegs_requested_additions,
//   requested_gaussian_cap,
//   the densify selection             topology/mod.rs:634-765
//   refine_decay_for_schedule         topology/mod.rs:767-787
//   plan_brush_refine_mutation,
//   topology_mutation_aftermath       topology/mod.rs:895-981
//   rebuild_host_snapshot             topology/apply.rs:29-188
//
// Two pieces differ on purpose (the GPU step does the same):
//  * Random draws.  The reference draws from rand 0.8.5's StdRng (ChaCha12)
//    and samples with rand::seq::index::sample_weighted (Efraimidis and
//    Spirakis' keys u^(1/w), largest first).  That crate is not available
//    here; the draws come from SplitMix64 (rgs_uniform01), the keys are the
//    equivalent log(u) / w, and the sampled indices are taken in descending
//    key order (ties: lower index first).  The Box-Muller normal of the
//    refine offset is the reference's formula.
//  * The scene centre is summed in double in a fixed tree (blocks of 1024
//    positions, 4 per lane, pairwise within the block, blocks in order)
//    instead of the reference's sequential f32 sum, so that the GPU step can
//    reproduce it bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../rustscan_amd/csrc/rgs_math.h"
#include "rgs_oracle.h"

namespace {

constexpr float kBrushMinOpacity = 1.0f / 255.0f;  // mod.rs:25
constexpr float kOpacityResetCap = 0.01f;          // mod.rs:26
constexpr float kBrushMinScale = 1e-10f;           // mod.rs:27
constexpr float kRefineProgressLimit = 0.95f;      // mod.rs:28
constexpr float kPercentDense = 0.01f;             // mod.rs:24
constexpr uint64_t kSelectionSalt = 0x6c69746567737365ull;  // mod.rs:29
constexpr uint64_t kRefineSalt = 0x62727573685f7266ull;     // mod.rs:30

float fmax_r(float a, float b) { return std::fmax(a, b); }  // Rust f32::max: NaN loses
float fmin_r(float a, float b) { return std::fmin(a, b); }
float clamp_r(float x, float lo, float hi) { return fmin_r(fmax_r(x, lo), hi); }

// core/splats.rs:322-324
float sigmoid_core(float x) { return 1.0f / (1.0f + rgs_exp(-x)); }
// topology/apply.rs:179-187
float sigmoid_apply(float x) {
  if (x >= 0.0f) {
    const float e = rgs_exp(-x);
    return 1.0f / (1.0f + e);
  }
  const float e = rgs_exp(x);
  return e / (1.0f + e);
}

float sanitize_count(float v) { return std::isfinite(v) ? fmax_r(v, 0.0f) : 0.0f; }  // bridge.rs:82-128

// f32::round (half away from zero) as usize
uint64_t round_usize(float v) {
  const float r = std::round(v);
  return r <= 0.0f ? 0ull : (uint64_t)r;
}

}  // namespace

extern "C" {

void orc_topology_schedule(const orc_topology_cfg* cfg, uint32_t iteration, uint32_t frame_count,
                           orc_topology_schedule_t* out) {
  std::memset(out, 0, sizeof(*out));
  out->completed_epoch = -1;
  if (frame_count == 0 || iteration == 0) return;  // schedule.rs:128-133
  const uint32_t epoch = (iteration - 1) / frame_count;
  const uint32_t phase_iter = iteration - 1;
  const bool topology_frozen = cfg->topology_freeze_after_epoch >= 0 && epoch >= (uint32_t)cfg->topology_freeze_after_epoch;
  const bool growth_frozen =
      topology_frozen || (cfg->growth_freeze_after_epoch >= 0 && epoch >= (uint32_t)cfg->growth_freeze_after_epoch);
  const uint32_t refine_every = std::max<uint32_t>(cfg->refine_every, 1);
  const float progress = cfg->iterations == 0 ? 1.0f : (float)phase_iter / (float)cfg->iterations;
  const bool cadence = phase_iter > 0 && phase_iter % refine_every == 0;
  const bool growth_window = progress <= kRefineProgressLimit;
  const bool prune_window = cfg->prune_until_epoch >= 0 ? epoch < (uint32_t)cfg->prune_until_epoch : growth_window;
  out->completed_epoch = (int32_t)epoch;
  out->densify = cadence && !growth_frozen && growth_window;
  out->prune = cadence && !topology_frozen && prune_window;
  out->reset_opacity = cadence && !topology_frozen && cfg->opacity_reset_mode == 1 &&
                       cfg->opacity_reset_interval > 0 && epoch > 0 && epoch % cfg->opacity_reset_interval == 0;
  out->allow_extra_growth = out->densify && phase_iter < cfg->growth_stop_iter;
}

uint64_t orc_topology_seed(uint64_t base_seed, uint32_t iteration) {  // mod.rs:1031-1036
  const uint64_t x = base_seed * 0x9e3779b97f4a7c15ull + (uint64_t)iteration;
  return (x << 17) | (x >> 47);
}

float orc_topology_refine_scalar(uint64_t seed, uint64_t j) {  // mod.rs:1024-1029 (Box-Muller)
  float u1 = rgs_uniform01(seed ^ kRefineSalt, 2 * j);
  const float u2 = rgs_uniform01(seed ^ kRefineSalt, 2 * j + 1);
  u1 = clamp_r(u1, 1e-6f, 1.0f - 1e-6f);
  return rgs_sqrt(-2.0f * rgs_log(u1)) * rgs_cos_2pi(u2);
}

// Efraimidis-Spirakis key of candidate i in pass `pass`: log(u) / w, u in (0, 1].
float orc_topology_key(uint64_t seed, int pass, uint64_t i, float w) {
  const float u = rgs_uniform01(seed ^ kSelectionSalt, ((uint64_t)pass << 40) | i) + 5.9604644775390625e-08f;
  if (!(std::isfinite(w) && w > 0.0f)) return -INFINITY;
  return rgs_log(u) / w;
}

int orc_topology_step(const orc_topology_cfg* cfg, int64_t n, int deg, const float* tf, const float* sh,
                      const float* op, const float* dens, const uint32_t* birth, const uint32_t* invis,
                      uint32_t iteration, uint32_t frame_count, orc_topology_out* out) {
  if (n < 0 || deg < 0 || deg > 4) return 1;
  const int nk = (deg + 1) * (deg + 1);
  const int shw = 3 * nk;
  std::memset(out, 0, sizeof(*out));
  orc_topology_schedule_t sch;
  orc_topology_schedule(cfg, iteration, frame_count, &sch);
  out->schedule = sch;
  out->before = n;
  const uint64_t seed = orc_topology_seed(cfg->frame_shuffle_seed, iteration);

  // ---- visibility state (trainer.rs:877-922): windows updated, ages
  std::vector<uint32_t> win(n), age(n);
  for (int64_t i = 0; i < n; ++i) {
    const float vis = sanitize_count(dens[8 * n + i]);
    const bool visible = std::isfinite(vis) && vis > 0.0f;
    win[i] = visible ? 0u : (invis[i] == 0xffffffffu ? invis[i] : invis[i] + 1u);
    age[i] = iteration > birth[i] ? iteration - birth[i] : 0u;
  }

  // ---- centre (fixed double tree), scene extent, brush extent
  double sum[3] = {0.0, 0.0, 0.0};
  const int64_t blocks = (n + 1023) / 1024;
  for (int64_t b = 0; b < blocks; ++b) {
    for (int a = 0; a < 3; ++a) {
      double lane[256];
      for (int t = 0; t < 256; ++t) {
        double s = 0.0;
        for (int q = 0; q < 4; ++q) {
          const int64_t i = b * 1024 + t * 4 + q;
          if (i < n) s += (double)tf[i * 10 + a];
        }
        lane[t] = s;
      }
      for (int stride = 128; stride >= 1; stride >>= 1)
        for (int t = 0; t < stride; ++t) lane[t] += lane[t + stride];
      sum[a] += lane[0];
    }
  }
  float center[3] = {0.0f, 0.0f, 0.0f};
  float scene_extent = 1.0f, brush_extent = 1.0f;  // empty snapshot (core/splats.rs:259-261, splat_metrics.rs:87-89)
  if (n > 0) {
    for (int a = 0; a < 3; ++a) center[a] = (float)(sum[a] / (double)n);
    float max_dist = 0.0f, ext = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
      const float dx = tf[i * 10] - center[0], dy = tf[i * 10 + 1] - center[1], dz = tf[i * 10 + 2] - center[2];
      max_dist = fmax_r(max_dist, rgs_sqrt((dx * dx + dy * dy) + dz * dz));
      ext = fmax_r(ext, std::fabs(dx));
      ext = fmax_r(ext, std::fabs(dy));
      ext = fmax_r(ext, std::fabs(dz));
    }
    scene_extent = fmax_r(max_dist, 1e-3f);
    brush_extent = fmax_r(ext, 1e-6f);
  }
  for (int a = 0; a < 3; ++a) out->center[a] = center[a];
  out->scene_extent = scene_extent;
  out->brush_extent = brush_extent;
  const float clone_thr = fmax_r(scene_extent * kPercentDense, 1e-4f);      // mod.rs:185-187
  const float max_bounds = fmax_r(brush_extent, fmax_r(scene_extent, 1e-6f)) * 100.0f;  // mod.rs:306

  // ---- analysis (mod.rs:293-436) and prune test (mod.rs:1079-1122)
  std::vector<uint8_t> flags(n, 0);
  std::vector<float> opac(n), gw(n);
  int64_t P = 0, G = 0;
  const bool opacity_prune_enabled = sch.densify;
  for (int64_t i = 0; i < n; ++i) {
    const float* t = tf + i * 10;
    float s[3];
    for (int a = 0; a < 3; ++a) s[a] = rgs_exp(t[7 + a]);
    const float max_scale = fmax_r(fmax_r(s[0], s[1]), s[2]);
    const float opacity = sigmoid_core(op[i]);
    opac[i] = opacity;
    bool retainable = true;
    for (int a = 0; a < 7; ++a)
      if (!std::isfinite(t[a])) retainable = false;  // position and rotation
    for (int k = 0; k < shw; ++k)
      if (!std::isfinite(sh[i * shw + k])) retainable = false;
    // build_host_snapshot_stats (mod.rs:1124-1235)
    auto d = [&](int k) { return dens[(int64_t)k * n + i]; };
    const float obs = fmax_r(sanitize_count(d(7)), 0.0f);
    const float denom = fmax_r(obs, 1.0f);
    const float vis_obs = fmax_r(sanitize_count(d(8)), 0.0f);
    const uint64_t visible_count = round_usize(vis_obs);
    const float vis_denom = fmax_r(vis_obs, 1.0f);
    const float act = fmax_r(sanitize_count(d(9)), 0.0f);
    const float ratio = obs > 0.0f ? clamp_r(act / obs, 0.0f, 1.0f) : 0.0f;
    const float pixel_coverage = fmax_r(d(4), 0.0f);
    const float coverage_denom = fmax_r(pixel_coverage, 1e-6f);
    const float mean2d = d(0) / denom;
    const float abs_mean = d(2) / vis_denom;
    const float abs_pixel = d(3) / coverage_denom;
    const float camera_depth = d(5) / coverage_denom;
    // depth_scale_factor (mod.rs:606-613)
    float depth_scale = 0.0f;
    if (std::isfinite(camera_depth) && camera_depth > 0.0f) {
      const float dn = fmax_r(cfg->depth_scale_gamma * fmax_r(scene_extent, 1e-6f), 1e-6f);
      const float r = camera_depth / dn;
      depth_scale = clamp_r(r * r, 0.0f, 1.0f);
    }
    const uint32_t mode = cfg->split_score_mode;
    const float split_score = mode == 0 ? mean2d : mode == 1 ? abs_mean : mode == 2 ? abs_pixel : abs_pixel * depth_scale;
    const float gthr = cfg->growth_grad_threshold;
    const float sthr = mode == 0 ? gthr : cfg->split_grad_threshold;
    const bool clone = max_scale <= clone_thr && std::isfinite(mean2d) && mean2d >= gthr;
    const bool split_opacity_ready = std::isfinite(opacity) && opacity > kBrushMinOpacity;
    const bool baseline_split = max_scale > clone_thr && std::isfinite(mean2d) && mean2d >= gthr;
    const bool score_split = mode != 0 && std::isfinite(split_score) && split_score >= sthr && split_opacity_ready;
    const bool split = max_scale > clone_thr && (mode == 3 ? score_split : (mode == 1 || mode == 2) ? (baseline_split || score_split) : baseline_split);
    float weight = 0.0f;
    if (clone) weight = fmax_r(mean2d, 0.0f);
    else if (split) weight = mode == 0 ? fmax_r(mean2d, 0.0f) : mode == 3 ? fmax_r(split_score, 0.0f)
                                                                          : fmax_r(fmax_r(split_score, mean2d), 0.0f);
    const bool growth = clone || split;
    gw[i] = weight;
    // prune (mod.rs:1079-1122)
    const float othr = fmax_r(cfg->prune_opacity_threshold, kBrushMinOpacity);
    const bool old_enough = age[i] >= cfg->prune_min_age;
    const bool opacity_prune = old_enough && (!std::isfinite(opacity) || opacity < othr);
    const bool history_invisible = old_enough && visible_count == 0 && win[i] >= cfg->prune_invisible_epochs;
    bool contribution;
    if (cfg->prune_mode == 0) contribution = opacity_prune || history_invisible;
    else if (cfg->prune_mode == 1) contribution = (opacity_prune_enabled && opacity_prune) || (!opacity_prune_enabled && history_invisible);
    else {
      const bool visibility_prune = old_enough && std::isfinite(ratio) && ratio < cfg->prune_visibility_threshold &&
                                    opacity < cfg->prune_high_opacity_threshold;
      contribution = (opacity_prune_enabled && opacity_prune) || (!opacity_prune_enabled && history_invisible) || visibility_prune;
    }
    bool scale_small = false, scale_large = false, oob = false;
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(s[a]) || s[a] < kBrushMinScale) scale_small = true;
      if (std::isfinite(s[a]) && s[a] > max_bounds) scale_large = true;
      if (std::fabs(t[a] - center[a]) > max_bounds) oob = true;
    }
    const bool prune = contribution || scale_small || scale_large || oob || !retainable;
    flags[i] = (uint8_t)((clone ? 1 : 0) | (split ? 2 : 0) | (growth ? 4 : 0) | (prune ? 8 : 0));
    P += prune;
    G += growth;
    out->clone_candidates += clone;
    out->split_candidates += split;
  }
  out->growth_candidates = G;
  out->prune_candidates = P;

  // ---- selection (mod.rs:634-765)
  const uint64_t grow_count = round_usize((float)G * cfg->growth_select_fraction);
  const uint64_t requested = sch.allow_extra_growth ? (uint64_t)P + (grow_count > (uint64_t)P ? grow_count - P : 0)
                                                    : (uint64_t)P;
  const uint64_t budget = std::max<uint64_t>(cfg->max_initial_gaussians, 1);
  const uint64_t max_g = std::max<uint64_t>(budget, (uint64_t)n + requested);
  const uint64_t max_new = max_g - (uint64_t)n;
  std::vector<int64_t> selected;
  std::vector<uint8_t> used(n, 0);
  auto weighted_pick = [&](int pass, const std::vector<int64_t>& src, const std::vector<float>& w, uint64_t k) {
    std::vector<std::pair<uint32_t, int64_t>> keyed;
    keyed.reserve(src.size());
    for (size_t j = 0; j < src.size(); ++j) {
      const float key = orc_topology_key(seed, pass, (uint64_t)src[j], w[j]);
      keyed.push_back({rgs_f2u(key) & 0x7fffffffu, src[j]});  // ascending |key| = descending key
    }
    std::stable_sort(keyed.begin(), keyed.end(), [](auto& a, auto& b) { return a.first < b.first; });
    std::vector<int64_t> picked;
    for (uint64_t j = 0; j < k && j < keyed.size(); ++j) picked.push_back(keyed[j].second);
    return picked;
  };
  if (max_new > 0 && n > 0) {
    std::vector<int64_t> rsrc;
    std::vector<float> rw;
    for (int64_t i = 0; i < n; ++i)
      if (!(flags[i] & 8) && std::isfinite(opac[i]) && opac[i] > 0.0f) {
        rsrc.push_back(i);
        rw.push_back(fmax_r(opac[i], 0.0f));
      }
    const uint64_t rcount = std::min<uint64_t>(std::min<uint64_t>((uint64_t)P, max_new), rsrc.size());
    if (rcount > 0)
      for (int64_t i : weighted_pick(1, rsrc, rw, rcount)) {
        selected.push_back(i);
        used[i] = 1;
        out->replacement_count += 1;
      }
    if (sch.allow_extra_growth && selected.size() < max_new) {
      const uint64_t sample_high = grow_count > (uint64_t)P ? grow_count - P : 0;
      const uint64_t limit = std::min<uint64_t>(sample_high, max_new - selected.size());
      if (limit > 0) {
        std::vector<int64_t> gsrc;
        std::vector<float> gwv;
        for (int64_t i = 0; i < n; ++i)
          if (!(flags[i] & 8) && (flags[i] & 4) && !used[i]) {
            gsrc.push_back(i);
            gwv.push_back(fmax_r(gw[i], 0.0f));
          }
        for (int64_t i : weighted_pick(2, gsrc, gwv, std::min<uint64_t>(limit, gsrc.size()))) {
          selected.push_back(i);
          out->extra_growth_count += 1;
        }
      }
    }
  }

  // ---- plan (mod.rs:895-981)
  std::vector<int64_t> survivors;
  for (int64_t i = 0; i < n; ++i)
    if (!(flags[i] & 8)) survivors.push_back(i);
  if (survivors.empty())
    for (int64_t i = 0; i < n; ++i) survivors.push_back(i);
  std::vector<int64_t> row_of(n, -1);
  std::vector<int64_t> rsrc(survivors);
  std::vector<int> rkind(survivors.size(), 0);
  std::vector<float> rscalar(survivors.size(), 0.0f);
  for (size_t r = 0; r < survivors.size(); ++r) row_of[survivors[r]] = (int64_t)r;
  const int64_t pruned = n - (int64_t)survivors.size();
  int64_t added = 0;
  if (sch.densify && !selected.empty()) {
    for (size_t j = 0; j < selected.size(); ++j) {
      const int64_t src = selected[j];
      const int64_t r = row_of[src];
      if (r < 0) continue;
      const float sc = orc_topology_refine_scalar(seed, (uint64_t)j);
      rkind[r] = 1;
      rscalar[r] = sc;
      rsrc.push_back(src);
      rkind.push_back(2);
      rscalar.push_back(sc);
      ++added;
    }
  }
  const bool decay = (sch.densify || sch.prune) && (cfg->opacity_decay != 0.0f || cfg->scale_decay != 0.0f);
  const bool changed = added > 0 || pruned > 0;
  const bool mutates = changed || decay || sch.reset_opacity;
  out->added = added;
  out->pruned = pruned;
  out->mutated = mutates;
  out->adam_reset = mutates;  // requires_adam_rebuild || apply_opacity_reset (trainer.rs:851-853)
  const int64_t nn = mutates ? (int64_t)rsrc.size() : n;
  out->after = nn;
  for (size_t j = 0; j < selected.size() && j < (size_t)ORC_TOPOLOGY_MAX_REPORT; ++j) out->selected[j] = selected[j];
  out->selected_count = (int64_t)selected.size();

  // ---- rebuilt splats (apply.rs:29-188) and remapped state (trainer.rs:924-944)
  float* ntf = new float[(size_t)std::max<int64_t>(nn, 1) * 10];
  float* nsh = new float[(size_t)std::max<int64_t>(nn, 1) * shw];
  float* nop = new float[(size_t)std::max<int64_t>(nn, 1)];
  uint32_t* nbirth = new uint32_t[(size_t)std::max<int64_t>(nn, 1)];
  uint32_t* nwin = new uint32_t[(size_t)std::max<int64_t>(nn, 1)];
  int64_t* nsrc = new int64_t[(size_t)std::max<int64_t>(nn, 1)];
  int* nkind = new int[(size_t)std::max<int64_t>(nn, 1)];
  const float train_t = cfg->iterations == 0 ? 1.0f : (float)iteration / (float)cfg->iterations;
  for (int64_t r = 0; r < nn; ++r) {
    const int64_t src = mutates ? rsrc[r] : r;
    const int kind = mutates ? rkind[r] : 0;
    const float* t = tf + src * 10;
    float pos[3] = {t[0], t[1], t[2]};
    const float rot[4] = {t[3], t[4], t[5], t[6]};
    float ls[3] = {t[7], t[8], t[9]};
    float logit = op[src];
    if (mutates && kind != 0) {
      const float sc[3] = {rgs_exp(ls[0]), rgs_exp(ls[1]), rgs_exp(ls[2])};
      float off[3];
      orc_brush_refine_offset(rot, sc, rscalar[r], off);
      // brush_refine_scale (apply.rs:139-149): the largest axis (first on ties) halved
      float refined[3] = {sc[0], sc[1], sc[2]};
      int ax = 0;
      for (int a = 1; a < 3; ++a)
        if (refined[a] > refined[ax]) ax = a;
      refined[ax] = refined[ax] * 0.5f;
      for (int a = 0; a < 3; ++a) {
        pos[a] = kind == 1 ? pos[a] - off[a] : pos[a] + off[a];
        ls[a] = rgs_log(fmax_r(refined[a], 1e-6f));
      }
      // brush_refine_opacity_logit (apply.rs:151-155)
      const float o = clamp_r(sigmoid_apply(logit), 1.0f / 255.0f, 1.0f - 1.0f / 255.0f);
      const float rr = clamp_r(1.0f - rgs_sqrt(1.0f - o), 1.0f / 255.0f, 1.0f - 1.0f / 255.0f);
      logit = rgs_log(rr / (1.0f - rr));
    }
    if (mutates && decay) {  // apply_refine_decay (apply.rs:163-177)
      const float shrink = 1.0f - clamp_r(train_t, 0.0f, 1.0f);
      if (cfg->scale_decay > 0.0f) {
        const float scaling = fmax_r(1.0f - cfg->scale_decay * shrink, 1e-6f);
        for (int a = 0; a < 3; ++a) ls[a] = rgs_log(fmax_r(rgs_exp(ls[a]) * scaling, 1e-12f));
      }
      if (cfg->opacity_decay > 0.0f) {
        const float o = sigmoid_apply(logit);
        const float dd = clamp_r(o - cfg->opacity_decay * shrink, 1e-12f, 1.0f - 1e-12f);
        logit = rgs_log(dd / (1.0f - dd));
      }
    }
    if (mutates && sch.reset_opacity) {  // opacity_reset_logit (apply.rs:157-161)
      const float o = clamp_r(sigmoid_apply(logit), 1e-12f, 1.0f - 1e-12f);
      const float rr = fmin_r(o, kOpacityResetCap);
      logit = rgs_log(rr / (1.0f - rr));
    }
    float* o10 = ntf + r * 10;
    o10[0] = pos[0]; o10[1] = pos[1]; o10[2] = pos[2];
    o10[3] = rot[0]; o10[4] = rot[1]; o10[5] = rot[2]; o10[6] = rot[3];
    o10[7] = ls[0]; o10[8] = ls[1]; o10[9] = ls[2];
    std::memcpy(nsh + r * shw, sh + src * shw, sizeof(float) * shw);
    nop[r] = logit;
    nbirth[r] = kind == 2 ? iteration : birth[src];
    nwin[r] = kind == 2 ? 0u : win[src];
    nsrc[r] = src;
    nkind[r] = kind;
  }
  out->transforms = ntf;
  out->sh = nsh;
  out->raw_opac = nop;
  out->birth = nbirth;
  out->invisible_windows = nwin;
  out->row_src = nsrc;
  out->row_kind = nkind;
  out->flags = new uint8_t[(size_t)std::max<int64_t>(n, 1)];
  std::memcpy(out->flags, flags.data(), (size_t)n);
  return 0;
}

float orc_cos_2pi(float u) { return rgs_cos_2pi(u); }

void orc_topology_free(orc_topology_out* out) {
  delete[] out->transforms;
  delete[] out->sh;
  delete[] out->raw_opac;
  delete[] out->birth;
  delete[] out->invisible_windows;
  delete[] out->row_src;
  delete[] out->row_kind;
  delete[] out->flags;
  out->transforms = out->sh = out->raw_opac = nullptr;
  out->birth = out->invisible_windows = nullptr;
  out->row_src = nullptr;
  out->row_kind = nullptr;
  out->flags = nullptr;
}

// brush_refine_offset (apply.rs:126-137): the normalised quaternion (identity
// if its squared length is 0) applied to scale * sample_scalar, with glam's
// Quat * Vec3: v (w^2 - |b|^2) + b 2 (v.b) + (b x v) 2w.
void orc_brush_refine_offset(const float* rot_wxyz, const float* scale, float sample_scalar, float* out) {
  float w = rot_wxyz[0], x = rot_wxyz[1], y = rot_wxyz[2], z = rot_wxyz[3];
  const float len2 = ((x * x + y * y) + z * z) + w * w;
  if (len2 > 0.0f) {
    const float inv = 1.0f / rgs_sqrt(len2);
    w = w * inv; x = x * inv; y = y * inv; z = z * inv;
  } else {
    w = 1.0f; x = y = z = 0.0f;
  }
  const float v[3] = {scale[0] * sample_scalar, scale[1] * sample_scalar, scale[2] * sample_scalar};
  const float b2 = (x * x + y * y) + z * z;
  const float vb = (v[0] * x + v[1] * y) + v[2] * z;
  const float c[3] = {y * v[2] - z * v[1], z * v[0] - x * v[2], x * v[1] - y * v[0]};
  const float ww = w * w - b2, t2 = vb * 2.0f, w2 = w * 2.0f;
  out[0] = (v[0] * ww + x * t2) + c[0] * w2;
  out[1] = (v[1] * ww + y * t2) + c[1] * w2;
  out[2] = (v[2] * ww + z * t2) + c[2] * w2;
}

}  // extern "C"
