// Topology step kernels (densify / prune / opacity reset) of
// RustGS/src/training/topology/, run on the splat store in place of the
// reference's device -> host -> device round trip (bridge.rs:25-146,
// apply.rs:10-27):
//
//   k_topo_center / _final   scene centre: double sums over blocks of 1024
//                            positions (4 per lane, pairwise in the block),
//                            blocks added in order: the same value on every run
//   k_topo_extent            largest distance to the centre and largest
//                            per-axis offset (core/splats.rs:258-284,
//                            splat_metrics.rs:86-113)
//   k_topo_analyze           visibility windows and ages (trainer.rs:877-922),
//                            the per-Gaussian statistics (mod.rs:1124-1235) and
//                            the clone / split / growth / prune tests
//                            (mod.rs:293-436, 1079-1122)
//   k_topo_keys              weighted-sampling keys log(u) / w (Efraimidis and
//                            Spirakis) of the replacement or growth sources;
//                            the keys are then radix-sorted and the first k taken
//   k_topo_rows*             the rows of the rebuilt store (mod.rs:895-945):
//                            survivors in order, selected sources refined, new
//                            rows appended in selection order
//   k_topo_apply             rebuild_host_snapshot (apply.rs:29-188): gather,
//                            refine offset / scale / opacity, refine decay,
//                            opacity reset; remapped birth and windows
//                            (trainer.rs:924-944)
#include <cmath>

#include "rgs_device.h"

namespace rgs {

constexpr uint64_t kTopoSelectionSalt = 0x6c69746567737365ull;  // topology/mod.rs:29
constexpr uint64_t kTopoRefineSalt = 0x62727573685f7266ull;     // topology/mod.rs:30

__global__ __launch_bounds__(256) void k_topo_center(const float* __restrict__ P, uint32_t npad, uint32_t n,
                                                     double* __restrict__ partial) {
  __shared__ double red[256];
  const uint32_t t = threadIdx.x;
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
    for (int q = 0; q < 4; ++q) {
      const uint32_t i = blockIdx.x * 1024u + t * 4u + (uint32_t)q;
      if (i < n) s += (double)P[(size_t)a * npad + i];
    }
    red[t] = s;
    __syncthreads();
    for (uint32_t stride = 128; stride >= 1; stride >>= 1) {
      if (t < stride) red[t] += red[t + stride];
      __syncthreads();
    }
    if (t == 0) partial[(size_t)blockIdx.x * 3 + a] = red[0];
    __syncthreads();
  }
}

// ctr: [0..2] centre, [3] largest distance, [4] largest per-axis offset (bits, atomicMax)
__global__ void k_topo_center_final(const double* __restrict__ partial, uint32_t blocks, uint32_t n,
                                    float* __restrict__ ctr) {
  if (threadIdx.x != 0) return;
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[(size_t)b * 3 + a];
    ctr[a] = (float)(s / (double)n);
  }
  ctr[3] = 0.0f;
  ctr[4] = 0.0f;
}

__global__ __launch_bounds__(256) void k_topo_extent(const float* __restrict__ P, uint32_t npad, uint32_t n,
                                                     float* __restrict__ ctr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float dist = 0.0f, ext = 0.0f;
  if (i < n) {
    const float dx = P[i] - ctr[0], dy = P[(size_t)npad + i] - ctr[1], dz = P[2 * (size_t)npad + i] - ctr[2];
    dist = fmaxf(dist, rgs_sqrt((dx * dx + dy * dy) + dz * dz));
    ext = fmaxf(ext, fabsf(dx));
    ext = fmaxf(ext, fabsf(dy));
    ext = fmaxf(ext, fabsf(dz));
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    dist = fmaxf(dist, __shfl_xor(dist, o, 64));
    ext = fmaxf(ext, __shfl_xor(ext, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {  // non-negative floats order as their bits
    atomicMax(reinterpret_cast<uint32_t*>(ctr + 3), __float_as_uint(dist));
    atomicMax(reinterpret_cast<uint32_t*>(ctr + 4), __float_as_uint(ext));
  }
}

__device__ __forceinline__ float topo_sanitize(float v) { return isfinite(v) ? fmaxf(v, 0.0f) : 0.0f; }

// flags: bit 0 clone, 1 split, 2 growth, 3 prune
__global__ __launch_bounds__(256) void k_topo_analyze(const float* __restrict__ P, const float* __restrict__ dens,
                                                      const uint32_t* __restrict__ birth, uint32_t* __restrict__ invis,
                                                      const float* __restrict__ ctr, TopoParams T,
                                                      uint8_t* __restrict__ flags, float* __restrict__ opac,
                                                      float* __restrict__ gw, uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t c_prune = 0, c_growth = 0, c_clone = 0, c_split = 0, c_src = 0;
  if (i < T.n) {
    const size_t np = T.npad;
    auto d = [&](int k) { return dens[(size_t)k * np + i]; };
    // visibility windows (trainer.rs:877-895) and age (trainer.rs:911-922)
    const float vis_s = topo_sanitize(d(8));
    const uint32_t w_old = invis[i];
    const uint32_t win = vis_s > 0.0f ? 0u : (w_old == 0xffffffffu ? w_old : w_old + 1u);
    invis[i] = win;
    const uint32_t age = T.iteration > birth[i] ? T.iteration - birth[i] : 0u;
    // extents (core/splats.rs:258-284, splat_metrics.rs:86-113), policy thresholds (mod.rs:185-187, 306)
    const float scene_extent = fmaxf(ctr[3], 1e-3f);
    const float brush_extent = fmaxf(ctr[4], 1e-6f);
    const float clone_thr = fmaxf(scene_extent * 0.01f, 1e-4f);
    const float max_bounds = fmaxf(brush_extent, fmaxf(scene_extent, 1e-6f)) * 100.0f;
    float pos[3], s[3];
    bool retainable = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pos[a] = P[(size_t)a * np + i];
      s[a] = rgs_exp(P[(size_t)(7 + a) * np + i]);
      if (!isfinite(pos[a])) retainable = false;
    }
#pragma unroll
    for (int a = 3; a < 7; ++a)
      if (!isfinite(P[(size_t)a * np + i])) retainable = false;
    for (uint32_t k = 0; k < T.nsh; ++k)
      if (!isfinite(P[(size_t)(11 + k) * np + i])) retainable = false;
    const float max_scale = fmaxf(fmaxf(s[0], s[1]), s[2]);
    const float opacity = 1.0f / (1.0f + rgs_exp(-P[10 * np + i]));  // core/splats.rs:322-324
    // statistics (mod.rs:1124-1235)
    const float obs = fmaxf(topo_sanitize(d(7)), 0.0f);
    const float denom = fmaxf(obs, 1.0f);
    const float vis_obs = fmaxf(vis_s, 0.0f);
    const bool visible_zero = roundf(vis_obs) <= 0.0f;
    const float vis_denom = fmaxf(vis_obs, 1.0f);
    const float act = fmaxf(topo_sanitize(d(9)), 0.0f);
    const float ratio = obs > 0.0f ? fminf(fmaxf(act / obs, 0.0f), 1.0f) : 0.0f;
    const float coverage = fmaxf(d(4), 0.0f);
    const float cov_denom = fmaxf(coverage, 1e-6f);
    const float mean2d = d(0) / denom;
    const float abs_mean = d(2) / vis_denom;
    const float abs_pixel = d(3) / cov_denom;
    const float cam_depth = d(5) / cov_denom;
    float depth_scale = 0.0f;  // depth_scale_factor (mod.rs:606-613)
    if (isfinite(cam_depth) && cam_depth > 0.0f) {
      const float dn = fmaxf(T.depth_scale_gamma * fmaxf(scene_extent, 1e-6f), 1e-6f);
      const float r = cam_depth / dn;
      depth_scale = fminf(fmaxf(r * r, 0.0f), 1.0f);
    }
    const uint32_t mode = T.split_score_mode;
    const float split_score = mode == 0 ? mean2d : (mode == 1 ? abs_mean : (mode == 2 ? abs_pixel : abs_pixel * depth_scale));
    const float gthr = T.growth_grad_threshold;
    const float sthr = mode == 0 ? gthr : T.split_grad_threshold;
    const bool clone = max_scale <= clone_thr && isfinite(mean2d) && mean2d >= gthr;
    const bool split_ready = isfinite(opacity) && opacity > (1.0f / 255.0f);
    const bool baseline_split = max_scale > clone_thr && isfinite(mean2d) && mean2d >= gthr;
    const bool score_split = mode != 0 && isfinite(split_score) && split_score >= sthr && split_ready;
    const bool split = max_scale > clone_thr &&
                       (mode == 3 ? score_split : ((mode == 1 || mode == 2) ? (baseline_split || score_split) : baseline_split));
    float weight = 0.0f;
    if (clone)
      weight = fmaxf(mean2d, 0.0f);
    else if (split)
      weight = mode == 0 ? fmaxf(mean2d, 0.0f) : (mode == 3 ? fmaxf(split_score, 0.0f) : fmaxf(fmaxf(split_score, mean2d), 0.0f));
    const bool growth = clone || split;
    // prune test (mod.rs:1079-1122); opacity pruning is enabled by densification (mod.rs:259)
    const float othr = fmaxf(T.prune_opacity_threshold, 1.0f / 255.0f);
    const bool old_enough = age >= T.prune_min_age;
    const bool opacity_prune = old_enough && (!isfinite(opacity) || opacity < othr);
    const bool hist_invisible = old_enough && visible_zero && win >= T.prune_invisible_epochs;
    const bool en = T.densify != 0;
    bool contribution;
    if (T.prune_mode == 0)
      contribution = opacity_prune || hist_invisible;
    else if (T.prune_mode == 1)
      contribution = (en && opacity_prune) || (!en && hist_invisible);
    else
      contribution = (en && opacity_prune) || (!en && hist_invisible) ||
                     (old_enough && isfinite(ratio) && ratio < T.prune_visibility_threshold &&
                      opacity < T.prune_high_opacity_threshold);
    bool small = false, large = false, oob = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!isfinite(s[a]) || s[a] < 1e-10f) small = true;
      if (isfinite(s[a]) && s[a] > max_bounds) large = true;
      if (fabsf(pos[a] - ctr[a]) > max_bounds) oob = true;
    }
    const bool prune = contribution || small || large || oob || !retainable;
    flags[i] = (uint8_t)((clone ? 1 : 0) | (split ? 2 : 0) | (growth ? 4 : 0) | (prune ? 8 : 0));
    opac[i] = opacity;
    gw[i] = weight;
    c_prune = prune;
    c_growth = growth;
    c_clone = clone;
    c_split = split;
    c_src = !prune && isfinite(opacity) && opacity > 0.0f;
  }
  // one atomic per wave per counter
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    c_prune += __shfl_xor(c_prune, o, 64);
    c_growth += __shfl_xor(c_growth, o, 64);
    c_clone += __shfl_xor(c_clone, o, 64);
    c_split += __shfl_xor(c_split, o, 64);
    c_src += __shfl_xor(c_src, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (c_prune) atomicAdd(&counts[kTopoPrune], c_prune);
    if (c_growth) atomicAdd(&counts[kTopoGrowth], c_growth);
    if (c_clone) atomicAdd(&counts[kTopoClone], c_clone);
    if (c_split) atomicAdd(&counts[kTopoSplit], c_split);
    if (c_src) atomicAdd(&counts[kTopoReplacementSources], c_src);
  }
}

// Sort keys of one sampling pass: |log(u) / w| bits (ascending = largest key
// first) for eligible sources, 0xffffffff otherwise; counts the eligible ones.
// pass 1: replacement sources (not pruned, finite opacity > 0; weight opacity)
// pass 2: growth sources (not pruned, growth candidate, not picked in pass 1;
//         weight the growth weight)
__global__ __launch_bounds__(256) void k_topo_keys(const uint8_t* __restrict__ flags, const float* __restrict__ opac,
                                                   const float* __restrict__ gw, const uint8_t* __restrict__ used,
                                                   uint32_t n, uint64_t seed, int pass, uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t elig = 0;
  if (i < n) {
    const uint8_t f = flags[i];
    float w = 0.0f;
    if (pass == 1) {
      elig = !(f & 8) && isfinite(opac[i]) && opac[i] > 0.0f;
      w = fmaxf(opac[i], 0.0f);
    } else {
      elig = !(f & 8) && (f & 4) && !used[i];
      w = fmaxf(gw[i], 0.0f);
    }
    uint32_t key = 0xffffffffu;
    if (elig) {
      const float u = rgs_uniform01(seed ^ kTopoSelectionSalt, ((uint64_t)pass << 40) | i) + 5.9604644775390625e-08f;
      const float k = (isfinite(w) && w > 0.0f) ? rgs_log(u) / w : -__builtin_huge_valf();
      key = __float_as_uint(k) & 0x7fffffffu;
    }
    keys[i] = key;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) elig += __shfl_xor(elig, o, 64);
  if ((threadIdx.x & 63) == 0 && elig) atomicAdd(&counts[kTopoEligible], elig);
}

// The first k sorted sources are selected (appended at sel + offset) and marked used.
__global__ void k_topo_take(const uint32_t* __restrict__ sorted, uint32_t k, uint32_t* __restrict__ sel,
                            uint32_t offset, uint8_t* __restrict__ used) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const uint32_t src = sorted[j];
  sel[offset + j] = src;
  used[src] = 1;
}

__global__ void k_topo_survive(const uint8_t* __restrict__ flags, uint32_t n, uint32_t all,
                               uint32_t* __restrict__ surv) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) surv[i] = (all || !(flags[i] & 8)) ? 1u : 0u;
}

// rows of the survivors (offs: exclusive scan of surv)
__global__ void k_topo_rows_existing(const uint32_t* __restrict__ surv, const uint32_t* __restrict__ offs, uint32_t n,
                                     uint32_t* __restrict__ rsrc, uint8_t* __restrict__ rkind,
                                     float* __restrict__ rsc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !surv[i]) return;
  const uint32_t r = offs[i];
  rsrc[r] = i;
  rkind[r] = 0;
  rsc[r] = 0.0f;
}

// refined pairs of the selected sources (mod.rs:926-943; the normal sample of
// selection j is Box-Muller on draws 2j and 2j+1, mod.rs:1024-1029)
__global__ void k_topo_rows_selected(const uint32_t* __restrict__ sel, uint32_t K, const uint32_t* __restrict__ offs,
                                     uint32_t S, uint64_t seed, uint32_t* __restrict__ rsrc,
                                     uint8_t* __restrict__ rkind, float* __restrict__ rsc) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const uint32_t src = sel[j];
  float u1 = rgs_uniform01(seed ^ kTopoRefineSalt, 2ull * j);
  const float u2 = rgs_uniform01(seed ^ kTopoRefineSalt, 2ull * j + 1ull);
  u1 = fminf(fmaxf(u1, 1e-6f), 1.0f - 1e-6f);
  const float sc = rgs_sqrt(-2.0f * rgs_log(u1)) * rgs_cos_2pi(u2);
  const uint32_t r = offs[src];
  rkind[r] = 1;
  rsc[r] = sc;
  rsrc[S + j] = src;
  rkind[S + j] = 2;
  rsc[S + j] = sc;
}

__device__ __forceinline__ float sigmoid_apply(float x) {  // topology/apply.rs:179-187
  if (x >= 0.0f) {
    const float e = rgs_exp(-x);
    return 1.0f / (1.0f + e);
  }
  const float e = rgs_exp(x);
  return e / (1.0f + e);
}

__global__ __launch_bounds__(256) void k_topo_apply(const float* __restrict__ P, const uint32_t* __restrict__ rsrc,
                                                    const uint8_t* __restrict__ rkind, const float* __restrict__ rsc,
                                                    const uint32_t* __restrict__ birth,
                                                    const uint32_t* __restrict__ invis, TopoApply A,
                                                    float* __restrict__ Pn, uint32_t* __restrict__ birth_n,
                                                    uint32_t* __restrict__ invis_n) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.nn) return;
  const uint32_t src = rsrc[r];
  const int kind = rkind[r];
  const size_t so = A.old_npad, no = A.new_npad;
  float pos[3], ls[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    pos[a] = P[(size_t)a * so + src];
    ls[a] = P[(size_t)(7 + a) * so + src];
  }
  float q[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) q[a] = P[(size_t)(3 + a) * so + src];
  float logit = P[10 * so + src];
  if (kind != 0) {
    const float sc[3] = {rgs_exp(ls[0]), rgs_exp(ls[1]), rgs_exp(ls[2])};
    // brush_refine_offset (apply.rs:126-137): normalised quaternion (identity
    // when its squared length is 0) applied to scale * sample
    float w = q[0], x = q[1], y = q[2], z = q[3];
    const float len2 = ((x * x + y * y) + z * z) + w * w;
    if (len2 > 0.0f) {
      const float inv = 1.0f / rgs_sqrt(len2);
      w = w * inv;
      x = x * inv;
      y = y * inv;
      z = z * inv;
    } else {
      w = 1.0f;
      x = y = z = 0.0f;
    }
    const float smp = rsc[r];
    const float v[3] = {sc[0] * smp, sc[1] * smp, sc[2] * smp};
    const float b2 = (x * x + y * y) + z * z;
    const float vb = (v[0] * x + v[1] * y) + v[2] * z;
    const float cx = y * v[2] - z * v[1], cy = z * v[0] - x * v[2], cz = x * v[1] - y * v[0];
    const float ww = w * w - b2, t2 = vb * 2.0f, w2 = w * 2.0f;
    const float off[3] = {(v[0] * ww + x * t2) + cx * w2, (v[1] * ww + y * t2) + cy * w2,
                          (v[2] * ww + z * t2) + cz * w2};
    // brush_refine_scale (apply.rs:139-149)
    float refined[3] = {sc[0], sc[1], sc[2]};
    int ax = 0;
    if (refined[1] > refined[ax]) ax = 1;
    if (refined[2] > refined[ax]) ax = 2;
    refined[ax] = refined[ax] * 0.5f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      pos[a] = kind == 1 ? pos[a] - off[a] : pos[a] + off[a];
      ls[a] = rgs_log(fmaxf(refined[a], 1e-6f));
    }
    // brush_refine_opacity_logit (apply.rs:151-155)
    const float o = fminf(fmaxf(sigmoid_apply(logit), 1.0f / 255.0f), 1.0f - 1.0f / 255.0f);
    const float rr = fminf(fmaxf(1.0f - rgs_sqrt(1.0f - o), 1.0f / 255.0f), 1.0f - 1.0f / 255.0f);
    logit = rgs_log(rr / (1.0f - rr));
  }
  if (A.decay) {  // apply_refine_decay (apply.rs:163-177)
    const float shrink = 1.0f - fminf(fmaxf(A.train_t, 0.0f), 1.0f);
    if (A.scale_decay > 0.0f) {
      const float scaling = fmaxf(1.0f - A.scale_decay * shrink, 1e-6f);
#pragma unroll
      for (int a = 0; a < 3; ++a) ls[a] = rgs_log(fmaxf(rgs_exp(ls[a]) * scaling, 1e-12f));
    }
    if (A.opacity_decay > 0.0f) {
      const float o = sigmoid_apply(logit);
      const float dd = fminf(fmaxf(o - A.opacity_decay * shrink, 1e-12f), 1.0f - 1e-12f);
      logit = rgs_log(dd / (1.0f - dd));
    }
  }
  if (A.reset_opacity) {  // opacity_reset_logit (apply.rs:157-161)
    const float o = fminf(fmaxf(sigmoid_apply(logit), 1e-12f), 1.0f - 1e-12f);
    const float rr = fminf(o, 0.01f);
    logit = rgs_log(rr / (1.0f - rr));
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    Pn[(size_t)a * no + r] = pos[a];
    Pn[(size_t)(7 + a) * no + r] = ls[a];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) Pn[(size_t)(3 + a) * no + r] = q[a];
  Pn[10 * no + r] = logit;
  for (uint32_t p = 11; p < A.nplanes; ++p) Pn[(size_t)p * no + r] = P[(size_t)p * so + src];
  birth_n[r] = kind == 2 ? A.iteration : birth[src];
  invis_n[r] = kind == 2 ? 0u : invis[src];
}

// ------------------------------------------------------------------ launch
static uint32_t topo_blocks(uint32_t n) { return (n + 255) / 256; }

void launch_topo_analyze(hipStream_t s, const float* P, const float* dens, const uint32_t* birth, uint32_t* invis,
                         const TopoParams& T, double* partial, float* ctr, uint8_t* flags, float* opac, float* gw,
                         uint32_t* counts) {
  const uint32_t blocks = (T.n + 1023) / 1024;
  k_topo_center<<<blocks, 256, 0, s>>>(P, T.npad, T.n, partial);
  k_topo_center_final<<<1, 64, 0, s>>>(partial, blocks, T.n, ctr);
  k_topo_extent<<<topo_blocks(T.n), 256, 0, s>>>(P, T.npad, T.n, ctr);
  k_topo_analyze<<<topo_blocks(T.n), 256, 0, s>>>(P, dens, birth, invis, ctr, T, flags, opac, gw, counts);
}

void launch_topo_keys(hipStream_t s, const uint8_t* flags, const float* opac, const float* gw, const uint8_t* used,
                      uint32_t n, uint64_t seed, int pass, uint32_t* keys, uint32_t* counts) {
  k_topo_keys<<<topo_blocks(n), 256, 0, s>>>(flags, opac, gw, used, n, seed, pass, keys, counts);
}

void launch_topo_take(hipStream_t s, const uint32_t* sorted, uint32_t k, uint32_t* sel, uint32_t offset,
                      uint8_t* used) {
  if (k) k_topo_take<<<topo_blocks(k), 256, 0, s>>>(sorted, k, sel, offset, used);
}

void launch_topo_survive(hipStream_t s, const uint8_t* flags, uint32_t n, uint32_t all, uint32_t* surv) {
  k_topo_survive<<<topo_blocks(n), 256, 0, s>>>(flags, n, all, surv);
}

void launch_topo_rows(hipStream_t s, const uint32_t* surv, const uint32_t* offs, uint32_t n, const uint32_t* sel,
                      uint32_t K, uint32_t S, uint64_t seed, uint32_t* rsrc, uint8_t* rkind, float* rsc) {
  k_topo_rows_existing<<<topo_blocks(n), 256, 0, s>>>(surv, offs, n, rsrc, rkind, rsc);
  if (K) k_topo_rows_selected<<<topo_blocks(K), 256, 0, s>>>(sel, K, offs, S, seed, rsrc, rkind, rsc);
}

void launch_topo_apply(hipStream_t s, const float* P, const uint32_t* rsrc, const uint8_t* rkind, const float* rsc,
                       const uint32_t* birth, const uint32_t* invis, const TopoApply& A, float* Pn, uint32_t* birth_n,
                       uint32_t* invis_n) {
  if (A.nn) k_topo_apply<<<topo_blocks(A.nn), 256, 0, s>>>(P, rsrc, rkind, rsc, birth, invis, A, Pn, birth_n, invis_n);
}

}  // namespace rgs
