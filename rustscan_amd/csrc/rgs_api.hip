// Host runtime behind include/rgs.h: context, SoA splat store, workspace,
// forward/backward pipelines (per-view and batched), HIP-graph capture of the
// training step, autotuning, and the debug accessors used by parity tests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rgs.h"
#include "rgs_internal.h"
#include "rgs_host.h"

namespace rgs {
void launch_project_fwd(int deg, hipStream_t s, const float* P, uint32_t npad, uint32_t n, const ViewDev* views,
                        int v0, int nv, float blur, int tw, ProjSplat* proj, uint32_t* keys, uint32_t* vals,
                        uint32_t* cnt, uint32_t* counters, int count_ib);
void launch_project_bwd(int deg, hipStream_t s, const float* P, uint32_t npad, uint32_t n, const ViewDev* views,
                        int v0, int nv, float blur, float* acc, const uint32_t* keys_by_inst, float* grads,
                        float* stats, int accumulate, const uint32_t* counters, float* dinc, const uint8_t* vis);
void launch_densify_commit(hipStream_t s, float* dens, const float* dinc, const float* stats, uint32_t npad,
                           uint32_t n, int nviews, int flags, const uint32_t* counters);
uint32_t radix_grid(uint64_t cap);
uint64_t radix_scratch_words(uint64_t cap);
int radix_sort_pairs(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                     uint32_t* kc, uint32_t* vc, const uint32_t* n_dev, uint32_t n_host, uint32_t cap, int key_bits,
                     uint32_t* scratch, const uint32_t* gsrc, uint32_t* gout, int count_ib);
int radix_sort_pairs_packed(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                            uint32_t* kout, uint32_t* vout, const uint32_t* n_dev, uint32_t n_host, uint32_t cap,
                            int key_bits, int value_bits, uint32_t* scratch);
void scan_u32(hipStream_t s, const uint32_t* in, const uint32_t* gather, const uint32_t* n_dev, uint32_t n_host,
              uint32_t cap, uint32_t* out, int inclusive, uint32_t* total_dev, uint32_t* max_dev,
              uint32_t* scratch);
void launch_emit(hipStream_t s, const ProjSplat* proj, const uint32_t* order, const uint32_t* offs,
                 const uint32_t* cnt, uint32_t* counters, uint32_t n_per_view, const ViewDev* views, int tw,
                 uint32_t cap_inst, uint32_t cap, uint32_t* okeys, uint32_t* ovals);
void launch_tile_ranges(hipStream_t s, const uint32_t* keys, const uint32_t* counters, uint32_t cap,
                        uint32_t num_tiles, uint32_t* ranges);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
void launch_step_begin(hipStream_t s, uint32_t* counters, float* fscal, int nviews);
void launch_fwd_begin(hipStream_t s, uint32_t* counters);
void launch_step_end(hipStream_t s, uint32_t* counters);
void launch_set_u32(hipStream_t s, uint32_t* p, uint32_t v);
int launch_raster_fwd(hipStream_t s, int tw, int batch, const RasterArgs& A, uint32_t grid);
int launch_raster_bwd(hipStream_t s, int tw, int batch, const RasterArgs& A, uint32_t grid, bool refine);
int raster_persistent_grid(int tw, int batch, bool bwd, int num_cus);
uint32_t raster_items(int tw, uint32_t tiles);
void launch_adam(hipStream_t s, float* P, const float* G, float* M, float* V, uint32_t npad, uint32_t n, int nplanes,
                 const float* lr_planes, float b1, float b2, float eps, float wd, float* fscal, uint32_t* counters,
                 uint32_t* step, int nviews, int check_loss);
void launch_l1(hipStream_t s, const float* pred, const float* tgt, uint32_t np, float scale, float* v_out,
               float* loss_sum);
void launch_aos_to_soa(hipStream_t s, const float* aos, int width, uint32_t n, float* soa, uint32_t npad, int plane0);
void launch_soa_to_aos(hipStream_t s, const float* soa, uint32_t npad, int plane0, int width, uint32_t n, float* aos);
void launch_fill_f32(hipStream_t s, float* p, float v, size_t n);
hipError_t noalloc_setup(int device);
void make_loss_coef(const float* w6, int residual, int degenerate_nan, LossCoef* L);
uint32_t loss_blocks(uint32_t W, uint32_t H);
void launch_loss(hipStream_t s, const ViewDev* views, int v0, int nv, uint32_t max_blocks, const LossCoef& L,
                 float* loss_sums, float* loss_parts);
size_t loss_parts_floats(int views);
int noalloc_sort_pairs(int device, hipStream_t stream, const uint32_t* keys, const uint32_t* values,
                       uint32_t* out_keys, uint32_t* out_values, uint32_t n, const char** err);
}  // namespace rgs

using namespace rgs;

namespace {

int check_camera(rgs_ctx* c, const rgs_camera& cam) {
  if (cam.width == 0 || cam.height == 0 || cam.width > 16384 || cam.height > 16384)
    return fail(c, RGS_ERR_INVALID, "camera size %ux%u out of range", cam.width, cam.height);
  if (!(cam.fx > 0.0f) || !(cam.fy > 0.0f)) return fail(c, RGS_ERR_INVALID, "focal lengths must be positive");
  return RGS_OK;
}

// ViewDev table for a batch of cameras (ProjectUniforms, forward/mod.rs:31-75).
int build_views(rgs_ctx* c, const rgs_camera* cams, int nv, const float* bg, float loss_weight, int tw,
                ViewSet& out) {
  if (nv <= 0 || nv > kMaxViews) return fail(c, RGS_ERR_INVALID, "n_views must be in [1, %d]", kMaxViews);
  out.v.assign(nv, ViewDev{});
  out.tw = tw;
  uint32_t tile_base = 0;
  uint64_t pix = 0;
  for (int i = 0; i < nv; ++i) {
    int rc = check_camera(c, cams[i]);
    if (rc) return rc;
    ViewDev& V = out.v[i];
    std::memcpy(V.vm, cams[i].viewmat, sizeof(V.vm));
    V.fx = cams[i].fx;
    V.fy = cams[i].fy;
    V.cx = cams[i].cx;
    V.cy = cams[i].cy;
    V.W = cams[i].width;
    V.H = cams[i].height;
    for (int k = 0; k < 3; ++k) V.campos[k] = cams[i].cam_pos[k];
    V.tiles_x = (V.W + tw - 1) / tw;
    V.tiles_y = (V.H + tw - 1) / tw;
    V.tile_base = tile_base;
    V.pixel_base = (uint32_t)pix;
    // calc_cam_j limits, helpers.wgsl:144-145, in f32 as the shader computes them
    V.lim_px = ((float)1.15 * (float)V.W - V.cx) / V.fx;
    V.lim_py = ((float)1.15 * (float)V.H - V.cy) / V.fy;
    V.lim_nx = ((float)-0.15 * (float)V.W - V.cx) / V.fx;
    V.lim_ny = ((float)-0.15 * (float)V.H - V.cy) / V.fy;
    for (int k = 0; k < 3; ++k) V.bg[k] = bg ? bg[k] : 0.0f;
    V.loss_scale = loss_weight / (float)(3ull * V.W * V.H);
    V.img = nullptr;
    V.target = nullptr;
    V.vout = nullptr;
    V.lgrad = nullptr;
    V.loss_weight = loss_weight;
    tile_base += V.tiles_x * V.tiles_y;
    pix += (uint64_t)V.W * V.H;
  }
  out.tiles_total = tile_base;
  if (tile_base >= (1u << 24)) return fail(c, RGS_ERR_INVALID, "too many tiles (%u)", tile_base);
  return RGS_OK;
}

int ensure_workspace(rgs_ctx* c, const rgs_splats* s, int view_slots, uint32_t tiles, size_t cap_isect) {
  Workspace& w = c->ws;
  const size_t n = (size_t)s->n;
  const size_t inst = n * (size_t)view_slots;
  if (inst >= (1ull << 32)) return fail(c, RGS_ERR_INVALID, "N * views must stay below 2^32");
  bool grew = false;
  if (inst > w.cap_inst || w.n != n || w.view_slots < view_slots) {
    const size_t ci = std::max(inst, w.cap_inst);
    int rc = 0;
    rc |= ensure(c, w.proj, ci * sizeof(ProjSplat));
    rc |= ensure(c, w.keys, ci * 4);
    rc |= ensure(c, w.vals, ci * 4);
    rc |= ensure(c, w.cnt, ci * 4);
    rc |= ensure(c, w.scnt, ci * 4);
    rc |= ensure(c, w.sk1, ci * 4);
    rc |= ensure(c, w.sv1, ci * 4);
    rc |= ensure(c, w.sk2, ci * 4);
    rc |= ensure(c, w.sv2, ci * 4);
    rc |= ensure(c, w.offs, ci * 4);
    rc |= ensure(c, w.visi, ci + 4);
    bool acc_new = w.acc.bytes < ci * kAcc * 4;
    rc |= ensure(c, w.acc, ci * kAcc * 4);
    if (rc) return RGS_ERR_HIP;
    if (acc_new) launch_fill_f32(c->stream, w.acc.as<float>(), 0.0f, ci * kAcc);
    w.cap_inst = ci;
    w.n = (uint32_t)n;
    w.view_slots = std::max(view_slots, w.view_slots);
    grew = true;
  }
  if (cap_isect > w.cap_isect) {
    const size_t cm = std::max<size_t>(cap_isect, 1 << 16);
    if (cm >= (1ull << 32)) return fail(c, RGS_ERR_CAPACITY, "intersection count exceeds 2^32");
    int rc = 0;
    rc |= ensure(c, w.ik0, cm * 4);
    rc |= ensure(c, w.iv0, cm * 4);
    rc |= ensure(c, w.ik1, cm * 4);
    rc |= ensure(c, w.iv1, cm * 4);
    if (rc) return RGS_ERR_HIP;
    w.cap_isect = cm;
    grew = true;
  }
  if (tiles > w.tiles_cap) {
    if (ensure(c, w.ranges, (size_t)tiles * 2 * 4)) return RGS_ERR_HIP;
    w.tiles_cap = tiles;
    grew = true;
  }
  const size_t scratch = std::max<size_t>(radix_scratch_words(std::max(w.cap_inst, w.cap_isect)), 4096);
  if (ensure(c, w.scratch, scratch * 4)) return RGS_ERR_HIP;
  if (!w.views.p) {
    if (ensure(c, w.views, sizeof(ViewDev) * kMaxViews)) return RGS_ERR_HIP;
    if (ensure(c, w.counters, sizeof(uint32_t) * kCntNum)) return RGS_ERR_HIP;
    if (ensure(c, w.fscal, sizeof(float) * kFNum)) return RGS_ERR_HIP;
    launch_fill_u32(c->stream, w.counters.as<uint32_t>(), 0u, kCntNum);
    launch_fill_f32(c->stream, w.fscal.as<float>(), 0.0f, kFNum);
  }
  if (grew && !c->ckpt_active) invalidate_graph(c);
  return RGS_OK;
}

// Attach per-view image pointers; internal buffers for views without one.
int attach_images(rgs_ctx* c, ViewSet& vs, float* const* out_rgba) {
  Workspace& w = c->ws;
  if ((int)w.own_img.size() < (int)vs.v.size()) w.own_img.resize(vs.v.size());
  for (size_t i = 0; i < vs.v.size(); ++i) {
    if (out_rgba && out_rgba[i]) {
      vs.v[i].img = out_rgba[i];
    } else {
      if (ensure(c, w.own_img[i], (size_t)vs.v[i].W * vs.v[i].H * 16)) return RGS_ERR_HIP;
      vs.v[i].img = w.own_img[i].as<float>();
    }
  }
  return RGS_OK;
}

// TrainingLossConfig -> kernel coefficients; the residual variant follows
// reconstruction_residual_loss (loss.rs:153-176).
LossCoef loss_coef(const rgs_loss_cfg& cfg, bool degenerate_nan = true) {
  int mode = 0;
  if (std::isfinite(cfg.robust_delta) && cfg.robust_delta > 0.0f)
    mode = 1;
  else if (std::isfinite(cfg.outlier_threshold) && cfg.outlier_threshold > 0.0f &&
           std::isfinite(cfg.outlier_weight) && cfg.outlier_weight < 1.0f)
    mode = 2;
  const float w6[6] = {cfg.l1_weight, cfg.ssim_weight, cfg.gradient_weight, cfg.robust_delta,
                       cfg.outlier_threshold, cfg.outlier_weight};
  LossCoef L;
  make_loss_coef(w6, mode, degenerate_nan ? 1 : 0, &L);
  return L;
}

int check_loss_cfg(rgs_ctx* c, const rgs_loss_cfg& cfg) {
  const float w[6] = {cfg.l1_weight, cfg.ssim_weight, cfg.gradient_weight, cfg.robust_delta,
                      cfg.outlier_threshold, cfg.outlier_weight};
  for (float x : w)
    if (!std::isfinite(x) || x < 0.0f) return fail(c, RGS_ERR_INVALID, "loss weights must be finite and >= 0");
  return RGS_OK;
}

// Per-view buffers of the combined loss: SSIM partial-derivative maps and the
// image gradient the compositing backward reads.
int attach_loss_buffers(rgs_ctx* c, ViewSet& vs) {
  Workspace& w = c->ws;
  if (!w.loss_parts.p) {
    if (ensure(c, w.loss_parts, loss_parts_floats(kMaxViews) * 4)) return RGS_ERR_HIP;
    launch_fill_f32(c->stream, w.loss_parts.as<float>(), 0.0f, loss_parts_floats(kMaxViews));
  }
  if (w.loss_grad.size() < vs.v.size()) w.loss_grad.resize(vs.v.size());
  if (w.loss_vout.size() < vs.v.size()) w.loss_vout.resize(vs.v.size());
  for (size_t i = 0; i < vs.v.size(); ++i) {
    const size_t np = (size_t)vs.v[i].W * vs.v[i].H;
    if (ensure(c, w.loss_grad[i], np * 9 * 4) || ensure(c, w.loss_vout[i], np * 16)) return RGS_ERR_HIP;
    vs.v[i].lgrad = w.loss_grad[i].as<float>();
    vs.v[i].vout = w.loss_vout[i].as<float>();
  }
  return RGS_OK;
}

int upload_views(rgs_ctx* c, const ViewSet& vs) {
  HIPCHK(c, hipMemcpyAsync(c->ws.views.p, vs.v.data(), sizeof(ViewDev) * vs.v.size(), hipMemcpyHostToDevice,
                           c->stream));
  return RGS_OK;
}

uint32_t raster_grid(rgs_ctx* c, bool bwd, uint32_t tiles) {
  tiles = raster_items((int)c->raster.tile_width, tiles);  // workgroup items: (tile, sub-tile)
  if (!c->raster.persistent) return tiles;
  const uint32_t g = (uint32_t)raster_persistent_grid((int)c->raster.tile_width, (int)c->raster.batch_depth, bwd,
                                                      c->num_cus);
  return std::max<uint32_t>(1, std::min(g, tiles));
}

// Enqueue the forward of views [v0, v0+nv) (render_forward, forward/mod.rs:156-294).
int enqueue_forward(rgs_ctx* c, const rgs_splats* s, const ViewSet& vs, int v0, int nv, float blur, bool with_loss,
                    bool vis_inst = false) {
  Workspace& w = c->ws;
  hipStream_t st = c->stream;
  const uint32_t n = (uint32_t)s->n;
  uint32_t* counters = w.counters.as<uint32_t>();
  const uint32_t tile_first = vs.v[v0].tile_base;
  const ViewDev& last = vs.v[v0 + nv - 1];
  const uint32_t tile_end = last.tile_base + last.tiles_x * last.tiles_y;
  const uint32_t tile_count = tile_end - tile_first;
  const uint32_t nitems = n * (uint32_t)nv;
  const size_t ibase = (size_t)v0 * n;
  const ViewDev* views = w.views.as<ViewDev>();

  // Instance ids are below (v0 + nv) * n; with at most 26 bits the depth
  // sort's values also carry the instance's tile count in the bits above them.
  const int inst_bits = (int)ceil_log2((uint64_t)(v0 + nv) * n);
  const int count_ib = (inst_bits <= 26 && !c->plain_sorts) ? std::max(inst_bits, 1) : 0;

  launch_fwd_begin(st, counters);
  launch_fill_u32(st, w.ranges.as<uint32_t>() + 2 * (size_t)tile_first, 0u, 2 * (size_t)tile_count);
  {
    StageTimer t(c, RGS_STAGE_PROJECT_FWD);
    launch_project_fwd(s->deg, st, s->P.as<float>(), s->npad, n, views, v0, nv, blur, vs.tw, w.proj.as<ProjSplat>(),
                       w.keys.as<uint32_t>(), w.vals.as<uint32_t>(), w.cnt.as<uint32_t>(), counters, count_ib);
  }
  // depth sort of all instances of the batch (sorting.rs:34-61; canonical ties: instance order)
  int wd;
  {
    StageTimer t(c, RGS_STAGE_DEPTH_SORT);
    // the last pass also puts the tile counts in depth order (scnt)
    wd = radix_sort_pairs(st, w.keys.as<uint32_t>() + ibase, w.vals.as<uint32_t>() + ibase, w.sk1.as<uint32_t>(),
                          w.sv1.as<uint32_t>(), w.sk2.as<uint32_t>(), w.sv2.as<uint32_t>(), nullptr, nitems, nitems,
                          32, w.scratch.as<uint32_t>(), w.cnt.as<uint32_t>(), w.scnt.as<uint32_t>(), count_ib);
  }
  uint32_t* order = wd == 1 ? w.sv1.as<uint32_t>() : w.sv2.as<uint32_t>();
  // exclusive scan of per-instance tile counts in depth order (prefix_sum.rs:53-136)
  {
    StageTimer t(c, RGS_STAGE_COUNT_SCAN);
    scan_u32(st, w.scnt.as<uint32_t>(), nullptr, counters + kCntVisible, 0, nitems, w.offs.as<uint32_t>(), 0,
             counters + kCntIsect, counters + kCntIsectMax, w.scratch.as<uint32_t>());
  }
  // emission (map_gaussian_to_intersects.wgsl) and stable tile sort (tile_mapping.rs:206-215)
  {
    StageTimer t(c, RGS_STAGE_EMIT);
    launch_emit(st, w.proj.as<ProjSplat>(), order, w.offs.as<uint32_t>(), w.cnt.as<uint32_t>(), counters, n, views,
                vs.tw, nitems, (uint32_t)w.cap_isect, w.ik0.as<uint32_t>(), w.iv0.as<uint32_t>());
  }
  const int bits = (int)ceil_log2(tile_end);
  int wt;
  {
    StageTimer t(c, RGS_STAGE_TILE_SORT);
    if (c->plain_sorts)
      wt = radix_sort_pairs(st, w.ik0.as<uint32_t>(), w.iv0.as<uint32_t>(), w.ik1.as<uint32_t>(), w.iv1.as<uint32_t>(),
                            w.ik0.as<uint32_t>(), w.iv0.as<uint32_t>(), counters + kCntIsect, 0,
                            (uint32_t)w.cap_isect, bits, w.scratch.as<uint32_t>(), nullptr, nullptr, 0);
    else
      wt = radix_sort_pairs_packed(st, w.ik0.as<uint32_t>(), w.iv0.as<uint32_t>(), w.ik1.as<uint32_t>(),
                                   w.iv1.as<uint32_t>(), w.ik0.as<uint32_t>(), w.iv0.as<uint32_t>(),
                                   counters + kCntIsect, 0, (uint32_t)w.cap_isect, bits,
                                   std::max(1, (int)ceil_log2(n)), w.scratch.as<uint32_t>());
  }
  uint32_t* skeys = wt == 1 ? w.ik1.as<uint32_t>() : w.ik0.as<uint32_t>();
  uint32_t* svals = wt == 1 ? w.iv1.as<uint32_t>() : w.iv0.as<uint32_t>();
  {
    StageTimer t(c, RGS_STAGE_TILE_RANGES);
    launch_tile_ranges(st, skeys, counters, (uint32_t)w.cap_isect, tile_end, w.ranges.as<uint32_t>());
  }
  // compositing (rasterize.rs:67-119)
  RasterArgs A{};
  A.proj = w.proj.as<ProjSplat>();
  A.isect_vals = svals;
  A.ranges = w.ranges.as<uint32_t>();
  A.views = views;
  A.v0 = v0;
  A.nv = nv;
  A.tile_first = tile_first;
  A.tile_count = tile_count;
  A.queue = c->raster.persistent ? counters + kCntQueueFwd : nullptr;
  A.cull = c->no_cull ? 0 : 1;
  A.n = n;
  A.visible = s->visible.as<float>();
  A.vis_inst = vis_inst ? w.visi.as<uint8_t>() : nullptr;
  A.loss = with_loss ? w.fscal.as<float>() + kFLossBase : nullptr;
  A.acc = nullptr;
  A.frag = (c->count_frags && !c->capturing) ? c->frag.as<unsigned long long>() : nullptr;
  StageTimer tr(c, RGS_STAGE_RASTER_FWD);
  if (tile_count > 0 &&
      launch_raster_fwd(st, (int)c->raster.tile_width, (int)c->raster.batch_depth, A, raster_grid(c, false, tile_count)))
    return fail(c, RGS_ERR_UNSUPPORTED, "raster configuration %ux%u/%u not built", c->raster.tile_width,
                c->raster.tile_width, c->raster.batch_depth);
  c->fwd_order = order;
  c->fwd_isect_keys = skeys;
  c->fwd_isect_vals = svals;
  c->fwd_v0 = v0;
  c->fwd_nv = nv;
  return RGS_OK;
}

int enqueue_backward(rgs_ctx* c, rgs_splats* s, const ViewSet& vs, int v0, int nv, float blur, bool accumulate,
                     bool refine, bool densify = false) {
  Workspace& w = c->ws;
  hipStream_t st = c->stream;
  const uint32_t n = (uint32_t)s->n;
  uint32_t* counters = w.counters.as<uint32_t>();
  const uint32_t tile_first = vs.v[v0].tile_base;
  const ViewDev& last = vs.v[v0 + nv - 1];
  const uint32_t tile_count = last.tile_base + last.tiles_x * last.tiles_y - tile_first;
  RasterArgs A{};
  A.proj = w.proj.as<ProjSplat>();
  A.isect_vals = c->fwd_isect_vals;
  A.ranges = w.ranges.as<uint32_t>();
  A.views = w.views.as<ViewDev>();
  A.v0 = v0;
  A.nv = nv;
  A.tile_first = tile_first;
  A.tile_count = tile_count;
  A.queue = c->raster.persistent ? counters + kCntQueueBwd : nullptr;
  A.cull = c->no_cull ? 0 : 1;
  A.n = n;
  A.visible = nullptr;
  A.loss = nullptr;
  A.acc = w.acc.as<float>();
  A.frag = (c->count_frags && !c->capturing) ? c->frag.as<unsigned long long>() : nullptr;
  launch_set_u32(st, counters + kCntQueueBwd, 0u);
  {
    StageTimer t(c, RGS_STAGE_RASTER_BWD);
    if (tile_count > 0 && launch_raster_bwd(st, (int)c->raster.tile_width, (int)c->raster.batch_depth, A,
                                            raster_grid(c, true, tile_count), refine))
      return fail(c, RGS_ERR_UNSUPPORTED, "raster configuration not built");
  }
  if (c->keep_raster_grads) {
    const size_t bytes = (size_t)n * nv * kAcc * 4;
    if (ensure(c, w.acc_dbg, (size_t)w.cap_inst * kAcc * 4)) return RGS_ERR_HIP;
    HIPCHK(c, hipMemcpyAsync(w.acc_dbg.as<float>() + (size_t)v0 * n * kAcc, w.acc.as<float>() + (size_t)v0 * n * kAcc,
                             bytes, hipMemcpyDeviceToDevice, st));
  }
  StageTimer tp(c, RGS_STAGE_PROJECT_BWD);
  launch_project_bwd(s->deg, st, s->P.as<float>(), s->npad, n, w.views.as<ViewDev>(), v0, nv, blur, w.acc.as<float>(),
                     w.keys.as<uint32_t>(), s->G.as<float>(), s->stats.as<float>(), accumulate ? 1 : 0, counters,
                     densify ? s->dinc.as<float>() : nullptr, densify ? w.visi.as<uint8_t>() : nullptr);
  return RGS_OK;
}

void set_lr(rgs_splats* s, const rgs_adam_cfg* cfg, float* host_planes) {
  for (int p = 0; p < 10; ++p) host_planes[p] = cfg->lr_transforms[p];
  host_planes[10] = cfg->lr_opacity;
  for (int k = 0; k < s->nk; ++k)
    for (int ch = 0; ch < 3; ++ch) host_planes[11 + 3 * k + ch] = cfg->lr_sh[k];
}

int enqueue_adam(rgs_ctx* c, rgs_splats* s, const rgs_adam_cfg* cfg, int nviews, int check_loss) {
  StageTimer t(c, RGS_STAGE_ADAM);
  launch_adam(c->stream, s->P.as<float>(), s->G.as<float>(), s->M.as<float>(), s->V.as<float>(), s->npad,
              (uint32_t)s->n, s->nplanes, s->lr.as<float>(), cfg->beta1, cfg->beta2, cfg->eps, cfg->weight_decay,
              c->ws.fscal.as<float>(),
              c->ws.counters.as<uint32_t>(), s->step.as<uint32_t>(), nviews, check_loss);
  return RGS_OK;
}

int enqueue_step(rgs_ctx* c, rgs_splats* s, const ViewSet& vs, const rgs_step_cfg* cfg,
                 const std::vector<int>* groups) {
  const int nv = (int)vs.v.size();
  const float blur = cfg->cov_blur > 0.0f ? cfg->cov_blur : 0.0f;
  const bool dens = (cfg->densify & RGS_DENSIFY_ACCUMULATE) != 0;
  const bool combined = cfg->loss.combined != 0;
  const LossCoef lcoef = loss_coef(cfg->loss);
  // the combined loss runs between a group's forward and its backward
  auto enqueue_loss = [&](int v0, int cnt) {
    if (!combined) return;
    uint32_t blocks = 0;
    for (int v = v0; v < v0 + cnt; ++v) blocks = std::max(blocks, loss_blocks(vs.v[v].W, vs.v[v].H));
    launch_loss(c->stream, c->ws.views.as<ViewDev>(), v0, cnt, blocks, lcoef, c->ws.fscal.as<float>() + kFLossBase,
                c->ws.loss_parts.as<float>());
  };
  launch_step_begin(c->stream, c->ws.counters.as<uint32_t>(), c->ws.fscal.as<float>(), nv);
  launch_fill_f32(c->stream, s->visible.as<float>(), 0.0f, s->npad);
  if (dens) launch_fill_u32(c->stream, c->ws.visi.as<uint32_t>(), 0u, ((size_t)s->n * nv + 3) / 4);
  int rc = RGS_OK;
  if (cfg->mode == RGS_MODE_BATCHED) {
    const std::vector<int> all = {0, nv};
    const std::vector<int>& g = (groups && groups->size() >= 2) ? *groups : all;
    for (size_t i = 0; i + 1 < g.size(); ++i) {
      const int v0 = g[i], cnt = g[i + 1] - g[i];
      if ((rc = enqueue_forward(c, s, vs, v0, cnt, blur, !combined, dens))) return rc;
      enqueue_loss(v0, cnt);
      if ((rc = enqueue_backward(c, s, vs, v0, cnt, blur, i > 0, false, dens))) return rc;
    }
  } else {
    for (int v = 0; v < nv; ++v) {
      if ((rc = enqueue_forward(c, s, vs, v, 1, blur, !combined, dens))) return rc;
      enqueue_loss(v, 1);
      if ((rc = enqueue_backward(c, s, vs, v, 1, blur, v > 0, false, dens))) return rc;
    }
  }
  enqueue_adam(c, s, &cfg->adam, nv, cfg->check_loss);
  if (dens)
    launch_densify_commit(c->stream, s->dens.as<float>(), s->dinc.as<float>(), s->stats.as<float>(), s->npad,
                          (uint32_t)s->n, nv, cfg->densify, c->ws.counters.as<uint32_t>());
  launch_step_end(c->stream, c->ws.counters.as<uint32_t>());
  HIPCHK(c, hipMemcpyAsync(c->h_step, c->ws.counters.p, sizeof(uint32_t) * kCntNum, hipMemcpyDeviceToHost,
                           c->stream));
  return RGS_OK;
}

// The sticky overflow counters were handled: clear them on the device (in
// stream order) and in the host copy.
void clear_overflow_sticky(rgs_ctx* c) {
  launch_set_u32(c->stream, c->ws.counters.as<uint32_t>() + kCntOverflowSticky, 0u);
  launch_set_u32(c->stream, c->ws.counters.as<uint32_t>() + kCntIsectMaxSticky, 0u);
  c->h_step[kCntOverflowSticky] = 0u;
  c->h_step[kCntIsectMaxSticky] = 0u;
}

int read_counters(rgs_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->h_counters, c->ws.counters.p, sizeof(uint32_t) * kCntNum, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->h_fscal, c->ws.fscal.p, sizeof(float) * kFNum, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

// Run forwards only (no parameter change) to learn the intersection counts,
// then grow the capacity with headroom.  In batched mode the views are first
// rendered one at a time to learn each view's count, and grouped so that no
// launch exceeds c->launch_limit intersections (a batch's count is the sum
// of its views' counts); *groups receives the group starts and the view count.

int size_for(rgs_ctx* c, rgs_splats* s, ViewSet& vs, int mode, float blur, std::vector<int>* groups) {
  const int nv = (int)vs.v.size();
  int rc = 0;
  uint64_t need = 0;
  if (mode == RGS_MODE_BATCHED && nv > 1) {
    std::vector<uint64_t> per(nv, 0);
    for (int attempt = 0; attempt < 3; ++attempt) {
      if ((rc = upload_views(c, vs))) return rc;
      uint64_t worst = 0;
      bool over = false;
      for (int v = 0; v < nv; ++v) {
        launch_step_begin(c->stream, c->ws.counters.as<uint32_t>(), c->ws.fscal.as<float>(), nv);
        if ((rc = enqueue_forward(c, s, vs, v, 1, blur, false))) return rc;
        if ((rc = read_counters(c))) return rc;
        per[v] = c->h_counters[kCntIsect];
        worst = std::max<uint64_t>(worst, per[v]);
        over = over || c->h_counters[kCntOverflow] != 0;
      }
      if (!over) break;
      if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, worst * 5 / 4 + 1024))) return rc;
      if (attempt == 2) return fail(c, RGS_ERR_CAPACITY, "could not size the intersection buffers");
    }
    groups->assign(1, 0);
    uint64_t run = 0;
    for (int v = 0; v < nv; ++v) {
      if (v > groups->back() && run + per[v] > c->launch_limit) {
        need = std::max(need, run);
        groups->push_back(v);
        run = 0;
      }
      run += per[v];
    }
    need = std::max(need, run);
    groups->push_back(nv);
    if (need * 5 / 4 + 1024 > c->ws.cap_isect || need * 5 / 4 > c->ws.cap_isect) {
      if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, need * 5 / 4 + 1024))) return rc;
    }
    return RGS_OK;
  }
  groups->assign({0, nv});
  for (int attempt = 0; attempt < 3; ++attempt) {
    if ((rc = upload_views(c, vs))) return rc;
    launch_step_begin(c->stream, c->ws.counters.as<uint32_t>(), c->ws.fscal.as<float>(), nv);
    if (mode == RGS_MODE_BATCHED) {
      if ((rc = enqueue_forward(c, s, vs, 0, nv, blur, false))) return rc;
    } else {
      for (int v = 0; v < nv; ++v)
        if ((rc = enqueue_forward(c, s, vs, v, 1, blur, false))) return rc;
    }
    if ((rc = read_counters(c))) return rc;
    need = c->h_counters[kCntIsectMax];
    if (need <= c->ws.cap_isect) {
      if (need * 5 / 4 > c->ws.cap_isect) {
        if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, need * 5 / 4 + 1024))) return rc;
      }
      return RGS_OK;
    }
    if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, need * 5 / 4 + 1024))) return rc;
  }
  return fail(c, RGS_ERR_CAPACITY, "could not size the intersection buffers");
}

StepPlanKey plan_key(rgs_ctx* c, const rgs_splats* s, const ViewSet& vs, int mode) {
  StepPlanKey k;
  k.n = s->n;
  k.deg = s->deg;
  k.nviews = (int)vs.v.size();
  k.mode = mode;
  k.tw = (int)c->raster.tile_width;
  k.batch = (int)c->raster.batch_depth;
  k.persistent = (int)c->raster.persistent;
  k.densify = 0;
  for (auto& v : vs.v) {
    k.dims.push_back(v.W);
    k.dims.push_back(v.H);
  }
  return k;
}

}  // namespace

// ======================================================================== ABI
extern "C" {

const char* rgs_version(void) { return "rgs-hip 0.1 (gfx950)"; }

int rgs_ctx_create(int device, rgs_ctx** out) {
  if (!out) return RGS_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return RGS_ERR_HIP;
  if (device < 0 || device >= count) return RGS_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return RGS_ERR_HIP;
  // this device's copy of the no-allocation sort's scratch array
  if (noalloc_setup(device) != hipSuccess) return RGS_ERR_HIP;
  rgs_ctx* c = new rgs_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return RGS_ERR_HIP;
  }
  hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device);
  if (hipHostMalloc((void**)&c->h_counters, sizeof(uint32_t) * kCntNum, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_fscal, sizeof(float) * kFNum, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_step, sizeof(uint32_t) * kCntNum, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->step_done, hipEventDisableTiming) != hipSuccess) {
    if (c->h_counters) hipHostFree(c->h_counters);
    if (c->h_fscal) hipHostFree(c->h_fscal);
    if (c->h_step) hipHostFree(c->h_step);
    hipStreamDestroy(c->stream);
    delete c;
    return RGS_ERR_HIP;
  }
  std::memset(c->h_step, 0, sizeof(uint32_t) * kCntNum);
  *out = c;
  return RGS_OK;
}

void rgs_ctx_destroy(rgs_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  invalidate_graph(c);
  Workspace& w = c->ws;
  release(c->frag);
  DevBuf* bufs[] = {&w.views, &w.proj, &w.keys, &w.vals, &w.cnt, &w.scnt, &w.sk1, &w.sv1, &w.sk2, &w.sv2, &w.offs, &w.ik0,
                    &w.iv0, &w.ik1, &w.iv1, &w.ranges, &w.acc, &w.acc_dbg, &w.counters, &w.fscal, &w.scratch, &w.visi};
  for (DevBuf* b : bufs) release(*b);
  for (auto& b : w.own_img) release(b);
  for (auto& b : w.loss_grad) release(b);
  for (auto& b : w.loss_vout) release(b);
  release(w.loss_parts);
  release(c->loss_view);
  c->eval.release_all();
  hipHostFree(c->h_counters);
  hipHostFree(c->h_fscal);
  hipHostFree(c->h_step);
  hipEventDestroy(c->step_done);
  hipStreamDestroy(c->stream);
  delete c;
}

const char* rgs_last_error(const rgs_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rgs_sync(rgs_ctx* c) {
  if (!c) return RGS_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

void* rgs_get_stream(rgs_ctx* c) { return c ? (void*)c->stream : nullptr; }

int rgs_set_raster_config(rgs_ctx* c, const rgs_raster_cfg* cfg) {
  if (!c || !cfg) return RGS_ERR_INVALID;
  const uint32_t tw = cfg->tile_width, b = cfg->batch_depth;
  if (!(tw == 8 || tw == 16 || tw == 32) || !(b == 128 || b == 256 || b == 512))
    return fail(c, RGS_ERR_INVALID, "tile width must be 8/16/32 and batch depth 128/256/512");
  if (tw != c->raster.tile_width || b != c->raster.batch_depth || (cfg->persistent != 0) != (c->raster.persistent != 0)) {
    invalidate_graph(c);
    c->fwd_valid = false;
  }
  c->raster = *cfg;
  c->raster.persistent = cfg->persistent ? 1 : 0;
  return RGS_OK;
}

int rgs_get_raster_config(rgs_ctx* c, rgs_raster_cfg* cfg) {
  if (!c || !cfg) return RGS_ERR_INVALID;
  *cfg = c->raster;
  return RGS_OK;
}

int rgs_set_launch_limit(rgs_ctx* c, int64_t max_intersections) {
  if (!c || max_intersections < 1 || max_intersections > (int64_t)3000000000ll)
    return c ? fail(c, RGS_ERR_INVALID, "launch limit out of range") : RGS_ERR_INVALID;
  c->launch_limit = (uint64_t)max_intersections;
  c->sized.clear();
  c->groups.clear();
  invalidate_graph(c);
  return RGS_OK;
}

int rgs_splats_create(rgs_ctx* c, int64_t n, int deg, rgs_splats** out) {
  if (!c || !out) return RGS_ERR_INVALID;
  *out = nullptr;
  if (n < 1 || n > (1ll << 28)) return fail(c, RGS_ERR_INVALID, "N must be in [1, 2^28]");
  if (deg < 0 || deg > kMaxShDegree) return fail(c, RGS_ERR_UNSUPPORTED, "SH degree must be in [0, 4]");
  rgs_splats* s = new rgs_splats();
  s->n = n;
  s->npad = (uint32_t)((n + 63) / 64 * 64);
  s->deg = deg;
  s->nk = sh_coeffs(deg);
  s->nplanes = 11 + 3 * s->nk;
  const size_t plane_bytes = (size_t)s->npad * 4 * s->nplanes;
  int rc = 0;
  rc |= ensure(c, s->P, plane_bytes);
  rc |= ensure(c, s->G, plane_bytes);
  rc |= ensure(c, s->M, plane_bytes);
  rc |= ensure(c, s->V, plane_bytes);
  rc |= ensure(c, s->stats, (size_t)s->npad * 4 * 7);
  rc |= ensure(c, s->visible, (size_t)s->npad * 4);
  rc |= ensure(c, s->lr, 4 * (size_t)s->nplanes);
  rc |= ensure(c, s->step, 4 * kAdamGroups);
  rc |= ensure(c, s->dens, (size_t)s->npad * 4 * 10);
  rc |= ensure(c, s->dinc, (size_t)s->npad * 4 * 5);
  rc |= ensure(c, s->birth, (size_t)s->npad * 4);
  rc |= ensure(c, s->invis, (size_t)s->npad * 4);
  if (rc) {
    rgs_splats_destroy(c, s);
    return RGS_ERR_HIP;
  }
  const size_t nf = (size_t)s->npad * s->nplanes;
  launch_fill_f32(c->stream, s->P.as<float>(), 0.0f, nf);
  launch_fill_f32(c->stream, s->G.as<float>(), 0.0f, nf);
  launch_fill_f32(c->stream, s->M.as<float>(), 0.0f, nf);
  launch_fill_f32(c->stream, s->V.as<float>(), 0.0f, nf);
  launch_fill_f32(c->stream, s->stats.as<float>(), 0.0f, (size_t)s->npad * 7);
  launch_fill_f32(c->stream, s->visible.as<float>(), 0.0f, s->npad);
  launch_fill_f32(c->stream, s->lr.as<float>(), 0.0f, s->nplanes);
  launch_fill_u32(c->stream, s->step.as<uint32_t>(), 0u, kAdamGroups);
  launch_fill_f32(c->stream, s->dens.as<float>(), 0.0f, (size_t)s->npad * 10);
  launch_fill_f32(c->stream, s->dinc.as<float>(), 0.0f, (size_t)s->npad * 5);
  launch_fill_u32(c->stream, s->birth.as<uint32_t>(), 0u, s->npad);
  launch_fill_u32(c->stream, s->invis.as<uint32_t>(), 0u, s->npad);
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    rgs_splats_destroy(c, s);
    return fail(c, RGS_ERR_HIP, "stream error");
  }
  *out = s;
  return RGS_OK;
}

void rgs_splats_destroy(rgs_ctx* c, rgs_splats* s) {
  if (!s) return;
  if (c) {
    hipStreamSynchronize(c->stream);
    if (c->gsplats == s) invalidate_graph(c);
    if (c->fwd_splats == s) c->fwd_valid = false;
  }
  DevBuf* bufs[] = {&s->P, &s->G, &s->M, &s->V, &s->stats, &s->visible, &s->lr, &s->step, &s->dens, &s->dinc,
                    &s->birth, &s->invis};
  for (DevBuf* b : bufs) release(*b);
  delete s;
}

int64_t rgs_splats_count(const rgs_splats* s) { return s ? s->n : 0; }
int rgs_splats_sh_degree(const rgs_splats* s) { return s ? s->deg : -1; }

int rgs_adam_reset(rgs_ctx* c, rgs_splats* s) {
  if (!c || !s) return RGS_ERR_INVALID;
  int rc = ensure_workspace(c, s, 1, 1, 0);
  if (rc) return rc;
  const size_t nf = (size_t)s->npad * s->nplanes;
  launch_fill_f32(c->stream, s->M.as<float>(), 0.0f, nf);
  launch_fill_f32(c->stream, s->V.as<float>(), 0.0f, nf);
  launch_fill_u32(c->stream, s->step.as<uint32_t>(), 0u, kAdamGroups);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_adam_get_steps(rgs_ctx* c, const rgs_splats* s, uint32_t* steps) {
  if (!c || !s || !steps) return RGS_ERR_INVALID;
  HIPCHK(c, hipMemcpyAsync(steps, s->step.p, 4 * kAdamGroups, hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_adam_set_steps(rgs_ctx* c, rgs_splats* s, const uint32_t* steps) {
  if (!c || !s || !steps) return RGS_ERR_INVALID;
  HIPCHK(c, hipMemcpyAsync(s->step.p, steps, 4 * kAdamGroups, hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_splats_import_aos(rgs_ctx* c, rgs_splats* s, const float* tf, const float* sh, const float* op) {
  if (!c || !s || !tf || !sh || !op) return RGS_ERR_INVALID;
  const uint32_t n = (uint32_t)s->n;
  DevBuf tmp;
  const size_t floats = (size_t)n * std::max(10, 3 * s->nk);
  if (ensure(c, tmp, floats * 4)) return RGS_ERR_HIP;
  auto put = [&](const float* src, int width, int plane0) -> int {
    HIPCHK(c, hipMemcpyAsync(tmp.p, src, (size_t)n * width * 4, hipMemcpyDefault, c->stream));
    launch_aos_to_soa(c->stream, tmp.as<float>(), width, n, s->P.as<float>(), s->npad, plane0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RGS_OK;
  };
  int rc = put(tf, 10, 0);
  if (!rc) rc = put(op, 1, 10);
  if (!rc) rc = put(sh, 3 * s->nk, 11);
  release(tmp);
  if (rc) return rc;
  if (c->fwd_splats == s) c->fwd_valid = false;
  // new splats: their topology state starts over (trainer.rs:171-172)
  launch_fill_u32(c->stream, s->birth.as<uint32_t>(), 0u, s->npad);
  launch_fill_u32(c->stream, s->invis.as<uint32_t>(), 0u, s->npad);
  return rgs_adam_reset(c, s);
}

int rgs_splats_export_aos(rgs_ctx* c, const rgs_splats* s, float* tf, float* sh, float* op) {
  if (!c || !s) return RGS_ERR_INVALID;
  const uint32_t n = (uint32_t)s->n;
  DevBuf tmp;
  if (ensure(c, tmp, (size_t)n * std::max(10, 3 * s->nk) * 4)) return RGS_ERR_HIP;
  auto get = [&](float* dst, int width, int plane0) -> int {
    if (!dst) return RGS_OK;
    launch_soa_to_aos(c->stream, s->P.as<float>(), s->npad, plane0, width, n, tmp.as<float>());
    HIPCHK(c, hipMemcpyAsync(dst, tmp.p, (size_t)n * width * 4, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RGS_OK;
  };
  int rc = get(tf, 10, 0);
  if (!rc) rc = get(op, 1, 10);
  if (!rc) rc = get(sh, 3 * s->nk, 11);
  release(tmp);
  return rc;
}

int rgs_splats_export_grads(rgs_ctx* c, const rgs_splats* s, float* vt, float* vsh, float* vop, float* stats) {
  if (!c || !s) return RGS_ERR_INVALID;
  const uint32_t n = (uint32_t)s->n;
  DevBuf tmp;
  if (ensure(c, tmp, (size_t)n * std::max(10, 3 * s->nk) * 4)) return RGS_ERR_HIP;
  auto get = [&](const float* planes, float* dst, int width, int plane0) -> int {
    if (!dst) return RGS_OK;
    launch_soa_to_aos(c->stream, planes, s->npad, plane0, width, n, tmp.as<float>());
    HIPCHK(c, hipMemcpyAsync(dst, tmp.p, (size_t)n * width * 4, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RGS_OK;
  };
  int rc = get(s->G.as<float>(), vt, 10, 0);
  if (!rc) rc = get(s->G.as<float>(), vop, 1, 10);
  if (!rc) rc = get(s->G.as<float>(), vsh, 3 * s->nk, 11);
  if (!rc) rc = get(s->stats.as<float>(), stats, 7, 0);
  release(tmp);
  return rc;
}

int rgs_adam_export_moments(rgs_ctx* c, const rgs_splats* s, float* m_tf, float* m_sh, float* m_op, float* v_tf,
                            float* v_sh, float* v_op) {
  if (!c || !s) return RGS_ERR_INVALID;
  const uint32_t n = (uint32_t)s->n;
  DevBuf tmp;
  if (ensure(c, tmp, (size_t)n * std::max(10, 3 * s->nk) * 4)) return RGS_ERR_HIP;
  auto get = [&](const float* planes, float* dst, int width, int plane0) -> int {
    if (!dst) return RGS_OK;
    launch_soa_to_aos(c->stream, planes, s->npad, plane0, width, n, tmp.as<float>());
    HIPCHK(c, hipMemcpyAsync(dst, tmp.p, (size_t)n * width * 4, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RGS_OK;
  };
  int rc = get(s->M.as<float>(), m_tf, 10, 0);
  if (!rc) rc = get(s->M.as<float>(), m_op, 1, 10);
  if (!rc) rc = get(s->M.as<float>(), m_sh, 3 * s->nk, 11);
  if (!rc) rc = get(s->V.as<float>(), v_tf, 10, 0);
  if (!rc) rc = get(s->V.as<float>(), v_op, 1, 10);
  if (!rc) rc = get(s->V.as<float>(), v_sh, 3 * s->nk, 11);
  release(tmp);
  return rc;
}

int rgs_adam_import_moments(rgs_ctx* c, rgs_splats* s, const float* m_tf, const float* m_sh, const float* m_op,
                            const float* v_tf, const float* v_sh, const float* v_op) {
  if (!c || !s || !m_tf || !m_sh || !m_op || !v_tf || !v_sh || !v_op) return RGS_ERR_INVALID;
  const uint32_t n = (uint32_t)s->n;
  DevBuf tmp;
  if (ensure(c, tmp, (size_t)n * std::max(10, 3 * s->nk) * 4)) return RGS_ERR_HIP;
  auto put = [&](float* planes, const float* src, int width, int plane0) -> int {
    HIPCHK(c, hipMemcpyAsync(tmp.p, src, (size_t)n * width * 4, hipMemcpyDefault, c->stream));
    launch_aos_to_soa(c->stream, tmp.as<float>(), width, n, planes, s->npad, plane0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RGS_OK;
  };
  int rc = put(s->M.as<float>(), m_tf, 10, 0);
  if (!rc) rc = put(s->M.as<float>(), m_op, 1, 10);
  if (!rc) rc = put(s->M.as<float>(), m_sh, 3 * s->nk, 11);
  if (!rc) rc = put(s->V.as<float>(), v_tf, 10, 0);
  if (!rc) rc = put(s->V.as<float>(), v_op, 1, 10);
  if (!rc) rc = put(s->V.as<float>(), v_sh, 3 * s->nk, 11);
  release(tmp);
  return rc;
}

int rgs_densify_export(rgs_ctx* c, const rgs_splats* s, float* out) {
  if (!c || !s || !out) return RGS_ERR_INVALID;
  for (int k = 0; k < 10; ++k)
    HIPCHK(c, hipMemcpyAsync(out + (size_t)k * s->n, s->dens.as<float>() + (size_t)k * s->npad, 4 * (size_t)s->n,
                             hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_densify_import(rgs_ctx* c, rgs_splats* s, const float* in) {
  if (!c || !s || !in) return RGS_ERR_INVALID;
  for (int k = 0; k < 10; ++k)
    HIPCHK(c, hipMemcpyAsync(s->dens.as<float>() + (size_t)k * s->npad, in + (size_t)k * s->n, 4 * (size_t)s->n,
                             hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_densify_reset(rgs_ctx* c, rgs_splats* s) {
  if (!c || !s) return RGS_ERR_INVALID;
  launch_fill_f32(c->stream, s->dens.as<float>(), 0.0f, (size_t)s->npad * 10);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_splats_export_visible(rgs_ctx* c, const rgs_splats* s, float* visible) {
  if (!c || !s || !visible) return RGS_ERR_INVALID;
  HIPCHK(c, hipMemcpyAsync(visible, s->visible.p, (size_t)s->n * 4, hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_render_forward(rgs_ctx* c, rgs_splats* s, const rgs_camera* cams, int nv, const float* bg, float blur,
                       float* const* out_rgba) {
  if (!c || !s || !cams) return RGS_ERR_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  ViewSet vs;
  int rc = build_views(c, cams, nv, bg, 1.0f, (int)c->raster.tile_width, vs);
  if (rc) return rc;
  // Render into library-owned images: the backward reads them later, and a
  // caller's buffer may be released or reused in between.
  if ((rc = attach_images(c, vs, nullptr))) return rc;
  if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, 0))) return rc;
  blur = blur > 0.0f ? blur : 0.0f;
  for (int attempt = 0; attempt < 3; ++attempt) {
    if ((rc = upload_views(c, vs))) return rc;
    launch_step_begin(c->stream, c->ws.counters.as<uint32_t>(), c->ws.fscal.as<float>(), nv);
    launch_fill_f32(c->stream, s->visible.as<float>(), 0.0f, s->npad);
    if ((rc = enqueue_forward(c, s, vs, 0, nv, blur, false))) return rc;
    if ((rc = read_counters(c))) return rc;
    if (c->h_counters[kCntOverflow] == 0) break;
    if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, (size_t)c->h_counters[kCntIsectMax] * 5 / 4 + 1024)))
      return rc;
    if (attempt == 2) return fail(c, RGS_ERR_CAPACITY, "intersection buffer could not be sized");
  }
  if (out_rgba) {
    for (int i = 0; i < nv; ++i)
      if (out_rgba[i])
        HIPCHK(c, hipMemcpyAsync(out_rgba[i], vs.v[i].img, (size_t)vs.v[i].W * vs.v[i].H * 16,
                                 hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->fwd_valid = true;
  c->fwd_splats = s;
  c->fwd_views = vs;
  c->fwd_blur = blur;
  c->bwd_done = false;
  return RGS_OK;
}

int rgs_render_backward(rgs_ctx* c, rgs_splats* s, const float* const* v_out) {
  if (!c || !s || !v_out) return RGS_ERR_INVALID;
  if (!c->fwd_valid || c->fwd_splats != s)
    return fail(c, RGS_ERR_STATE, "rgs_render_backward needs a preceding rgs_render_forward on these splats");
  HIPCHK(c, hipSetDevice(c->device));
  ViewSet& vs = c->fwd_views;
  for (size_t i = 0; i < vs.v.size(); ++i) {
    if (!v_out[i]) return fail(c, RGS_ERR_INVALID, "v_out[%zu] is null", i);
    vs.v[i].vout = v_out[i];
  }
  int rc = upload_views(c, vs);
  if (rc) return rc;
  rc = enqueue_backward(c, s, vs, 0, (int)vs.v.size(), c->fwd_blur, false, true);
  for (auto& v : vs.v) v.vout = nullptr;
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->bwd_done = true;
  c->fwd_valid = false;  // the rasterizer accumulators were consumed
  return RGS_OK;
}

// ---- checkpoints: forward states of their own -----------------------------
namespace {
void swap_forward_state(rgs_ctx* c, rgs_checkpoint* k) {
  std::swap(c->ws, k->ws);
  std::swap(c->fwd_valid, k->fwd_valid);
  std::swap(c->fwd_splats, k->fwd_splats);
  std::swap(c->fwd_v0, k->fwd_v0);
  std::swap(c->fwd_nv, k->fwd_nv);
  std::swap(c->fwd_order, k->fwd_order);
  std::swap(c->fwd_isect_keys, k->fwd_isect_keys);
  std::swap(c->fwd_isect_vals, k->fwd_isect_vals);
  std::swap(c->fwd_views, k->fwd_views);
  std::swap(c->fwd_blur, k->fwd_blur);
  std::swap(c->bwd_done, k->bwd_done);
  c->ckpt_active = !c->ckpt_active;
}

void release_workspace(Workspace& w) {
  DevBuf* bufs[] = {&w.views, &w.proj, &w.keys, &w.vals, &w.cnt, &w.scnt, &w.sk1, &w.sv1, &w.sk2, &w.sv2, &w.offs, &w.ik0,
                    &w.iv0, &w.ik1, &w.iv1, &w.ranges, &w.acc, &w.acc_dbg, &w.counters, &w.fscal, &w.scratch, &w.visi};
  for (DevBuf* b : bufs) release(*b);
  for (auto& b : w.own_img) release(b);
  for (auto& b : w.loss_grad) release(b);
  for (auto& b : w.loss_vout) release(b);
  release(w.loss_parts);
}
}  // namespace

int rgs_checkpoint_create(rgs_ctx* c, rgs_checkpoint** out) {
  if (!c || !out) return RGS_ERR_INVALID;
  *out = new rgs_checkpoint();
  return RGS_OK;
}

void rgs_checkpoint_destroy(rgs_ctx* c, rgs_checkpoint* k) {
  if (!k) return;
  if (c) {
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
  }
  release_workspace(k->ws);
  release(k->params);
  delete k;
}

int rgs_render_forward_checkpoint(rgs_ctx* c, rgs_splats* s, const rgs_camera* cams, int nv, const float* bg,
                                  float blur, float* const* out_rgba, rgs_checkpoint* k) {
  if (!c || !s || !cams || !k) return RGS_ERR_INVALID;
  if (c->ckpt_active) return fail(c, RGS_ERR_STATE, "a checkpoint is already active on this context");
  swap_forward_state(c, k);
  int rc = rgs_render_forward(c, s, cams, nv, bg, blur, out_rgba);
  swap_forward_state(c, k);
  if (rc) return rc;
  const size_t bytes = (size_t)s->npad * 4 * s->nplanes;
  if (ensure(c, k->params, bytes)) return RGS_ERR_HIP;
  HIPCHK(c, hipMemcpyAsync(k->params.p, s->P.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  k->n = s->n;
  k->npad = s->npad;
  k->deg = s->deg;
  return RGS_OK;
}

int rgs_render_backward_checkpoint(rgs_ctx* c, rgs_splats* s, rgs_checkpoint* k, const float* const* v_out) {
  if (!c || !s || !k || !v_out) return RGS_ERR_INVALID;
  if (c->ckpt_active) return fail(c, RGS_ERR_STATE, "a checkpoint is already active on this context");
  if (!k->fwd_valid || k->fwd_splats != s || k->n != s->n || k->npad != s->npad || k->deg != s->deg)
    return fail(c, RGS_ERR_STATE,
                "rgs_render_backward_checkpoint needs a forward into this checkpoint on these splats (same count)");
  std::swap(s->P, k->params);  // the parameters the forward read
  swap_forward_state(c, k);
  const int rc = rgs_render_backward(c, s, v_out);
  swap_forward_state(c, k);
  std::swap(s->P, k->params);
  return rc;
}

int rgs_l1_loss_grad(rgs_ctx* c, const float* pred, const float* target, uint32_t W, uint32_t H, float weight,
                     float* v_out, float* loss_out) {
  if (!c || !pred || !target || W == 0 || H == 0) return RGS_ERR_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  DevBuf acc;
  if (ensure(c, acc, 4)) return RGS_ERR_HIP;
  launch_fill_f32(c->stream, acc.as<float>(), 0.0f, 1);
  const uint32_t np = W * H;
  launch_l1(c->stream, pred, target, np, weight / (float)(3ull * np), v_out, acc.as<float>());
  if (loss_out) {
    float sum = 0.0f;
    hipError_t e = hipMemcpyAsync(&sum, acc.p, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      release(acc);
      return fail(c, RGS_ERR_HIP, "loss readback: %s", hipGetErrorString(e));
    }
    *loss_out = weight * sum / (float)(3ull * np);
  } else {
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  release(acc);
  return RGS_OK;
}

int rgs_image_loss_grad(rgs_ctx* c, const float* pred, const float* target, uint32_t W, uint32_t H, float weight,
                        const rgs_loss_cfg* cfg, float* v_out, float* loss_out) {
  if (!c || !pred || !target || !cfg || W == 0 || H == 0 || W > 16384 || H > 16384) return RGS_ERR_INVALID;
  if ((reinterpret_cast<uintptr_t>(pred) & 15u) != 0)
    return fail(c, RGS_ERR_INVALID, "pred must be 16-byte aligned (RGBA float rows)");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = check_loss_cfg(c, *cfg);
  if (rc) return rc;
  rgs_loss_cfg lc = *cfg;
  const bool plain = !lc.combined;
  if (plain) lc = rgs_loss_cfg{1, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};  // plain L1 through the same kernels
  const size_t np = (size_t)W * H;
  DevBuf grad, vtmp, acc, parts;
  if (ensure(c, grad, np * 9 * 4) || ensure(c, acc, 4) || (!v_out && ensure(c, vtmp, np * 16)) ||
      ensure(c, parts, loss_parts_floats(1) * 4) || ensure(c, c->loss_view, sizeof(ViewDev))) {
    release(grad);
    release(vtmp);
    release(acc);
    release(parts);
    return RGS_ERR_HIP;
  }
  launch_fill_f32(c->stream, parts.as<float>(), 0.0f, loss_parts_floats(1));
  ViewDev V{};
  V.W = W;
  V.H = H;
  V.img = const_cast<float*>(pred);
  V.target = target;
  V.vout = v_out ? v_out : vtmp.as<float>();
  V.lgrad = grad.as<float>();
  V.loss_weight = weight;
  hipError_t e = hipMemcpyAsync(c->loss_view.p, &V, sizeof(V), hipMemcpyHostToDevice, c->stream);
  launch_fill_f32(c->stream, acc.as<float>(), 0.0f, 1);
  if (e == hipSuccess) {
    launch_loss(c->stream, c->loss_view.as<ViewDev>(), 0, 1, loss_blocks(W, H), loss_coef(lc, !plain),
                acc.as<float>(), parts.as<float>());
    float sum = 0.0f;
    e = hipMemcpyAsync(&sum, acc.p, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && loss_out) *loss_out = weight * sum / (float)(3ull * np);
  }
  release(grad);
  release(vtmp);
  release(acc);
  release(parts);
  if (e != hipSuccess) return fail(c, RGS_ERR_HIP, "loss: %s", hipGetErrorString(e));
  return RGS_OK;
}

int rgs_adam_step(rgs_ctx* c, rgs_splats* s, const rgs_adam_cfg* cfg) {
  if (!c || !s || !cfg) return RGS_ERR_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  int rc = ensure_workspace(c, s, 1, 1, 0);
  if (rc) return rc;
  std::vector<float> lr(s->nplanes);
  set_lr(s, cfg, lr.data());
  HIPCHK(c, hipMemcpyAsync(s->lr.p, lr.data(), 4 * lr.size(), hipMemcpyHostToDevice, c->stream));
  launch_step_begin(c->stream, c->ws.counters.as<uint32_t>(), c->ws.fscal.as<float>(), 0);
  enqueue_adam(c, s, cfg, 0, 0);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_train_step(rgs_ctx* c, rgs_splats* s, const rgs_camera* cams, const float* const* targets, int nv,
                   const rgs_step_cfg* cfg, float* loss_out) {
  if (!c || !s || !cams || !targets || !cfg) return RGS_ERR_INVALID;
  if (cfg->mode != RGS_MODE_PER_VIEW && cfg->mode != RGS_MODE_BATCHED) return fail(c, RGS_ERR_INVALID, "bad mode");
  HIPCHK(c, hipSetDevice(c->device));
  ViewSet vs;
  int rc = build_views(c, cams, nv, cfg->background, cfg->loss_weight, (int)c->raster.tile_width, vs);
  if (rc) return rc;
  for (int i = 0; i < nv; ++i) {
    if (!targets[i]) return fail(c, RGS_ERR_INVALID, "targets[%d] is null", i);
    vs.v[i].target = targets[i];
  }
  if ((rc = attach_images(c, vs, nullptr))) return rc;
  if (cfg->loss.combined) {
    if ((rc = check_loss_cfg(c, cfg->loss))) return rc;
    if ((rc = attach_loss_buffers(c, vs))) return rc;
  }
  if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, 0))) return rc;
  const float blur = cfg->cov_blur > 0.0f ? cfg->cov_blur : 0.0f;
  // learning rates (device table; updated every call so schedules can change them)
  std::vector<float> lr(s->nplanes);
  set_lr(s, &cfg->adam, lr.data());
  HIPCHK(c, hipMemcpyAsync(s->lr.p, lr.data(), 4 * lr.size(), hipMemcpyHostToDevice, c->stream));
  StepPlanKey key = plan_key(c, s, vs, cfg->mode);
  key.densify = cfg->densify;
  {
    const float lw[6] = {cfg->loss.l1_weight, cfg->loss.ssim_weight, cfg->loss.gradient_weight,
                         cfg->loss.robust_delta, cfg->loss.outlier_threshold, cfg->loss.outlier_weight};
    key.loss.assign(1, (uint32_t)cfg->loss.combined);
    for (float x : lw) {
      uint32_t b;
      std::memcpy(&b, &x, 4);
      key.loss.push_back(cfg->loss.combined ? b : 0u);
    }
  }
  // An asynchronous step leaves its counters for the next call.  Waiting for
  // the previous step here keeps one step in flight, so a capacity overflow
  // (which skips that step's update) is seen before the next step is
  // enqueued: the plan is re-sized and at most one step is lost.
  if (c->step_pending) {
    HIPCHK(c, hipEventSynchronize(c->step_done));
    c->step_pending = false;
  }
  if (c->h_step[kCntOverflowSticky]) {
    const size_t need = (size_t)c->h_step[kCntIsectMaxSticky];
    clear_overflow_sticky(c);
    if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, need * 5 / 4 + 1024))) return rc;
    c->sized.clear();  // every plan re-sizes (and re-groups its views) on its next call
    c->groups.clear();
  }
  if (!c->sized[key]) {
    if ((rc = size_for(c, s, vs, cfg->mode, blur, &c->groups[key]))) return rc;
    c->sized[key] = true;
  }
  const std::vector<int>* groups = &c->groups[key];
  for (int attempt = 0; attempt < 3; ++attempt) {
    if ((rc = upload_views(c, vs))) return rc;
    if (cfg->graph == RGS_GRAPH_ON) {
      const bool reuse = c->graph_valid && c->gkey == key && c->gsplats == s;
      if (!reuse) {
        invalidate_graph(c);
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        rc = enqueue_step(c, s, vs, cfg, groups);
        c->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
          if (g) hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) return fail(c, RGS_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
        c->graph = g;
        e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
          invalidate_graph(c);
          return fail(c, RGS_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
        }
        c->graph_valid = true;
        c->gkey = key;
        c->gsplats = s;
      }
      HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    } else {
      if ((rc = enqueue_step(c, s, vs, cfg, groups))) return rc;
    }
    if (!loss_out) {
      HIPCHK(c, hipEventRecord(c->step_done, c->stream));
      c->step_pending = true;
      break;
    }
    // synchronous call: check the step and retry after growing on overflow
    if ((rc = read_counters(c))) return rc;
    if (c->h_counters[kCntOverflow] == 0) {
      for (int v = 0; v < nv; ++v)
        loss_out[v] = cfg->loss_weight * c->h_fscal[kFLossBase + v] / (float)(3ull * vs.v[v].W * vs.v[v].H);
      if (c->h_step[kCntOverflowSticky]) clear_overflow_sticky(c);  // an overflowed attempt was redone
      break;
    }
    if ((rc = ensure_workspace(c, s, nv, vs.tiles_total, (size_t)c->h_counters[kCntIsectMax] * 5 / 4 + 1024)))
      return rc;
    if (attempt == 2) return fail(c, RGS_ERR_CAPACITY, "intersection buffer could not be sized");
    // the overflowed attempt counted itself as skipped; it is redone now
    if (c->h_counters[kCntSkipped] > 0)
      launch_set_u32(c->stream, c->ws.counters.as<uint32_t>() + kCntSkipped, c->h_counters[kCntSkipped] - 1u);
  }
  c->fwd_valid = true;
  c->fwd_splats = s;
  c->fwd_views = vs;
  c->fwd_blur = blur;
  c->bwd_done = true;
  return RGS_OK;
}

int rgs_autotune(rgs_ctx* c, rgs_splats* s, const rgs_camera* cams, const float* const* targets, int nv,
                 const rgs_step_cfg* cfg, rgs_raster_cfg* chosen) {
  if (!c || !s || !cams || !targets || !cfg || nv <= 0) return RGS_ERR_INVALID;
  std::vector<uint32_t> tkey{(uint32_t)s->n, (uint32_t)s->deg, (uint32_t)cfg->mode};
  for (int i = 0; i < nv; ++i) {
    tkey.push_back(cams[i].width);
    tkey.push_back(cams[i].height);
  }
  auto it = c->tuned.find(tkey);
  if (it != c->tuned.end()) {
    int rc = rgs_set_raster_config(c, &it->second);
    if (chosen) *chosen = it->second;
    return rc;
  }
  // save parameters and optimizer state; the tuning steps must not train
  const size_t bytes = (size_t)s->npad * 4 * s->nplanes;
  DevBuf sp, sm, sv;
  int rc = 0;
  rc |= ensure(c, sp, bytes);
  rc |= ensure(c, sm, bytes);
  rc |= ensure(c, sv, bytes);
  if (rc) return RGS_ERR_HIP;
  HIPCHK(c, hipMemcpyAsync(sp.p, s->P.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(sm.p, s->M.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(sv.p, s->V.p, bytes, hipMemcpyDeviceToDevice, c->stream));
  uint32_t saved_step[kAdamGroups] = {};
  HIPCHK(c, hipMemcpyAsync(saved_step, s->step.p, 4 * kAdamGroups, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int64_t saved_skipped = rgs_skipped_steps(c);
  // parameters, moments and step counter as the caller left them: every
  // candidate is timed from this state, and it is what the caller gets back
  auto restore = [&]() -> int {
    HIPCHK(c, hipMemcpyAsync(s->P.p, sp.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->M.p, sm.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->V.p, sv.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->step.p, saved_step, 4 * kAdamGroups, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return RGS_OK;
  };
  const rgs_raster_cfg orig = c->raster;
  rgs_raster_cfg best = orig;
  float best_ms = 1e30f;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const uint32_t tws[3] = {8, 16, 32}, bs[3] = {128, 256, 512};
  for (uint32_t tw : tws)
    for (uint32_t b : bs) {
      rgs_raster_cfg rc_cfg{tw, b, 1};
      if (rgs_set_raster_config(c, &rc_cfg)) continue;
      if ((rc = restore())) return rc;
      int r = rgs_train_step(c, s, cams, targets, nv, cfg, nullptr);  // warm-up (sizes, captures)
      if (r) continue;
      hipEventRecord(e0, c->stream);
      const int reps = 5;
      for (int k = 0; k < reps && !r; ++k) r = rgs_train_step(c, s, cams, targets, nv, cfg, nullptr);
      hipEventRecord(e1, c->stream);
      hipEventSynchronize(e1);
      if (r) continue;
      float ms = 0.0f;
      hipEventElapsedTime(&ms, e0, e1);
      ms /= reps;
      // a later candidate must win by 1% (configurations within timing
      // noise keep the earlier one: smaller tile, then smaller batch)
      if (ms < best_ms * 0.99f) {
        best_ms = ms;
        best = rc_cfg;
      }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (c->step_pending) {
    HIPCHK(c, hipEventSynchronize(c->step_done));
    c->step_pending = false;
  }
  if ((rc = restore())) return rc;
  // the tuning steps count in neither the skipped steps nor the overflow
  // state the next step reacts to (each candidate sized its own plan)
  if (c->ws.counters.p) {
    if (saved_skipped >= 0)
      launch_set_u32(c->stream, c->ws.counters.as<uint32_t>() + kCntSkipped, (uint32_t)saved_skipped);
    clear_overflow_sticky(c);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  release(sp);
  release(sm);
  release(sv);
  c->fwd_valid = false;
  if (best_ms >= 1e29f) {
    rgs_set_raster_config(c, &orig);
    return fail(c, RGS_ERR_STATE, "autotune: no configuration ran");
  }
  c->tuned[tkey] = best;
  if (chosen) *chosen = best;
  return rgs_set_raster_config(c, &best);
}

int rgs_get_counts(rgs_ctx* c, rgs_counts* out, int nv) {
  if (!c || !out) return RGS_ERR_INVALID;
  if (!c->ws.counters.p || c->fwd_views.v.empty()) return fail(c, RGS_ERR_STATE, "no forward yet");
  // Per view: visible instances from the depth keys (culled ones hold
  // 0xffffffff), intersections from the tile ranges of the view's tiles.
  // Both stay valid for every view of the last step, also when its views were
  // launched in groups.
  const ViewSet& vs = c->fwd_views;
  if (nv > (int)vs.v.size()) return fail(c, RGS_ERR_INVALID, "n_views larger than the last forward");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const uint32_t n = c->ws.n;
  std::vector<uint32_t> keys(n);
  for (int v = 0; v < nv; ++v) {
    const ViewDev& V = vs.v[v];
    const uint32_t T = V.tiles_x * V.tiles_y;
    std::vector<uint32_t> rg(2ull * T);
    if (T) HIPCHK(c, hipMemcpy(rg.data(), c->ws.ranges.as<uint32_t>() + 2ull * V.tile_base, 8ull * T,
                               hipMemcpyDeviceToHost));
    int64_t m = 0;
    for (uint32_t t = 0; t < T; ++t) m += (int64_t)rg[2 * t + 1] - (int64_t)rg[2 * t];
    if (n) HIPCHK(c, hipMemcpy(keys.data(), c->ws.keys.as<uint32_t>() + (size_t)v * n, 4ull * n,
                               hipMemcpyDeviceToHost));
    int64_t vis = 0;
    for (uint32_t i = 0; i < n; ++i) vis += keys[i] != 0xffffffffu;
    out[v].num_visible = vis;
    out[v].num_intersections = m;
    out[v].num_tiles = T;
  }
  return RGS_OK;
}

int64_t rgs_skipped_steps(rgs_ctx* c) {
  if (!c || !c->ws.counters.p) return 0;
  if (read_counters(c)) return -1;
  return c->h_counters[kCntSkipped];
}

int rgs_sort_pairs_u32(rgs_ctx* c, const uint32_t* keys, const uint32_t* values, uint32_t* ok, uint32_t* ov,
                       int64_t n) {
  if (!c || n < 0 || (n > 0 && (!keys || !values || !ok || !ov))) return RGS_ERR_INVALID;
  if (n == 0) return RGS_OK;
  if (n >= (1ll << 31)) return fail(c, RGS_ERR_INVALID, "n too large");
  HIPCHK(c, hipSetDevice(c->device));
  DevBuf k2, v2, scratch;
  int rc = 0;
  rc |= ensure(c, k2, (size_t)n * 4);
  rc |= ensure(c, v2, (size_t)n * 4);
  rc |= ensure(c, scratch, radix_scratch_words((uint64_t)n) * 4);
  if (rc) return RGS_ERR_HIP;
  const int where = radix_sort_pairs(c->stream, keys, values, ok, ov, k2.as<uint32_t>(), v2.as<uint32_t>(), nullptr,
                                     (uint32_t)n, (uint32_t)n, 32, scratch.as<uint32_t>(), nullptr, nullptr, 0);
  if (where == 2) {
    HIPCHK(c, hipMemcpyAsync(ok, k2.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(ov, v2.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  hipError_t e = hipStreamSynchronize(c->stream);
  release(k2);
  release(v2);
  release(scratch);
  if (e != hipSuccess) return fail(c, RGS_ERR_HIP, "sort: %s", hipGetErrorString(e));
  return RGS_OK;
}

int rgs_sort_pairs_u32_noalloc(rgs_ctx* c, const uint32_t* keys, const uint32_t* values, uint32_t* ok, uint32_t* ov,
                               int64_t n) {
  if (!c || n < 0 || (n > 0 && (!keys || !values || !ok || !ov))) return RGS_ERR_INVALID;
  if (n > RGS_SORT_NOALLOC_MAX_PAIRS)
    return fail(c, RGS_ERR_INVALID, "rgs_sort_pairs_u32_noalloc: n = %lld exceeds the maximum of %d pairs",
                (long long)n, RGS_SORT_NOALLOC_MAX_PAIRS);
  if (n == 0) return RGS_OK;
  HIPCHK(c, hipSetDevice(c->device));
  const char* err = "";
  const int rc = noalloc_sort_pairs(c->device, c->stream, keys, values, ok, ov, (uint32_t)n, &err);
  if (rc) return fail(c, rc, "rgs_sort_pairs_u32_noalloc: %s", err);
  return RGS_OK;
}

int rgs_inclusive_scan_u32(rgs_ctx* c, const uint32_t* in, uint32_t* out, int64_t n) {
  if (!c || n < 0 || (n > 0 && (!in || !out))) return RGS_ERR_INVALID;
  if (n == 0) return RGS_OK;
  if (n >= (1ll << 31)) return fail(c, RGS_ERR_INVALID, "n too large");
  HIPCHK(c, hipSetDevice(c->device));
  DevBuf scratch;
  if (ensure(c, scratch, 4096 * 4)) return RGS_ERR_HIP;
  scan_u32(c->stream, in, nullptr, nullptr, (uint32_t)n, (uint32_t)n, out, 1, nullptr, nullptr, scratch.as<uint32_t>());
  hipError_t e = hipStreamSynchronize(c->stream);
  release(scratch);
  if (e != hipSuccess) return fail(c, RGS_ERR_HIP, "scan: %s", hipGetErrorString(e));
  return RGS_OK;
}

int rgs_debug_get(rgs_ctx* c, int what, int view, void* dst, int64_t* count) {
  if (!c || !count) return RGS_ERR_INVALID;
  if (!c->fwd_splats) return fail(c, RGS_ERR_STATE, "no forward state");
  const ViewSet& vs = c->fwd_views;
  if (view < 0 || view >= (int)vs.v.size()) return fail(c, RGS_ERR_INVALID, "bad view");
  const ViewDev& V = vs.v[view];
  const uint32_t n = c->ws.n;
  const uint32_t T = V.tiles_x * V.tiles_y;
  const bool batched_state = view >= c->fwd_v0 && view < c->fwd_v0 + c->fwd_nv;
  int rc = read_counters(c);
  if (rc) return rc;
  const uint32_t Vtot = c->h_counters[kCntVisible];
  const uint32_t M = std::min<uint64_t>(c->h_counters[kCntIsect], c->ws.cap_isect);
  auto need_state = [&]() -> int {
    if (!batched_state) return fail(c, RGS_ERR_STATE, "view %d is not in the last forward batch", view);
    return RGS_OK;
  };
  switch (what) {
    case RGS_DEBUG_DEPTH_ORDER: {
      if ((rc = need_state())) return rc;
      std::vector<uint32_t> order(Vtot), res;
      if (Vtot) HIPCHK(c, hipMemcpy(order.data(), c->fwd_order, 4ull * Vtot, hipMemcpyDeviceToHost));
      for (uint32_t i = 0; i < Vtot; ++i)
        if (order[i] / n == (uint32_t)view) res.push_back(order[i] - view * n);
      *count = (int64_t)res.size();
      if (dst && !res.empty()) HIPCHK(c, hipMemcpy(dst, res.data(), 4 * res.size(), hipMemcpyDefault));
      return RGS_OK;
    }
    case RGS_DEBUG_ISECT_TILES:
    case RGS_DEBUG_ISECT_GIDS:
    case RGS_DEBUG_TILE_RANGES: {
      if ((rc = need_state())) return rc;
      std::vector<uint32_t> keys(M), vals(M);
      if (M) {
        HIPCHK(c, hipMemcpy(keys.data(), c->fwd_isect_keys, 4ull * M, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(vals.data(), c->fwd_isect_vals, 4ull * M, hipMemcpyDeviceToHost));
      }
      const uint32_t lo = (uint32_t)(std::lower_bound(keys.begin(), keys.end(), V.tile_base) - keys.begin());
      const uint32_t hi = (uint32_t)(std::lower_bound(keys.begin(), keys.end(), V.tile_base + T) - keys.begin());
      std::vector<uint32_t> res;
      if (what == RGS_DEBUG_TILE_RANGES) {
        std::vector<uint32_t> rg(2ull * T);
        HIPCHK(c, hipMemcpy(rg.data(), c->ws.ranges.as<uint32_t>() + 2ull * V.tile_base, 8ull * T,
                            hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < T; ++t)
          if (rg[2 * t + 1] > rg[2 * t]) {
            rg[2 * t] -= lo;
            rg[2 * t + 1] -= lo;
          } else {
            rg[2 * t] = rg[2 * t + 1] = 0;
          }
        res.swap(rg);
      } else {
        for (uint32_t i = lo; i < hi; ++i)
          res.push_back(what == RGS_DEBUG_ISECT_TILES ? keys[i] - V.tile_base : vals[i]);
      }
      *count = (int64_t)res.size();
      if (dst && !res.empty()) HIPCHK(c, hipMemcpy(dst, res.data(), 4 * res.size(), hipMemcpyDefault));
      return RGS_OK;
    }
    case RGS_DEBUG_PROJECTED: {
      if ((rc = need_state())) return rc;
      std::vector<ProjSplat> pr(n);
      std::vector<uint32_t> keys(n);
      HIPCHK(c, hipMemcpy(pr.data(), c->ws.proj.as<ProjSplat>() + (size_t)view * n, sizeof(ProjSplat) * n,
                          hipMemcpyDeviceToHost));
      HIPCHK(c, hipMemcpy(keys.data(), c->ws.keys.as<uint32_t>() + (size_t)view * n, 4ull * n, hipMemcpyDeviceToHost));
      std::vector<float> res(9ull * n, 0.0f);
      for (uint32_t g = 0; g < n; ++g) {
        if (keys[g] == 0xffffffffu) continue;
        const ProjSplat& p = pr[g];
        const float r9[9] = {p.p0.x, p.p0.y, 2.0f * p.p0.z, p.p0.w, 2.0f * p.p1.x, p.p1.z, p.p1.w, p.p2.x, p.p1.y};
        std::memcpy(&res[9ull * g], r9, sizeof(r9));
      }
      *count = (int64_t)res.size();
      if (dst) HIPCHK(c, hipMemcpy(dst, res.data(), 4 * res.size(), hipMemcpyDefault));
      return RGS_OK;
    }
    case RGS_DEBUG_TILE_COUNTS: {
      if ((rc = need_state())) return rc;
      *count = n;
      if (dst) HIPCHK(c, hipMemcpy(dst, c->ws.cnt.as<uint32_t>() + (size_t)view * n, 4ull * n, hipMemcpyDefault));
      return RGS_OK;
    }
    case RGS_DEBUG_IMAGE: {
      *count = (int64_t)V.W * V.H * 4;
      if (dst) HIPCHK(c, hipMemcpy(dst, V.img, 4ull * (*count), hipMemcpyDefault));
      return RGS_OK;
    }
    case RGS_DEBUG_RASTER_GRADS: {
      if (!c->ws.acc_dbg.p) return fail(c, RGS_ERR_STATE, "raster gradients were not kept (enable with rgs_set_debug)");
      *count = (int64_t)n * kAcc;
      if (dst)
        HIPCHK(c, hipMemcpy(dst, c->ws.acc_dbg.as<float>() + (size_t)view * n * kAcc, 4ull * n * kAcc,
                            hipMemcpyDefault));
      return RGS_OK;
    }
    default:
      return fail(c, RGS_ERR_INVALID, "unknown debug item %d", what);
  }
}

int rgs_set_timing(rgs_ctx* c, int enable) {
  if (!c) return RGS_ERR_INVALID;
  c->timing = enable != 0;
  return RGS_OK;
}

int rgs_get_timing(rgs_ctx* c, double* ms, int64_t* calls, int reset) {
  if (!c) return RGS_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& p : c->pending) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
      c->stage_ms[p.stage] += t;
      c->stage_calls[p.stage] += 1;
    }
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  c->pending.clear();
  for (int i = 0; i < RGS_STAGE_COUNT; ++i) {
    if (ms) ms[i] = c->stage_ms[i];
    if (calls) calls[i] = c->stage_calls[i];
    if (reset) {
      c->stage_ms[i] = 0.0;
      c->stage_calls[i] = 0;
    }
  }
  return RGS_OK;
}

int rgs_set_debug(rgs_ctx* c, int flags) {
  if (!c) return RGS_ERR_INVALID;
  c->keep_raster_grads = (flags & 1) != 0;
  c->count_frags = (flags & 2) != 0;
  c->plain_sorts = (flags & 4) != 0;
  c->no_cull = (flags & 8) != 0;
  if (c->count_frags && !c->frag.p) {
    if (ensure(c, c->frag, kFragCounters * sizeof(unsigned long long))) return RGS_ERR_HIP;
    HIPCHK(c, hipMemsetAsync(c->frag.p, 0, kFragCounters * sizeof(unsigned long long), c->stream));
  }
  invalidate_graph(c);
  return RGS_OK;
}

int rgs_get_fragment_counts(rgs_ctx* c, int64_t* out8, int reset) {
  if (!c || !out8) return RGS_ERR_INVALID;
  if (!c->frag.p) {
    for (int i = 0; i < kFragCounters; ++i) out8[i] = 0;
    return RGS_OK;
  }
  unsigned long long h[kFragCounters];
  HIPCHK(c, hipMemcpyAsync(h, c->frag.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < kFragCounters; ++i) out8[i] = (int64_t)h[i];
  if (reset) HIPCHK(c, hipMemsetAsync(c->frag.p, 0, sizeof(h), c->stream));
  return RGS_OK;
}

}  // extern "C"
