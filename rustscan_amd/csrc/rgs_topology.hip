// Host side of the topology step (densify / prune / opacity reset) behind
// rgs_topology_step: the trainer's apply_topology_mutations
// (RustGS/src/training/engine/trainer.rs:792-859) with its planning
// (topology/bridge.rs, topology/mod.rs, topology/schedule.rs) and rebuild
// (topology/apply.rs) on the device.  The host keeps only the schedule and
// the handful of integer decisions that need the candidate counts (two small
// read-backs per step); every per-Gaussian computation, the weighted
// sampling (a key per source, radix-sorted) and the rebuilt store stay in HBM.
#include <algorithm>
#include <cstring>

#include "rgs_host.h"

namespace rgs {

void launch_topo_analyze(hipStream_t s, const float* P, const float* dens, const uint32_t* birth, uint32_t* invis,
                         const TopoParams& T, double* partial, float* ctr, uint8_t* flags, float* opac, float* gw,
                         uint32_t* counts);
void launch_topo_keys(hipStream_t s, const uint8_t* flags, const float* opac, const float* gw, const uint8_t* used,
                      uint32_t n, uint64_t seed, int pass, uint32_t* keys, uint32_t* counts);
void launch_topo_take(hipStream_t s, const uint32_t* sorted, uint32_t k, uint32_t* sel, uint32_t offset,
                      uint8_t* used);
void launch_topo_survive(hipStream_t s, const uint8_t* flags, uint32_t n, uint32_t all, uint32_t* surv);
void launch_topo_rows(hipStream_t s, const uint32_t* surv, const uint32_t* offs, uint32_t n, const uint32_t* sel,
                      uint32_t K, uint32_t S, uint64_t seed, uint32_t* rsrc, uint8_t* rkind, float* rsc);
void launch_topo_apply(hipStream_t s, const float* P, const uint32_t* rsrc, const uint8_t* rkind, const float* rsc,
                       const uint32_t* birth, const uint32_t* invis, const TopoApply& A, float* Pn, uint32_t* birth_n,
                       uint32_t* invis_n);
int radix_sort_pairs(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                     uint32_t* kc, uint32_t* vc, const uint32_t* n_dev, uint32_t n_host, uint32_t cap, int key_bits,
                     uint32_t* scratch, const uint32_t* gsrc, uint32_t* gout, int count_ib);
uint64_t radix_scratch_words(uint64_t cap);
void scan_u32(hipStream_t s, const uint32_t* in, const uint32_t* gather, const uint32_t* n_dev, uint32_t n_host,
              uint32_t cap, uint32_t* out, int inclusive, uint32_t* total_dev, uint32_t* max_dev, uint32_t* scratch);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
void launch_fill_f32(hipStream_t s, float* p, float v, size_t n);

// schedule_topology (topology/schedule.rs:38-95)
rgs_topology_schedule topology_schedule(const rgs_topology_cfg& cfg, uint32_t iteration, uint32_t frame_count) {
  rgs_topology_schedule o{};
  o.completed_epoch = -1;
  if (frame_count == 0 || iteration == 0) return o;
  const uint32_t epoch = (iteration - 1) / frame_count, phase = iteration - 1;
  const bool topo_frozen = cfg.topology_freeze_after_epoch >= 0 && epoch >= (uint32_t)cfg.topology_freeze_after_epoch;
  const bool growth_frozen =
      topo_frozen || (cfg.growth_freeze_after_epoch >= 0 && epoch >= (uint32_t)cfg.growth_freeze_after_epoch);
  const uint32_t every = std::max<uint32_t>(cfg.refine_every, 1);
  const float progress = cfg.iterations == 0 ? 1.0f : (float)phase / (float)cfg.iterations;
  const bool cadence = phase > 0 && phase % every == 0;
  const bool growth_window = progress <= 0.95f;  // BRUSH_REFINE_PROGRESS_LIMIT (mod.rs:28)
  const bool prune_window = cfg.prune_until_epoch >= 0 ? epoch < (uint32_t)cfg.prune_until_epoch : growth_window;
  o.completed_epoch = (int32_t)epoch;
  o.densify = cadence && !growth_frozen && growth_window;
  o.prune = cadence && !topo_frozen && prune_window;
  o.reset_opacity = cadence && !topo_frozen && cfg.opacity_reset_mode == 1 && cfg.opacity_reset_interval > 0 &&
                    epoch > 0 && epoch % cfg.opacity_reset_interval == 0;
  o.allow_extra_growth = o.densify && phase < cfg.growth_stop_iter;
  return o;
}

}  // namespace rgs

using namespace rgs;

namespace {

uint64_t topology_seed(uint64_t base, uint32_t iteration) {  // topology/mod.rs:1031-1036
  const uint64_t x = base * 0x9e3779b97f4a7c15ull + (uint64_t)iteration;
  return (x << 17) | (x >> 47);
}

// Scratch of one topology step (released at its end: the step runs every
// few hundred iterations).
struct TopoScratch {
  DevBuf partial, ctr, flags, opac, gw, used, keys, kb, vb, kc, vc, sort, sel, surv, offs, scan, counts, rsrc, rkind,
      rsc;
  void release_all() {
    DevBuf* b[] = {&partial, &ctr, &flags, &opac, &gw, &used, &keys, &kb, &vb, &kc, &vc, &sort, &sel, &surv, &offs,
                   &scan, &counts, &rsrc, &rkind, &rsc};
    for (DevBuf* x : b) release(*x);
  }
};

int read_u32(rgs_ctx* c, const DevBuf& b, uint32_t* out, int count) {
  HIPCHK(c, hipMemcpyAsync(out, b.p, 4ull * count, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

// Weighted sampling pass: keys of the eligible sources, sorted; the first
// min(k, eligible) become selected[offset ...].  Returns the number taken.
int sample_pass(rgs_ctx* c, TopoScratch& S, uint32_t n, uint64_t seed, int pass, uint64_t k, uint32_t offset,
                uint32_t* taken) {
  *taken = 0;
  launch_fill_u32(c->stream, S.counts.as<uint32_t>() + kTopoEligible, 0u, 1);
  launch_topo_keys(c->stream, S.flags.as<uint8_t>(), S.opac.as<float>(), S.gw.as<float>(), S.used.as<uint8_t>(), n,
                   seed, pass, S.keys.as<uint32_t>(), S.counts.as<uint32_t>());
  uint32_t cnt[kTopoCountNum];
  int rc = read_u32(c, S.counts, cnt, kTopoCountNum);
  if (rc) return rc;
  const uint64_t take = std::min<uint64_t>(k, cnt[kTopoEligible]);
  if (take == 0) return RGS_OK;
  const int where = radix_sort_pairs(c->stream, S.keys.as<uint32_t>(), nullptr, S.kb.as<uint32_t>(),
                                     S.vb.as<uint32_t>(), S.kc.as<uint32_t>(), S.vc.as<uint32_t>(), nullptr, n, n, 32,
                                     S.sort.as<uint32_t>(), nullptr, nullptr, 0);
  const uint32_t* vals = where == 1 ? S.vb.as<uint32_t>() : S.vc.as<uint32_t>();
  launch_topo_take(c->stream, vals, (uint32_t)take, S.sel.as<uint32_t>(), offset, S.used.as<uint8_t>());
  *taken = (uint32_t)take;
  return RGS_OK;
}

}  // namespace

extern "C" {

int rgs_topology_should_apply(const rgs_topology_cfg* cfg, uint32_t iteration, uint32_t frame_count) {
  if (!cfg) return 0;
  const rgs_topology_schedule s = topology_schedule(*cfg, iteration, frame_count);
  return (s.densify || s.prune || s.reset_opacity) ? 1 : 0;
}

int rgs_topology_get_schedule(const rgs_topology_cfg* cfg, uint32_t iteration, uint32_t frame_count,
                              rgs_topology_schedule* out) {
  if (!cfg || !out) return RGS_ERR_INVALID;
  *out = topology_schedule(*cfg, iteration, frame_count);
  return RGS_OK;
}

int rgs_topology_export_state(rgs_ctx* c, const rgs_splats* s, uint32_t* birth, uint32_t* invisible_windows) {
  if (!c || !s) return RGS_ERR_INVALID;
  if (birth) HIPCHK(c, hipMemcpyAsync(birth, s->birth.p, 4ull * s->n, hipMemcpyDefault, c->stream));
  if (invisible_windows)
    HIPCHK(c, hipMemcpyAsync(invisible_windows, s->invis.p, 4ull * s->n, hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_topology_import_state(rgs_ctx* c, rgs_splats* s, const uint32_t* birth, const uint32_t* invisible_windows) {
  if (!c || !s) return RGS_ERR_INVALID;
  if (birth) HIPCHK(c, hipMemcpyAsync(s->birth.p, birth, 4ull * s->n, hipMemcpyDefault, c->stream));
  if (invisible_windows)
    HIPCHK(c, hipMemcpyAsync(s->invis.p, invisible_windows, 4ull * s->n, hipMemcpyDefault, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return RGS_OK;
}

int rgs_topology_step(rgs_ctx* c, rgs_splats* s, const rgs_topology_cfg* cfg, uint32_t iteration,
                      uint32_t frame_count, rgs_topology_result* res) {
  if (!c || !s || !cfg || !res) return RGS_ERR_INVALID;
  if (cfg->split_score_mode > 3 || cfg->prune_mode > 2 || cfg->opacity_reset_mode > 1)
    return fail(c, RGS_ERR_INVALID, "topology: unknown split score, prune or opacity reset mode");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->step_pending) {
    HIPCHK(c, hipEventSynchronize(c->step_done));
    c->step_pending = false;
  }
  std::memset(res, 0, sizeof(*res));
  const rgs_topology_schedule sch = topology_schedule(*cfg, iteration, frame_count);
  res->schedule = sch;
  const uint32_t n = (uint32_t)s->n;
  res->before = n;
  const uint64_t seed = topology_seed(cfg->frame_shuffle_seed, iteration);
  hipStream_t st = c->stream;
  TopoScratch S;
  auto cleanup = [&](int rc) {
    S.release_all();
    return rc;
  };
  {
    int rc = 0;
    const uint32_t blocks = (n + 1023) / 1024;
    rc |= ensure(c, S.partial, 8ull * 3 * std::max<uint32_t>(blocks, 1));
    rc |= ensure(c, S.ctr, 4 * 8);
    rc |= ensure(c, S.flags, n);
    rc |= ensure(c, S.opac, 4ull * n);
    rc |= ensure(c, S.gw, 4ull * n);
    rc |= ensure(c, S.used, n + 4);
    rc |= ensure(c, S.keys, 4ull * n);
    rc |= ensure(c, S.kb, 4ull * n);
    rc |= ensure(c, S.vb, 4ull * n);
    rc |= ensure(c, S.kc, 4ull * n);
    rc |= ensure(c, S.vc, 4ull * n);
    rc |= ensure(c, S.sort, 4 * std::max<uint64_t>(radix_scratch_words(n), 4096));
    rc |= ensure(c, S.sel, 4ull * n);
    rc |= ensure(c, S.counts, 4 * kTopoCountNum);
    if (rc) return cleanup(RGS_ERR_HIP);
  }
  launch_fill_u32(st, S.counts.as<uint32_t>(), 0u, kTopoCountNum);
  launch_fill_u32(st, S.used.as<uint32_t>(), 0u, (n + 3) / 4);
  TopoParams T{};
  T.n = n;
  T.npad = s->npad;
  T.nsh = 3u * (uint32_t)s->nk;
  T.iteration = iteration;
  T.densify = sch.densify;
  T.prune = sch.prune;
  T.reset_opacity = sch.reset_opacity;
  T.split_score_mode = cfg->split_score_mode;
  T.prune_mode = cfg->prune_mode;
  T.prune_min_age = cfg->prune_min_age;
  T.prune_invisible_epochs = cfg->prune_invisible_epochs;
  T.growth_grad_threshold = cfg->growth_grad_threshold;
  T.split_grad_threshold = cfg->split_grad_threshold;
  T.depth_scale_gamma = cfg->depth_scale_gamma;
  T.prune_opacity_threshold = cfg->prune_opacity_threshold;
  T.prune_visibility_threshold = cfg->prune_visibility_threshold;
  T.prune_high_opacity_threshold = cfg->prune_high_opacity_threshold;
  T.seed = seed;
  launch_topo_analyze(st, s->P.as<float>(), s->dens.as<float>(), s->birth.as<uint32_t>(), s->invis.as<uint32_t>(), T,
                      S.partial.as<double>(), S.ctr.as<float>(), S.flags.as<uint8_t>(), S.opac.as<float>(),
                      S.gw.as<float>(), S.counts.as<uint32_t>());
  uint32_t cnt[kTopoCountNum];
  if (int rc = read_u32(c, S.counts, cnt, kTopoCountNum)) return cleanup(rc);
  const uint64_t P = cnt[kTopoPrune], G = cnt[kTopoGrowth];
  res->prune_candidates = (int64_t)P;
  res->growth_candidates = (int64_t)G;
  res->clone_candidates = cnt[kTopoClone];
  res->split_candidates = cnt[kTopoSplit];

  // selection (topology/mod.rs:634-765)
  const uint64_t grow = (uint64_t)std::max(0.0f, std::round((float)G * cfg->growth_select_fraction));
  const uint64_t requested = sch.allow_extra_growth ? P + (grow > P ? grow - P : 0) : P;
  const uint64_t budget = std::max<uint64_t>(cfg->max_initial_gaussians, 1);
  const uint64_t max_new = std::max<uint64_t>(budget, (uint64_t)n + requested) - n;
  uint32_t k1 = 0, k2 = 0;
  if (max_new > 0) {
    const uint64_t rcount = std::min<uint64_t>(std::min<uint64_t>(P, max_new), cnt[kTopoReplacementSources]);
    if (rcount > 0)
      if (int rc = sample_pass(c, S, n, seed, 1, rcount, 0, &k1)) return cleanup(rc);
    if (sch.allow_extra_growth && k1 < max_new) {
      const uint64_t high = grow > P ? grow - P : 0;
      const uint64_t limit = std::min<uint64_t>(high, max_new - k1);
      if (limit > 0)
        if (int rc = sample_pass(c, S, n, seed, 2, limit, k1, &k2)) return cleanup(rc);
    }
  }
  res->replacement_count = k1;
  res->extra_growth_count = k2;
  res->selected_count = k1 + k2;

  // plan (topology/mod.rs:895-981)
  const bool all_survive = P == n;
  const uint32_t survivors = all_survive ? n : n - (uint32_t)P;
  const uint32_t added = sch.densify ? k1 + k2 : 0;
  const uint32_t pruned = n - survivors;
  const bool decay = (sch.densify || sch.prune) && (cfg->opacity_decay != 0.0f || cfg->scale_decay != 0.0f);
  const bool mutates = added > 0 || pruned > 0 || decay || sch.reset_opacity;
  res->added = added;
  res->pruned = pruned;
  res->mutated = mutates;
  res->adam_reset = mutates;  // requires_adam_rebuild || apply_opacity_reset (trainer.rs:851-853)
  if (!mutates) {
    // reset_accumulators (trainer.rs:854-871); the splats and Adam state stay
    launch_fill_f32(st, s->dens.as<float>(), 0.0f, (size_t)s->npad * 10);
    HIPCHK(c, hipStreamSynchronize(st));
    res->after = n;
    return cleanup(RGS_OK);
  }
  const uint32_t nn = survivors + added;
  {
    int rc = 0;
    rc |= ensure(c, S.surv, 4ull * n);
    rc |= ensure(c, S.offs, 4ull * n);
    rc |= ensure(c, S.scan, 4 * 4096);
    rc |= ensure(c, S.rsrc, 4ull * nn);
    rc |= ensure(c, S.rkind, nn);
    rc |= ensure(c, S.rsc, 4ull * nn);
    if (rc) return cleanup(RGS_ERR_HIP);
  }
  launch_topo_survive(st, S.flags.as<uint8_t>(), n, all_survive ? 1u : 0u, S.surv.as<uint32_t>());
  scan_u32(st, S.surv.as<uint32_t>(), nullptr, nullptr, n, n, S.offs.as<uint32_t>(), 0, nullptr, nullptr,
           S.scan.as<uint32_t>());
  launch_topo_rows(st, S.surv.as<uint32_t>(), S.offs.as<uint32_t>(), n, S.sel.as<uint32_t>(), added, survivors, seed,
                   S.rsrc.as<uint32_t>(), S.rkind.as<uint8_t>(), S.rsc.as<float>());

  // the rebuilt store (apply.rs:29-188): new planes, Adam state and
  // accumulators cleared, birth and windows remapped
  rgs_splats* t = nullptr;
  if (int rc = rgs_splats_create(c, nn, s->deg, &t)) return cleanup(rc);
  TopoApply A{};
  A.nn = nn;
  A.old_npad = s->npad;
  A.new_npad = t->npad;
  A.nplanes = (uint32_t)s->nplanes;
  A.iteration = iteration;
  A.decay = decay;
  A.reset_opacity = sch.reset_opacity;
  A.scale_decay = cfg->scale_decay;
  A.opacity_decay = cfg->opacity_decay;
  A.train_t = cfg->iterations == 0 ? 1.0f : (float)iteration / (float)cfg->iterations;
  launch_topo_apply(st, s->P.as<float>(), S.rsrc.as<uint32_t>(), S.rkind.as<uint8_t>(), S.rsc.as<float>(),
                    s->birth.as<uint32_t>(), s->invis.as<uint32_t>(), A, t->P.as<float>(), t->birth.as<uint32_t>(),
                    t->invis.as<uint32_t>());
  hipError_t e = hipMemcpyAsync(t->lr.p, s->lr.p, 4ull * s->nplanes, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    rgs_splats_destroy(c, t);
    return cleanup(fail(c, RGS_ERR_HIP, "topology rebuild: %s", hipGetErrorString(e)));
  }
  // the caller's handle keeps its identity: swap the storage
  invalidate_graph(c);
  c->sized.clear();
  c->groups.clear();
  c->fwd_valid = false;
  std::swap(*s, *t);
  rgs_splats_destroy(c, t);
  res->after = nn;
  return cleanup(RGS_OK);
}

}  // extern "C"
