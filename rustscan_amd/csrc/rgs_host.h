// Host-side state of the library (context, splat store, workspace) shared by
// the host source files behind include/rgs.h.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rgs.h"
#include "rgs_internal.h"

namespace rgs {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

inline uint32_t ceil_log2(uint64_t x) {
  uint32_t b = 0;
  while ((1ull << b) < x) ++b;
  return b;
}

}  // namespace rgs

using rgs::DevBuf;
using rgs::ViewDev;

struct rgs_splats {
  int64_t n = 0;
  uint32_t npad = 0;
  int deg = 0, nk = 1, nplanes = 11;
  DevBuf P, G, M, V, stats, visible, lr;
  // densification accumulators [10][npad] (trainer.rs:66-75) and one step's
  // increments [5][npad] (see k_densify_commit)
  DevBuf dens, dinc;
  DevBuf step;  // Adam step counters of the three groups (AdamState::step, optimizer.rs:27-32)
  // topology state per Gaussian (trainer.rs:76-77): birth iteration and
  // consecutive topology steps without visibility
  DevBuf birth, invis;
};

struct Workspace {
  uint32_t n = 0;
  int view_slots = 0;
  size_t cap_inst = 0;   // n * view_slots
  size_t cap_isect = 0;  // intersection capacity
  uint32_t tiles_cap = 0;
  DevBuf views, proj, keys, vals, cnt, scnt, sk1, sv1, sk2, sv2, offs, ik0, iv0, ik1, iv1, ranges, acc, acc_dbg, counters,
      fscal, scratch, visi;
  std::vector<DevBuf> own_img;  // images kept internal when the caller passes none
  // combined loss: per view, the SSIM partial-derivative maps and the image gradient
  std::vector<DevBuf> loss_grad, loss_vout;
  DevBuf loss_parts;  // partial loss sums per view (kept zero between uses)
};

struct ViewSet {
  std::vector<ViewDev> v;
  uint32_t tiles_total = 0;
  int tw = 16;
};

struct StepPlanKey {
  int64_t n;
  int deg, nviews, mode, tw, batch, persistent, densify;
  std::vector<uint32_t> dims;
  std::vector<uint32_t> loss;  // bits of the loss configuration (captured as kernel arguments)
  bool operator<(const StepPlanKey& o) const {
    return std::tie(n, deg, nviews, mode, tw, batch, persistent, densify, dims, loss) <
           std::tie(o.n, o.deg, o.nviews, o.mode, o.tw, o.batch, o.persistent, o.densify, o.dims, o.loss);
  }
  bool operator==(const StepPlanKey& o) const { return !(*this < o) && !(o < *this); }
};

// Scratch of the evaluation renderer and metrics (grown on demand, kept
// between calls: evaluation renders one frame per call).
struct EvalScratch {
  DevBuf rec, box, live, ntl, dkey, pos, ck, cv, kb, vb, offs, rec_s, box_s, ik0, iv0, ik1, iv1, ranges, sort, scan,
      counts, parts, sums, img8, brank;
  void release_all() {
    DevBuf* b[] = {&rec, &box, &live, &ntl, &dkey, &pos, &ck, &cv, &kb, &vb, &offs, &rec_s, &box_s,
                   &ik0, &iv0, &ik1, &iv1, &ranges, &sort, &scan, &counts, &parts, &sums, &img8, &brank};
    for (DevBuf* x : b) {
      if (x->p) hipFree(x->p);
      x->p = nullptr;
      x->bytes = 0;
    }
  }
};

// A forward state of its own (RenderCheckpoint, backward/autodiff.rs:56-71):
// the workspace that holds the projected splats, tile lists and images of one
// rgs_render_forward_checkpoint call, the forward-state fields, and the
// parameter planes the forward read.  Swapped into the context for the calls
// that use it.
struct rgs_checkpoint {
  Workspace ws;
  bool fwd_valid = false;
  const rgs_splats* fwd_splats = nullptr;
  int fwd_v0 = 0, fwd_nv = 0;
  uint32_t* fwd_order = nullptr;
  uint32_t* fwd_isect_keys = nullptr;
  uint32_t* fwd_isect_vals = nullptr;
  ViewSet fwd_views;
  float fwd_blur = 0.0f;
  bool bwd_done = false;
  DevBuf params;
  int64_t n = 0;
  uint32_t npad = 0;
  int deg = -1;
};

struct rgs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int num_cus = 256;
  rgs_raster_cfg raster{16, 256, 1};
  Workspace ws;
  // last forward state
  bool fwd_valid = false;
  const rgs_splats* fwd_splats = nullptr;
  int fwd_v0 = 0, fwd_nv = 0;
  uint32_t* fwd_order = nullptr;
  uint32_t* fwd_isect_keys = nullptr;
  uint32_t* fwd_isect_vals = nullptr;
  ViewSet fwd_views;
  float fwd_blur = 0.0f;
  bool bwd_done = false;
  bool keep_raster_grads = false;
  // graph
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  bool graph_valid = false;
  StepPlanKey gkey{};
  const rgs_splats* gsplats = nullptr;
  std::map<StepPlanKey, bool> sized;
  // batched mode: view groups [starts[i], starts[i+1]) launched together, so
  // that no launch holds more than launch_limit intersections
  std::map<StepPlanKey, std::vector<int>> groups;
  uint64_t launch_limit = 1600000000ull;
  // host mirrors
  uint32_t* h_counters = nullptr;
  float* h_fscal = nullptr;
  // counters block as the last completed training step left it: every step
  // ends with a device-to-host copy into this pinned block (a node of the
  // captured graph too), read by the next call after step_done
  uint32_t* h_step = nullptr;
  hipEvent_t step_done = nullptr;
  bool step_pending = false;
  std::map<std::vector<uint32_t>, rgs_raster_cfg> tuned;
  // the single view of rgs_image_loss_grad (kept apart from the forward state)
  DevBuf loss_view;
  EvalScratch eval;
  // per-stage timing
  bool timing = false;
  bool capturing = false;
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  double stage_ms[RGS_STAGE_COUNT] = {};
  int64_t stage_calls[RGS_STAGE_COUNT] = {};
  // fragment counters (device, kFragCounters x u64) when enabled
  DevBuf frag;
  bool count_frags = false;
  bool ckpt_active = false;  // a checkpoint's workspace is swapped in (the train graph's stays valid)
  bool plain_sorts = false;  // debug: unpacked tile sort, count gather in the depth sort
  bool no_cull = false;      // debug: no per-block culling in the compositing kernels
};

namespace rgs {

inline int fail(rgs_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(ctx, RGS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

inline int ensure(rgs_ctx* c, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return RGS_OK;
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (bytes == 0) return RGS_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(c, RGS_ERR_HIP, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return RGS_OK;
}

inline void release(DevBuf& b) {
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Records HIP events around one stage when timing is on (never while capturing).
struct StageTimer {
  rgs_ctx* c;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(rgs_ctx* ctx, int st) : c(ctx), stage(st) {
    if (c->timing && !c->capturing) {
      hipEventCreate(&a);
      hipEventCreate(&b);
      hipEventRecord(a, c->stream);
    }
  }
  ~StageTimer() {
    if (a) {
      hipEventRecord(b, c->stream);
      c->pending.push_back({stage, a, b});
    }
  }
};

inline void invalidate_graph(rgs_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->graph_valid = false;
}


}  // namespace rgs
