// Host side of the evaluation renderer and the per-frame metrics:
// rgs_tiled_render (TiledRenderer::render_camera_splats,
// RustGS/src/render/tiled_renderer.rs:174-492), rgs_render_splats
// (GaussianRenderer::render_splats, render/renderer.rs:72-98),
// rgs_image_metrics (compute_psnr_f32 and the sharpness measures,
// training/evaluation/core.rs:238-319) and rgs_evaluate_frame
// (SplatEvaluationRenderer::render + the metrics of one frame,
// evaluation/core.rs:362-377, 431-449).  The kernels are in k_eval.hip.
// One small read-back per render (the live-Gaussian and instance counts)
// sizes the instance buffers; everything else stays on the device.
#include <cmath>

#include "rgs_host.h"

namespace rgs {

void launch_eval_project(int deg, hipStream_t s, const float* P, uint32_t npad, uint32_t n, const EvalCam& C,
                         float4* rec, uint2* box, uint32_t* live, uint32_t* ntl, uint32_t* dkey,
                         unsigned long long* total);
void launch_eval_compact(hipStream_t s, const uint32_t* live, const uint32_t* pos, const uint32_t* dkey, uint32_t n,
                         uint32_t* ck, uint32_t* cv);
void launch_eval_gather(hipStream_t s, const uint32_t* order, uint32_t V, const float4* rec, const uint2* box,
                        float4* rec_s, uint2* box_s);
void launch_eval_expand(hipStream_t s, const uint32_t* offs, const uint2* box_s, uint32_t V, uint32_t T, uint32_t nbx,
                        uint32_t* brank, uint32_t* keys, uint32_t* vals);
void launch_tile_ranges_n(hipStream_t s, const uint32_t* keys, uint32_t n, uint32_t num_tiles, uint32_t* ranges);
void launch_eval_raster(hipStream_t s, const float4* rec_s, const uint2* box_s, const uint32_t* vals,
                        const uint32_t* ranges, const EvalCam& C, const float bg[3], float* color, uint8_t* color8,
                        float* depth, int depth_zero);
uint32_t eval_metrics_blocks(uint32_t W, uint32_t H);
void launch_eval_metrics(hipStream_t s, const float* rend, const uint8_t* rend8, const float* tgt, uint32_t W,
                         uint32_t H, double* parts, double* out);
int radix_sort_pairs(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                     uint32_t* kc, uint32_t* vc, const uint32_t* n_dev, uint32_t n_host, uint32_t cap, int key_bits,
                     uint32_t* scratch, const uint32_t* gsrc, uint32_t* gout, int count_ib);
int radix_sort_pairs_packed(hipStream_t s, const uint32_t* kin, const uint32_t* vin, uint32_t* kb, uint32_t* vb,
                            uint32_t* kout, uint32_t* vout, const uint32_t* n_dev, uint32_t n_host, uint32_t cap,
                            int key_bits, int value_bits, uint32_t* scratch);
uint64_t radix_scratch_words(uint64_t cap);
void scan_u32(hipStream_t s, const uint32_t* in, const uint32_t* gather, const uint32_t* n_dev, uint32_t n_host,
              uint32_t cap, uint32_t* out, int inclusive, uint32_t* total_dev, uint32_t* max_dev, uint32_t* scratch);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);

}  // namespace rgs

using namespace rgs;

namespace {

// Instance limit of one evaluation render (the sort's index range).
constexpr uint64_t kEvalMaxInstances = 1ull << 31;

int check_camera(rgs_ctx* c, const rgs_camera* cam, const char* who) {
  if (cam->width == 0 || cam->height == 0 || cam->width > kEvalMaxSide || cam->height > kEvalMaxSide)
    return fail(c, RGS_ERR_INVALID, "%s: image size %ux%u outside 1..%u", who, cam->width, cam->height, kEvalMaxSide);
  return RGS_OK;
}

EvalCam eval_camera(const rgs_camera* cam, float cov_blur) {
  EvalCam C{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) C.r[i * 3 + j] = cam->viewmat[j * 4 + i];  // camera_rotation_rows
    C.t[i] = cam->viewmat[12 + i];
    C.pos[i] = cam->cam_pos[i];
  }
  C.fx = cam->fx;
  C.fy = cam->fy;
  C.cx = cam->cx;
  C.cy = cam->cy;
  C.blur = std::fmax(cov_blur, 0.0f);  // with_raster_cov_blur: raster_cov_blur.max(0.0)
  C.W = cam->width;
  C.H = cam->height;
  C.ntx = (C.W + kEvalTile - 1) / kEvalTile;
  C.nty = (C.H + kEvalTile - 1) / kEvalTile;
  C.nbx = (C.ntx + 1) / 2;
  C.nby = (C.nty + 1) / 2;
  return C;
}

int eval_render(rgs_ctx* c, const rgs_splats* s, const rgs_camera* cam, const float* background, float cov_blur,
                float* color, uint8_t* color8, float* depth, int depth_zero, const char* who) {
  if (!c || !s || !cam) return RGS_ERR_INVALID;
  int rc = check_camera(c, cam, who);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  EvalScratch& E = c->eval;
  const hipStream_t st = c->stream;
  const EvalCam C = eval_camera(cam, cov_blur);
  const float bg[3] = {background ? background[0] : 0.0f, background ? background[1] : 0.0f,
                       background ? background[2] : 0.0f};
  const uint32_t nbins = C.nbx * C.nby;  // the per-tile lists are kept per binning tile
  const uint32_t n = (uint32_t)s->n;
  rc |= ensure(c, E.ranges, (size_t)nbins * 8);
  rc |= ensure(c, E.counts, 16);
  if (n) {
    rc |= ensure(c, E.rec, (size_t)n * 48);
    rc |= ensure(c, E.box, (size_t)n * 8);
    rc |= ensure(c, E.live, (size_t)n * 4);
    rc |= ensure(c, E.ntl, (size_t)n * 4);
    rc |= ensure(c, E.dkey, (size_t)n * 4);
    rc |= ensure(c, E.pos, (size_t)n * 4);
    rc |= ensure(c, E.ck, (size_t)n * 4);
    rc |= ensure(c, E.cv, (size_t)n * 4);
    rc |= ensure(c, E.scan, 4096 * 4);
  }
  if (rc) return RGS_ERR_HIP;
  launch_fill_u32(st, E.ranges.as<uint32_t>(), 0u, (size_t)nbins * 2);
  uint32_t V = 0;
  uint64_t T = 0;
  if (n) {
    uint32_t* counts = E.counts.as<uint32_t>();  // [0, 1] instance total (u64), [2] live Gaussians
    launch_fill_u32(st, counts, 0u, 4);
    launch_eval_project(s->deg, st, s->P.as<float>(), s->npad, n, C, E.rec.as<float4>(), E.box.as<uint2>(),
                        E.live.as<uint32_t>(), E.ntl.as<uint32_t>(), E.dkey.as<uint32_t>(),
                        reinterpret_cast<unsigned long long*>(counts));
    scan_u32(st, E.live.as<uint32_t>(), nullptr, nullptr, n, n, E.pos.as<uint32_t>(), 0, counts + 2, nullptr,
             E.scan.as<uint32_t>());
    launch_eval_compact(st, E.live.as<uint32_t>(), E.pos.as<uint32_t>(), E.dkey.as<uint32_t>(), n,
                        E.ck.as<uint32_t>(), E.cv.as<uint32_t>());
    uint32_t h[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(h, counts, 16, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    T = (uint64_t)h[0] | ((uint64_t)h[1] << 32);
    V = h[2];
    if (T >= kEvalMaxInstances)
      return fail(c, RGS_ERR_CAPACITY, "%s: %llu tile instances exceed the limit of %llu", who,
                  (unsigned long long)T, (unsigned long long)kEvalMaxInstances);
  }
  const uint32_t* vals = nullptr;
  if (T) {
    const uint32_t Ti = (uint32_t)T;
    rc |= ensure(c, E.kb, (size_t)V * 4);
    rc |= ensure(c, E.vb, (size_t)V * 4);
    rc |= ensure(c, E.offs, (size_t)V * 4);
    rc |= ensure(c, E.rec_s, (size_t)V * 48);
    rc |= ensure(c, E.box_s, (size_t)V * 8);
    rc |= ensure(c, E.ik0, (size_t)Ti * 4);
    rc |= ensure(c, E.iv0, (size_t)Ti * 4);
    rc |= ensure(c, E.ik1, (size_t)Ti * 4);
    rc |= ensure(c, E.iv1, (size_t)Ti * 4);
    rc |= ensure(c, E.sort, radix_scratch_words(std::max<uint64_t>(V, T)) * 4);
    rc |= ensure(c, E.brank, ((size_t)Ti + 255) / 256 * 4);
    if (rc) return RGS_ERR_HIP;
    uint32_t* sort = E.sort.as<uint32_t>();
    // depth order of the live Gaussians: positive float depths sort as their bits
    const int w1 = radix_sort_pairs(st, E.ck.as<uint32_t>(), E.cv.as<uint32_t>(), E.kb.as<uint32_t>(),
                                    E.vb.as<uint32_t>(), E.ck.as<uint32_t>(), E.cv.as<uint32_t>(), nullptr, V, V, 31,
                                    sort, nullptr, nullptr, 0);
    const uint32_t* order = w1 == 1 ? E.vb.as<uint32_t>() : E.cv.as<uint32_t>();
    scan_u32(st, E.ntl.as<uint32_t>(), order, nullptr, V, V, E.offs.as<uint32_t>(), 0, nullptr, nullptr,
             E.scan.as<uint32_t>());
    launch_eval_gather(st, order, V, E.rec.as<float4>(), E.box.as<uint2>(), E.rec_s.as<float4>(),
                       E.box_s.as<uint2>());
    launch_eval_expand(st, E.offs.as<uint32_t>(), E.box_s.as<uint2>(), V, Ti, C.nbx, E.brank.as<uint32_t>(),
                       E.ik0.as<uint32_t>(), E.iv0.as<uint32_t>());
    // per-tile lists: stable sort by tile id keeps the depth order inside a
    // tile (two packed passes when the tile id and the record index fit one
    // word, as in the training step's tile sort)
    const int w2 = radix_sort_pairs_packed(st, E.ik0.as<uint32_t>(), E.iv0.as<uint32_t>(), E.ik1.as<uint32_t>(),
                                           E.iv1.as<uint32_t>(), E.ik0.as<uint32_t>(), E.iv0.as<uint32_t>(), nullptr,
                                           Ti, Ti, (int)ceil_log2(nbins), std::max(1, (int)ceil_log2(V)), sort);
    const uint32_t* keys = w2 == 1 ? E.ik1.as<uint32_t>() : E.ik0.as<uint32_t>();
    vals = w2 == 1 ? E.iv1.as<uint32_t>() : E.iv0.as<uint32_t>();
    launch_tile_ranges_n(st, keys, Ti, nbins, E.ranges.as<uint32_t>());
  }
  launch_eval_raster(st, E.rec_s.as<float4>(), E.box_s.as<uint2>(), vals, E.ranges.as<uint32_t>(), C, bg, color,
                     color8, depth, depth_zero);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(st));
  return RGS_OK;
}

// Final float32 arithmetic of compute_psnr_f32, compute_gradient_sharpness_f32,
// compute_laplacian_sharpness_f32 and sharpness_ratio on the float64 sums.
void finish_metrics(const double* S, uint32_t W, uint32_t H, rgs_frame_metrics* out) {
  const uint64_t n3 = (uint64_t)W * H * 3;
  rgs_frame_metrics m{};
  if (n3) {
    const float mse = (float)S[0] / (float)n3;
    m.psnr_db = mse <= 1e-12f ? 100.0f : 10.0f * std::log10(1.0f / mse);
  }
  auto grad = [&](double sum) -> float {
    if (W < 2 || H < 2) return 0.0f;
    const uint64_t cnt = (uint64_t)(W - 1) * H + (uint64_t)W * (H - 1);
    return (float)sum / (float)cnt;
  };
  auto lap = [&](double sum, double sum_sq) -> float {
    if (W < 3 || H < 3) return 0.0f;
    const float cnt = (float)((uint64_t)(W - 2) * (H - 2));
    const float mean = (float)sum / cnt;
    return std::fmax((float)sum_sq / cnt - mean * mean, 0.0f);
  };
  auto ratio = [](float r, float t) { return t <= 1e-12f ? 0.0f : r / t; };
  m.render_grad = grad(S[1]);
  m.target_grad = grad(S[2]);
  m.render_lap = lap(S[3], S[4]);
  m.target_lap = lap(S[5], S[6]);
  m.sharpness_grad_ratio = ratio(m.render_grad, m.target_grad);
  m.sharpness_lap_ratio = ratio(m.render_lap, m.target_lap);
  *out = m;
}

int eval_metrics(rgs_ctx* c, const float* rend, const uint8_t* rend8, const float* tgt, uint32_t W, uint32_t H,
                 rgs_frame_metrics* out) {
  double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if ((uint64_t)W * H > 0) {
    int rc = ensure(c, c->eval.parts, (size_t)eval_metrics_blocks(W, H) * kEvalSums * 8);
    rc |= ensure(c, c->eval.sums, kEvalSums * 8);
    if (rc) return RGS_ERR_HIP;
    launch_eval_metrics(c->stream, rend, rend8, tgt, W, H, c->eval.parts.as<double>(), c->eval.sums.as<double>());
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(S, c->eval.sums.p, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  finish_metrics(S, W, H, out);
  return RGS_OK;
}

}  // namespace

int rgs_tiled_render(rgs_ctx* c, const rgs_splats* s, const rgs_camera* cam, const float* background, float cov_blur,
                     float* color, float* depth) {
  return eval_render(c, s, cam, background, cov_blur, color, nullptr, depth, 0, "rgs_tiled_render");
}

int rgs_render_splats(rgs_ctx* c, const rgs_splats* s, const rgs_camera* cam, const float* background,
                      float cov_blur, uint8_t* color, float* depth) {
  return eval_render(c, s, cam, background, cov_blur, nullptr, color, depth, 1, "rgs_render_splats");
}

int rgs_image_metrics(rgs_ctx* c, const float* rendered, const float* target, uint32_t width, uint32_t height,
                      rgs_frame_metrics* out) {
  if (!c || !out || ((uint64_t)width * height > 0 && (!rendered || !target))) return RGS_ERR_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  return eval_metrics(c, rendered, nullptr, target, width, height, out);
}

int rgs_evaluate_frame(rgs_ctx* c, const rgs_splats* s, const rgs_camera* cam, float cov_blur, const float* target,
                       uint8_t* color, rgs_frame_metrics* out) {
  if (!c || !s || !cam || !target || !out) return RGS_ERR_INVALID;
  int rc = check_camera(c, cam, "rgs_evaluate_frame");
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  uint8_t* img = color;
  if (!img) {
    if (ensure(c, c->eval.img8, (size_t)cam->width * cam->height * 3)) return RGS_ERR_HIP;
    img = c->eval.img8.as<uint8_t>();
  }
  rc = eval_render(c, s, cam, nullptr, cov_blur, nullptr, img, nullptr, 1, "rgs_evaluate_frame");
  if (rc) return rc;
  return eval_metrics(c, nullptr, img, target, cam->width, cam->height, out);
}
