// Evaluation renderer and per-frame metrics on the device.
//
// The evaluation / `rustgs render` path of the reference renders with its CPU
// renderer (TiledRenderer::render_camera_splats,
// RustGS/src/render/tiled_renderer.rs:174-492, wrapped by
// GaussianRenderer::render_splats, render/renderer.rs:72-98), whose
// semantics differ from the training rasterizer: colour from sh.rs's
// evaluate_sh_rgb, no EWA clip, tile lists from a 3-sigma box of the blurred
// covariance with its own rounding, a per-tile stable depth sort, and an
// alpha-accumulation rule (alpha += w (1 - alpha) while alpha < 0.999,
// weight cut at 0.001, Mahalanobis cut at 9, depth from the first splat with
// alpha > 0.01).  These kernels reproduce that arithmetic operation for
// operation (no contraction), with the C library's expf bits for exp
// (rgs_expf_libm), so that they match the renderer with the libm it links,
// and the oracle's restatement of it, bit for bit.
//
// Pipeline (one view):
//   k_eval_project<DEG>  per Gaussian: colour, projection, culling, tile box
//   (scan + k_eval_compact + radix sort of depth bits)  depth order of the
//                        live Gaussians, ties by index (stable)
//   k_eval_gather        records and tile boxes into depth order
//   k_eval_block_ranks, k_eval_expand  one thread per (Gaussian, tile) instance
//   (radix sort of binning-tile ids, stable)  per-tile lists in depth order
//   (tile ranges)        [start, end) of every tile's list
//   k_eval_raster        one 16x16 workgroup per tile, an 8x8 block per wave
// Metrics (evaluation/core.rs:238-319): k_eval_metrics sums the float32
// per-pixel terms of PSNR and the two sharpness measures in float64 per
// workgroup, k_eval_metrics_final adds the workgroup sums in a fixed order.
#include <cfloat>

#include "rgs_device.h"
#include "rgs_math.h"

namespace rgs {

namespace {

// Rust `f as usize` of a non-negative clamp result, saturated to 32 bits
// (every tile bound that matters is far below 2^32).
__device__ __forceinline__ uint32_t as_index(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xffffffffu;
  return (uint32_t)f;
}

// sh.rs:35-112 (evaluate_sh_rgb), f32 literals, left to right.
template <int DEG>
__device__ __forceinline__ void sh_rgb(const float* __restrict__ P, uint32_t npad, uint32_t g, float x, float y,
                                       float z, float col[3]) {
  auto co = [&](int k, int ch) { return P[(size_t)(11 + 3 * k + ch) * npad + g]; };
  auto add = [&](int k, float s) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) col[ch] += co(k, ch) * s;
  };
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) col[ch] = 0.2820948f * co(0, ch);
  if constexpr (DEG >= 1) {
    add(1, 0.48860252f * -y);
    add(2, 0.48860252f * z);
    add(3, 0.48860252f * -x);
  }
  if constexpr (DEG >= 2) {
    const float z2 = z * z, f0b = -1.0925485f * z, f1a = 0.54627424f, fc1 = x * x - y * y, fs1 = 2.0f * x * y;
    add(4, f1a * fs1);
    add(5, f0b * y);
    add(6, 0.9461747f * z2 - 0.31539157f);
    add(7, f0b * x);
    add(8, f1a * fc1);
    if constexpr (DEG >= 3) {
      const float f0c = -2.285229f * z2 + 0.4570458f, f1b = 1.4453057f * z, f2a = -0.5900436f;
      const float fc2 = x * fc1 - y * fs1, fs2 = x * fs1 + y * fc1;
      add(9, f2a * fs2);
      add(10, f1b * fs1);
      add(11, f0c * y);
      add(12, z * (1.8658817f * z2 - 1.119529f));
      add(13, f0c * x);
      add(14, f1b * fc1);
      add(15, f2a * fc2);
      if constexpr (DEG >= 4) {
        const float f0d = z * (-4.683326f * z2 + 2.0071397f), f1c = 3.3116114f * z2 - 0.47308734f;
        const float f2b = -1.7701308f * z, f3a = 0.6258357f;
        const float fc3 = x * fc2 - y * fs2, fs3 = x * fs2 + y * fc2;
        const float p12 = z * (1.8658817f * z2 - 1.119529f), p6 = 0.9461747f * z2 - 0.31539157f;
        add(16, f3a * fs3);
        add(17, f2b * fs2);
        add(18, f1c * fs1);
        add(19, f0d * y);
        add(20, 1.9843135f * z * p12 - 1.0062306f * p6);
        add(21, f0d * x);
        add(22, f1c * fc1);
        add(23, f2b * fc2);
        add(24, f3a * fc3);
      }
    }
  }
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) col[ch] += 0.5f;
}

// Per Gaussian (render_camera_splats' projection loop, tiled_renderer.rs:184-245,
// project_projected_gaussian :389-492, compute_tile_bounds :127-150 with
// tile_alpha 4): the record {x, y, opacity, depth}, {cxx, 2 cxy, cyy,
// 1/det}, {r, g, b, 0} and the tile box, for every Gaussian that can change
// a pixel.  Beyond the reference's culls (depth <= 0, raw determinant <= 0,
// opacity < 1/255) and its per-splat skip in the raster loop (det < 1e-10),
// Gaussians whose centre, determinant or opacity is not finite are dropped:
// in the reference their Mahalanobis distance or weight is NaN at every pixel,
// so they never contribute.  live[g] = 1 iff the box covers a tile.
template <int DEG>
__global__ __launch_bounds__(256) void k_eval_project(const float* __restrict__ P, uint32_t npad, uint32_t n,
                                                      EvalCam C, float4* __restrict__ rec, uint2* __restrict__ box,
                                                      uint32_t* __restrict__ live, uint32_t* __restrict__ ntl,
                                                      uint32_t* __restrict__ dkey,
                                                      unsigned long long* __restrict__ total) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t count = 0;
  if (g < n) {
    const float wx = P[g], wy = P[npad + g], wz = P[2 * (size_t)npad + g];
    float vd[3] = {wx - C.pos[0], wy - C.pos[1], wz - C.pos[2]};
    const float len = sqrtf(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
    if (len <= 1e-12f) {  // normalize3 (tiled_renderer.rs:507-513)
      vd[0] = 0.0f;
      vd[1] = 0.0f;
      vd[2] = 1.0f;
    } else {
      vd[0] = vd[0] / len;
      vd[1] = vd[1] / len;
      vd[2] = vd[2] / len;
    }
    float col[3];
    sh_rgb<DEG>(P, npad, g, vd[0], vd[1], vd[2], col);
    const float* r = C.r;
    const float ccx = r[0] * wx + r[1] * wy + r[2] * wz + C.t[0];
    const float ccy = r[3] * wx + r[4] * wy + r[5] * wz + C.t[1];
    const float ccz = r[6] * wx + r[7] * wy + r[8] * wz + C.t[2];
    bool keep = !(ccz <= 0.0f);
    const float px = C.fx * ccx / ccz + C.cx, py = C.fy * ccy / ccz + C.cy;
    const float qw = P[3 * (size_t)npad + g], qx = P[4 * (size_t)npad + g], qy = P[5 * (size_t)npad + g],
                qz = P[6 * (size_t)npad + g];
    const float r00 = 1.0f - 2.0f * (qy * qy + qz * qz), r01 = 2.0f * (qx * qy - qw * qz),
                r02 = 2.0f * (qx * qz + qw * qy);
    const float r10 = 2.0f * (qx * qy + qw * qz), r11 = 1.0f - 2.0f * (qx * qx + qz * qz),
                r12 = 2.0f * (qy * qz - qw * qx);
    const float r20 = 2.0f * (qx * qz - qw * qy), r21 = 2.0f * (qy * qz + qw * qx),
                r22 = 1.0f - 2.0f * (qx * qx + qy * qy);
    const float sx = fabsf(rgs_expf_libm(P[7 * (size_t)npad + g])), sy = fabsf(rgs_expf_libm(P[8 * (size_t)npad + g])),
                sz = fabsf(rgs_expf_libm(P[9 * (size_t)npad + g]));
    const float sxx = sx * sx, syy = sy * sy, szz = sz * sz;
    const float cw00 = r00 * r00 * sxx + r01 * r01 * syy + r02 * r02 * szz;
    const float cw01 = r00 * r10 * sxx + r01 * r11 * syy + r02 * r12 * szz;
    const float cw02 = r00 * r20 * sxx + r01 * r21 * syy + r02 * r22 * szz;
    const float cw11 = r10 * r10 * sxx + r11 * r11 * syy + r12 * r12 * szz;
    const float cw12 = r10 * r20 * sxx + r11 * r21 * syy + r12 * r22 * szz;
    const float cw22 = r20 * r20 * sxx + r21 * r21 * syy + r22 * r22 * szz;
    const float m00 = r[0] * cw00 + r[1] * cw01 + r[2] * cw02;
    const float m01 = r[0] * cw01 + r[1] * cw11 + r[2] * cw12;
    const float m02 = r[0] * cw02 + r[1] * cw12 + r[2] * cw22;
    const float m10 = r[3] * cw00 + r[4] * cw01 + r[5] * cw02;
    const float m11 = r[3] * cw01 + r[4] * cw11 + r[5] * cw12;
    const float m12 = r[3] * cw02 + r[4] * cw12 + r[5] * cw22;
    const float m20 = r[6] * cw00 + r[7] * cw01 + r[8] * cw02;
    const float m21 = r[6] * cw01 + r[7] * cw11 + r[8] * cw12;
    const float m22 = r[6] * cw02 + r[7] * cw12 + r[8] * cw22;
    const float cc00 = m00 * r[0] + m01 * r[1] + m02 * r[2];
    const float cc01 = m00 * r[3] + m01 * r[4] + m02 * r[5];
    const float cc02 = m00 * r[6] + m01 * r[7] + m02 * r[8];
    const float cc11 = m10 * r[3] + m11 * r[4] + m12 * r[5];
    const float cc12 = m10 * r[6] + m11 * r[7] + m12 * r[8];
    const float cc22 = m20 * r[6] + m21 * r[7] + m22 * r[8];
    const float iz = 1.0f / ccz, iz2 = iz * iz;
    const float jx0 = C.fx * iz, jx2 = -C.fx * ccx * iz2, jy1 = C.fy * iz, jy2 = -C.fy * ccy * iz2;
    float cxx = jx0 * (cc00 * jx0 + cc02 * jx2) + jx2 * (cc02 * jx0 + cc22 * jx2);
    const float cxy = jx0 * (cc01 * jy1 + cc02 * jy2) + jx2 * (cc12 * jy1 + cc22 * jy2);
    float cyy = jy1 * (cc11 * jy1 + cc12 * jy2) + jy2 * (cc12 * jy1 + cc22 * jy2);
    const float det_raw = cxx * cyy - cxy * cxy;
    cxx = cxx + C.blur;
    cyy = cyy + C.blur;
    const float det_bl = fmaxf(cxx * cyy - cxy * cxy, 1e-12f);
    keep = keep && !(det_raw <= 0.0f);
    const float opacity = 1.0f / (1.0f + rgs_expf_libm(-P[10 * (size_t)npad + g])) * sqrtf(det_raw / det_bl);
    keep = keep && !(opacity < 1.0f / 255.0f);
    const float det = cxx * cyy - cxy * cxy;  // the raster loop's determinant
    keep = keep && det >= 1e-10f && __builtin_isfinite(det) && __builtin_isfinite(px) &&
           __builtin_isfinite(py) && __builtin_isfinite(ccz) && opacity == opacity;
    uint32_t txmin = 0, txmax = 0, tymin = 0, tymax = 0;
    if (keep) {
      const float sgx = fmaxf(sqrtf(cxx * 4.0f), 1.0f), sgy = fmaxf(sqrtf(cyy * 4.0f), 1.0f);
      const uint32_t xmin = as_index(fmaxf(px - 3.0f * sgx, 0.0f));
      const uint32_t xmax = as_index(fminf(px + 3.0f * sgx, (float)C.W - 1.0f));
      const uint32_t ymin = as_index(fmaxf(py - 3.0f * sgy, 0.0f));
      const uint32_t ymax = as_index(fminf(py + 3.0f * sgy, (float)C.H - 1.0f));
      txmin = xmin / kEvalTile;
      txmax = (xmax + kEvalTile - 1) / kEvalTile;
      tymin = ymin / kEvalTile;
      tymax = (ymax + kEvalTile - 1) / kEvalTile;
      // instances: the binning tiles (2x2 reference tiles) the box touches
      if (txmin < txmax && tymin < tymax) count = (((txmax + 1) >> 1) - (txmin >> 1)) * (((tymax + 1) >> 1) - (tymin >> 1));
    }
    live[g] = count ? 1u : 0u;
    ntl[g] = count;
    if (count) {
      dkey[g] = __float_as_uint(ccz);
      rec[3 * (size_t)g] = make_float4(px, py, opacity, ccz);
      rec[3 * (size_t)g + 1] = make_float4(cxx, 2.0f * cxy, cyy, 1.0f / det);
      rec[3 * (size_t)g + 2] = make_float4(col[0], col[1], col[2], 0.0f);
      box[g] = make_uint2(txmin | (txmax << 16), tymin | (tymax << 16));
    }
  }
  // instance total: wave sums (a tile count is below 2^20), one atomic per wave
  uint32_t s = count;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += (uint32_t)__shfl_xor((int)s, o, 64);
  if ((threadIdx.x & 63u) == 0 && s) atomicAdd(total, (unsigned long long)s);
}

__global__ __launch_bounds__(256) void k_eval_compact(const uint32_t* __restrict__ live,
                                                      const uint32_t* __restrict__ pos,
                                                      const uint32_t* __restrict__ dkey, uint32_t n,
                                                      uint32_t* __restrict__ ck, uint32_t* __restrict__ cv) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n || !live[g]) return;
  ck[pos[g]] = dkey[g];
  cv[pos[g]] = g;
}

__global__ __launch_bounds__(256) void k_eval_gather(const uint32_t* __restrict__ order, uint32_t V,
                                                     const float4* __restrict__ rec, const uint2* __restrict__ box,
                                                     float4* __restrict__ rec_s, uint2* __restrict__ box_s) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const uint32_t g = order[r];
  rec_s[3 * (size_t)r] = rec[3 * (size_t)g];
  rec_s[3 * (size_t)r + 1] = rec[3 * (size_t)g + 1];
  rec_s[3 * (size_t)r + 2] = rec[3 * (size_t)g + 2];
  box_s[r] = box[g];
}

// Instance i belongs to the depth rank r with offs[r] <= i < offs[r + 1]
// (every live Gaussian has at least one tile, so offs is strictly
// increasing); its k-th binning tile in row-major order over the binning
// tiles its reference-tile box touches.  Expansion works on
// 256-instance blocks; k_eval_block_ranks writes the rank of each block's
// first instance (rank r writes it for the blocks that start in its range).
__global__ __launch_bounds__(256) void k_eval_block_ranks(const uint32_t* __restrict__ offs, uint32_t V, uint32_t T,
                                                          uint32_t* __restrict__ brank) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= V) return;
  const uint32_t s = offs[r], e = r + 1 < V ? offs[r + 1] : T;
  for (uint32_t b = (s + 255u) / 256u; b * 256u < e; ++b) brank[b] = r;
}

// The 256 instances of a block belong to at most 256 consecutive ranks
// (rank r0 + 256 starts past the block): the workgroup stages their offsets
// in LDS and each thread finds its own rank there.
__global__ __launch_bounds__(256) void k_eval_expand(const uint32_t* __restrict__ offs,
                                                     const uint32_t* __restrict__ brank,
                                                     const uint2* __restrict__ box_s, uint32_t V, uint32_t T,
                                                     uint32_t nbx, uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
  __shared__ uint32_t s_off[256];
  const uint32_t r0 = brank[blockIdx.x];
  s_off[threadIdx.x] = r0 + threadIdx.x < V ? offs[r0 + threadIdx.x] : 0xffffffffu;
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T) return;
  // largest q with s_off[q] <= i (s_off[0] <= i)
  uint32_t q = 0;
#pragma unroll
  for (uint32_t step = 128; step > 0; step >>= 1)
    if (s_off[q + step] <= i) q += step;
  const uint32_t lo = r0 + q;
  const uint32_t k = i - s_off[q];
  const uint2 b = box_s[lo];
  const uint32_t bx0 = (b.x & 0xffffu) >> 1, bx1 = ((b.x >> 16) + 1) >> 1, by0 = (b.y & 0xffffu) >> 1;
  const uint32_t wt = bx1 - bx0;  // >= 1: the box of a live Gaussian covers a tile
  uint32_t row, col;
  bbox_rc(k, wt, __builtin_amdgcn_rcpf((float)wt), &row, &col);  // k < 2^20: exact
  keys[i] = (by0 + row) * nbx + bx0 + col;
  vals[i] = lo;
}

// One workgroup per 16x16 tile, one pixel per thread; each 64-lane wave owns
// an 8x8 block of the tile.  The list of the tile's binning tile is staged
// through LDS 256 records at a time; a record whose reference-tile box does
// not hold this tile is not in the reference's list for it and is masked out
// (the filter keeps the list's depth order), the others get a mask of the
// blocks their 3-sigma ellipse
// can reach (block_touch_mask on the Mahalanobis form, bound d^2/2 < 4.5:
// every pixel centre of an unmarked block provably has d^2 >= 9, where the
// reference skips the splat).  The pixel's d^2 below rounds each of its
// three terms at most five times and the mask's coefficients a, b, c once
// each: together within 1e-6 (a dx^2 + c dy^2) of the exact form, inside the
// mask's 2e-6 margin.  A wave walks only the records that reach its
// block, in list order, with the reference's per-pixel arithmetic.  A pixel
// stops once its alpha reaches 0.999 (the reference skips it for every later
// splat, and alpha never decreases); a wave stops when all its pixels have,
// the workgroup when every pixel has.
__global__ __launch_bounds__(256) void k_eval_raster(const float4* __restrict__ rec_s, const uint2* __restrict__ box_s,
                                                     const uint32_t* __restrict__ vals,
                                                     const uint32_t* __restrict__ ranges, uint32_t W, uint32_t H,
                                                     uint32_t ntx, uint32_t nbx, float bg0, float bg1, float bg2,
                                                     float* __restrict__ color, uint8_t* __restrict__ color8,
                                                     float* __restrict__ depth, int depth_zero) {
  __shared__ float4 sa[256], sb[256], sc[256];
  __shared__ uint64_t swords[4][4];  // [block][record / 64]: bit record % 64 reaches the block
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % ntx, ty = tile / ntx;
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t tox = tx * kEvalTile, toy = ty * kEvalTile;
  const uint32_t px = tox + (wv & 1u) * 8u + (lane & 7u), py = toy + (wv >> 1) * 8u + (lane >> 3);
  const bool inside = px < W && py < H;
  const float fpx = (float)px + 0.5f, fpy = (float)py + 0.5f;
  float alpha = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, dep = FLT_MAX;
  bool done = !inside;
  const uint32_t bin = (ty >> 1) * nbx + (tx >> 1);
  const uint32_t start = ranges[2 * (size_t)bin], end = ranges[2 * (size_t)bin + 1];
  for (uint32_t base = start; base < end; base += 256) {
    if (__syncthreads_count(done ? 0 : 1) == 0) break;
    const uint32_t i = base + threadIdx.x;
    uint32_t bits = 0;
    if (i < end) {
      // the box first: only records held by this tile are loaded and staged
      // (a slot left unstaged has an empty mask and is never read)
      const uint32_t r = vals[i];
      const uint2 bx = box_s[r];
      const bool held = tx >= (bx.x & 0xffffu) && tx < (bx.x >> 16) && ty >= (bx.y & 0xffffu) && ty < (bx.y >> 16);
      if (held) {
        const float4 A = rec_s[3 * (size_t)r], B = rec_s[3 * (size_t)r + 1];
        sa[threadIdx.x] = A;
        sb[threadIdx.x] = B;
        sc[threadIdx.x] = rec_s[3 * (size_t)r + 2];
        // d^2 = a dx^2 + 2 b dx dy + c dy^2 with the conic of {cxx, 2 cxy, cyy, 1/det}
        const float a = B.z * B.w, b = -0.5f * B.y * B.w, c = B.x * B.w;
        bits = block_touch_mask<kEvalTile>(make_float4(A.x, A.y, 0.5f * a, b), 0.5f * c, 4.5f, tox, toy);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t m = __builtin_amdgcn_ballot_w64((bits >> k) & 1u);
      if (lane == 0) swords[k][wv] = m;
    }
    __syncthreads();
    const uint32_t nw = (min(256u, end - base) + 63u) / 64u;
    for (uint32_t wi = 0; wi < nw; ++wi) {
      const uint64_t wm = swords[wv][wi];
      uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(wm >> 32)) << 32) |
                   (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)wm);
      while (m != 0) {
        if (__builtin_amdgcn_ballot_w64(!done) == 0) break;
        const uint32_t j = wi * 64u + (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const float4 A = sa[j], B = sb[j];
        const float dx = A.x - fpx, dy = A.y - fpy;
        const float d2 = (B.z * dx * dx - B.y * dx * dy + B.x * dy * dy) * B.w;
        if (!done && d2 < 9.0f) {
          const float w = rgs_expf_libm(-0.5f * d2) * A.z;
          if (w > 0.001f) {
            const float a = w * (1.0f - alpha);
            if (a > 0.0f) {
              const float4 Cc = sc[j];
              c0 = c0 + Cc.x * a;
              c1 = c1 + Cc.y * a;
              c2 = c2 + Cc.z * a;
              alpha = alpha + a;
              if (dep == FLT_MAX && a > 0.01f) dep = A.w;
              if (alpha >= 0.999f) done = true;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  if (!inside) return;
  const float t = 1.0f - alpha;
  float v[3] = {c0 + bg0 * t, c1 + bg1 * t, c2 + bg2 * t};
  const size_t id = (size_t)py * W + px;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float x = v[ch] < 0.0f ? 0.0f : (v[ch] > 1.0f ? 1.0f : v[ch]);  // f32::clamp keeps NaN
    if (color) color[id * 3 + ch] = x;
    if (color8) color8[id * 3 + ch] = x == x ? (uint8_t)(x * 255.0f) : (uint8_t)0;
  }
  if (depth) depth[id] = (depth_zero && dep == FLT_MAX) ? 0.0f : dep;
}

template <typename TR>
__device__ __forceinline__ float pix(const TR* p, size_t i);
template <>
__device__ __forceinline__ float pix<float>(const float* p, size_t i) {
  return p[i];
}
template <>
__device__ __forceinline__ float pix<uint8_t>(const uint8_t* p, size_t i) {
  return (float)p[i] / 255.0f;  // `value as f32 / 255.0` (evaluation/core.rs:372-376)
}

template <typename TR>
__device__ __forceinline__ float lum(const TR* p, uint32_t W, uint32_t x, uint32_t y) {
  const size_t b = ((size_t)y * W + x) * 3;
  return 0.2126f * pix(p, b) + 0.7152f * pix(p, b + 1) + 0.0722f * pix(p, b + 2);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Per-pixel float32 terms of compute_psnr_f32 (core.rs:238-256),
// compute_gradient_sharpness_f32 (:258-280) and
// compute_laplacian_sharpness_f32 (:282-306), summed in float64:
// parts[block][k], k = 0 squared error, 1/2 gradient terms of rendered /
// target, 3/4 Laplacian and its square (rendered), 5/6 the same (target).
template <typename TR>
__global__ __launch_bounds__(256) void k_eval_metrics(const TR* __restrict__ rend, const float* __restrict__ tgt,
                                                      uint32_t W, uint32_t H, double* __restrict__ parts) {
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  const size_t npx = (size_t)W * H;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < npx; id += (size_t)gridDim.x * blockDim.x) {
    const uint32_t x = (uint32_t)(id % W), y = (uint32_t)(id / W);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float d = pix(rend, id * 3 + ch) - tgt[id * 3 + ch];
      s[0] += (double)(d * d);
    }
    const float lr = lum(rend, W, x, y), lt = lum(tgt, W, x, y);
    if (x + 1 < W) {
      s[1] += (double)fabsf(lum(rend, W, x + 1, y) - lr);
      s[2] += (double)fabsf(lum(tgt, W, x + 1, y) - lt);
    }
    if (y + 1 < H) {
      s[1] += (double)fabsf(lum(rend, W, x, y + 1) - lr);
      s[2] += (double)fabsf(lum(tgt, W, x, y + 1) - lt);
    }
    if (x >= 1 && y >= 1 && x + 1 < W && y + 1 < H) {
      const float ar = -4.0f * lr + lum(rend, W, x - 1, y) + lum(rend, W, x + 1, y) + lum(rend, W, x, y - 1) +
                       lum(rend, W, x, y + 1);
      const float at = -4.0f * lt + lum(tgt, W, x - 1, y) + lum(tgt, W, x + 1, y) + lum(tgt, W, x, y - 1) +
                       lum(tgt, W, x, y + 1);
      s[3] += (double)ar;
      s[4] += (double)(ar * ar);
      s[5] += (double)at;
      s[6] += (double)(at * at);
    }
  }
  __shared__ double red[4][7];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const double v = wave_sum(s[k]);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 7)
    parts[(size_t)blockIdx.x * kEvalSums + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// out[k] = sum over blocks of parts[block][k], in a fixed order.
__global__ __launch_bounds__(256) void k_eval_metrics_final(const double* __restrict__ parts, uint32_t blocks,
                                                            double* __restrict__ out) {
  __shared__ double red[4][7];
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < 7; ++k) s[k] += parts[(size_t)b * kEvalSums + k];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const double v = wave_sum(s[k]);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 7) out[threadIdx.x] = ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                                          red[3][threadIdx.x];
}

}  // namespace

void launch_eval_project(int deg, hipStream_t s, const float* P, uint32_t npad, uint32_t n, const EvalCam& C,
                         float4* rec, uint2* box, uint32_t* live, uint32_t* ntl, uint32_t* dkey,
                         unsigned long long* total) {
  if (n == 0) return;
  const dim3 grid((n + 255) / 256);
  switch (deg) {
    case 0: k_eval_project<0><<<grid, 256, 0, s>>>(P, npad, n, C, rec, box, live, ntl, dkey, total); break;
    case 1: k_eval_project<1><<<grid, 256, 0, s>>>(P, npad, n, C, rec, box, live, ntl, dkey, total); break;
    case 2: k_eval_project<2><<<grid, 256, 0, s>>>(P, npad, n, C, rec, box, live, ntl, dkey, total); break;
    case 3: k_eval_project<3><<<grid, 256, 0, s>>>(P, npad, n, C, rec, box, live, ntl, dkey, total); break;
    default: k_eval_project<4><<<grid, 256, 0, s>>>(P, npad, n, C, rec, box, live, ntl, dkey, total); break;
  }
}

void launch_eval_compact(hipStream_t s, const uint32_t* live, const uint32_t* pos, const uint32_t* dkey, uint32_t n,
                         uint32_t* ck, uint32_t* cv) {
  if (n) k_eval_compact<<<(n + 255) / 256, 256, 0, s>>>(live, pos, dkey, n, ck, cv);
}

void launch_eval_gather(hipStream_t s, const uint32_t* order, uint32_t V, const float4* rec, const uint2* box,
                        float4* rec_s, uint2* box_s) {
  if (V) k_eval_gather<<<(V + 255) / 256, 256, 0, s>>>(order, V, rec, box, rec_s, box_s);
}

// brank: (T + 255) / 256 words.
void launch_eval_expand(hipStream_t s, const uint32_t* offs, const uint2* box_s, uint32_t V, uint32_t T, uint32_t nbx,
                        uint32_t* brank, uint32_t* keys, uint32_t* vals) {
  if (T == 0 || V == 0) return;
  k_eval_block_ranks<<<(V + 255) / 256, 256, 0, s>>>(offs, V, T, brank);
  k_eval_expand<<<(T + 255) / 256, 256, 0, s>>>(offs, brank, box_s, V, T, nbx, keys, vals);
}


void launch_eval_raster(hipStream_t s, const float4* rec_s, const uint2* box_s, const uint32_t* vals,
                        const uint32_t* ranges, const EvalCam& C, const float bg[3], float* color, uint8_t* color8,
                        float* depth, int depth_zero) {
  k_eval_raster<<<C.ntx * C.nty, 256, 0, s>>>(rec_s, box_s, vals, ranges, C.W, C.H, C.ntx, C.nbx, bg[0], bg[1],
                                              bg[2], color, color8, depth, depth_zero);
}

uint32_t eval_metrics_blocks(uint32_t W, uint32_t H) {
  const uint64_t px = (uint64_t)W * H;
  const uint64_t b = (px + 255) / 256;
  return (uint32_t)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

void launch_eval_metrics(hipStream_t s, const float* rend, const uint8_t* rend8, const float* tgt, uint32_t W,
                         uint32_t H, double* parts, double* out) {
  const uint32_t G = eval_metrics_blocks(W, H);
  if (rend8)
    k_eval_metrics<uint8_t><<<G, 256, 0, s>>>(rend8, tgt, W, H, parts);
  else
    k_eval_metrics<float><<<G, 256, 0, s>>>(rend, tgt, W, H, parts);
  k_eval_metrics_final<<<1, 256, 0, s>>>(parts, G, out);
}

}  // namespace rgs
