// Device-side building blocks shared by the projection, binning and raster
// kernels.  The arithmetic of every function is written out operation by
// operation (files are compiled without fp contraction) because the binning
// decisions must be reproducible bit for bit; each cites the WGSL it computes.
#pragma once

#include "rgs_internal.h"

namespace rgs {

// Constants are the reference's WGSL abstract-float literals rounded to f32.
#define RGS_F(x) ((float)(x))

// Margin of the compositing kernels' exp skip: a fragment with
// sigma > log(255 * opacity) + kSkipMargin has alpha < 1/255 whatever exp
// rounds to, so it is rejected without evaluating exp.
constexpr float kSkipMargin = 1e-3f;

__device__ __forceinline__ float fminf_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmaxf_(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clampf_(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }

// 2x2 matrix, column-major: c{col}{row}
struct M2 {
  float c00, c01, c10, c11;
};

__device__ __forceinline__ float det2(const M2& m) { return m.c00 * m.c11 - m.c01 * m.c10; }

// helpers.wgsl:185-198 inverse2x2 (with the +1e-6 diagonal regulariser)
__device__ __forceinline__ M2 inverse2x2(const M2& m) {
  M2 r = m;
  r.c00 = r.c00 + RGS_F(1e-6);
  r.c11 = r.c11 + RGS_F(1e-6);
  const float det = det2(r);
  if (det <= RGS_F(1e-10)) return M2{0.f, 0.f, 0.f, 0.f};
  const float inv = 1.0f / det;
  return M2{r.c11 * inv, -r.c01 * inv, -r.c01 * inv, r.c00 * inv};
}

// helpers.wgsl:204-216 get_tile_bbox: (min_x, min_y, max_x, max_y) in tiles
__device__ __forceinline__ void tile_bbox(float cx_, float cy_, float ex, float ey, uint32_t tbx, uint32_t tby,
                                          int tw, uint32_t bb[4]) {
  // tw is a power of two, so x * (1/tw) rounds exactly as x / tw does
  const float inv = 1.0f / (float)tw;
  const float c0 = cx_ * inv, c1 = cy_ * inv;
  const float e0 = ex * inv, e1 = ey * inv;
  const float bx = (float)tbx, by = (float)tby;
  bb[0] = (uint32_t)clampf_(c0 - e0, 0.0f, bx);
  bb[1] = (uint32_t)clampf_(c1 - e1, 0.0f, by);
  bb[2] = (uint32_t)clampf_(c0 + e0 + 1.0f, 0.0f, bx);
  bb[3] = (uint32_t)clampf_(c1 + e1 + 1.0f, 0.0f, by);
}

// helpers.wgsl:218-224 compute_bbox_extent
__device__ __forceinline__ void bbox_extent(float cov00, float cov11, float thr, float* ex, float* ey) {
  const float safe = fmaxf_(thr, 0.0f);
  *ex = rgs_sqrt(fmaxf_(0.0f, 2.0f * safe * cov00));
  *ey = rgs_sqrt(fmaxf_(0.0f, 2.0f * safe * cov11));
}

// helpers.wgsl:232-235 calc_sigma (delta = pixel - mean)
__device__ __forceinline__ float calc_sigma(float px, float py, float a, float b, float c, float mx, float my) {
  const float dx = px - mx, dy = py - my;
  return 0.5f * (a * dx * dx + 2.0f * b * dx * dy + c * dy * dy);
}

// helpers.wgsl:237-283 will_primitive_contribute for tile (tx, ty)
__device__ __forceinline__ bool tile_contributes(uint32_t tx, uint32_t ty, int tw, float mx, float my, float a,
                                                 float b, float c, float thr) {
  const float rx0 = (float)(tx * (uint32_t)tw), ry0 = (float)(ty * (uint32_t)tw);
  const float rx1 = rx0 + (float)tw, ry1 = ry0 + (float)tw;
  const bool x_left = mx < rx0, x_right = mx > rx1;
  const bool in_x = !(x_left || x_right);
  const bool y_above = my < ry0, y_below = my > ry1;
  const bool in_y = !(y_above || y_below);
  if (in_x && in_y) return true;
  const float ccx = x_left ? rx0 : rx1, ccy = y_above ? ry0 : ry1;
  const float width = rx1 - rx0, height = ry1 - ry0;
  const float dxr = x_left ? width : -width, dyr = y_above ? height : -height;
  const float diffx = mx - ccx, diffy = my - ccy;
  float t0 = 0.0f, t1 = 0.0f;
  if (!in_y)
    t0 = clampf_((dxr * a * diffx + dxr * b * diffy) / fmaxf_(dxr * a * dxr, RGS_F(1e-12)), 0.0f, 1.0f);
  if (!in_x)
    t1 = clampf_((dyr * b * diffx + dyr * c * diffy) / fmaxf_(dyr * c * dyr, RGS_F(1e-12)), 0.0f, 1.0f);
  const float mpx = ccx + t0 * dxr, mpy = ccy + t1 * dyr;
  return calc_sigma(mpx, mpy, a, b, c, mx, my) <= thr;
}

__device__ __forceinline__ float pack_u16x2(uint32_t lo, uint32_t hi) {
  return __uint_as_float((lo & 0xffffu) | (hi << 16));
}
__device__ __forceinline__ void unpack_u16x2(float f, uint32_t* lo, uint32_t* hi) {
  const uint32_t u = __float_as_uint(f);
  *lo = u & 0xffffu;
  *hi = u >> 16;
}

// Tiles of bbox bb passing the tile test (the K5 emission rule), visited in
// the reference's row-major order (map_gaussian_to_intersects.wgsl:37-40).
// For bboxes of at most 64 tiles the passing set is returned as a bit mask.
__device__ __forceinline__ uint32_t count_tiles(const uint32_t bb[4], int tw, float mx, float my, float a, float b,
                                                float c, float thr, uint64_t* mask) {
  const uint32_t bw = bb[2] - bb[0];
  const uint32_t nb = (bb[3] - bb[1]) * bw;
  uint32_t hit = 0;
  uint64_t m = 0;
  uint32_t k = 0;
  for (uint32_t ty = bb[1]; ty < bb[3]; ++ty)
    for (uint32_t tx = bb[0]; tx < bb[2]; ++tx, ++k)
      if (tile_contributes(tx, ty, tw, mx, my, a, b, c, thr)) {
        ++hit;
        if (k < 64) m |= 1ull << k;
      }
  *mask = nb <= 64 ? m : 0ull;
  return hit;
}

// Row and column of linear index k in a bbox of width bw (row-major), with
// inv_bw = 1/bw: a float estimate of k / bw corrected to the exact quotient.
__device__ __forceinline__ void bbox_rc(uint32_t k, uint32_t bw, float inv_bw, uint32_t* row, uint32_t* col) {
  int32_t q = (int32_t)((float)k * inv_bw);
  int32_t r = (int32_t)k - q * (int32_t)bw;
  if (r < 0) {
    --q;
    r += (int32_t)bw;
  } else if (r >= (int32_t)bw) {
    ++q;
    r -= (int32_t)bw;
  }
  *row = (uint32_t)q;
  *col = (uint32_t)r;
}

// Tiles of a bbox passing the tile test, counted by the whole wave (all 64
// lanes active, arguments wave-uniform): 64 bbox tiles per step, in the
// reference's row-major order.  Used for bboxes of more than 64 tiles, whose
// per-lane loop would stall the wave.
__device__ __forceinline__ uint32_t wave_count_tiles(const uint32_t bb[4], int tw, float mx, float my, float a,
                                                     float b, float c, float thr) {
  const uint32_t bw = bb[2] - bb[0];
  const uint32_t nb = (bb[3] - bb[1]) * bw;
  const float inv = 1.0f / (float)bw;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t total = 0;
  for (uint32_t k0 = 0; k0 < nb; k0 += 64) {
    const uint32_t k = k0 + lane;
    bool hit = false;
    if (k < nb) {
      uint32_t row, col;
      bbox_rc(k, bw, inv, &row, &col);
      hit = tile_contributes(bb[0] + col, bb[1] + row, tw, mx, my, a, b, c, thr);
    }
    total += (uint32_t)__popcll(__ballot(hit));
  }
  return total;
}

// Inclusive scans over the 64 lanes of a wave (every lane active): DPP row
// shifts within each 16-lane row, then row 0's and row 2's totals into rows 1
// and 3 (row_bcast:15), then lane 31's into rows 2 and 3 (row_bcast:31).
// Disabled rows and out-of-row sources read 0, the identity of both ops.
#define RGS_DPP_U32(v, ctrl, rm) (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(v), ctrl, rm, 0xf, true)
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t v) {
  v += RGS_DPP_U32(v, 0x111, 0xf);
  v += RGS_DPP_U32(v, 0x112, 0xf);
  v += RGS_DPP_U32(v, 0x114, 0xf);
  v += RGS_DPP_U32(v, 0x118, 0xf);
  v += RGS_DPP_U32(v, 0x142, 0xa);
  v += RGS_DPP_U32(v, 0x143, 0xc);
  return v;
}
__device__ __forceinline__ uint32_t umax_(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
  v = umax_(v, RGS_DPP_U32(v, 0x111, 0xf));
  v = umax_(v, RGS_DPP_U32(v, 0x112, 0xf));
  v = umax_(v, RGS_DPP_U32(v, 0x114, 0xf));
  v = umax_(v, RGS_DPP_U32(v, 0x118, 0xf));
  v = umax_(v, RGS_DPP_U32(v, 0x142, 0xa));
  v = umax_(v, RGS_DPP_U32(v, 0x143, 0xc));
  return v;
}
#undef RGS_DPP_U32

// Orders this wave's LDS accesses before and after it (the LDS executes one
// wave's accesses in issue order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront"); }

__device__ __forceinline__ int find_view(const ViewDev* __restrict__ views, int nv, uint32_t tile) {
  int v = 0;
  for (int i = 1; i < nv; ++i)
    if (tile >= views[i].tile_base) v = i;
  return v;
}

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Which 8x8 pixel blocks of a workgroup a staged splat can reach.  Bit
// bx + by*(RW/8) is cleared only when every pixel centre of block (bx, by)
// provably has sigma > bound, so the wave of that block would find no pixel
// to composite and may skip the splat without changing any result.
// The bound is the exact minimum of q(dx, dy) = (a dx^2 + 2b dx dy + c dy^2)/2
// over the block's rectangle of (dx, dy) (rounding is monotone, so the float
// dx, dy of every pixel centre lie between those of the extreme centres).
// For a positive-definite conic (a, c > 0, ac - b^2 > 0) q is convex and
// 2-homogeneous, so when the rectangle misses the origin its minimum lies on
// an edge facing the origin (scaling any point towards the origin lowers q
// and meets such an edge inside the rectangle); along the edge dx = X the
// minimum is at dy = clamp(-bX/c, Y0, Y1).  The argmin is found with the
// hardware reciprocal: its relative error moves q by second-order amounts.
// The float sigma of a pixel and the float edge minimum each lie within
// 4.8e-7*(a dx^2 + c dy^2) of their exact values (four roundings per term,
// 2|b dx dy| <= a dx^2 + c dy^2); E = 2e-6*(a xm^2 + c ym^2), with xm, ym the
// largest |dx|, |dy| over the workgroup, covers both.  NaN or
// non-positive-definite conics keep every bit.
template <int RW>
__device__ __forceinline__ uint32_t block_touch_mask(const float4 q0, float hc, float bound, uint32_t tox,
                                                     uint32_t toy) {
  constexpr int BPR = RW / 8;
  constexpr uint32_t kAll = (1u << (BPR * BPR)) - 1u;
  const float a = 2.0f * q0.z, b = q0.w, c = 2.0f * hc;
  const float det_lo = (a * c) * (1.0f - 1e-6f) - (b * b) * (1.0f + 1e-6f);
  if (!(a > 0.0f && c > 0.0f && det_lo > 0.0f)) return kAll;
  const float ra = __builtin_amdgcn_rcpf(a), rc = __builtin_amdgcn_rcpf(c);
  // dx = x - pcx over the pixel centres pcx = tox + 8i + {0.5 .. 7.5}
  float x0[BPR], x1[BPR], y0[BPR], y1[BPR];
#pragma unroll
  for (int i = 0; i < BPR; ++i) {
    x0[i] = q0.x - ((float)(tox + 8 * i + 7) + 0.5f);
    x1[i] = q0.x - ((float)(tox + 8 * i) + 0.5f);
    y0[i] = q0.y - ((float)(toy + 8 * i + 7) + 0.5f);
    y1[i] = q0.y - ((float)(toy + 8 * i) + 0.5f);
  }
  const float xm = fmaxf_(fabsf(x0[0]), fabsf(x1[BPR - 1])), ym = fmaxf_(fabsf(y0[0]), fabsf(y1[BPR - 1]));
  const float lim = bound + 2e-6f * (a * xm * xm + c * ym * ym);
  uint32_t bits = 0;
#pragma unroll
  for (int by = 0; by < BPR; ++by)
#pragma unroll
    for (int bx = 0; bx < BPR; ++bx) {
      const float X0 = x0[bx], X1 = x1[bx], Y0 = y0[by], Y1 = y1[by];
      const bool fx = X0 > 0.0f || X1 < 0.0f, fy = Y0 > 0.0f || Y1 < 0.0f;
      const float Xe = X0 > 0.0f ? X0 : X1, Ye = Y0 > 0.0f ? Y0 : Y1;
      const float dy = __builtin_amdgcn_fmed3f(-b * Xe * rc, Y0, Y1);
      const float dx = __builtin_amdgcn_fmed3f(-b * Ye * ra, X0, X1);
      const float qx = 0.5f * ((a * Xe) * Xe + (2.0f * b * Xe + c * dy) * dy);
      const float qy = 0.5f * ((c * Ye) * Ye + (2.0f * b * Ye + a * dx) * dx);
      const float m = fx ? (fy ? fminf_(qx, qy) : qx) : (fy ? qy : 0.0f);
      if (!(m > lim)) bits |= 1u << (bx + by * BPR);
    }
  return bits;
}

}  // namespace rgs
