// The training loss of RustGS (engine/loss.rs) and its image gradient, for
// the views of a step:
//   combined_loss_with_kernel (loss.rs:112-142) =
//     l1 * residual (L1, saturating robust or soft-outlier, loss.rs:144-196)
//   + ssim * (1 - mean SSIM) (11-tap separable Gaussian, sigma 1.5, edge
//     padding; loss.rs:42-76, 233-276)
//   + gradient * image-gradient difference (loss.rs:213-227)
//
// k_loss_fwd: one 256-thread workgroup per 16x16 pixel block of a view.  For
// each channel it stages the 26x26 block with its 5-pixel halo in LDS
// (coordinates clamped to the image: that is the edge padding of both blur
// passes), blurs x, y, x^2, y^2 and xy horizontally (26 rows) and then
// vertically, evaluates the SSIM map, and writes the three per-pixel partial
// derivatives of the SSIM map with respect to mu_x, E[x^2] and E[xy]
// (scaled by -ssim / (3HW)).  It also sums the view's loss.
// k_loss_bwd applies the adjoint of the blur to those maps (a separable
// 11-tap pass whose taps outside the image fold onto the edge pixel) and adds
// the residual and gradient-difference derivatives, writing the [H][W][4]
// image gradient the compositing backward consumes.
#include <cmath>

#include "rgs_device.h"

namespace rgs {

constexpr int kLB = 16;            // output block edge
constexpr int kLR = 5;             // blur radius (11 taps)
constexpr int kLS = kLB + 2 * kLR; // staged edge, 26
constexpr int kLossParts = 64;     // partial loss sums per view

__device__ __forceinline__ void residual_terms(const LossCoef& L, float d, float* r, float* dr) {
  const float a = fabsf(d);
  const float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
  if (L.residual == 1) {  // a * delta / (a + delta)
    *r = (a * L.delta) / (a + L.delta);
    *dr = sg * (L.delta * L.delta) / ((a + L.delta) * (a + L.delta));
  } else if (L.residual == 2) {  // a * (floor + (1 - floor) * thr / (a + thr))
    const float w = (a * 0.0f + L.floor_w) + ((a * 0.0f + (1.0f - L.floor_w)) * L.delta) / (a + L.delta);
    *r = a * w;
    *dr = sg * (L.floor_w + (1.0f - L.floor_w) * L.delta * L.delta / ((a + L.delta) * (a + L.delta)));
  } else {
    *r = a;
    *dr = sg;
  }
}

__device__ __forceinline__ bool view_block(const ViewDev& V, uint32_t* ox, uint32_t* oy) {
  const uint32_t bx = (V.W + kLB - 1) / kLB, by = (V.H + kLB - 1) / kLB;
  if (blockIdx.x >= bx * by) return false;
  *ox = (blockIdx.x % bx) * kLB;
  *oy = (blockIdx.x / bx) * kLB;
  return true;
}

__global__ __launch_bounds__(256) void k_loss_fwd(const ViewDev* __restrict__ views, int v0, LossCoef L,
                                                  float* __restrict__ loss_parts) {
  const int v = v0 + (int)blockIdx.y;
  const ViewDev& V = views[v];
  uint32_t ox, oy;
  if (!view_block(V, &ox, &oy)) return;
  const int W = (int)V.W, H = (int)V.H;
  const size_t np = (size_t)W * H;
  const float* __restrict__ pred = V.img;
  const float* __restrict__ tgt = V.target;
  // the staged block of all three channels (one 16-byte load per pixel of
  // the prediction), then one channel at a time through the blur passes
  __shared__ float sxa[3][kLS][kLS + 1], sya[3][kLS][kLS + 1];
  __shared__ float hb[5][kLS][kLB + 1];
  const int tid = threadIdx.x;
  const int px = tid % kLB, py = tid / kLB;
  const int gx = (int)ox + px, gy = (int)oy + py;
  const bool inside = gx < W && gy < H;
  const float N = 3.0f * (float)np;
  float lsum = 0.0f;
  for (int i = tid; i < kLS * kLS; i += 256) {
    const int r = i / kLS, col = i % kLS;
    const int yy = min(max((int)oy + r - kLR, 0), H - 1), xx = min(max((int)ox + col - kLR, 0), W - 1);
    const size_t p = (size_t)yy * W + xx;
    const float4 q = reinterpret_cast<const float4*>(pred)[p];
    sxa[0][r][col] = q.x;
    sxa[1][r][col] = q.y;
    sxa[2][r][col] = q.z;
    sya[0][r][col] = tgt[p * 3];
    sya[1][r][col] = tgt[p * 3 + 1];
    sya[2][r][col] = tgt[p * 3 + 2];
  }
  __syncthreads();
  for (int c = 0; c < 3; ++c) {
#pragma clang fp contract(fast)
    float (*sx)[kLS + 1] = sxa[c];
    float (*sy)[kLS + 1] = sya[c];
    for (int i = tid; i < kLS * kLB; i += 256) {
      const int r = i / kLB, col = i % kLB;
      float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f, m4 = 0.f;
#pragma unroll
      for (int t = 0; t < 11; ++t) {
        const float a = sx[r][col + t], b = sy[r][col + t], w = L.k[t];
        m0 = m0 + a * w;
        m1 = m1 + b * w;
        m2 = m2 + (a * a) * w;
        m3 = m3 + (b * b) * w;
        m4 = m4 + (a * b) * w;
      }
      hb[0][r][col] = m0; hb[1][r][col] = m1; hb[2][r][col] = m2; hb[3][r][col] = m3; hb[4][r][col] = m4;
    }
    __syncthreads();
    if (inside) {
      float mu_x = 0.f, mu_y = 0.f, exx = 0.f, eyy = 0.f, exy = 0.f;
#pragma unroll
      for (int t = 0; t < 11; ++t) {
        const float w = L.k[t];
        mu_x = mu_x + hb[0][py + t][px] * w;
        mu_y = mu_y + hb[1][py + t][px] * w;
        exx = exx + hb[2][py + t][px] * w;
        eyy = eyy + hb[3][py + t][px] * w;
        exy = exy + hb[4][py + t][px] * w;
      }
      const float mxx = mu_x * mu_x, myy = mu_y * mu_y, mxy = mu_x * mu_y;
      const float sxx = exx - mxx, syy = eyy - myy, sxy = exy - mxy;
      const float a1 = mxy * 2.0f + L.c1, a2 = sxy * 2.0f + L.c2;
      const float b1r = (mxx + myy) + L.c1, b2r = (sxx + syy) + L.c2;
      const float b1 = fmaxf_(b1r, 1e-8f), b2 = fmaxf_(b2r, 1e-8f);
      // one division: the quotients below are products with 1 / (b1 b2)
      const float inv = 1.0f / (b1 * b2);
      const float s = (a1 * a2) * inv;
      // partial derivatives of s (clamp_min passes no gradient where it clamps)
      const float d_a1 = a2 * inv, d_a2 = a1 * inv;
      const float d_b1 = b1r > 1e-8f ? -s * (b2 * inv) : 0.0f, d_b2 = b2r > 1e-8f ? -s * (b1 * inv) : 0.0f;
      const float g_mu = 2.0f * mu_y * (d_a1 - d_a2) + 2.0f * mu_x * (d_b1 - d_b2);
      const float scale = -L.ssim / N;
      float* gm = V.lgrad + (size_t)(3 * c) * np + (size_t)gy * W + gx;
      gm[0] = g_mu * scale;
      gm[np] = d_b2 * scale;
      gm[2 * np] = 2.0f * d_a2 * scale;
      // loss, in units of 1 / (3HW)
      const float x0 = sx[py + kLR][px + kLR], y0 = sy[py + kLR][px + kLR];
      float r, dr;
      residual_terms(L, x0 - y0, &r, &dr);
      lsum += L.l1 * r + L.ssim * (1.0f - s);
      if (L.grad != 0.0f) {
        if (gx + 1 < W) {
          const float e = (sx[py + kLR][px + kLR + 1] - x0) - (sy[py + kLR][px + kLR + 1] - y0);
          lsum += L.grad * 0.5f * fabsf(e) * (N / (3.0f * (float)H * (float)(W - 1)));
        }
        if (gy + 1 < H) {
          const float e = (sx[py + kLR + 1][px + kLR] - x0) - (sy[py + kLR + 1][px + kLR] - y0);
          lsum += L.grad * 0.5f * fabsf(e) * (N / (3.0f * (float)(H - 1) * (float)W));
        }
      }
    }
    __syncthreads();
  }
  if (L.degenerate_nan && (W == 1 || H == 1) && blockIdx.x == 0 && tid == 0) lsum += __builtin_nanf("");
  // one partial per block, spread over kLossParts slots per view (one
  // address per view would serialise the atomics of every block on it)
  __shared__ float wsum[4];
  const float s = wave_sum(lsum);
  if ((tid & 63) == 0) wsum[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    const float b = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    if (b != 0.0f) atomicAdd(&loss_parts[(size_t)v * kLossParts + (blockIdx.x % kLossParts)], b);
  }
}

// Adds each view's partial sums to its loss and clears them for the next use.
__global__ __launch_bounds__(kLossParts) void k_loss_finish(float* __restrict__ loss_parts, int v0,
                                                            float* __restrict__ loss_sums) {
  const int v = v0 + (int)blockIdx.x;
  float x = loss_parts[(size_t)v * kLossParts + threadIdx.x];
  loss_parts[(size_t)v * kLossParts + threadIdx.x] = 0.0f;
  x = wave_sum(x);
  if (threadIdx.x == 0) loss_sums[v] += x;
}

// Adjoint of one clamped 11-tap pass at position q of a line of n samples:
// taps from p in [q-5, q+5] that land on q, plus (at the two ends) the taps
// that the clamp folded onto q.  at(p) returns the gradient at sample p, and
// 0 for p outside [0, n) (the callers stage zeros there).
// (Tap indices are compile-time constants: the coefficients are kernel
// arguments, which a run-time index would have to fetch from memory.)
// The loss and its gradient are compared within a relative tolerance, so
// the blur sums here and in k_loss_fwd use fused multiply-adds.
template <class F>
__device__ __forceinline__ float adjoint_tap(const LossCoef& L, int q, int n, F at) {
#pragma clang fp contract(fast)
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < 11; ++t) acc += at(q + kLR - t) * L.k[t];  // sample q + 5 - t reaches q through tap t
  if (q == 0) {  // taps of samples 0..4 that fell below the image
#pragma unroll
    for (int p = 0; p < kLR; ++p)
      if (p < n) acc += at(p) * L.kcum[kLR - 1 - p];
  }
  if (q == n - 1) {  // taps of samples n-1..n-5 that fell beyond it
#pragma unroll
    for (int j = 0; j < kLR; ++j) {
      const int p = n - 1 - j;
      if (p >= 0) acc += at(p) * L.kdown[kLR + 1 + j];
    }
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_loss_bwd(const ViewDev* __restrict__ views, int v0, LossCoef L) {
  const int v = v0 + (int)blockIdx.y;
  const ViewDev& V = views[v];
  uint32_t ox, oy;
  if (!view_block(V, &ox, &oy)) return;
  const int W = (int)V.W, H = (int)V.H;
  const size_t np = (size_t)W * H;
  __shared__ float G[3][kLS][kLS + 1];
  __shared__ float Vt[3][kLB][kLS + 1];
  const int tid = threadIdx.x;
  const int px = tid % kLB, py = tid / kLB;
  const int gx = (int)ox + px, gy = (int)oy + py;
  const bool inside = gx < W && gy < H;
  const float N = 3.0f * (float)np;
  float out[3] = {0.f, 0.f, 0.f};
  for (int c = 0; c < 3; ++c) {
    for (int i = tid; i < 3 * kLS * kLS; i += 256) {
      const int m = i / (kLS * kLS), j = i % (kLS * kLS);
      const int r = j / kLS, col = j % kLS;
      const int yy = (int)oy + r - kLR, xx = (int)ox + col - kLR;
      G[m][r][col] = (yy >= 0 && yy < H && xx >= 0 && xx < W)
                         ? V.lgrad[(size_t)(3 * c + m) * np + (size_t)yy * W + xx]
                         : 0.0f;
    }
    __syncthreads();
    // vertical adjoint for the block's 16 rows, all 26 staged columns
    for (int i = tid; i < 3 * kLB * kLS; i += 256) {
      const int m = i / (kLB * kLS), j = i % (kLB * kLS);
      const int r = j / kLS, col = j % kLS;
      const int q = (int)oy + r;
      Vt[m][r][col] = q < H ? adjoint_tap(L, q, H, [&](int p) { return G[m][p - (int)oy + kLR][col]; }) : 0.0f;
    }
    __syncthreads();
    if (inside) {
      float d[3];
#pragma unroll
      for (int m = 0; m < 3; ++m)
        d[m] = adjoint_tap(L, gx, W, [&](int p) { return Vt[m][py][p - (int)ox + kLR]; });
      const size_t pix = (size_t)gy * W + gx;
      const float x = V.img[pix * 4 + c], y = V.target[pix * 3 + c];
      float g = d[0] + 2.0f * x * d[1] + y * d[2];
      float r, dr;
      residual_terms(L, x - y, &r, &dr);
      g += dr * (L.l1 / N);
      if (L.grad != 0.0f) {
        const float wx = L.grad * 0.5f / (3.0f * (float)H * (float)(W - 1));
        const float wy = L.grad * 0.5f / (3.0f * (float)(H - 1) * (float)W);
        auto sgn = [](float e) { return e > 0.0f ? 1.0f : (e < 0.0f ? -1.0f : 0.0f); };
        auto diff = [&](size_t a, size_t b) {  // (pred[a] - pred[b]) - (tgt[a] - tgt[b])
          return (V.img[a * 4 + c] - V.img[b * 4 + c]) - (V.target[a * 3 + c] - V.target[b * 3 + c]);
        };
        if (gx + 1 < W) g -= wx * sgn(diff(pix + 1, pix));
        if (gx > 0) g += wx * sgn(diff(pix, pix - 1));
        if (gy + 1 < H) g -= wy * sgn(diff(pix + W, pix));
        if (gy > 0) g += wy * sgn(diff(pix, pix - W));
      }
      out[c] = g * V.loss_weight;
    }
    __syncthreads();
  }
  if (inside)
    reinterpret_cast<float4*>(const_cast<float*>(V.vout))[(size_t)gy * W + gx] =
        make_float4(out[0], out[1], out[2], 0.0f);
}

// ------------------------------------------------------------------ launch
void make_loss_coef(const float* w6, int residual, int degenerate_nan, LossCoef* L) {
  L->degenerate_nan = degenerate_nan;
  L->l1 = w6[0];
  L->ssim = w6[1];
  L->grad = w6[2];
  L->residual = residual;
  L->delta = residual == 1 ? fmaxf(w6[3], 1e-6f) : fmaxf(w6[4], 1e-6f);
  L->floor_w = fminf(fmaxf(w6[5], 0.0f), 1.0f);
  // SsimConfig defaults (loss.rs:18-27): k1 0.01, k2 0.03, data range 1
  L->c1 = (float)((0.01 * 1.0) * (0.01 * 1.0));
  L->c2 = (float)((0.03 * 1.0) * (0.03 * 1.0));
  // gaussian_kernel_1d (loss.rs:233-253): exp in double, f32 sum and division
  float vals[11], sum = 0.0f;
  for (int o = -kLR; o <= kLR; ++o) {
    vals[o + kLR] = (float)std::exp(-(double)(o * o) / (2.0 * 1.5 * 1.5));
    sum += vals[o + kLR];
  }
  const float den = fmaxf(sum, 1e-8f);
  for (int t = 0; t < 11; ++t) L->k[t] = vals[t] / den;
  float run = 0.0f;
  for (int t = 0; t < 11; ++t) {
    run += L->k[t];
    L->kcum[t] = run;
  }
  run = 0.0f;
  for (int t = 10; t >= 0; --t) {
    run += L->k[t];
    L->kdown[t] = run;
  }
}

uint32_t loss_blocks(uint32_t W, uint32_t H) { return ((W + kLB - 1) / kLB) * ((H + kLB - 1) / kLB); }

// loss_parts: kLossParts floats per view of the batch (all zero on entry;
// left zero), indexed by global view id like loss_sums.
void launch_loss(hipStream_t s, const ViewDev* views, int v0, int nv, uint32_t max_blocks, const LossCoef& L,
                 float* loss_sums, float* loss_parts) {
  if (nv <= 0 || max_blocks == 0) return;
  const dim3 grid(max_blocks, (uint32_t)nv);
  k_loss_fwd<<<grid, 256, 0, s>>>(views, v0, L, loss_parts);
  k_loss_finish<<<(uint32_t)nv, kLossParts, 0, s>>>(loss_parts, v0, loss_sums);
  k_loss_bwd<<<grid, 256, 0, s>>>(views, v0, L);
}

size_t loss_parts_floats(int views) { return (size_t)views * kLossParts; }

}  // namespace rgs
