// Fused Adam, standalone L1 loss, and layout conversion kernels.
#include "rgs_device.h"

namespace rgs {

// compiler-rt __powisf2, the routine Rust's f32::powi lowers to for a runtime
// exponent (engine/optimizer.rs:146-147 computes the bias corrections with it).
__host__ __device__ inline float powi_f32(float a, int b) {
  const bool recip = b < 0;
  float r = 1.0f;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

// Advances the Adam step counters of the three parameter groups of
// AdamScaled (transforms, SH, opacity: AdamState::step, optimizer.rs:27-32,
// 144) and computes each group's f32 bias corrections, unless this step was
// invalidated (capacity overflow or non-finite loss).
__global__ void k_adam_prologue(uint32_t* __restrict__ counters, float* __restrict__ fscal,
                                uint32_t* __restrict__ step_ctr, float b1, float b2, int nviews, int check_loss) {
  if (threadIdx.x != 0) return;
  if (check_loss) {
    for (int v = 0; v < nviews; ++v) {
      const float l = fscal[kFLossBase + v];
      if (!(l == l) || fabsf(l) > 3.0e38f) counters[kCntStepValid] = 0u;  // trainer.rs:297-302
    }
  }
  if (counters[kCntStepValid] == 0u) return;
  for (int g = 0; g < kAdamGroups; ++g) {
    const uint32_t step = step_ctr[g] + 1u;  // saturating_add in the reference; 2^32 steps are not reached
    step_ctr[g] = step;
    fscal[kFBc1 + 2 * g] = 1.0f - powi_f32(b1, (int)step);
    fscal[kFBc1 + 2 * g + 1] = 1.0f - powi_f32(b2, (int)step);
  }
}

// engine/optimizer.rs:122-162 for all groups at once: SoA planes, one
// learning rate per plane (lr * scaling column).  28 bytes of HBM traffic per
// parameter (read p, g, m, v; write p, m, v).  Grid: (column blocks, planes);
// each thread updates four consecutive columns with 16-byte accesses (planes
// are npad = multiple of 64 floats long, so every row is 16-byte aligned).
// Columns past n are written back unchanged.
__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, float b1, float b2, float omb1,
                                         float omb2, float bc1, float bc2, float eps, float lr, float wd) {
  if (wd != 0.0f) g = g + p * wd;  // optimizer.rs:128-130
  const float gsq = g * g;
  const float m1 = m * b1 + g * omb1;
  const float m2 = v * b2 + gsq * omb2;
  const float mh = m1 / bc1;
  const float vh = m2 / bc2;
  const float upd = mh / (rgs_sqrt(vh) + eps);
  p = p - upd * lr;
  m = m1;
  v = m2;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ G,
                                              float* __restrict__ M, float* __restrict__ Vv, uint32_t npad,
                                              uint32_t n, const float* __restrict__ lr_planes, float b1, float b2,
                                              float eps, float wd, const float* __restrict__ fscal,
                                              const uint32_t* __restrict__ counters) {
  if (counters[kCntStepValid] == 0u) return;
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;  // float4 column
  const uint32_t c0 = 4 * q;
  if (c0 >= n) return;
  const float omb1 = 1.0f - b1, omb2 = 1.0f - b2;
  const int group = adam_group_of_plane((int)blockIdx.y);
  const float bc1 = fscal[kFBc1 + 2 * group], bc2 = fscal[kFBc1 + 2 * group + 1];
  const float lr = lr_planes[blockIdx.y];
  const size_t i4 = ((size_t)blockIdx.y * npad) / 4 + q;
  float4 p = reinterpret_cast<float4*>(P)[i4];
  const float4 g = reinterpret_cast<const float4*>(G)[i4];
  float4 m = reinterpret_cast<float4*>(M)[i4];
  float4 v = reinterpret_cast<float4*>(Vv)[i4];
  adam_one(p.x, g.x, m.x, v.x, b1, b2, omb1, omb2, bc1, bc2, eps, lr, wd);
  if (c0 + 1 < n) adam_one(p.y, g.y, m.y, v.y, b1, b2, omb1, omb2, bc1, bc2, eps, lr, wd);
  if (c0 + 2 < n) adam_one(p.z, g.z, m.z, v.z, b1, b2, omb1, omb2, bc1, bc2, eps, lr, wd);
  if (c0 + 3 < n) adam_one(p.w, g.w, m.w, v.w, b1, b2, omb1, omb2, bc1, bc2, eps, lr, wd);
  reinterpret_cast<float4*>(P)[i4] = p;
  reinterpret_cast<float4*>(M)[i4] = m;
  reinterpret_cast<float4*>(Vv)[i4] = v;
}

// L1 loss and its image gradient for one image (loss.rs:144-196 with the RGB
// slice of trainer.rs:278): v_out = sign(pred - target) * scale on RGB, 0 on A.
__global__ __launch_bounds__(256) void k_l1(const float* __restrict__ pred, const float* __restrict__ tgt,
                                            uint32_t np, float scale, float* __restrict__ v_out,
                                            float* __restrict__ loss_sum) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float s = 0.0f;
  if (i < np) {
    float g[4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float d = pred[(size_t)i * 4 + c] - tgt[(size_t)i * 3 + c];
      s += fabsf(d);
      g[c] = d > 0.0f ? scale : (d < 0.0f ? -scale : 0.0f);
    }
    g[3] = 0.0f;
    if (v_out) reinterpret_cast<float4*>(v_out)[i] = make_float4(g[0], g[1], g[2], g[3]);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0 && s != 0.0f && loss_sum) atomicAdd(loss_sum, s);
}

// AoS [n][width] (row stride width) <-> SoA planes [width][npad] at plane offset.
__global__ __launch_bounds__(256) void k_aos_to_soa(const float* __restrict__ aos, int width, uint32_t n,
                                                    float* __restrict__ soa, uint32_t npad, int plane0) {
  const size_t total = (size_t)n * width;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t row = (uint32_t)(i / width);
    const int c = (int)(i - (size_t)row * width);
    soa[(size_t)(plane0 + c) * npad + row] = aos[i];
  }
}

__global__ __launch_bounds__(256) void k_soa_to_aos(const float* __restrict__ soa, uint32_t npad, int plane0,
                                                    int width, uint32_t n, float* __restrict__ aos) {
  const size_t total = (size_t)n * width;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t row = (uint32_t)(i / width);
    const int c = (int)(i - (size_t)row * width);
    aos[i] = soa[(size_t)(plane0 + c) * npad + row];
  }
}

__global__ void k_fill_f32(float* __restrict__ p, float v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ launch
static uint32_t blocks_for(size_t n, uint32_t cap = 16384) {
  size_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  return (uint32_t)(b < cap ? b : cap);
}

void launch_adam(hipStream_t s, float* P, const float* G, float* M, float* V, uint32_t npad, uint32_t n, int nplanes,
                 const float* lr_planes, float b1, float b2, float eps, float wd, float* fscal, uint32_t* counters,
                 uint32_t* step, int nviews, int check_loss) {
  k_adam_prologue<<<1, 64, 0, s>>>(counters, fscal, step, b1, b2, nviews, check_loss);
  if (n == 0 || nplanes <= 0) return;
  const dim3 grid((npad / 4 + 255) / 256, (uint32_t)nplanes);
  k_adam<<<grid, 256, 0, s>>>(P, G, M, V, npad, n, lr_planes, b1, b2, eps, wd, fscal, counters);
}

void launch_l1(hipStream_t s, const float* pred, const float* tgt, uint32_t np, float scale, float* v_out,
               float* loss_sum) {
  if (np == 0) return;
  k_l1<<<(np + 255) / 256, 256, 0, s>>>(pred, tgt, np, scale, v_out, loss_sum);
}

void launch_aos_to_soa(hipStream_t s, const float* aos, int width, uint32_t n, float* soa, uint32_t npad, int plane0) {
  if (n == 0 || width == 0) return;
  k_aos_to_soa<<<blocks_for((size_t)n * width), 256, 0, s>>>(aos, width, n, soa, npad, plane0);
}

void launch_soa_to_aos(hipStream_t s, const float* soa, uint32_t npad, int plane0, int width, uint32_t n, float* aos) {
  if (n == 0 || width == 0) return;
  k_soa_to_aos<<<blocks_for((size_t)n * width), 256, 0, s>>>(soa, npad, plane0, width, n, aos);
}

void launch_fill_f32(hipStream_t s, float* p, float v, size_t n) {
  if (n == 0) return;
  k_fill_f32<<<blocks_for(n, 8192), 256, 0, s>>>(p, v, n);
}

}  // namespace rgs
