// Projection kernels.
//
// k_project_fwd fuses three reference kernels into one pass per Gaussian:
//   project_forward.wgsl:12-97      cull (K1), depth
//   project_visible.wgsl:129-214    conic, screen mean, SH colour (K3)
//   map_gaussian_to_intersects.wgsl:22-48 (counting half)  tile count (K5 rule)
// One thread owns one Gaussian and loops over the views of the batch, so the
// 11 geometry floats and 3K SH floats are read from HBM once per step
// instead of once per view.  Parameters are SoA planes (coalesced loads).
//
// k_project_bwd is project_backwards.wgsl:181-320 (K9), again one thread per
// Gaussian looping over views: the per-view gradients are summed in
// registers (view order) and written once.
#include "rgs_device.h"

namespace rgs {

// project_visible.wgsl:141-224 sh_coeffs_to_color, operation order preserved.
// s[k*3 + ch] = coefficient k of channel ch.
template <int DEG>
__device__ __forceinline__ void sh_color(const float* s, float x, float y, float z, float col[3]) {
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) col[ch] = RGS_F(0.28209479177387814) * s[ch];
  if (DEG == 0) return;
  const float fTmp0A = RGS_F(0.48860251190292);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float t = ((-y) * s[3 + ch] + z * s[6 + ch]) - x * s[9 + ch];
    col[ch] = col[ch] + fTmp0A * t;
  }
  if (DEG == 1) return;
  const float z2 = z * z;
  const float fTmp0B = RGS_F(-1.092548430592079) * z;
  const float fTmp1A = RGS_F(0.5462742152960395);
  const float fC1 = x * x - y * y;
  const float fS1 = 2.0f * x * y;
  const float pSH6 = RGS_F(0.9461746957575601) * z2 - RGS_F(0.3153915652525201);
  const float pSH7 = fTmp0B * x;
  const float pSH5 = fTmp0B * y;
  const float pSH8 = fTmp1A * fC1;
  const float pSH4 = fTmp1A * fS1;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float t = (((pSH4 * s[12 + ch] + pSH5 * s[15 + ch]) + pSH6 * s[18 + ch]) + pSH7 * s[21 + ch]) +
                    pSH8 * s[24 + ch];
    col[ch] = col[ch] + t;
  }
  if (DEG == 2) return;
  const float fTmp0C = RGS_F(-2.285228997322329) * z2 + RGS_F(0.4570457994644658);
  const float fTmp1B = RGS_F(1.445305721320277) * z;
  const float fTmp2A = RGS_F(-0.5900435899266435);
  const float fC2 = x * fC1 - y * fS1;
  const float fS2 = x * fS1 + y * fC1;
  const float pSH12 = z * (RGS_F(1.865881662950577) * z2 - RGS_F(1.119528997770346));
  const float pSH13 = fTmp0C * x;
  const float pSH11 = fTmp0C * y;
  const float pSH14 = fTmp1B * fC1;
  const float pSH10 = fTmp1B * fS1;
  const float pSH15 = fTmp2A * fC2;
  const float pSH9 = fTmp2A * fS2;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float t = (((((pSH9 * s[27 + ch] + pSH10 * s[30 + ch]) + pSH11 * s[33 + ch]) + pSH12 * s[36 + ch]) +
                      pSH13 * s[39 + ch]) +
                     pSH14 * s[42 + ch]) +
                    pSH15 * s[45 + ch];
    col[ch] = col[ch] + t;
  }
  if (DEG == 3) return;
  const float fTmp0D = z * (RGS_F(-4.683325804901025) * z2 + RGS_F(2.007139630671868));
  const float fTmp1C = RGS_F(3.31161143515146) * z2 - RGS_F(0.47308734787878);
  const float fTmp2B = RGS_F(-1.770130769779931) * z;
  const float fTmp3A = RGS_F(0.6258357354491763);
  const float fC3 = x * fC2 - y * fS2;
  const float fS3 = x * fS2 + y * fC2;
  const float pSH20 = RGS_F(1.984313483298443) * z * pSH12 - RGS_F(1.006230589874905) * pSH6;
  const float pSH21 = fTmp0D * x;
  const float pSH19 = fTmp0D * y;
  const float pSH22 = fTmp1C * fC1;
  const float pSH18 = fTmp1C * fS1;
  const float pSH23 = fTmp2B * fC2;
  const float pSH17 = fTmp2B * fS2;
  const float pSH24 = fTmp3A * fC3;
  const float pSH16 = fTmp3A * fS3;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float t = (((((((pSH16 * s[48 + ch] + pSH17 * s[51 + ch]) + pSH18 * s[54 + ch]) + pSH19 * s[57 + ch]) +
                        pSH20 * s[60 + ch]) +
                       pSH21 * s[63 + ch]) +
                      pSH22 * s[66 + ch]) +
                     pSH23 * s[69 + ch]) +
                    pSH24 * s[72 + ch];
    col[ch] = col[ch] + t;
  }
}

// Rotation matrix of a unit quaternion (w, x, y, z), helpers.wgsl:56-80; r[c][row].
__device__ __forceinline__ void quat_rot(float w, float x, float y, float z, float r[3][3]) {
  const float x2 = x * x, y2 = y * y, z2 = z * z;
  const float xy = x * y, xz = x * z, yz = y * z;
  const float wx = w * x, wy = w * y, wz = w * z;
  r[0][0] = 1.0f - 2.0f * (y2 + z2); r[0][1] = 2.0f * (xy + wz); r[0][2] = 2.0f * (xz - wy);
  r[1][0] = 2.0f * (xy - wz); r[1][1] = 1.0f - 2.0f * (x2 + z2); r[1][2] = 2.0f * (yz + wx);
  r[2][0] = 2.0f * (xz + wy); r[2][1] = 2.0f * (yz - wx); r[2][2] = 1.0f - 2.0f * (x2 + y2);
}

// World covariance of one Gaussian: m = R*diag(s), cov = m * m^T (helpers.wgsl:128-136).
struct Geom {
  float m[3][3];    // m[c][r]
  float cov[3][3];  // cov[c][r]
};

__device__ __forceinline__ void make_geom(float w, float x, float y, float z, const float s[3], Geom& g) {
  float r[3][3];
  quat_rot(w, x, y, z, r);
  // R * S with S diagonal: the literal product adds signed zeros; keep them.
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      const float s0 = (c == 0) ? s[0] : 0.0f, s1 = (c == 1) ? s[1] : 0.0f, s2 = (c == 2) ? s[2] : 0.0f;
      g.m[c][rr] = (r[0][rr] * s0 + r[1][rr] * s1) + r[2][rr] * s2;
    }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int rr = 0; rr < 3; ++rr)
      g.cov[c][rr] = (g.m[0][rr] * g.m[0][c] + g.m[1][rr] * g.m[1][c]) + g.m[2][rr] * g.m[2][c];
}

// Camera-space covariance (R * cov * R^T) and the 2D covariance J * cc * J^T
// (helpers.wgsl:138-168), literal matrix products.
struct Proj2 {
  float cc[3][3];      // camera covariance [c][r]
  float j20, j21;      // J column 2; J = [[fx*rz, 0], [0, fy*rz], [j20, j21]]
  float jx, jy;        // fx*rz, fy*rz
  float rz;            // 1 / max(z, 0.01)
  M2 cov2d;            // before blur
};

__device__ __forceinline__ void project_cov(const Geom& g, const ViewDev& V, float mcx, float mcy, float mcz,
                                            Proj2& o) {
  float A[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      A[c][r] = (V.vm[0 * 4 + r] * g.cov[c][0] + V.vm[1 * 4 + r] * g.cov[c][1]) + V.vm[2 * 4 + r] * g.cov[c][2];
  // (A * R^T)[c][r] = sum_k A[k][r] * R^T[c][k] = sum_k A[k][r] * R[k][c], R[k][c] = vm[k*4 + c]
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r)
      o.cc[c][r] = (A[0][r] * V.vm[0 * 4 + c] + A[1][r] * V.vm[1 * 4 + c]) + A[2][r] * V.vm[2 * 4 + c];
  const float rz = 1.0f / fmaxf_(mcz, RGS_F(0.01));
  o.rz = rz;
  const float ux = clampf_(mcx * rz, V.lim_nx, V.lim_px);
  const float uy = clampf_(mcy * rz, V.lim_ny, V.lim_py);
  o.jx = V.fx * rz;
  o.jy = V.fy * rz;
  o.j20 = -o.jx * ux;
  o.j21 = -o.jy * uy;
  const float J[3][2] = {{o.jx, 0.0f}, {0.0f, o.jy}, {o.j20, o.j21}};
  float JC[3][2];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 2; ++r) JC[c][r] = (J[0][r] * o.cc[c][0] + J[1][r] * o.cc[c][1]) + J[2][r] * o.cc[c][2];
  float C2[2][2];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int r = 0; r < 2; ++r) C2[c][r] = (JC[0][r] * J[0][c] + JC[1][r] * J[1][c]) + JC[2][r] * J[2][c];
  o.cov2d = M2{C2[0][0], C2[0][1], C2[1][0], C2[1][1]};
}

template <int DEG>
__global__ __launch_bounds__(256) void k_project_fwd(const float* __restrict__ P, uint32_t npad, uint32_t n,
                                                     const ViewDev* __restrict__ views, int v0, int nv, float blur,
                                                     int tw, ProjSplat* __restrict__ proj,
                                                     uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                     uint32_t* __restrict__ cnt, uint32_t* __restrict__ counters,
                                                     int count_ib) {
  constexpr int NK = (DEG + 1) * (DEG + 1);
  // per wave: the tile-test parameters of its 64 lanes, and the start
  // markers of their bboxes in the wave's joint tile list
  __shared__ float4 s_par[4][64 * 3];
  __shared__ __align__(16) uint8_t s_mark[4][64 * 64];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = gid < n;
  const uint32_t gl = live ? gid : 0u;  // lanes past n run the view loop with pass = false
  uint32_t nvis = 0;
  const float mean[3] = {P[gl], P[npad + gl], P[2 * npad + gl]};
  const float q0 = P[3 * npad + gl], q1 = P[4 * npad + gl], q2 = P[5 * npad + gl], q3 = P[6 * npad + gl];
  const float sc[3] = {rgs_exp(P[7 * npad + gl]), rgs_exp(P[8 * npad + gl]), rgs_exp(P[9 * npad + gl])};
  const float sig = rgs_sigmoid(P[10 * npad + gl]);
  const float qn = ((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3;
  const bool qok = live && !(qn < RGS_F(1e-6));
  const float is = 1.0f / rgs_sqrt(qn);
  Geom g;
  make_geom(q0 * is, q1 * is, q2 * is, q3 * is, sc, g);
  for (int v = v0; v < v0 + nv; ++v) {
    const ViewDev& V = views[v];
    const size_t inst = (size_t)v * n + gid;
    const float mcx = ((V.vm[0] * mean[0] + V.vm[4] * mean[1]) + V.vm[8] * mean[2]) + V.vm[12];
    const float mcy = ((V.vm[1] * mean[0] + V.vm[5] * mean[1]) + V.vm[9] * mean[2]) + V.vm[13];
    const float mcz = ((V.vm[2] * mean[0] + V.vm[6] * mean[1]) + V.vm[10] * mean[2]) + V.vm[14];
    bool pass = qok && !(mcz < RGS_F(0.01) || mcz > RGS_F(1e10));
    uint32_t count = 0, nb = 0;
    bool big = false;
    float m2x = 0.f, m2y = 0.f, a = 0.f, b = 0.f, c = 0.f, thr = 0.f;
    uint32_t bb[4] = {0u, 0u, 0u, 0u};
    if (pass) {
      Proj2 pc;
      project_cov(g, V, mcx, mcy, mcz, pc);
      // compensate_cov2d (helpers.wgsl:170-183)
      M2 c2 = pc.cov2d;
      const float det_raw = fmaxf_(det2(c2), 0.0f);
      c2.c00 = c2.c00 + blur;
      c2.c11 = c2.c11 + blur;
      const float det_end = fmaxf_(det2(c2), RGS_F(1e-12));
      const float comp = det_raw <= 0.0f ? 0.0f : rgs_sqrt(det_raw / det_end);
      const float opacity = sig * comp;
      pass = !(opacity < RGS_F(1.0 / 255.0));
      if (pass) {
        const float rzm = pc.rz;  // = 1 / mcz: mcz >= 0.01 here
        m2x = V.fx * mcx * rzm + V.cx;
        m2y = V.fy * mcy * rzm + V.cy;
        const M2 con = inverse2x2(c2);
        thr = rgs_log(opacity * 255.0f);
        float ex, ey;
        bbox_extent(c2.c00, c2.c11, thr, &ex, &ey);
        if (ex < 0.0f || ey < 0.0f) pass = false;
        if (m2x + ex <= 0.0f || m2x - ex >= (float)V.W || m2y + ey <= 0.0f || m2y - ey >= (float)V.H)
          pass = false;
        if (pass) {
          a = con.c00;
          b = con.c01;
          c = con.c11;
          // K5 rule: extent from the inverse of the conic
          const M2 cv = inverse2x2(M2{a, b, b, c});
          float ex5, ey5;
          bbox_extent(cv.c00, cv.c11, thr, &ex5, &ey5);
          tile_bbox(m2x, m2y, ex5, ey5, V.tiles_x, V.tiles_y, tw, bb);
          nb = (bb[3] - bb[1]) * (bb[2] - bb[0]);
          big = nb > 64u;
          // colour
          const float dx = mean[0] - V.campos[0], dy = mean[1] - V.campos[1], dz = mean[2] - V.campos[2];
          const float dd = (dx * dx + dy * dy) + dz * dz;
          const float rs = 1.0f / rgs_sqrt(fmaxf_(dd, RGS_F(1e-12)));
          // The SH coefficients are read per passing view rather than held
          // across the view loop: 48 live registers at degree 3 would halve
          // the kernel's occupancy.  The empty asm hides the address from
          // the optimiser so it cannot hoist the loads out of the loop again.
          const float* shp = P + (size_t)11 * npad + gl;
          asm volatile("" : "+v"(shp));
          float sh[NK * 3];
#pragma unroll
          for (int i = 0; i < NK * 3; ++i) sh[i] = shp[(size_t)i * npad];
          float col[3];
          sh_color<DEG>(sh, dx * rs, dy * rs, dz * rs, col);
          ProjSplat& ps = proj[inst];
          ps.p0 = make_float4(m2x, m2y, 0.5f * a, b);
          ps.p1 = make_float4(0.5f * c, opacity, col[0] + 0.5f, col[1] + 0.5f);
          ps.p2 = make_float4(col[2] + 0.5f, thr, pack_u16x2(bb[0], bb[1]), pack_u16x2(bb[2], bb[3]));
          keys[inst] = __float_as_uint(mcz);
        }
      }
    }
    // bboxes of at most 64 tiles: the wave tests its lanes' tiles together.
    // The bboxes are laid end to end (exclusive scan of their sizes) and
    // each lane tests one tile of that joint list per step, so the steps
    // number total/64 instead of the largest bbox of the wave.  A lane finds
    // the owner of its tile as the last bbox start at or before it (max-scan
    // of the start markers) and reads the owner's parameters from LDS; each
    // owner collects its bits from the step's ballot.  Tiles are visited in
    // the reference's row-major bbox order: bit k of the mask is tile k.
    const bool small = pass && !big;
    const uint32_t nbs = small ? nb : 0u;
    const uint32_t incl = wave_incl_add(nbs);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const uint32_t excl = incl - nbs;
    uint64_t mask = 0;
    if (total) {
      float4* par = s_par[wv];
      uint8_t* mark = s_mark[wv];
      const uint32_t bw = bb[2] - bb[0];
      par[lane * 3 + 0] = make_float4(m2x, m2y, a, b);
      par[lane * 3 + 1] = make_float4(c, thr, __uint_as_float(bb[0] | (bb[1] << 16)), __uint_as_float(bw));
      par[lane * 3 + 2] = make_float4(__uint_as_float(excl), 1.0f / (float)(bw ? bw : 1u), 0.0f, 0.0f);
      for (uint32_t z = lane * 16u; z < total; z += 1024u) *reinterpret_cast<uint4*>(mark + z) = make_uint4(0, 0, 0, 0);
      wave_lds_fence();
      if (nbs) mark[excl] = (uint8_t)(lane + 1u);
      wave_lds_fence();
      uint32_t carry = 0;
      for (uint32_t c0 = 0; c0 < total; c0 += 64u) {
        const uint32_t g = c0 + lane;
        const bool valid = g < total;
        uint32_t o = wave_incl_max(valid ? (uint32_t)mark[g] : 0u);
        o = o > carry ? o : carry;
        carry = (uint32_t)__builtin_amdgcn_readlane((int)o, 63);
        const uint32_t ow = valid ? o - 1u : 0u;
        const float4 q0 = par[ow * 3 + 0], q1 = par[ow * 3 + 1], q2 = par[ow * 3 + 2];
        const uint32_t obb = __float_as_uint(q1.z);
        uint32_t row, col;
        bbox_rc(g - __float_as_uint(q2.x), __float_as_uint(q1.w), q2.y, &row, &col);
        const bool hit = valid && tile_contributes((obb & 0xffffu) + col, (obb >> 16) + row, tw, q0.x, q0.y, q0.z,
                                                   q0.w, q1.x, q1.y);
        const uint64_t B = __ballot(hit);
        const uint32_t lo = excl > c0 ? excl : c0;
        const uint32_t hi = excl + nbs < c0 + 64u ? excl + nbs : c0 + 64u;
        if (lo < hi) {
          const uint32_t len = hi - lo;
          const uint64_t bits = (B >> (lo - c0)) & (len == 64u ? ~0ull : ((1ull << len) - 1ull));
          mask |= bits << (lo - excl);
        }
      }
      wave_lds_fence();
    }
    if (small) {
      count = (uint32_t)__popcll(mask);
      proj[inst].p3 = make_float4(__uint_as_float((uint32_t)mask), __uint_as_float((uint32_t)(mask >> 32)), 0.0f, 0.0f);
    } else if (pass) {
      proj[inst].p3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // bboxes of more than 64 tiles: the wave counts them together
    uint64_t bigm = __ballot(big);
    while (bigm) {
      const int j = __builtin_ctzll(bigm);
      bigm &= bigm - 1;
      uint32_t jb[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) jb[e] = (uint32_t)__shfl((int)bb[e], j, 64);
      const uint32_t tot = wave_count_tiles(jb, tw, __shfl(m2x, j, 64), __shfl(m2y, j, 64), __shfl(a, j, 64),
                                            __shfl(b, j, 64), __shfl(c, j, 64), __shfl(thr, j, 64));
      if ((int)(threadIdx.x & 63) == j) count = tot;
    }
    if (live) {
      if (!pass) keys[inst] = 0xffffffffu;
      // with count_ib > 0 the sort value carries the count above the
      // instance id (all ones when it does not fit: the sort looks it up)
      const uint32_t cesc = count_ib > 0 ? 0xffffffffu >> count_ib : 0u;
      vals[inst] = count_ib > 0 ? (uint32_t)inst | ((count < cesc ? count : cesc) << count_ib) : (uint32_t)inst;
      cnt[inst] = count;
    }
    nvis += pass ? 1u : 0u;
  }
  // one atomic per wave
  uint32_t tot = nvis;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o, 64);
  if ((threadIdx.x & 63) == 0 && tot) atomicAdd(&counters[kCntVisible], tot);
}

// ---------------------------------------------------------------- backward
// SH vector-Jacobian product (helpers.wgsl:289-380): g[k*3+ch] += basis_k * vc[ch];
// *abs_basis receives the sum of |basis_k| over k (the colour-gradient
// statistic of the densification accumulators uses it).
template <int DEG>
__device__ __forceinline__ void sh_vjp_acc(float x, float y, float z, const float vc[3], float* g, float* abs_basis) {
  float ab = 0.0f;
  auto put = [&](int k, float b) {
    ab += fabsf(b);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) g[k * 3 + ch] += b * vc[ch];
  };
  put(0, RGS_F(0.28209479177387814));
  if (DEG == 0) {
    *abs_basis = ab;
    return;
  }
  const float fTmp0A = RGS_F(0.48860251190292);
  put(1, -fTmp0A * y);
  put(2, fTmp0A * z);
  put(3, -fTmp0A * x);
  if (DEG == 1) {
    *abs_basis = ab;
    return;
  }
  const float z2 = z * z;
  const float fTmp0B = RGS_F(-1.092548430592079) * z;
  const float fTmp1A = RGS_F(0.5462742152960395);
  const float fC1 = x * x - y * y;
  const float fS1 = 2.0f * x * y;
  put(4, fTmp1A * fS1);
  put(5, fTmp0B * y);
  put(6, RGS_F(0.9461746957575601) * z2 - RGS_F(0.3153915652525201));
  put(7, fTmp0B * x);
  put(8, fTmp1A * fC1);
  if (DEG == 2) {
    *abs_basis = ab;
    return;
  }
  const float fTmp0C = RGS_F(-2.285228997322329) * z2 + RGS_F(0.4570457994644658);
  const float fTmp1B = RGS_F(1.445305721320277) * z;
  const float fTmp2A = RGS_F(-0.5900435899266435);
  const float fC2 = x * fC1 - y * fS1;
  const float fS2 = x * fS1 + y * fC1;
  const float pSH12 = z * (RGS_F(1.865881662950577) * z2 - RGS_F(1.119528997770346));
  put(9, fTmp2A * fS2);
  put(10, fTmp1B * fS1);
  put(11, fTmp0C * y);
  put(12, pSH12);
  put(13, fTmp0C * x);
  put(14, fTmp1B * fC1);
  put(15, fTmp2A * fC2);
  if (DEG == 3) {
    *abs_basis = ab;
    return;
  }
  const float fTmp0D = z * (RGS_F(-4.683325804901025) * z2 + RGS_F(2.007139630671868));
  const float fTmp1C = RGS_F(3.31161143515146) * z2 - RGS_F(0.47308734787878);
  const float fTmp2B = RGS_F(-1.770130769779931) * z;
  const float fTmp3A = RGS_F(0.6258357354491763);
  const float fC3 = x * fC2 - y * fS2;
  const float fS3 = x * fS2 + y * fC2;
  const float pSH6 = RGS_F(0.9461746957575601) * z2 - RGS_F(0.3153915652525201);
  put(16, fTmp3A * fS3);
  put(17, fTmp2B * fS2);
  put(18, fTmp1C * fS1);
  put(19, fTmp0D * y);
  put(20, RGS_F(1.984313483298443) * z * pSH12 - RGS_F(1.006230589874905) * pSH6);
  put(21, fTmp0D * x);
  put(22, fTmp1C * fC1);
  put(23, fTmp2B * fC2);
  put(24, fTmp3A * fC3);
  *abs_basis = ab;
}

// Per-instance rasterizer gradient accumulator (written with atomics by the
// raster backward, consumed and cleared here), 16 floats = one 64-byte line:
//   [0..1] v_xy  [2..4] v_conic  [5..7] v_rgb  [8] v_opacity  [9] v_refine
//   [10..11] |v_xy|  [12] coverage  [13..15] unused
// The reference accumulates v_xy a second time into its screen-gradient
// buffer (rasterize_backwards.wgsl:158-159); the same sum is read from [0..1].

// ---- the covariance chain of the projection backward, in FP64 ----------
// project_backwards.wgsl:112-150 and helpers.wgsl:56-203,382-427 restated in
// double.  In FP32 this chain loses the gradient of elongated splats: the
// conic, the inverse VJP -C*V*C and the products with J, the view rotation
// and R*S are formed in pixel / world bases, where the large components along
// a splat's long axis cancel, and the rounding left along that axis is then
// scaled by the long axis' scale squared (quaternion and mean gradients of a
// splat with a 1:1000 axis ratio lose every digit).  Double arithmetic keeps
// the chain within the FP64 result; CDNA4 issues FP64 FMAs at the non-packed
// FP32 rate and this kernel is bound by memory, not by these operations.
// Rotation of a unit quaternion (w, x, y, z), helpers.wgsl:56-80, in double; r[c][row].
__device__ __forceinline__ void quat_rot64(const double q[4], double r[3][3]) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double x2 = x * x, y2 = y * y, z2 = z * z;
  const double xy = x * y, xz = x * z, yz = y * z;
  const double wx = w * x, wy = w * y, wz = w * z;
  r[0][0] = 1.0 - 2.0 * (y2 + z2); r[0][1] = 2.0 * (xy + wz); r[0][2] = 2.0 * (xz - wy);
  r[1][0] = 2.0 * (xy - wz); r[1][1] = 1.0 - 2.0 * (x2 + z2); r[1][2] = 2.0 * (yz + wx);
  r[2][0] = 2.0 * (xz + wy); r[2][1] = 2.0 * (yz - wx); r[2][2] = 1.0 - 2.0 * (x2 + y2);
}

// Symmetric 3x3 matrix: xx, yy, zz, xy, xz, yz.
struct S3 {
  double xx, yy, zz, xy, xz, yz;
  __device__ __forceinline__ double at(int i, int j) const {
    if (i == j) return i == 0 ? xx : (i == 1 ? yy : zz);
    const int k = i + j;  // 1: xy, 2: xz, 3: yz
    return k == 1 ? xy : (k == 2 ? xz : yz);
  }
};

// The Gaussian's frame in double: raw quaternion qu, |qu|^2, |qu|, the unit
// quaternion q, its rotation (columns rot[c]) and the squared scales
// (project_backwards.wgsl:57-63).
__device__ __forceinline__ void gauss_frame64(const float* P, uint32_t npad, uint32_t gid, double qu[4], double* qn,
                                              double* qlen, double q[4], double rot[3][3], double s2[3]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) qu[i] = (double)P[(3 + i) * npad + gid];
  *qn = qu[0] * qu[0] + qu[1] * qu[1] + qu[2] * qu[2] + qu[3] * qu[3];
  *qlen = sqrt(fmax(*qn, 1e-20));
  const double qis = 1.0 / *qlen;
#pragma unroll
  for (int i = 0; i < 4; ++i) q[i] = qu[i] * qis;
  quat_rot64(q, rot);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double sc = exp((double)P[(7 + c) * npad + gid]);
    s2[c] = sc * sc;
  }
}

// grads: SoA planes like the parameters (59 planes at SH3).
// stats: [7][npad] screen-gradient statistics (project_backwards.wgsl:33-36,106-108), summed over views.
// accumulate != 0 adds into grads/stats/dinc instead of overwriting (per-view launches).
// dinc (optional): [5][npad] this step's increments of the densification
// accumulators that are not plain sums of the screen statistics
// (accumulate_gradients, trainer.rs:651-728, applied once per view):
//   [0] mean |v_mean| over xyz  [1] |screen v_xy|  [2] |sum |v_xy||
//   [3] mean |v_sh| over coefficients and channels  [4] views where a
//   fragment blended the Gaussian (from vis: one byte per instance).
// They are added to the accumulators by k_densify_commit once the step is
// known to be valid.
//
// Symmetric matrices of the chain (covariances and their gradients) are
// handled as symmetric: the reference forms them as literal products, whose
// asymmetric rounding is below FP64's.
// Three waves per SIMD (168 VGPRs, a 12-byte spill): the kernel streams the
// accumulator lines and parameter planes, and the third wave hides more of
// that latency than the spill costs (0.40 -> 0.32 ms on 8 views x 1M).
template <int DEG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_project_bwd(const float* __restrict__ P, uint32_t npad, uint32_t n,
                                                     const ViewDev* __restrict__ views, int v0, int nv, float blur,
                                                     float* __restrict__ acc, const uint32_t* __restrict__ keys_by_inst,
                                                     float* __restrict__ grads, float* __restrict__ stats,
                                                     int accumulate, const uint32_t* __restrict__ counters,
                                                     float* __restrict__ dinc, const uint8_t* __restrict__ vis) {
  constexpr int NK = (DEG + 1) * (DEG + 1);
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const float mean[3] = {P[gid], P[npad + gid], P[2 * npad + gid]};
  auto out = [&](int plane, float val) {
    float* p = grads + (size_t)plane * npad + gid;
    *p = accumulate ? (*p + val) : val;
  };
  // densification increments of this launch's views
  float d_g2 = 0.f, d_sg = 0.f, d_ag = 0.f, d_gc = 0.f, d_vis = 0.f;
  if (dinc && vis)
    for (int v = v0; v < v0 + nv; ++v) d_vis += vis[(size_t)v * n + gid] ? 1.0f : 0.0f;
  auto out_dinc = [&]() {
    if (!dinc) return;
    const float d[5] = {d_g2, d_sg, d_ag, d_gc, d_vis};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      float* p = dinc + (size_t)k * npad + gid;
      *p = accumulate ? (*p + d[k]) : d[k];
    }
  };

  // Pass 1: the SH colour VJP (FP32: the basis is bounded and its products
  // do not cancel), with the view direction from the unperturbed mean and no
  // mean gradient through it (project_backwards.wgsl:68-79).  Its planes are
  // written before pass 2 so the two passes' accumulators are never live
  // together.
  bool any = false;
  {
    float gsh[NK * 3];
#pragma unroll
    for (int i = 0; i < NK * 3; ++i) gsh[i] = 0.f;
    for (int v = v0; v < v0 + nv; ++v) {
      const size_t inst = (size_t)v * n + gid;
      if (keys_by_inst[inst] == 0xffffffffu) continue;  // culled in this view
      const float4* ap = reinterpret_cast<const float4*>(acc + inst * kAcc);
      // coverage 0: the rasterizer never flushed into this line, which is
      // still all zero, and every output term is zero
      if (ap[3].x == 0.0f) continue;
      const float4 a1 = ap[1];
      any = true;
      const ViewDev& V = views[v];
      const float vcol[3] = {a1.y, a1.z, a1.w};
      const float dx = mean[0] - V.campos[0], dy = mean[1] - V.campos[1], dz = mean[2] - V.campos[2];
      const float dd = (dx * dx + dy * dy) + dz * dz;
      const float rs = 1.0f / rgs_sqrt(fmaxf_(dd, RGS_F(1e-12)));
      float abs_basis;
      sh_vjp_acc<DEG>(dx * rs, dy * rs, dz * rs, vcol, gsh, &abs_basis);
      // mean over coefficients and channels of |basis_k * vcol_c|
      d_gc += (abs_basis * ((fabsf(vcol[0]) + fabsf(vcol[1])) + fabsf(vcol[2]))) / (float)(3 * NK);
    }
    if (!any && accumulate) {
      out_dinc();
      return;
    }
#pragma unroll
    for (int i = 0; i < NK * 3; ++i) out(11 + i, gsh[i]);
  }

  // Pass 2: opacity, screen statistics and the covariance chain in FP64;
  // consumes and clears the accumulator lines.  The covariance gradient is
  // linear in each view's v_cov2d, so it is summed over the views in world
  // space (v_cov3d = Rv^T vcc Rv) and the scale and rotation VJPs are taken
  // once, after the loop; the loop itself needs only the world covariance.
  const float opac = rgs_sigmoid(P[10 * npad + gid]);
  S3 cw;  // world covariance rot diag(s^2) rot^T
  {
    double qu[4], qn, qlen, q[4], rot[3][3], s2[3];
    gauss_frame64(P, npad, gid, qu, &qn, &qlen, q, rot, s2);
    auto cov = [&](int i, int j) { return rot[0][i] * rot[0][j] * s2[0] + rot[1][i] * rot[1][j] * s2[1] + rot[2][i] * rot[2][j] * s2[2]; };
    cw.xx = cov(0, 0); cw.yy = cov(1, 1); cw.zz = cov(2, 2);
    cw.xy = cov(0, 1); cw.xz = cov(0, 2); cw.yz = cov(1, 2);
  }
  float gmean[3] = {0.f, 0.f, 0.f};
  float gop = 0.f;
  float st[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  S3 w3 = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // v_cov3d summed over views
  bool chain = false;
  for (int v = v0; v < v0 + nv; ++v) {
    const size_t inst = (size_t)v * n + gid;
    if (keys_by_inst[inst] == 0xffffffffu) continue;
    float4* ap = reinterpret_cast<float4*>(acc + inst * kAcc);
    const float4 a3 = ap[3];
    if (a3.x == 0.0f) continue;
    const float4 a0 = ap[0], a1 = ap[1], a2 = ap[2];
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    ap[0] = z4; ap[1] = z4; ap[2] = z4; ap[3] = z4;
    const ViewDev& V = views[v];
    const float vca = a2.x;
    // screen statistics [0..3]
    st[0] += a0.x; st[1] += a0.y; st[2] += a2.z; st[3] += a2.w;
    d_sg += rgs_sqrt(a0.x * a0.x + a0.y * a0.y);
    d_ag += rgs_sqrt(a2.z * a2.z + a2.w * a2.w);
    const float mcz_f = ((V.vm[2] * mean[0] + V.vm[6] * mean[1]) + V.vm[10] * mean[2]) + V.vm[14];
    if (mcz_f < RGS_F(0.01)) continue;
    {
      const float abs_score = rgs_sqrt(a2.z * a2.z + a2.w * a2.w);
      st[4] += a3.x * abs_score;
      st[5] += a3.x;
      st[6] += a3.x * mcz_f;
    }
    chain = true;
    // view rotation Rv[i][k] = vm[4k + i]
    auto Rv = [&](int i, int k) { return (double)V.vm[4 * k + i]; };
    // camera-space covariance cc = Rv cw Rv^T
    S3 cc;
    {
      double t[3][3];  // t[i][l] = (Rv cw)[i][l]
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int l = 0; l < 3; ++l) t[i][l] = Rv(i, 0) * cw.at(0, l) + Rv(i, 1) * cw.at(1, l) + Rv(i, 2) * cw.at(2, l);
      auto ct = [&](int i, int j) { return t[i][0] * Rv(j, 0) + t[i][1] * Rv(j, 1) + t[i][2] * Rv(j, 2); };
      cc.xx = ct(0, 0); cc.yy = ct(1, 1); cc.zz = ct(2, 2);
      cc.xy = ct(0, 1); cc.xz = ct(0, 2); cc.yz = ct(1, 2);
    }
    // camera-space mean
    double mc[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
      mc[r] = (double)V.vm[r] * mean[0] + (double)V.vm[4 + r] * mean[1] + (double)V.vm[8 + r] * mean[2] +
              (double)V.vm[12 + r];
    // J = [[jx, 0, j20], [0, jy, j21]] with the EWA clip (helpers.wgsl:138-155).
    // mc[2] >= 0.01 up to the rounding of the float test above, so the
    // reference's 1/max(mc_z, 0.01) and 1/max(|mc_z|, 0.01) are one value.
    const double fx = V.fx, fy = V.fy;
    const double lpx = V.lim_px, lpy = V.lim_py, lnx = V.lim_nx, lny = V.lim_ny;
    const double rz = 1.0 / fmax(fabs(mc[2]), 0.01);
    const double jx = fx * rz, jy = fy * rz;
    const double j20 = -jx * fmin(fmax(mc[0] * rz, lnx), lpx), j21 = -jy * fmin(fmax(mc[1] * rz, lny), lpy);
    // cov2d = J cc J^T
    const double c2xx = jx * (jx * cc.xx + j20 * cc.xz) + j20 * (jx * cc.xz + j20 * cc.zz);
    const double c2yy = jy * (jy * cc.yy + j21 * cc.yz) + j21 * (jy * cc.yz + j21 * cc.zz);
    const double c2xy = jx * (jy * cc.xy + j21 * cc.xz) + j20 * (jy * cc.yz + j21 * cc.zz);
    // blur compensation (helpers.wgsl:170-183); its derivative is not taken (project_backwards.wgsl:110)
    const double det_raw = fmax(c2xx * c2yy - c2xy * c2xy, 0.0);
    const double bxx = c2xx + (double)blur, byy = c2yy + (double)blur;
    const double det_end = fmax(bxx * byy - c2xy * c2xy, 1e-12);
    const double comp = det_raw <= 0.0 ? 0.0 : sqrt(det_raw / det_end);
    gop += (float)comp * vca * opac * (1.0f - opac);
    // conic = inverse2x2(cov2d) with the +1e-6 regulariser (helpers.wgsl:185-198)
    double cixx, ciyy, cixy;
    {
      const double r00 = bxx + 1e-6, r11 = byy + 1e-6;
      const double det = r00 * r11 - c2xy * c2xy;
      const double inv = det <= 1e-10 ? 0.0 : 1.0 / det;
      cixx = r11 * inv; ciyy = r00 * inv; cixy = -c2xy * inv;
    }
    // v_cov2d = -C V C with V = [[vc0, vc1/2], [vc1/2, vc2]] (helpers.wgsl:200-202)
    double vxx, vyy, vxy;
    {
      const double w0 = a0.z, wh = 0.5 * (double)a0.w, w2 = a1.x;
      const double t00 = cixx * w0 + cixy * wh, t01 = cixy * w0 + ciyy * wh;  // (C V) column 0
      const double t10 = cixx * wh + cixy * w2, t11 = cixy * wh + ciyy * w2;  // (C V) column 1
      vxx = -(t00 * cixx + t10 * cixy);
      vxy = -(t01 * cixx + t11 * cixy);
      vyy = -(t01 * cixy + t11 * ciyy);
    }
    // persp_proj_vjp (helpers.wgsl:382-427); v_J = 2 vc J cc for symmetric vc and cc
    const double x = mc[0], y = mc[1];
    const double rz2 = rz * rz;
    double vm0 = fx * rz * (double)a0.x;
    double vm1 = fy * rz * (double)a0.y;
    double vm2 = -(fx * x * (double)a0.x + fy * y * (double)a0.y) * rz2;
    // (vc J) columns: vc (jx, 0), vc (0, jy), vc (j20, j21)
    const double p00 = vxx * jx, p01 = vxy * jx;
    const double p10 = vxy * jy, p11 = vyy * jy;
    const double p20 = vxx * j20 + vxy * j21, p21 = vxy * j20 + vyy * j21;
    const double vJ00 = 2.0 * (p00 * cc.xx + p10 * cc.xy + p20 * cc.xz);
    const double vJ11 = 2.0 * (p01 * cc.xy + p11 * cc.yy + p21 * cc.yz);
    const double vJ20 = 2.0 * (p00 * cc.xz + p10 * cc.yz + p20 * cc.zz);
    const double vJ21 = 2.0 * (p01 * cc.xz + p11 * cc.yz + p21 * cc.zz);
    if (x * rz <= lpx && x * rz >= lnx) {
      vm0 += -fx * rz2 * vJ20;
      vm2 += 2.0 * fx * x * rz2 * rz * vJ20;
    } else {
      vm2 += fx * fmin(fmax(x * rz, lnx), lpx) * rz2 * vJ20;
    }
    if (y * rz <= lpy && y * rz >= lny) {
      vm1 += -fy * rz2 * vJ21;
      vm2 += 2.0 * fy * y * rz2 * rz * vJ21;
    } else {
      vm2 += fy * fmin(fmax(y * rz, lny), lpy) * rz2 * vJ21;
    }
    vm2 += -fx * rz2 * vJ00 - fy * rz2 * vJ11;
    // v_mean = Rv^T v_mean_c
    {
      float vmv[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        vmv[r] = (float)((double)V.vm[r * 4] * vm0 + (double)V.vm[r * 4 + 1] * vm1 + (double)V.vm[r * 4 + 2] * vm2);
        gmean[r] += vmv[r];
      }
      d_g2 += ((fabsf(vmv[0]) + fabsf(vmv[1])) + fabsf(vmv[2])) / 3.0f;
    }
    // v_cov_cam = J^T vc J (symmetric)
    S3 vcc;
    vcc.xx = jx * jx * vxx;
    vcc.yy = jy * jy * vyy;
    vcc.xy = jx * jy * vxy;
    vcc.xz = jx * p20;
    vcc.yz = jy * p21;
    vcc.zz = j20 * p20 + j21 * p21;
    // v_cov3d = Rv^T vcc Rv, added to the world-space sum
    {
      double u[3][3];  // u[i][l] = (vcc Rv)[i][l]
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int l = 0; l < 3; ++l) u[i][l] = vcc.at(i, 0) * Rv(0, l) + vcc.at(i, 1) * Rv(1, l) + vcc.at(i, 2) * Rv(2, l);
      auto wt = [&](int k, int l) { return Rv(0, k) * u[0][l] + Rv(1, k) * u[1][l] + Rv(2, k) * u[2][l]; };
      w3.xx += wt(0, 0); w3.yy += wt(1, 1); w3.zz += wt(2, 2);
      w3.xy += wt(0, 1); w3.xz += wt(0, 2); w3.yz += wt(1, 2);
    }
  }

  // scale and rotation VJPs of the summed v_cov3d.  With m = rot diag(s):
  //   v_m[c] = 2 v_cov3d rot[c] s_c,  v_rot[c] = v_m[c] s_c = 2 s_c^2 v_cov3d rot[c],
  //   dL/dlog s_c = s_c rot[c] . v_m[c] = rot[c] . v_rot[c].
  // The frame is read again through an opaque pointer so that none of it is
  // held in registers across the view loop.
  float gq[4] = {0.f, 0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f};
  if (chain) {
    const float* Pf = P;
    asm volatile("" : "+s"(Pf));
    double qu[4], qn, qlen, q[4], rot[3][3], s2[3];
    gauss_frame64(Pf, npad, gid, qu, &qn, &qlen, q, rot, s2);
    double vr[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
        vr[c][r] = 2.0 * s2[c] * (w3.at(r, 0) * rot[c][0] + w3.at(r, 1) * rot[c][1] + w3.at(r, 2) * rot[c][2]);
      gls[c] = (float)(rot[c][0] * vr[c][0] + rot[c][1] * vr[c][1] + rot[c][2] * vr[c][2]);
    }
    // quat_to_mat3_vjp (helpers.wgsl:82-103)
    const double w = q[0], qx = q[1], qy = q[2], qz = q[3];
    double dq[4];
    dq[0] = 2.0 * (qx * (vr[1][2] - vr[2][1]) + qy * (vr[2][0] - vr[0][2]) + qz * (vr[0][1] - vr[1][0]));
    dq[1] = 2.0 * (-2.0 * qx * (vr[1][1] + vr[2][2]) + qy * (vr[0][1] + vr[1][0]) + qz * (vr[0][2] + vr[2][0]) +
                   w * (vr[1][2] - vr[2][1]));
    dq[2] = 2.0 * (qx * (vr[0][1] + vr[1][0]) - 2.0 * qy * (vr[0][0] + vr[2][2]) + qz * (vr[1][2] + vr[2][1]) +
                   w * (vr[2][0] - vr[0][2]));
    dq[3] = 2.0 * (qx * (vr[0][2] + vr[2][0]) + qy * (vr[1][2] + vr[2][1]) - 2.0 * qz * (vr[0][0] + vr[1][1]) +
                   w * (vr[0][1] - vr[1][0]));
    // normalize_vjp(qu) (helpers.wgsl:105-119): (|q|^2 I - q q^T) / |q|^3
    const double sN = 1.0 / (qlen * qn);
    const double qd = qu[0] * dq[0] + qu[1] * dq[1] + qu[2] * dq[2] + qu[3] * dq[3];
#pragma unroll
    for (int r = 0; r < 4; ++r) gq[r] = (float)((qn * dq[r] - qu[r] * qd) * sN);
  }

  out(0, gmean[0]); out(1, gmean[1]); out(2, gmean[2]);
  out(3, gq[0]); out(4, gq[1]); out(5, gq[2]); out(6, gq[3]);
  out(7, gls[0]); out(8, gls[1]); out(9, gls[2]);
  out(10, gop);
  if (stats) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      float* p = stats + (size_t)k * npad + gid;
      *p = accumulate ? (*p + st[k]) : st[k];
    }
  }
  out_dinc();
}

// Adds one step's densification increments to the accumulators
// (accumulate_gradients, trainer.rs:651-728), skipped when the step was
// invalidated (capacity overflow, non-finite loss: the reference returns
// before accumulating, trainer.rs:297-302).  dens: [10][npad]
//   [0] grad_2d  [1] screen_grad_2d  [2] abs_grad_2d  [3] abs_pixel_grad_2d
//   [4] pixel_coverage  [5] camera_depth  [6] grad_color  [7] num_observations
//   [8] visible_observations  [9] actual_visible_observations
// stats [4..6] are the step's sums over views of coverage * |abs v_xy|,
// coverage and coverage * depth, which accumulate linearly.
__global__ __launch_bounds__(256) void k_densify_commit(float* __restrict__ dens, const float* __restrict__ dinc,
                                                        const float* __restrict__ stats, uint32_t npad, uint32_t n,
                                                        float nviews, int flags,
                                                        const uint32_t* __restrict__ counters) {
  if (counters[kCntStepValid] == 0u) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  auto add = [&](int k, float x) { dens[(size_t)k * npad + g] += x; };
  const float vis = dinc[4 * (size_t)npad + g];
  add(0, dinc[g]);
  add(1, dinc[(size_t)npad + g]);
  add(2, dinc[2 * (size_t)npad + g]);
  add(3, stats[4 * (size_t)npad + g]);
  add(4, stats[5 * (size_t)npad + g]);
  add(5, stats[6 * (size_t)npad + g]);
  add(6, dinc[3 * (size_t)npad + g]);
  add(7, nviews);
  add(8, (flags & 2) ? vis : nviews);
  if (flags & 4) add(9, vis);
}

// ------------------------------------------------------------------ launch
void launch_project_fwd(int deg, hipStream_t s, const float* P, uint32_t npad, uint32_t n, const ViewDev* views,
                        int v0, int nv, float blur, int tw, ProjSplat* proj, uint32_t* keys, uint32_t* vals,
                        uint32_t* cnt, uint32_t* counters, int count_ib) {
  if (n == 0) return;
  const dim3 grid((n + 255) / 256), block(256);
  switch (deg) {
    case 0: k_project_fwd<0><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, tw, proj, keys, vals, cnt, counters, count_ib); break;
    case 1: k_project_fwd<1><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, tw, proj, keys, vals, cnt, counters, count_ib); break;
    case 2: k_project_fwd<2><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, tw, proj, keys, vals, cnt, counters, count_ib); break;
    case 3: k_project_fwd<3><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, tw, proj, keys, vals, cnt, counters, count_ib); break;
    default: k_project_fwd<4><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, tw, proj, keys, vals, cnt, counters, count_ib); break;
  }
}

void launch_project_bwd(int deg, hipStream_t s, const float* P, uint32_t npad, uint32_t n, const ViewDev* views,
                        int v0, int nv, float blur, float* acc, const uint32_t* keys_by_inst, float* grads,
                        float* stats, int accumulate, const uint32_t* counters, float* dinc, const uint8_t* vis) {
  if (n == 0) return;
  const dim3 grid((n + 255) / 256), block(256);
  switch (deg) {
    case 0: k_project_bwd<0><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters, dinc, vis); break;
    case 1: k_project_bwd<1><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters, dinc, vis); break;
    case 2: k_project_bwd<2><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters, dinc, vis); break;
    case 3: k_project_bwd<3><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters, dinc, vis); break;
    default: k_project_bwd<4><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters, dinc, vis); break;
  }
}

void launch_densify_commit(hipStream_t s, float* dens, const float* dinc, const float* stats, uint32_t npad,
                           uint32_t n, int nviews, int flags, const uint32_t* counters) {
  if (n == 0) return;
  k_densify_commit<<<(n + 255) / 256, 256, 0, s>>>(dens, dinc, stats, npad, n, (float)nviews, flags, counters);
}

}  // namespace rgs
