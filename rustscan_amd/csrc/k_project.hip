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
  float sh[NK * 3];
#pragma unroll
  for (int i = 0; i < NK * 3; ++i) sh[i] = P[(size_t)(11 + i) * npad + gl];
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
// SH vector-Jacobian product (helpers.wgsl:289-380): g[k*3+ch] += basis_k * vc[ch]
template <int DEG>
__device__ __forceinline__ void sh_vjp_acc(float x, float y, float z, const float vc[3], float* g) {
  auto put = [&](int k, float b) {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) g[k * 3 + ch] += b * vc[ch];
  };
  put(0, RGS_F(0.28209479177387814));
  if (DEG == 0) return;
  const float fTmp0A = RGS_F(0.48860251190292);
  put(1, -fTmp0A * y);
  put(2, fTmp0A * z);
  put(3, -fTmp0A * x);
  if (DEG == 1) return;
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
  if (DEG == 2) return;
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
  if (DEG == 3) return;
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
}

// Per-instance rasterizer gradient accumulator (written with atomics by the
// raster backward, consumed and cleared here), 16 floats = one 64-byte line:
//   [0..1] v_xy  [2..4] v_conic  [5..7] v_rgb  [8] v_opacity  [9] v_refine
//   [10..11] |v_xy|  [12] coverage  [13..15] unused
// The reference accumulates v_xy a second time into its screen-gradient
// buffer (rasterize_backwards.wgsl:158-159); the same sum is read from [0..1].

// grads: SoA planes like the parameters (59 planes at SH3).
// stats: [7][npad] screen-gradient statistics (project_backwards.wgsl:33-36,106-108), summed over views.
// visible: [npad] OR over views of the raster visibility (1.0/0.0), written when non-null.
// accumulate != 0 adds into grads/stats instead of overwriting (per-view launches).
template <int DEG>
__global__ __launch_bounds__(256) void k_project_bwd(const float* __restrict__ P, uint32_t npad, uint32_t n,
                                                     const ViewDev* __restrict__ views, int v0, int nv, float blur,
                                                     float* __restrict__ acc, const uint32_t* __restrict__ keys_by_inst,
                                                     float* __restrict__ grads, float* __restrict__ stats,
                                                     int accumulate, const uint32_t* __restrict__ counters) {
  constexpr int NK = (DEG + 1) * (DEG + 1);
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  const float mean[3] = {P[gid], P[npad + gid], P[2 * npad + gid]};
  const float qu[4] = {P[3 * npad + gid], P[4 * npad + gid], P[5 * npad + gid], P[6 * npad + gid]};
  const float sc[3] = {rgs_exp(P[7 * npad + gid]), rgs_exp(P[8 * npad + gid]), rgs_exp(P[9 * npad + gid])};
  const float opac = rgs_sigmoid(P[10 * npad + gid]);
  const float qn = ((qu[0] * qu[0] + qu[1] * qu[1]) + qu[2] * qu[2]) + qu[3] * qu[3];
  const float is = 1.0f / rgs_sqrt(qn);
  const float q[4] = {qu[0] * is, qu[1] * is, qu[2] * is, qu[3] * is};
  Geom g;
  make_geom(q[0], q[1], q[2], q[3], sc, g);

  float gmean[3] = {0.f, 0.f, 0.f}, gq[4] = {0.f, 0.f, 0.f, 0.f}, gls[3] = {0.f, 0.f, 0.f}, gop = 0.f;
  float gsh[NK * 3];
#pragma unroll
  for (int i = 0; i < NK * 3; ++i) gsh[i] = 0.f;
  float st[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  bool any = false;

  for (int v = v0; v < v0 + nv; ++v) {
    const size_t inst = (size_t)v * n + gid;
    if (keys_by_inst[inst] == 0xffffffffu) continue;  // culled in this view
    float4* ap = reinterpret_cast<float4*>(acc + inst * kAcc);
    const float4 a3 = ap[3];
    // coverage 0: the rasterizer never flushed into this line, which is
    // still all zero, and every output term is zero
    if (a3.x == 0.0f) continue;
    const float4 a0 = ap[0], a1 = ap[1], a2 = ap[2];
    // consume and clear so the accumulator starts at zero next step
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    ap[0] = z4; ap[1] = z4; ap[2] = z4; ap[3] = z4;
    any = true;
    const ViewDev& V = views[v];
    const float vxy[2] = {a0.x, a0.y};
    const float vcon[3] = {a0.z, a0.w, a1.x};
    const float vcol[3] = {a1.y, a1.z, a1.w};
    const float vca = a2.x;
    // screen statistics [0..3]
    st[0] += a0.x; st[1] += a0.y; st[2] += a2.z; st[3] += a2.w;
    // SH
    const float dx = mean[0] - V.campos[0], dy = mean[1] - V.campos[1], dz = mean[2] - V.campos[2];
    const float dd = (dx * dx + dy * dy) + dz * dz;
    const float rs = 1.0f / rgs_sqrt(fmaxf_(dd, RGS_F(1e-12)));
    sh_vjp_acc<DEG>(dx * rs, dy * rs, dz * rs, vcol, gsh);
    const float mcx = ((V.vm[0] * mean[0] + V.vm[4] * mean[1]) + V.vm[8] * mean[2]) + V.vm[12];
    const float mcy = ((V.vm[1] * mean[0] + V.vm[5] * mean[1]) + V.vm[9] * mean[2]) + V.vm[13];
    const float mcz = ((V.vm[2] * mean[0] + V.vm[6] * mean[1]) + V.vm[10] * mean[2]) + V.vm[14];
    if (mcz < RGS_F(0.01)) continue;
    Proj2 pc;
    project_cov(g, V, mcx, mcy, mcz, pc);
    M2 c2 = pc.cov2d;
    const float det_raw = fmaxf_(det2(c2), 0.0f);
    c2.c00 = c2.c00 + blur;
    c2.c11 = c2.c11 + blur;
    const float det_end = fmaxf_(det2(c2), RGS_F(1e-12));
    const float comp = det_raw <= 0.0f ? 0.0f : rgs_sqrt(det_raw / det_end);
    const float abs_score = rgs_sqrt(a2.z * a2.z + a2.w * a2.w);
    st[4] += a3.x * abs_score;
    st[5] += a3.x;
    st[6] += a3.x * mcz;
    gop += comp * vca * opac * (1.0f - opac);
    // v_cov2d = -Cinv * V * Cinv (inverse2x2_vjp) with V = [[vc0, vc1/2], [vc1/2, vc2]]
    const M2 ci = inverse2x2(c2);
    const float h = 0.5f * vcon[1];
    // (-ci) * V
    const float t00 = -ci.c00 * vcon[0] + -ci.c10 * h, t01 = -ci.c01 * vcon[0] + -ci.c11 * h;
    const float t10 = -ci.c00 * h + -ci.c10 * vcon[2], t11 = -ci.c01 * h + -ci.c11 * vcon[2];
    // (...) * ci
    M2 vc;
    vc.c00 = t00 * ci.c00 + t10 * ci.c01;
    vc.c01 = t01 * ci.c00 + t11 * ci.c01;
    vc.c10 = t00 * ci.c10 + t10 * ci.c11;
    vc.c11 = t01 * ci.c10 + t11 * ci.c11;
    // persp_proj_vjp (helpers.wgsl:382-427)
    const float x = mcx, y = mcy, zz = mcz;
    const float rz = 1.0f / fmaxf_(fabsf(zz), RGS_F(0.01));
    const float rz2 = rz * rz;
    float vm0 = V.fx * rz * vxy[0];
    float vm1 = V.fy * rz * vxy[1];
    float vm2 = -(V.fx * x * vxy[0] + V.fy * y * vxy[1]) * rz2;
    // v_J = vc * J * cc^T + vc^T * J * cc   (cc symmetric up to rounding: use both literally)
    const float J[3][2] = {{pc.jx, 0.0f}, {0.0f, pc.jy}, {pc.j20, pc.j21}};
    float vJ[3][2];
    {
      float VJ[3][2], VtJ[3][2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        VJ[c][0] = vc.c00 * J[c][0] + vc.c10 * J[c][1];
        VJ[c][1] = vc.c01 * J[c][0] + vc.c11 * J[c][1];
        VtJ[c][0] = vc.c00 * J[c][0] + vc.c01 * J[c][1];
        VtJ[c][1] = vc.c10 * J[c][0] + vc.c11 * J[c][1];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          // (VJ * cc^T)[c][r] = sum_k VJ[k][r] * cc[k][c];  (VtJ * cc)[c][r] = sum_k VtJ[k][r] * cc[c][k]
          const float p1 = (VJ[0][r] * pc.cc[0][c] + VJ[1][r] * pc.cc[1][c]) + VJ[2][r] * pc.cc[2][c];
          const float p2 = (VtJ[0][r] * pc.cc[c][0] + VtJ[1][r] * pc.cc[c][1]) + VtJ[2][r] * pc.cc[c][2];
          vJ[c][r] = p1 + p2;
        }
    }
    if (x * rz <= V.lim_px && x * rz >= V.lim_nx) {
      vm0 = vm0 + -V.fx * rz2 * vJ[2][0];
      vm2 = vm2 + 2.0f * V.fx * x * rz2 * rz * vJ[2][0];
    } else {
      const float cxl = clampf_(x * rz, V.lim_nx, V.lim_px);
      vm2 = vm2 + V.fx * cxl * rz2 * vJ[2][0];
    }
    if (y * rz <= V.lim_py && y * rz >= V.lim_ny) {
      vm1 = vm1 + -V.fy * rz2 * vJ[2][1];
      vm2 = vm2 + 2.0f * V.fy * y * rz2 * rz * vJ[2][1];
    } else {
      const float cyl = clampf_(y * rz, V.lim_ny, V.lim_py);
      vm2 = vm2 + V.fy * cyl * rz2 * vJ[2][1];
    }
    vm2 = vm2 + (-V.fx * rz2 * vJ[0][0] - V.fy * rz2 * vJ[1][1]);
    // v_cov_cam = J^T * vc * J  (3x3, [c][r] = sum_k,l J[r][k]... written out)
    float vcc[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        // (J^T vc)[col k][row r] = J[r][0]*vc[k][0] + J[r][1]*vc[k][1]; then * J: sum_k (J^T vc)[k][r] * J[c][k]
        const float a0_ = J[r][0] * vc.c00 + J[r][1] * vc.c01;
        const float a1_ = J[r][0] * vc.c10 + J[r][1] * vc.c11;
        vcc[c][r] = a0_ * J[c][0] + a1_ * J[c][1];
      }
    // v_cov3d = R^T * vcc * R, R[c][r] = vm[c*4+r]
    float T1[3][3], vc3[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)  // (R^T vcc)[c][r] = sum_k R^T[k][r] vcc[c][k] = sum_k R[r][k] vcc[c][k]
        T1[c][r] = (V.vm[r * 4 + 0] * vcc[c][0] + V.vm[r * 4 + 1] * vcc[c][1]) + V.vm[r * 4 + 2] * vcc[c][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r)  // (T1 R)[c][r] = sum_k T1[k][r] R[c][k]
        vc3[c][r] = (T1[0][r] * V.vm[c * 4 + 0] + T1[1][r] * V.vm[c * 4 + 1]) + T1[2][r] * V.vm[c * 4 + 2];
    // v_m = (vc3 + vc3^T) * m
    float vm_[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) s += (vc3[k][r] + vc3[r][k]) * g.m[c][k];
        vm_[c][r] = s;
      }
    float rot[3][3];
    quat_rot(q[0], q[1], q[2], q[3], rot);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float vs = (rot[c][0] * vm_[c][0] + rot[c][1] * vm_[c][1]) + rot[c][2] * vm_[c][2];
      gls[c] += vs * sc[c];
    }
    // quat_to_mat3_vjp(q, v_m * S)
    float vr[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) vr[c][r] = vm_[c][r] * sc[c];
    const float w = q[0], qx = q[1], qy = q[2], qz = q[3];
    float dq[4];
    dq[0] = 2.0f * (qx * (vr[1][2] - vr[2][1]) + qy * (vr[2][0] - vr[0][2]) + qz * (vr[0][1] - vr[1][0]));
    dq[1] = 2.0f * (-2.0f * qx * (vr[1][1] + vr[2][2]) + qy * (vr[0][1] + vr[1][0]) + qz * (vr[0][2] + vr[2][0]) +
                    w * (vr[1][2] - vr[2][1]));
    dq[2] = 2.0f * (qx * (vr[0][1] + vr[1][0]) - 2.0f * qy * (vr[0][0] + vr[2][2]) + qz * (vr[1][2] + vr[2][1]) +
                    w * (vr[2][0] - vr[0][2]));
    dq[3] = 2.0f * (qx * (vr[0][2] + vr[2][0]) + qy * (vr[1][2] + vr[2][1]) - 2.0f * qz * (vr[0][0] + vr[1][1]) +
                    w * (vr[0][1] - vr[1][0]));
    // normalize_vjp(qu): (|q|^2 I - q q^T) / |q|^3, symmetric
    {
      const float len = rgs_sqrt(fmaxf_(qn, RGS_F(1e-20)));
      const float sN = 1.0f / (len * qn);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const float e = (k == r) ? (qn - qu[k] * qu[k]) : -(qu[k] * qu[r]);
          s += e * sN * dq[k];
        }
        gq[r] += s;
      }
    }
    // v_mean = R^T v_mean_c
#pragma unroll
    for (int r = 0; r < 3; ++r)
      gmean[r] += (V.vm[r * 4 + 0] * vm0 + V.vm[r * 4 + 1] * vm1) + V.vm[r * 4 + 2] * vm2;
  }

  // write (or accumulate) the gradient planes
  auto out = [&](int plane, float val) {
    float* p = grads + (size_t)plane * npad + gid;
    *p = accumulate ? (*p + val) : val;
  };
  if (!any && accumulate) return;
  out(0, gmean[0]); out(1, gmean[1]); out(2, gmean[2]);
  out(3, gq[0]); out(4, gq[1]); out(5, gq[2]); out(6, gq[3]);
  out(7, gls[0]); out(8, gls[1]); out(9, gls[2]);
  out(10, gop);
#pragma unroll
  for (int i = 0; i < NK * 3; ++i) out(11 + i, gsh[i]);
  if (stats) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      float* p = stats + (size_t)k * npad + gid;
      *p = accumulate ? (*p + st[k]) : st[k];
    }
  }
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
                        float* stats, int accumulate, const uint32_t* counters) {
  if (n == 0) return;
  const dim3 grid((n + 255) / 256), block(256);
  switch (deg) {
    case 0: k_project_bwd<0><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters); break;
    case 1: k_project_bwd<1><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters); break;
    case 2: k_project_bwd<2><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters); break;
    case 3: k_project_bwd<3><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters); break;
    default: k_project_bwd<4><<<grid, block, 0, s>>>(P, npad, n, views, v0, nv, blur, acc, keys_by_inst, grads, stats, accumulate, counters); break;
  }
}

}  // namespace rgs
