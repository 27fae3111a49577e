// Oracle: CPU restatement of RustGS's Gaussian-splatting render/train path.
//
// TEST INFRASTRUCTURE ONLY (see rgs_oracle.h).  Every function below names the
// reference file:line it restates.  Arithmetic follows the WGSL expressions
// left to right with no contraction (the file is built with
// -ffp-contract=off); constants are WGSL abstract floats, i.e. binary64
// literals rounded to the working type.  exp/log/sqrt in FP32 use the
// deterministic scalar routines of rustscan_amd/csrc/rgs_math.h (WGSL leaves
// their rounding to the implementation); FP64 uses libm.
//
// Tie rule for the depth sort (the reference fills its presort list with an
// atomicAdd, project_forward.wgsl:94-96, so equal depths come out in a
// nondeterministic order): ascending (depth, global gid).
//
// Tile-count rule (the reference counts tiles in K1 from the 2D covariance,
// project_forward.wgsl:62-89, but emits them in K5 from the inverse of the
// conic, map_gaussian_to_intersects.wgsl:26-48; when the two disagree the
// reference leaves zero-filled or overwritten slots): both the count and the
// emission use K5's rule.  orc_fwd_stats.k1_count_mismatch reports how many
// Gaussians the reference would have mis-sized.
#include "rgs_oracle.h"

#include "../rustscan_amd/csrc/rgs_math.h"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- scalar math
template <class T> struct Mth;
template <> struct Mth<float> {
  static float exp(float x) { return rgs_exp(x); }
  static float log(float x) { return rgs_log(x); }
  static float sqrt(float x) { return rgs_sqrt(x); }
};
template <> struct Mth<double> {
  static double exp(double x) { return std::exp(x); }
  static double log(double x) { return std::log(x); }
  static double sqrt(double x) { return std::sqrt(x); }
};

template <class T> static inline T K(double c) { return (T)c; }
template <class T> static inline T mn(T a, T b) { return a < b ? a : b; }
template <class T> static inline T mx(T a, T b) { return a > b ? a : b; }
template <class T> static inline T clampv(T x, T lo, T hi) { return mn(mx(x, lo), hi); }

// Matrices are column-major like WGSL: m[col][row].
template <class T, int C, int R> struct Mat { T m[C][R]; };
template <class T> using M2 = Mat<T, 2, 2>;
template <class T> using M3 = Mat<T, 3, 3>;
template <class T> using M32 = Mat<T, 3, 2>;  // mat3x2<f32>: 3 columns of vec2
template <class T> using M23 = Mat<T, 2, 3>;  // mat2x3<f32>: 2 columns of vec3

// (A: K cols x R rows) * (B: C cols x K rows): o[c][r] = sum_k A[k][r]*B[c][k]
template <class T, int R, int KK, int C>
static Mat<T, C, R> mul(const Mat<T, KK, R>& a, const Mat<T, C, KK>& b) {
  Mat<T, C, R> o;
  for (int c = 0; c < C; ++c)
    for (int r = 0; r < R; ++r) {
      T s = a.m[0][r] * b.m[c][0];
      for (int k = 1; k < KK; ++k) s = s + a.m[k][r] * b.m[c][k];
      o.m[c][r] = s;
    }
  return o;
}
template <class T, int C, int R>
static Mat<T, R, C> tr(const Mat<T, C, R>& a) {
  Mat<T, R, C> o;
  for (int c = 0; c < C; ++c)
    for (int r = 0; r < R; ++r) o.m[r][c] = a.m[c][r];
  return o;
}
template <class T, int C, int R>
static Mat<T, C, R> add(const Mat<T, C, R>& a, const Mat<T, C, R>& b) {
  Mat<T, C, R> o;
  for (int c = 0; c < C; ++c)
    for (int r = 0; r < R; ++r) o.m[c][r] = a.m[c][r] + b.m[c][r];
  return o;
}
template <class T> static T det2(const M2<T>& m) {
  return m.m[0][0] * m.m[1][1] - m.m[0][1] * m.m[1][0];
}

// ------------------------------------------------------------ dual numbers
// Forward-mode derivative in 3 directions; used only by the exact-gradient
// mode to differentiate the SH colour with respect to the view direction.
template <class T> struct D3 {
  T v, d[3];
};
template <class T> static D3<T> dconst(T c) { return {c, {0, 0, 0}}; }
template <class T> static D3<T> operator+(D3<T> a, D3<T> b) {
  return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
template <class T> static D3<T> operator-(D3<T> a, D3<T> b) {
  return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
template <class T> static D3<T> operator-(D3<T> a) { return {-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
template <class T> static D3<T> operator*(D3<T> a, D3<T> b) {
  return {a.v * b.v,
          {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}

// --------------------------------------------------------------- SH colour
// project_visible.wgsl:141-224 (sh_coeffs_to_color).  S is the scalar type of
// the view direction (T, or D3<T> for derivatives); coefficients are T.
template <class T, class S> struct Sc {
  static S c(double v) { return S((T)v); }
  static S mulc(S a, T coef) { return a * S(coef); }
};
template <class T> struct Sc<T, D3<T>> {
  static D3<T> c(double v) { return dconst<T>((T)v); }
  static D3<T> mulc(D3<T> a, T coef) { return a * dconst<T>(coef); }
};

template <class T, class S>
static void sh_color_generic(int degree, const S vd[3], const T* co, S out[3]) {
  using C = Sc<T, S>;
  // co[k*3 + ch]
  auto coef = [&](int k, int ch) { return co[k * 3 + ch]; };
  S color[3];
  for (int ch = 0; ch < 3; ++ch) color[ch] = C::mulc(C::c(0.28209479177387814), coef(0, ch));
  if (degree == 0) {
    for (int ch = 0; ch < 3; ++ch) out[ch] = color[ch];
    return;
  }
  const S x = vd[0], y = vd[1], z = vd[2];
  const S fTmp0A = C::c(0.48860251190292);
  for (int ch = 0; ch < 3; ++ch) {
    S t = C::mulc(-y, coef(1, ch)) + C::mulc(z, coef(2, ch)) - C::mulc(x, coef(3, ch));
    color[ch] = color[ch] + fTmp0A * t;
  }
  if (degree == 1) {
    for (int ch = 0; ch < 3; ++ch) out[ch] = color[ch];
    return;
  }
  const S z2 = z * z;
  const S fTmp0B = C::c(-1.092548430592079) * z;
  const S fTmp1A = C::c(0.5462742152960395);
  const S fC1 = x * x - y * y;
  const S fS1 = C::c(2.0) * x * y;
  const S pSH6 = C::c(0.9461746957575601) * z2 - C::c(0.3153915652525201);
  const S pSH7 = fTmp0B * x;
  const S pSH5 = fTmp0B * y;
  const S pSH8 = fTmp1A * fC1;
  const S pSH4 = fTmp1A * fS1;
  for (int ch = 0; ch < 3; ++ch) {
    S t = C::mulc(pSH4, coef(4, ch)) + C::mulc(pSH5, coef(5, ch)) + C::mulc(pSH6, coef(6, ch)) +
          C::mulc(pSH7, coef(7, ch)) + C::mulc(pSH8, coef(8, ch));
    color[ch] = color[ch] + t;
  }
  if (degree == 2) {
    for (int ch = 0; ch < 3; ++ch) out[ch] = color[ch];
    return;
  }
  const S fTmp0C = C::c(-2.285228997322329) * z2 + C::c(0.4570457994644658);
  const S fTmp1B = C::c(1.445305721320277) * z;
  const S fTmp2A = C::c(-0.5900435899266435);
  const S fC2 = x * fC1 - y * fS1;
  const S fS2 = x * fS1 + y * fC1;
  const S pSH12 = z * (C::c(1.865881662950577) * z2 - C::c(1.119528997770346));
  const S pSH13 = fTmp0C * x;
  const S pSH11 = fTmp0C * y;
  const S pSH14 = fTmp1B * fC1;
  const S pSH10 = fTmp1B * fS1;
  const S pSH15 = fTmp2A * fC2;
  const S pSH9 = fTmp2A * fS2;
  for (int ch = 0; ch < 3; ++ch) {
    S t = C::mulc(pSH9, coef(9, ch)) + C::mulc(pSH10, coef(10, ch)) + C::mulc(pSH11, coef(11, ch)) +
          C::mulc(pSH12, coef(12, ch)) + C::mulc(pSH13, coef(13, ch)) + C::mulc(pSH14, coef(14, ch)) +
          C::mulc(pSH15, coef(15, ch));
    color[ch] = color[ch] + t;
  }
  if (degree == 3) {
    for (int ch = 0; ch < 3; ++ch) out[ch] = color[ch];
    return;
  }
  const S fTmp0D = z * (C::c(-4.683325804901025) * z2 + C::c(2.007139630671868));
  const S fTmp1C = C::c(3.31161143515146) * z2 - C::c(0.47308734787878);
  const S fTmp2B = C::c(-1.770130769779931) * z;
  const S fTmp3A = C::c(0.6258357354491763);
  const S fC3 = x * fC2 - y * fS2;
  const S fS3 = x * fS2 + y * fC2;
  const S pSH20 = C::c(1.984313483298443) * z * pSH12 - C::c(1.006230589874905) * pSH6;
  const S pSH21 = fTmp0D * x;
  const S pSH19 = fTmp0D * y;
  const S pSH22 = fTmp1C * fC1;
  const S pSH18 = fTmp1C * fS1;
  const S pSH23 = fTmp2B * fC2;
  const S pSH17 = fTmp2B * fS2;
  const S pSH24 = fTmp3A * fC3;
  const S pSH16 = fTmp3A * fS3;
  for (int ch = 0; ch < 3; ++ch) {
    S t = C::mulc(pSH16, coef(16, ch)) + C::mulc(pSH17, coef(17, ch)) + C::mulc(pSH18, coef(18, ch)) +
          C::mulc(pSH19, coef(19, ch)) + C::mulc(pSH20, coef(20, ch)) + C::mulc(pSH21, coef(21, ch)) +
          C::mulc(pSH22, coef(22, ch)) + C::mulc(pSH23, coef(23, ch)) + C::mulc(pSH24, coef(24, ch));
    color[ch] = color[ch] + t;
  }
  for (int ch = 0; ch < 3; ++ch) out[ch] = color[ch];
}

// helpers.wgsl:289-380 (sh_to_color_vjp): grads[k*3+ch]
template <class T>
static void sh_vjp(int degree, const T vd[3], const T vc[3], T* g) {
  auto put = [&](int k, T b) {
    for (int ch = 0; ch < 3; ++ch) g[k * 3 + ch] = b * vc[ch];
  };
  for (int ch = 0; ch < 3; ++ch) g[ch] = K<T>(0.28209479177387814) * vc[ch];
  if (degree == 0) return;
  const T x = vd[0], y = vd[1], z = vd[2];
  const T fTmp0A = K<T>(0.48860251190292);
  // grads[1] = -fTmp0A * y * v_color  -> ((-fTmp0A)*y)*v
  put(1, -fTmp0A * y);
  put(2, fTmp0A * z);
  put(3, -fTmp0A * x);
  if (degree == 1) return;
  const T z2 = z * z;
  const T fTmp0B = K<T>(-1.092548430592079) * z;
  const T fTmp1A = K<T>(0.5462742152960395);
  const T fC1 = x * x - y * y;
  const T fS1 = K<T>(2.0) * x * y;
  const T pSH6 = K<T>(0.9461746957575601) * z2 - K<T>(0.3153915652525201);
  const T pSH7 = fTmp0B * x;
  const T pSH5 = fTmp0B * y;
  const T pSH8 = fTmp1A * fC1;
  const T pSH4 = fTmp1A * fS1;
  put(4, pSH4); put(5, pSH5); put(6, pSH6); put(7, pSH7); put(8, pSH8);
  if (degree == 2) return;
  const T fTmp0C = K<T>(-2.285228997322329) * z2 + K<T>(0.4570457994644658);
  const T fTmp1B = K<T>(1.445305721320277) * z;
  const T fTmp2A = K<T>(-0.5900435899266435);
  const T fC2 = x * fC1 - y * fS1;
  const T fS2 = x * fS1 + y * fC1;
  const T pSH12 = z * (K<T>(1.865881662950577) * z2 - K<T>(1.119528997770346));
  const T pSH13 = fTmp0C * x;
  const T pSH11 = fTmp0C * y;
  const T pSH14 = fTmp1B * fC1;
  const T pSH10 = fTmp1B * fS1;
  const T pSH15 = fTmp2A * fC2;
  const T pSH9 = fTmp2A * fS2;
  put(9, pSH9); put(10, pSH10); put(11, pSH11); put(12, pSH12);
  put(13, pSH13); put(14, pSH14); put(15, pSH15);
  if (degree == 3) return;
  const T fTmp0D = z * (K<T>(-4.683325804901025) * z2 + K<T>(2.007139630671868));
  const T fTmp1C = K<T>(3.31161143515146) * z2 - K<T>(0.47308734787878);
  const T fTmp2B = K<T>(-1.770130769779931) * z;
  const T fTmp3A = K<T>(0.6258357354491763);
  const T fC3 = x * fC2 - y * fS2;
  const T fS3 = x * fS2 + y * fC2;
  const T pSH20 = K<T>(1.984313483298443) * z * pSH12 - K<T>(1.006230589874905) * pSH6;
  put(16, fTmp3A * fS3); put(17, fTmp2B * fS2); put(18, fTmp1C * fS1);
  put(19, fTmp0D * y); put(20, pSH20); put(21, fTmp0D * x);
  put(22, fTmp1C * fC1); put(23, fTmp2B * fC2); put(24, fTmp3A * fC3);
}

// ---------------------------------------------------------- helpers.wgsl
template <class T> static T sigmoid(T x) { return K<T>(1.0) / (K<T>(1.0) + Mth<T>::exp(-x)); }

// helpers.wgsl:56-80
template <class T> static M3<T> quat_to_mat3(const T q[4]) {
  const T w = q[0], x = q[1], y = q[2], z = q[3];
  const T x2 = x * x, y2 = y * y, z2 = z * z;
  const T xy = x * y, xz = x * z, yz = y * z;
  const T wx = w * x, wy = w * y, wz = w * z;
  const T one = K<T>(1.0), two = K<T>(2.0);
  M3<T> r;
  r.m[0][0] = one - two * (y2 + z2); r.m[0][1] = two * (xy + wz); r.m[0][2] = two * (xz - wy);
  r.m[1][0] = two * (xy - wz); r.m[1][1] = one - two * (x2 + z2); r.m[1][2] = two * (yz + wx);
  r.m[2][0] = two * (xz + wy); r.m[2][1] = two * (yz - wx); r.m[2][2] = one - two * (x2 + y2);
  return r;
}

// helpers.wgsl:82-103
template <class T> static void quat_to_mat3_vjp(const T q[4], const M3<T>& v, T out[4]) {
  const T w = q[0], x = q[1], y = q[2], z = q[3];
  const T two = K<T>(2.0);
  T a = x * (v.m[1][2] - v.m[2][1]) + y * (v.m[2][0] - v.m[0][2]) + z * (v.m[0][1] - v.m[1][0]);
  T b = -two * x * (v.m[1][1] + v.m[2][2]) + y * (v.m[0][1] + v.m[1][0]) + z * (v.m[0][2] + v.m[2][0]) +
        w * (v.m[1][2] - v.m[2][1]);
  T c = x * (v.m[0][1] + v.m[1][0]) - two * y * (v.m[0][0] + v.m[2][2]) + z * (v.m[1][2] + v.m[2][1]) +
        w * (v.m[2][0] - v.m[0][2]);
  T d = x * (v.m[0][2] + v.m[2][0]) + y * (v.m[1][2] + v.m[2][1]) - two * z * (v.m[0][0] + v.m[1][1]) +
        w * (v.m[0][1] - v.m[1][0]);
  out[0] = a * two; out[1] = b * two; out[2] = c * two; out[3] = d * two;
}

// helpers.wgsl:105-119: returns the 4x4 Jacobian of v/|v| (columns), already scaled
template <class T> static Mat<T, 4, 4> normalize_vjp(const T v[4]) {
  T vs[4] = {v[0] * v[0], v[1] * v[1], v[2] * v[2], v[3] * v[3]};
  const T lsq = ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3];
  const T len = Mth<T>::sqrt(mx(lsq, K<T>(1e-20)));
  const T cc[3] = {-v[0] * v[1], -v[1] * v[2], -v[2] * v[0]};
  const T cs[3] = {-v[0] * v[3], -v[1] * v[3], -v[2] * v[3]};
  Mat<T, 4, 4> m;
  m.m[0][0] = lsq - vs[0]; m.m[0][1] = cc[0]; m.m[0][2] = cc[2]; m.m[0][3] = cs[0];
  m.m[1][0] = cc[0]; m.m[1][1] = lsq - vs[1]; m.m[1][2] = cc[1]; m.m[1][3] = cs[1];
  m.m[2][0] = cc[2]; m.m[2][1] = cc[1]; m.m[2][2] = lsq - vs[2]; m.m[2][3] = cs[2];
  m.m[3][0] = cs[0]; m.m[3][1] = cs[1]; m.m[3][2] = cs[2]; m.m[3][3] = lsq - vs[3];
  const T s = K<T>(1.0) / (len * lsq);
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) m.m[c][r] = m.m[c][r] * s;
  return m;
}

template <class T> static M3<T> diag3(const T s[3]) {
  M3<T> o;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) o.m[c][r] = (c == r) ? s[c] : K<T>(0.0);
  return o;
}

// helpers.wgsl:128-136
template <class T> static M3<T> calc_cov3d(const T scale[3], const T q[4]) {
  const M3<T> m = mul(quat_to_mat3(q), diag3(scale));
  return mul(m, tr(m));
}

template <class T> struct View {
  T vm[4][4];  // vm[col][row]
  T fx, fy, cx, cy;
  uint32_t W, H;
  T campos[3];
  M3<T> rot() const {
    M3<T> r;
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < 3; ++k) r.m[c][k] = vm[c][k];
    return r;
  }
  explicit View(const orc_camera& c) {
    for (int col = 0; col < 4; ++col)
      for (int row = 0; row < 4; ++row) vm[col][row] = (T)c.viewmat[col * 4 + row];
    fx = (T)c.fx; fy = (T)c.fy; cx = (T)c.cx; cy = (T)c.cy;
    W = c.width; H = c.height;
    for (int i = 0; i < 3; ++i) campos[i] = (T)c.cam_pos[i];
  }
};

// helpers.wgsl:138-155
template <class T> static M32<T> calc_cam_j(const T mc[3], const View<T>& v) {
  const T lpx = (K<T>(1.15) * (T)v.W - v.cx) / v.fx;
  const T lpy = (K<T>(1.15) * (T)v.H - v.cy) / v.fy;
  const T lnx = (K<T>(-0.15) * (T)v.W - v.cx) / v.fx;
  const T lny = (K<T>(-0.15) * (T)v.H - v.cy) / v.fy;
  const T rz = K<T>(1.0) / mx(mc[2], K<T>(0.01));
  const T ux = clampv(mc[0] * rz, lnx, lpx);
  const T uy = clampv(mc[1] * rz, lny, lpy);
  const T dx = v.fx * rz, dy = v.fy * rz;
  M32<T> j;
  j.m[0][0] = dx; j.m[0][1] = K<T>(0.0);
  j.m[1][0] = K<T>(0.0); j.m[1][1] = dy;
  j.m[2][0] = -dx * ux; j.m[2][1] = -dy * uy;
  return j;
}

// helpers.wgsl:157-168
template <class T>
static M2<T> calc_cov2d(const M3<T>& cov3d, const T mc[3], const View<T>& v, M3<T>* cov_cam_out = nullptr,
                        M32<T>* j_out = nullptr) {
  const M3<T> R = v.rot();
  const M3<T> cc = mul(mul(R, cov3d), tr(R));
  const M32<T> j = calc_cam_j(mc, v);
  if (cov_cam_out) *cov_cam_out = cc;
  if (j_out) *j_out = j;
  return mul(mul(j, cc), tr(j));
}

// helpers.wgsl:170-183
template <class T> static T compensate_cov2d(M2<T>& cov, T blur) {
  const M2<T> start = cov;
  M2<T> end = cov;
  end.m[0][0] = end.m[0][0] + blur;
  end.m[1][1] = end.m[1][1] + blur;
  cov = end;
  const T det_raw = mx(det2(start), K<T>(0.0));
  const T det_end = mx(det2(end), K<T>(1e-12));
  if (det_raw <= K<T>(0.0)) return K<T>(0.0);
  return Mth<T>::sqrt(det_raw / det_end);
}

// helpers.wgsl:185-198
template <class T> static M2<T> inverse2x2(const M2<T>& m) {
  M2<T> r = m;
  r.m[0][0] = r.m[0][0] + K<T>(1e-6);
  r.m[1][1] = r.m[1][1] + K<T>(1e-6);
  const T det = det2(r);
  M2<T> o;
  if (det <= K<T>(1e-10)) {
    o.m[0][0] = o.m[0][1] = o.m[1][0] = o.m[1][1] = K<T>(0.0);
    return o;
  }
  const T inv = K<T>(1.0) / det;
  o.m[0][0] = r.m[1][1] * inv;
  o.m[0][1] = -r.m[0][1] * inv;
  o.m[1][0] = -r.m[0][1] * inv;
  o.m[1][1] = r.m[0][0] * inv;
  return o;
}

// helpers.wgsl:200-202: (-m_inv) * v * m_inv
template <class T> static M2<T> inverse2x2_vjp(const M2<T>& minv, const M2<T>& v) {
  M2<T> neg;
  for (int c = 0; c < 2; ++c)
    for (int r = 0; r < 2; ++r) neg.m[c][r] = -minv.m[c][r];
  return mul(mul(neg, v), minv);
}

// helpers.wgsl:204-216
template <class T>
static void get_tile_bbox(const T center[2], const T extent[2], uint32_t tbx, uint32_t tby, int tw,
                          uint32_t out[4]) {
  const T twf = (T)tw;
  const T c0 = center[0] / twf, c1 = center[1] / twf;
  const T e0 = extent[0] / twf, e1 = extent[1] / twf;
  const T bx = (T)tbx, by = (T)tby;
  out[0] = (uint32_t)clampv(c0 - e0, K<T>(0.0), bx);
  out[1] = (uint32_t)clampv(c1 - e1, K<T>(0.0), by);
  out[2] = (uint32_t)clampv(c0 + e0 + K<T>(1.0), K<T>(0.0), bx);
  out[3] = (uint32_t)clampv(c1 + e1 + K<T>(1.0), K<T>(0.0), by);
}

// helpers.wgsl:218-224
template <class T> static void compute_bbox_extent(const M2<T>& cov, T thr, T ext[2]) {
  const T safe = mx(thr, K<T>(0.0));
  ext[0] = Mth<T>::sqrt(mx(K<T>(0.0), K<T>(2.0) * safe * cov.m[0][0]));
  ext[1] = Mth<T>::sqrt(mx(K<T>(0.0), K<T>(2.0) * safe * cov.m[1][1]));
}

// helpers.wgsl:232-235
template <class T> static T calc_sigma(const T p[2], const T conic[3], const T mean[2]) {
  const T dx = p[0] - mean[0], dy = p[1] - mean[1];
  return K<T>(0.5) * (conic[0] * dx * dx + K<T>(2.0) * conic[1] * dx * dy + conic[2] * dy * dy);
}

// helpers.wgsl:237-283
template <class T>
static bool will_primitive_contribute(uint32_t tx, uint32_t ty, int tw, const T mean[2], const T conic[3],
                                      T threshold) {
  const T rx0 = (T)(tx * (uint32_t)tw), ry0 = (T)(ty * (uint32_t)tw);
  const T rx1 = rx0 + (T)tw, ry1 = ry0 + (T)tw;
  const bool x_left = mean[0] < rx0, x_right = mean[0] > rx1;
  const bool in_x = !(x_left || x_right);
  const bool y_above = mean[1] < ry0, y_below = mean[1] > ry1;
  const bool in_y = !(y_above || y_below);
  if (in_x && in_y) return true;
  const T cc[2] = {x_left ? rx0 : rx1, y_above ? ry0 : ry1};
  const T width = rx1 - rx0, height = ry1 - ry0;
  const T d[2] = {x_left ? width : -width, y_above ? height : -height};
  const T diff[2] = {mean[0] - cc[0], mean[1] - cc[1]};
  T t0 = K<T>(0.0), t1 = K<T>(0.0);
  if (!in_y)
    t0 = clampv((d[0] * conic[0] * diff[0] + d[0] * conic[1] * diff[1]) / mx(d[0] * conic[0] * d[0], K<T>(1e-12)),
                K<T>(0.0), K<T>(1.0));
  if (!in_x)
    t1 = clampv((d[1] * conic[1] * diff[0] + d[1] * conic[2] * diff[1]) / mx(d[1] * conic[2] * d[1], K<T>(1e-12)),
                K<T>(0.0), K<T>(1.0));
  const T mp[2] = {cc[0] + t0 * d[0], cc[1] + t1 * d[1]};
  return calc_sigma(mp, conic, mean) <= threshold;
}

// -------------------------------------------------------------- pipeline
template <class T> struct Proj {
  T x, y, a, b, c, r, g, bl, o;
};

struct Isect {
  uint32_t tile, cgid;
};

template <class T> struct Pipe {
  int64_t n = 0;
  int deg = 0, nk = 1;
  const T* tf = nullptr;  // [N*10]
  const T* sh = nullptr;  // [N*K*3]
  const T* op = nullptr;  // [N]
  View<T> v;
  T bg[3];
  T blur;
  int tw;
  uint32_t tbx, tby, ntiles;
  // frozen surrogate terms (FP64 finite differences)
  const T* frozen_vd = nullptr;
  const T* frozen_comp = nullptr;
  bool bwd_termination = false;
  // FP64 replay of an FP32 forward: the backward's skip and stop decisions are
  // taken from the FP32 splats with FP32 arithmetic, the gradient values are
  // computed in T.  Set only on a Pipe<double> built by orc_backward_f64.
  const std::vector<Proj<float>>* shadow = nullptr;
  // Frozen crossing splat (FP64 finite differences of the reference
  // gradient): at the pixels marked in frozen_pix, the splat with global id
  // frozen_g is composited with the projection frozen_proj (its unperturbed
  // values), so that its parameters do not act there.  The reference's
  // backward gives a splat no gradient at the pixels where it takes T to
  // 1e-4 or below (rasterize_backwards.wgsl:108-112), while its gradient for
  // the earlier splats is that of the forward that includes it
  // (rasterize.wgsl:70-80).  cross_gid (filled by every forward) records that
  // splat per pixel, -1 where none.
  int64_t frozen_g = -1;
  const uint8_t* frozen_pix = nullptr;
  Proj<T> frozen_proj{};
  std::vector<int64_t> cross_gid;

  std::vector<uint8_t> pass1;
  std::vector<T> depth;
  std::vector<uint32_t> k1cnt;
  std::vector<uint32_t> order;
  std::vector<Proj<T>> proj;
  std::vector<uint32_t> cnt;
  std::vector<uint32_t> isect_tile, isect_cgid;
  std::vector<uint32_t> offsets;
  std::vector<T> img;
  std::vector<uint8_t> visible;
  std::vector<uint8_t> crossing;  // by compact id
  int64_t mismatch = 0, tie_runs = 0, evaluated = 0, blended = 0;

  explicit Pipe(const orc_camera& c) : v(c) {}

  void mean_c(int64_t g, T out[3]) const {
    const T* p = tf + g * 10;
    for (int i = 0; i < 3; ++i)
      out[i] = ((v.vm[0][i] * p[0] + v.vm[1][i] * p[1]) + v.vm[2][i] * p[2]) + v.vm[3][i];
  }

  // project_forward.wgsl:12-97 (K1)
  bool k1(int64_t g, T* depth_out, uint32_t* count_out) const {
    const T* p = tf + g * 10;
    T mc[3];
    mean_c(g, mc);
    if (mc[2] < K<T>(0.01) || mc[2] > K<T>(1e10)) return false;
    T q[4] = {p[3], p[4], p[5], p[6]};
    const T qn = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    if (qn < K<T>(1e-6)) return false;
    const T is = K<T>(1.0) / Mth<T>::sqrt(qn);  // inverseSqrt
    for (int i = 0; i < 4; ++i) q[i] = q[i] * is;
    const T scale[3] = {Mth<T>::exp(p[7]), Mth<T>::exp(p[8]), Mth<T>::exp(p[9])};
    T opacity = sigmoid(op[g]);
    const M3<T> cov3d = calc_cov3d(scale, q);
    M2<T> cov2d = calc_cov2d(cov3d, mc, v);
    const T comp = compensate_cov2d(cov2d, blur);
    opacity = opacity * (frozen_comp ? frozen_comp[g] : comp);
    if (opacity < K<T>(1.0 / 255.0)) return false;
    const T rz = K<T>(1.0) / mc[2];
    const T mean2d[2] = {v.fx * mc[0] * rz + v.cx, v.fy * mc[1] * rz + v.cy};
    const M2<T> conic = inverse2x2(cov2d);
    const T thr = Mth<T>::log(opacity * K<T>(255.0));
    T ext[2];
    compute_bbox_extent(cov2d, thr, ext);
    if (ext[0] < K<T>(0.0) || ext[1] < K<T>(0.0)) return false;
    if (mean2d[0] + ext[0] <= K<T>(0.0) || mean2d[0] - ext[0] >= (T)v.W || mean2d[1] + ext[1] <= K<T>(0.0) ||
        mean2d[1] - ext[1] >= (T)v.H)
      return false;
    const T cp[3] = {conic.m[0][0], conic.m[0][1], conic.m[1][1]};
    uint32_t bb[4];
    get_tile_bbox(mean2d, ext, tbx, tby, tw, bb);
    const uint32_t bw = bb[2] - bb[0];
    const uint32_t nb = (bb[3] - bb[1]) * bw;
    uint32_t hit = 0;
    for (uint32_t i = 0; i < nb; ++i) {
      const uint32_t tx = bb[0] + i % bw, ty = bb[1] + i / bw;
      if (will_primitive_contribute(tx, ty, tw, mean2d, cp, thr)) ++hit;
    }
    *depth_out = mc[2];
    *count_out = hit;
    return true;
  }

  // project_visible.wgsl:129-214 (K3)
  Proj<T> k3(int64_t g, T vd_out[3]) const {
    const T* p = tf + g * 10;
    T mean[3] = {p[0], p[1], p[2]};
    const T scale[3] = {Mth<T>::exp(p[7]), Mth<T>::exp(p[8]), Mth<T>::exp(p[9])};
    T q[4] = {p[3], p[4], p[5], p[6]};
    const T qn = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    const T is = K<T>(1.0) / Mth<T>::sqrt(qn);  // normalize()
    for (int i = 0; i < 4; ++i) q[i] = q[i] * is;
    T mc[3];
    mean_c(g, mc);
    T opacity = sigmoid(op[g]);
    const M3<T> cov3d = calc_cov3d(scale, q);
    M2<T> cov2d = calc_cov2d(cov3d, mc, v);
    const T comp = compensate_cov2d(cov2d, blur);
    opacity = opacity * (frozen_comp ? frozen_comp[g] : comp);
    const M2<T> conic = inverse2x2(cov2d);
    const T rz = K<T>(1.0) / mc[2];
    Proj<T> o;
    o.x = v.fx * mc[0] * rz + v.cx;
    o.y = v.fy * mc[1] * rz + v.cy;
    T vd[3];
    if (frozen_vd) {
      for (int i = 0; i < 3; ++i) vd[i] = frozen_vd[g * 3 + i];
    } else {
      const T d[3] = {mean[0] - v.campos[0], mean[1] - v.campos[1], mean[2] - v.campos[2]};
      const T dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
      const T s = K<T>(1.0) / Mth<T>::sqrt(mx(dd, K<T>(1e-12)));
      for (int i = 0; i < 3; ++i) vd[i] = d[i] * s;
    }
    if (vd_out)
      for (int i = 0; i < 3; ++i) vd_out[i] = vd[i];
    T col[3];
    sh_color_generic<T, T>(deg, vd, sh + g * nk * 3, col);
    o.a = conic.m[0][0];
    o.b = conic.m[0][1];
    o.c = conic.m[1][1];
    o.r = col[0] + K<T>(0.5);
    o.g = col[1] + K<T>(0.5);
    o.bl = col[2] + K<T>(0.5);
    o.o = opacity;
    return o;
  }

  // map_gaussian_to_intersects.wgsl:22-33: tile bbox and threshold from a projected splat
  void k5_bbox(const Proj<T>& s, uint32_t bb[4], T* thr_out) const {
    const T mean2d[2] = {s.x, s.y};
    const T thr = Mth<T>::log(s.o * K<T>(255.0));
    M2<T> cm;
    cm.m[0][0] = s.a; cm.m[0][1] = s.b; cm.m[1][0] = s.b; cm.m[1][1] = s.c;
    const M2<T> cov = inverse2x2(cm);
    T ext[2];
    compute_bbox_extent(cov, thr, ext);
    get_tile_bbox(mean2d, ext, tbx, tby, tw, bb);
    *thr_out = thr;
  }

  uint32_t k5_count(const Proj<T>& s) const {
    uint32_t bb[4];
    T thr;
    k5_bbox(s, bb, &thr);
    const T mean2d[2] = {s.x, s.y};
    const T cp[3] = {s.a, s.b, s.c};
    const uint32_t bw = bb[2] - bb[0];
    const uint32_t nb = (bb[3] - bb[1]) * bw;
    uint32_t hit = 0;
    for (uint32_t i = 0; i < nb; ++i)
      if (will_primitive_contribute(bb[0] + i % bw, bb[1] + i / bw, tw, mean2d, cp, thr)) ++hit;
    return hit;
  }

  void forward() {
    tbx = (v.W + (uint32_t)tw - 1) / (uint32_t)tw;
    tby = (v.H + (uint32_t)tw - 1) / (uint32_t)tw;
    ntiles = tbx * tby;
    pass1.assign(n, 0);
    depth.assign(n, K<T>(0.0));
    k1cnt.assign(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t g = 0; g < n; ++g) {
      T d;
      uint32_t c;
      if (k1(g, &d, &c)) {
        pass1[g] = 1;
        depth[g] = d;
        k1cnt[g] = c;
      }
    }
    // presort in gid order, then stable sort by depth (sorting.rs:34-61 with the canonical tie rule)
    order.clear();
    for (int64_t g = 0; g < n; ++g)
      if (pass1[g]) order.push_back((uint32_t)g);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return depth[a] < depth[b]; });
    const int64_t V = (int64_t)order.size();
    tie_runs = 0;
    for (int64_t i = 1; i < V; ++i)
      if (depth[order[i]] == depth[order[i - 1]] && (i == 1 || depth[order[i - 1]] != depth[order[i - 2]]))
        ++tie_runs;
    proj.assign(V, Proj<T>{});
    cnt.assign(V, 0);
    int64_t mism = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : mism)
    for (int64_t i = 0; i < V; ++i) {
      proj[i] = k3(order[i], nullptr);
      cnt[i] = k5_count(proj[i]);
      if (cnt[i] != k1cnt[order[i]]) ++mism;
    }
    mismatch = mism;
    // prefix sum (prefix_sum.wgsl) + emission (map_gaussian_to_intersects.wgsl:34-48)
    std::vector<uint64_t> cum(V + 1, 0);
    for (int64_t i = 0; i < V; ++i) cum[i + 1] = cum[i] + cnt[i];
    const int64_t M = (int64_t)cum[V];
    std::vector<Isect> em(M);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < V; ++i) {
      uint32_t bb[4];
      T thr;
      k5_bbox(proj[i], bb, &thr);
      const T mean2d[2] = {proj[i].x, proj[i].y};
      const T cp[3] = {proj[i].a, proj[i].b, proj[i].c};
      const uint32_t bw = bb[2] - bb[0];
      const uint32_t nb = (bb[3] - bb[1]) * bw;
      uint64_t off = cum[i];
      for (uint32_t k = 0; k < nb; ++k) {
        const uint32_t tx = bb[0] + k % bw, ty = bb[1] + k / bw;
        if (will_primitive_contribute(tx, ty, tw, mean2d, cp, thr)) em[off++] = {tx + ty * tbx, (uint32_t)i};
      }
    }
    // stable sort by tile id (radix_sort.wgsl contract) as a counting sort; ranges (get_tile_offsets.wgsl)
    std::vector<uint64_t> tstart(ntiles + 1, 0);
    for (int64_t e = 0; e < M; ++e) tstart[em[e].tile + 1]++;
    for (uint32_t t = 0; t < ntiles; ++t) tstart[t + 1] += tstart[t];
    isect_tile.assign(M, 0);
    isect_cgid.assign(M, 0);
    {
      std::vector<uint64_t> pos(tstart.begin(), tstart.end() - 1);
      for (int64_t e = 0; e < M; ++e) {
        const uint64_t d = pos[em[e].tile]++;
        isect_tile[d] = em[e].tile;
        isect_cgid[d] = em[e].cgid;
      }
    }
    offsets.assign(2 * (size_t)ntiles, 0);
    for (uint32_t t = 0; t < ntiles; ++t)
      if (tstart[t + 1] > tstart[t]) {
        offsets[2 * t] = (uint32_t)tstart[t];
        offsets[2 * t + 1] = (uint32_t)tstart[t + 1];
      }
    raster();
  }

  // rasterize.wgsl:21-99 (K7)
  void raster() {
    img.assign((size_t)v.W * v.H * 4, K<T>(0.0));
    visible.assign(n, 0);
    crossing.assign(proj.size(), 0);
    cross_gid.assign((size_t)v.W * v.H, -1);
    int64_t ev = 0, bl = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : ev, bl)
    for (int64_t t = 0; t < (int64_t)ntiles; ++t) {
      const uint32_t tx = (uint32_t)t % tbx, ty = (uint32_t)t / tbx;
      const uint32_t s = offsets[2 * t], e = offsets[2 * t + 1];
      for (int ly = 0; ly < tw; ++ly)
        for (int lx = 0; lx < tw; ++lx) {
          const uint32_t px = tx * tw + lx, py = ty * tw + ly;
          if (px >= v.W || py >= v.H) continue;
          const T pc[2] = {(T)px + K<T>(0.5), (T)py + K<T>(0.5)};
          T Tr = K<T>(1.0);
          T acc[3] = {K<T>(0.0), K<T>(0.0), K<T>(0.0)};
          const size_t pix = (size_t)py * v.W + px;
          const bool frz = frozen_pix && frozen_pix[pix];
          for (uint32_t i = s; i < e; ++i) {
            const Proj<T>& sp = frz && (int64_t)order[isect_cgid[i]] == frozen_g ? frozen_proj : proj[isect_cgid[i]];
            ++ev;
            const T dx = sp.x - pc[0], dy = sp.y - pc[1];
            const T sigma = K<T>(0.5) * (sp.a * dx * dx + K<T>(2.0) * sp.b * dx * dy + sp.c * dy * dy);
            const T alpha = mn(K<T>(0.999), sp.o * Mth<T>::exp(-sigma));
            if (sigma >= K<T>(0.0) && alpha >= K<T>(1.0 / 255.0)) {
              if (bwd_termination && Tr * (K<T>(1.0) - alpha) <= K<T>(1e-4)) break;
              ++bl;
              const T vis = alpha * Tr;
              acc[0] = acc[0] + mx(sp.r, K<T>(0.0)) * vis;
              acc[1] = acc[1] + mx(sp.g, K<T>(0.0)) * vis;
              acc[2] = acc[2] + mx(sp.bl, K<T>(0.0)) * vis;
              Tr = Tr * (K<T>(1.0) - alpha);
              const uint32_t g = order[isect_cgid[i]];
              __atomic_store_n(&visible[g], (uint8_t)1, __ATOMIC_RELAXED);
              if (Tr < K<T>(1e-4)) {
                __atomic_store_n(&crossing[isect_cgid[i]], (uint8_t)1, __ATOMIC_RELAXED);
                cross_gid[pix] = (int64_t)g;
                break;
              }
            }
          }
          T* o = &img[((size_t)py * v.W + px) * 4];
          o[0] = acc[0] + Tr * bg[0];
          o[1] = acc[1] + Tr * bg[1];
          o[2] = acc[2] + Tr * bg[2];
          o[3] = K<T>(1.0) - Tr;
        }
    }
    evaluated = ev;
    blended = bl;
  }

  // rasterize_backwards.wgsl:24-169 (K8) -> per-compact v_splats[10], screen[5]
  // mode 0: reference; mode 1: exact gradient of raster() (forward sigma order, crossing splat included)
  void raster_bwd(const T* v_output, int mode, std::vector<T>& vs, std::vector<T>& ss) const {
    const int64_t M = (int64_t)isect_cgid.size();
    std::vector<T> buf((size_t)M * 15, K<T>(0.0));
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t t = 0; t < (int64_t)ntiles; ++t) {
      const uint32_t tx = (uint32_t)t % tbx, ty = (uint32_t)t / tbx;
      const uint32_t s = offsets[2 * t], e = offsets[2 * t + 1];
      for (int ly = 0; ly < tw; ++ly)
        for (int lx = 0; lx < tw; ++lx) {
          const uint32_t px = tx * tw + lx, py = ty * tw + ly;
          if (px >= v.W || py >= v.H) continue;
          const size_t pid = (size_t)py * v.W + px;
          const T pc[2] = {(T)px + K<T>(0.5), (T)py + K<T>(0.5)};
          const T* fc = &img[pid * 4];
          const T Tf = K<T>(1.0) - fc[3];
          const T rgbf[4] = {fc[0] - Tf * bg[0], fc[1] - Tf * bg[1], fc[2] - Tf * bg[2], fc[3]};
          const T* go = &v_output[pid * 4];
          const T vo[4] = {go[0], go[1], go[2],
                           (go[3] - ((bg[0] * go[0] + bg[1] * go[1]) + bg[2] * go[2])) * Tf};
          T po[4] = {K<T>(0.0), K<T>(0.0), K<T>(0.0), K<T>(1.0)};
          float shadow_T = 1.0f;  // FP32 transmittance of the replayed forward (shadow mode)
          for (uint32_t i = s; i < e; ++i) {
            const Proj<T>& sp = proj[isect_cgid[i]];
            const T col[3] = {sp.r, sp.g, sp.bl};
            const T cl[3] = {mx(sp.r, K<T>(0.0)), mx(sp.g, K<T>(0.0)), mx(sp.bl, K<T>(0.0))};
            const T dx = sp.x - pc[0], dy = sp.y - pc[1];
            T sigma;
            if (mode == 0)
              sigma = K<T>(0.5) * (sp.a * dx * dx + sp.c * dy * dy) + sp.b * dx * dy;
            else
              sigma = K<T>(0.5) * (sp.a * dx * dx + K<T>(2.0) * sp.b * dx * dy + sp.c * dy * dy);
            const T gauss = Mth<T>::exp(-sigma);
            const T alpha = mn(K<T>(0.999), sp.o * gauss);
            bool last = false;
            bool geo = sp.o * gauss <= K<T>(0.999);  // the alpha clamp gates the geometric terms
            if (shadow) {
              // the same decisions as the FP32 backward over the FP32 splats
              const Proj<float>& fs = (*shadow)[isect_cgid[i]];
              const float fdx = fs.x - ((float)px + 0.5f), fdy = fs.y - ((float)py + 0.5f);
              const float fsig = mode == 0 ? 0.5f * (fs.a * fdx * fdx + fs.c * fdy * fdy) + fs.b * fdx * fdy
                                           : 0.5f * (fs.a * fdx * fdx + 2.0f * fs.b * fdx * fdy + fs.c * fdy * fdy);
              const float fog = fs.o * rgs_exp(-fsig);
              const float falpha = std::min(0.999f, fog);
              geo = fog <= 0.999f;
              if (fsig < 0.0f || falpha < (float)(1.0 / 255.0)) continue;
              const float fnext = shadow_T * (1.0f - falpha);
              if (mode == 0) {
                if (fnext <= 1e-4f) break;
              } else {
                last = fnext < 1e-4f;
              }
              shadow_T = fnext;
            } else {
              if (sigma < K<T>(0.0) || alpha < K<T>(1.0 / 255.0)) continue;
              const T nt = po[3] * (K<T>(1.0) - alpha);
              if (mode == 0) {
                if (nt <= K<T>(1e-4)) break;
              } else {
                last = nt < K<T>(1e-4);
              }
            }
            const T next_T = po[3] * (K<T>(1.0) - alpha);
            const T vis = alpha * po[3];
            T vrgb[3];
            for (int c = 0; c < 3; ++c) vrgb[c] = col[c] >= K<T>(0.0) ? vis * vo[c] : K<T>(0.0);
            for (int c = 0; c < 3; ++c) po[c] = po[c] + vis * cl[c];
            const T ra = K<T>(1.0) / (K<T>(1.0) - alpha);
            T tv[3];
            for (int c = 0; c < 3; ++c) tv[c] = po[3] * cl[c] + (po[c] - rgbf[c]) * ra;
            const T v_alpha = ((tv[0] * vo[0] + tv[1] * vo[1]) + tv[2] * vo[2]) + vo[3] * ra;
            const T v_sigma = -alpha * v_alpha;
            const T vxl[2] = {v_sigma * (sp.a * dx + sp.b * dy), v_sigma * (sp.b * dx + sp.c * dy)};
            T vcon[3] = {K<T>(0.0), K<T>(0.0), K<T>(0.0)};
            T vxy[2] = {K<T>(0.0), K<T>(0.0)};
            T vat = K<T>(0.0), vref = K<T>(0.0);
            if (geo) {
              vcon[0] = K<T>(0.5) * v_sigma * dx * dx;
              vcon[1] = v_sigma * dx * dy;
              vcon[2] = K<T>(0.5) * v_sigma * dy * dy;
              vxy[0] = vxl[0];
              vxy[1] = vxl[1];
              vat = v_alpha * gauss;
              const T fa = mx(rgbf[3], K<T>(1e-5));
              const T lx2 = vxl[0] * (T)v.W, ly2 = vxl[1] * (T)v.H;
              vref = Mth<T>::sqrt(lx2 * lx2 + ly2 * ly2) / fa;
            }
            T* b = &buf[(size_t)i * 15];
            const T add10[10] = {vxy[0], vxy[1], vcon[0], vcon[1], vcon[2], vrgb[0], vrgb[1], vrgb[2], vat, vref};
            for (int k = 0; k < 10; ++k) b[k] = b[k] + add10[k];
            b[10] = b[10] + vxy[0];
            b[11] = b[11] + vxy[1];
            b[12] = b[12] + std::fabs(vxy[0]);
            b[13] = b[13] + std::fabs(vxy[1]);
            b[14] = b[14] + K<T>(1.0);
            po[3] = next_T;
            if (last) break;
          }
        }
    }
    const size_t V = proj.size();
    vs.assign(V * 10, K<T>(0.0));
    ss.assign(V * 5, K<T>(0.0));
    for (int64_t e = 0; e < M; ++e) {
      const uint32_t cg = isect_cgid[e];
      for (int k = 0; k < 10; ++k) vs[cg * 10 + k] = vs[cg * 10 + k] + buf[(size_t)e * 15 + k];
      for (int k = 0; k < 5; ++k) ss[cg * 5 + k] = ss[cg * 5 + k] + buf[(size_t)e * 15 + 10 + k];
    }
  }

  // project_backwards.wgsl:181-320 (K9); mode 1 adds the SH view-direction and
  // blur-compensation derivatives that the reference omits.
  void project_bwd(const std::vector<T>& vs, const std::vector<T>& ss, int mode, T* v_tf, T* v_sh, T* v_op,
                   T* stats) const {
    const int64_t V = (int64_t)proj.size();
    if (v_tf) std::fill(v_tf, v_tf + n * 10, K<T>(0.0));
    if (v_sh) std::fill(v_sh, v_sh + n * nk * 3, K<T>(0.0));
    if (v_op) std::fill(v_op, v_op + n, K<T>(0.0));
    if (stats) std::fill(stats, stats + n * 7, K<T>(0.0));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < V; ++i) {
      const int64_t g = order[i];
      const T* vsp = &vs[i * 10];
      const T* ssp = &ss[i * 5];
      if (stats) {
        for (int k = 0; k < 4; ++k) stats[g * 7 + k] = ssp[k];
      }
      const T vxy[2] = {vsp[0], vsp[1]};
      const T vcon[3] = {vsp[2], vsp[3], vsp[4]};
      const T vcol[3] = {vsp[5], vsp[6], vsp[7]};
      const T vca = vsp[8];
      const T* p = tf + g * 10;
      const T mean[3] = {p[0], p[1], p[2]};
      const T qu[4] = {p[3], p[4], p[5], p[6]};
      const T qn = ((qu[0] * qu[0] + qu[1] * qu[1]) + qu[2] * qu[2]) + qu[3] * qu[3];
      const T is = K<T>(1.0) / Mth<T>::sqrt(qn);
      const T q[4] = {qu[0] * is, qu[1] * is, qu[2] * is, qu[3] * is};
      const T scale[3] = {Mth<T>::exp(p[7]), Mth<T>::exp(p[8]), Mth<T>::exp(p[9])};
      const T d[3] = {mean[0] - v.campos[0], mean[1] - v.campos[1], mean[2] - v.campos[2]};
      const T dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
      const T rs = K<T>(1.0) / Mth<T>::sqrt(mx(dd, K<T>(1e-12)));
      const T vd[3] = {d[0] * rs, d[1] * rs, d[2] * rs};
      std::vector<T> shg((size_t)nk * 3);
      sh_vjp(deg, vd, vcol, shg.data());
      if (v_sh)
        for (int k = 0; k < nk * 3; ++k) v_sh[g * nk * 3 + k] = shg[k];
      const M3<T> R = v.rot();
      T mc[3];
      for (int r = 0; r < 3; ++r) mc[r] = ((R.m[0][r] * mean[0] + R.m[1][r] * mean[1]) + R.m[2][r] * mean[2]) + v.vm[3][r];
      if (mc[2] < K<T>(0.01)) continue;
      const M3<T> rot = quat_to_mat3(q);
      const M3<T> S = diag3(scale);
      const M3<T> m = mul(rot, S);
      const M3<T> cov3d = mul(m, tr(m));
      const M3<T> cov_cam = mul(mul(R, cov3d), tr(R));
      const M32<T> j = calc_cam_j(mc, v);
      M2<T> cov2d = mul(mul(j, cov_cam), tr(j));
      const M2<T> cov_raw = cov2d;
      const T comp = compensate_cov2d(cov2d, blur);
      const T opac = sigmoid(op[g]);
      if (stats) {
        const T abs_score = Mth<T>::sqrt(ssp[2] * ssp[2] + ssp[3] * ssp[3]);
        stats[g * 7 + 4] = ssp[4] * abs_score;
        stats[g * 7 + 5] = ssp[4];
        stats[g * 7 + 6] = ssp[4] * mc[2];
      }
      if (v_op) v_op[g] = comp * vca * opac * (K<T>(1.0) - opac);
      const M2<T> cinv = inverse2x2(cov2d);
      M2<T> vci;
      vci.m[0][0] = vcon[0]; vci.m[0][1] = K<T>(0.5) * vcon[1];
      vci.m[1][0] = K<T>(0.5) * vcon[1]; vci.m[1][1] = vcon[2];
      M2<T> vc2 = inverse2x2_vjp(cinv, vci);
      if (mode == 1) {
        // d comp / d Sigma_raw = 0.5 comp (Sigma_raw^-1 - Sigma_blur^-1); opacity = sigmoid * comp
        const T dr = det2(cov_raw), db = det2(cov2d);
        if (dr > K<T>(0.0) && db > K<T>(0.0)) {
          M2<T> ir, ib;
          ir.m[0][0] = cov_raw.m[1][1] / dr; ir.m[1][1] = cov_raw.m[0][0] / dr;
          ir.m[0][1] = -cov_raw.m[0][1] / dr; ir.m[1][0] = -cov_raw.m[1][0] / dr;
          ib.m[0][0] = cov2d.m[1][1] / db; ib.m[1][1] = cov2d.m[0][0] / db;
          ib.m[0][1] = -cov2d.m[0][1] / db; ib.m[1][0] = -cov2d.m[1][0] / db;
          const T f = vca * opac * K<T>(0.5) * comp;
          for (int c = 0; c < 2; ++c)
            for (int r = 0; r < 2; ++r) vc2.m[c][r] = vc2.m[c][r] + f * (ir.m[r][c] - ib.m[r][c]);
        }
      }
      T vmc[3];
      persp_proj_vjp(j, mc, cov_cam, vc2, vxy, vmc);
      const M3<T> vcc = mul(mul(tr(j), vc2), j);
      const M3<T> vc3 = mul(mul(tr(R), vcc), R);
      const M3<T> vm = mul(add(vc3, tr(vc3)), m);
      T vsc[3];
      for (int c = 0; c < 3; ++c)
        vsc[c] = (rot.m[c][0] * vm.m[c][0] + rot.m[c][1] * vm.m[c][1]) + rot.m[c][2] * vm.m[c][2];
      T vq[4];
      quat_to_mat3_vjp(q, mul(vm, S), vq);
      const Mat<T, 4, 4> nj = normalize_vjp(qu);
      T vqu[4];
      for (int r = 0; r < 4; ++r)
        vqu[r] = ((nj.m[0][r] * vq[0] + nj.m[1][r] * vq[1]) + nj.m[2][r] * vq[2]) + nj.m[3][r] * vq[3];
      T vmean[3];
      const M3<T> Rt = tr(R);
      for (int r = 0; r < 3; ++r) vmean[r] = (Rt.m[0][r] * vmc[0] + Rt.m[1][r] * vmc[1]) + Rt.m[2][r] * vmc[2];
      if (mode == 1) {
        // view direction -> mean: colour = SH(d/|d|); v_mean += (dcolour/dmean)^T v_colour
        D3<T> dv[3];
        for (int a = 0; a < 3; ++a) {
          dv[a].v = vd[a];
          for (int b = 0; b < 3; ++b) dv[a].d[b] = (a == b) ? K<T>(1.0) : K<T>(0.0);
        }
        D3<T> cd[3];
        sh_color_generic<T, D3<T>>(deg, dv, sh + g * nk * 3, cd);
        T gvd[3] = {K<T>(0.0), K<T>(0.0), K<T>(0.0)};
        for (int ch = 0; ch < 3; ++ch)
          for (int b = 0; b < 3; ++b) gvd[b] = gvd[b] + vcol[ch] * cd[ch].d[b];
        if (dd > K<T>(1e-12)) {
          const T vdg = (vd[0] * gvd[0] + vd[1] * gvd[1]) + vd[2] * gvd[2];
          for (int b = 0; b < 3; ++b) vmean[b] = vmean[b] + (gvd[b] - vd[b] * vdg) * rs;
        }
      }
      if (v_tf) {
        T* o = v_tf + g * 10;
        o[0] = vmean[0]; o[1] = vmean[1]; o[2] = vmean[2];
        o[3] = vqu[0]; o[4] = vqu[1]; o[5] = vqu[2]; o[6] = vqu[3];
        o[7] = vsc[0] * scale[0]; o[8] = vsc[1] * scale[1]; o[9] = vsc[2] * scale[2];
      }
    }
  }

  // helpers.wgsl:382-427
  void persp_proj_vjp(const M32<T>& j, const T m3[3], const M3<T>& cov3d, const M2<T>& vc2, const T vm2[2],
                      T out[3]) const {
    const T x = m3[0], y = m3[1], z = m3[2];
    const T fx = v.fx, fy = v.fy;
    const T rz = K<T>(1.0) / mx(std::fabs(z), K<T>(0.01));
    const T rz2 = rz * rz;
    T o0 = fx * rz * vm2[0];
    T o1 = fy * rz * vm2[1];
    T o2 = -(fx * x * vm2[0] + fy * y * vm2[1]) * rz2;
    const M32<T> vj = add(mul(mul(vc2, j), tr(cov3d)), mul(mul(tr(vc2), j), cov3d));
    const T lpx = (K<T>(1.15) * (T)v.W - v.cx) / fx, lpy = (K<T>(1.15) * (T)v.H - v.cy) / fy;
    const T lnx = (K<T>(-0.15) * (T)v.W - v.cx) / fx, lny = (K<T>(-0.15) * (T)v.H - v.cy) / fy;
    if (x * rz <= lpx && x * rz >= lnx) {
      o0 = o0 + -fx * rz2 * vj.m[2][0];
      o2 = o2 + K<T>(2.0) * fx * x * rz2 * rz * vj.m[2][0];
    } else {
      const T cx = clampv(x * rz, lnx, lpx);
      o2 = o2 + fx * cx * rz2 * vj.m[2][0];
    }
    if (y * rz <= lpy && y * rz >= lny) {
      o1 = o1 + -fy * rz2 * vj.m[2][1];
      o2 = o2 + K<T>(2.0) * fy * y * rz2 * rz * vj.m[2][1];
    } else {
      const T cy = clampv(y * rz, lny, lpy);
      o2 = o2 + fy * cy * rz2 * vj.m[2][1];
    }
    o2 = o2 + (-fx * rz2 * vj.m[0][0] - fy * rz2 * vj.m[1][1]);
    out[0] = o0; out[1] = o1; out[2] = o2;
  }
};

}  // namespace

// =================================================================== C ABI
struct orc_render {
  std::vector<float> tf, sh, op;
  orc_camera cam;
  Pipe<float>* p = nullptr;
  ~orc_render() { delete p; }
};

extern "C" {

orc_render* orc_forward(int64_t n, int sh_degree, const float* transforms, const float* sh, const float* raw_opac,
                        const orc_camera* cam, const float* background, float cov_blur, int tile_width) {
  if (n < 0 || sh_degree < 0 || sh_degree > 4 || !cam || tile_width <= 0) return nullptr;
  const int nk = (sh_degree + 1) * (sh_degree + 1);
  orc_render* r = new orc_render();
  r->tf.assign(transforms, transforms + n * 10);
  r->sh.assign(sh, sh + n * nk * 3);
  r->op.assign(raw_opac, raw_opac + n);
  r->cam = *cam;
  r->p = new Pipe<float>(*cam);
  Pipe<float>& p = *r->p;
  p.n = n;
  p.deg = sh_degree;
  p.nk = nk;
  p.tf = r->tf.data();
  p.sh = r->sh.data();
  p.op = r->op.data();
  for (int i = 0; i < 3; ++i) p.bg[i] = background ? background[i] : 0.0f;
  p.blur = cov_blur > 0.0f ? cov_blur : 0.0f;  // ProjectUniforms::new_with_cov_blur clamps at 0
  p.tw = tile_width;
  p.forward();
  return r;
}

void orc_free(orc_render* r) { delete r; }

void orc_stats(const orc_render* r, orc_fwd_stats* o) {
  const Pipe<float>& p = *r->p;
  o->num_visible = (int64_t)p.order.size();
  o->num_intersections = (int64_t)p.isect_cgid.size();
  o->k1_count_mismatch = p.mismatch;
  o->depth_tie_runs = p.tie_runs;
  o->evaluated_fragments = p.evaluated;
  o->blended_fragments = p.blended;
  o->num_tiles = p.ntiles;
}

void orc_get_image(const orc_render* r, float* out) { std::copy(r->p->img.begin(), r->p->img.end(), out); }
void orc_get_visible(const orc_render* r, float* out) {
  for (int64_t g = 0; g < r->p->n; ++g) out[g] = r->p->visible[g] ? 1.0f : 0.0f;
}
void orc_get_depth_order(const orc_render* r, uint32_t* gids) { std::copy(r->p->order.begin(), r->p->order.end(), gids); }
void orc_get_depths(const orc_render* r, float* out) { std::copy(r->p->depth.begin(), r->p->depth.end(), out); }
void orc_get_projected(const orc_render* r, float* out) {
  const auto& pr = r->p->proj;
  for (size_t i = 0; i < pr.size(); ++i) {
    const float v[9] = {pr[i].x, pr[i].y, pr[i].a, pr[i].b, pr[i].c, pr[i].r, pr[i].g, pr[i].bl, pr[i].o};
    std::memcpy(out + i * 9, v, sizeof(v));
  }
}
void orc_get_counts(const orc_render* r, uint32_t* counts) {
  std::fill(counts, counts + r->p->n, 0u);
  for (size_t i = 0; i < r->p->order.size(); ++i) counts[r->p->order[i]] = r->p->cnt[i];
}
void orc_get_isect(const orc_render* r, uint32_t* tiles, uint32_t* gids) {
  const auto& p = *r->p;
  for (size_t e = 0; e < p.isect_cgid.size(); ++e) {
    if (tiles) tiles[e] = p.isect_tile[e];
    if (gids) gids[e] = p.order[p.isect_cgid[e]];
  }
}
void orc_get_tile_offsets(const orc_render* r, uint32_t* o) { std::copy(r->p->offsets.begin(), r->p->offsets.end(), o); }

int orc_backward(orc_render* r, const float* v_out, int mode, float* v_tf, float* v_sh, float* v_op, float* stats,
                 float* vs_out, float* ss_out) {
  if (!r || !v_out || (mode != 0 && mode != 1)) return 1;
  Pipe<float>& p = *r->p;
  std::vector<float> vs, ss;
  p.raster_bwd(v_out, mode, vs, ss);
  if (vs_out) std::copy(vs.begin(), vs.end(), vs_out);
  if (ss_out) std::copy(ss.begin(), ss.end(), ss_out);
  p.project_bwd(vs, ss, mode, v_tf, v_sh, v_op, stats);
  return 0;
}

int orc_backward_f64(orc_render* r, const float* v_out, int mode, const float* vs_in, double* v_tf, double* v_sh,
                     double* v_op, double* vs_out) {
  if (!r || !v_out || (mode != 0 && mode != 1)) return 1;
  const Pipe<float>& p = *r->p;
  const int64_t n = p.n;
  std::vector<double> tf(r->tf.begin(), r->tf.end()), sh(r->sh.begin(), r->sh.end()), op(r->op.begin(), r->op.end());
  Pipe<double> d(r->cam);
  d.n = n;
  d.deg = p.deg;
  d.nk = p.nk;
  d.tf = tf.data();
  d.sh = sh.data();
  d.op = op.data();
  for (int i = 0; i < 3; ++i) d.bg[i] = p.bg[i];
  d.blur = p.blur;
  d.tw = p.tw;
  d.tbx = p.tbx;
  d.tby = p.tby;
  d.ntiles = p.ntiles;
  d.order = p.order;
  d.isect_cgid = p.isect_cgid;
  d.isect_tile = p.isect_tile;
  d.offsets = p.offsets;
  d.proj.resize(p.proj.size());
  for (size_t i = 0; i < p.proj.size(); ++i) {
    const Proj<float>& f = p.proj[i];
    d.proj[i] = Proj<double>{f.x, f.y, f.a, f.b, f.c, f.r, f.g, f.bl, f.o};
  }
  d.img.assign(p.img.begin(), p.img.end());
  d.shadow = &p.proj;
  const int64_t np = (int64_t)p.v.W * p.v.H;
  std::vector<double> vo(v_out, v_out + np * 4);
  std::vector<double> vs, ss;
  if (vs_in) {
    vs.assign(vs_in, vs_in + p.proj.size() * 10);
    ss.assign(p.proj.size() * 5, 0.0);
  } else {
    d.raster_bwd(vo.data(), mode, vs, ss);
  }
  if (vs_out) std::copy(vs.begin(), vs.end(), vs_out);
  d.project_bwd(vs, ss, mode, v_tf, v_sh, v_op, nullptr);
  return 0;
}

void orc_crossing_splats(const orc_render* r, uint8_t* flags) {
  const auto& p = *r->p;
  std::fill(flags, flags + p.n, (uint8_t)0);
  for (size_t i = 0; i < p.crossing.size(); ++i)
    if (p.crossing[i]) flags[p.order[i]] = 1;
}

void orc_render_bruteforce(const orc_render* r, float* out) {
  const auto& p = *r->p;
  const uint32_t W = p.v.W, H = p.v.H;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t py = 0; py < (int64_t)H; ++py)
    for (uint32_t px = 0; px < W; ++px) {
      const float pc[2] = {(float)px + 0.5f, (float)py + 0.5f};
      float Tr = 1.0f, acc[3] = {0, 0, 0};
      for (size_t i = 0; i < p.proj.size(); ++i) {
        const auto& sp = p.proj[i];
        const float dx = sp.x - pc[0], dy = sp.y - pc[1];
        const float sigma = 0.5f * (sp.a * dx * dx + 2.0f * sp.b * dx * dy + sp.c * dy * dy);
        const float alpha = std::min((float)0.999, sp.o * rgs_exp(-sigma));
        if (sigma >= 0.0f && alpha >= (float)(1.0 / 255.0)) {
          const float vis = alpha * Tr;
          acc[0] = acc[0] + std::max(sp.r, 0.0f) * vis;
          acc[1] = acc[1] + std::max(sp.g, 0.0f) * vis;
          acc[2] = acc[2] + std::max(sp.bl, 0.0f) * vis;
          Tr = Tr * (1.0f - alpha);
          if (Tr < (float)1e-4) break;
        }
      }
      float* o = out + ((size_t)py * W + px) * 4;
      o[0] = acc[0] + Tr * p.bg[0];
      o[1] = acc[1] + Tr * p.bg[1];
      o[2] = acc[2] + Tr * p.bg[2];
      o[3] = 1.0f - Tr;
    }
}

double orc_l1_loss_grad(const float* pred, const float* target, int64_t W, int64_t H, float weight, float* v_out) {
  // engine/loss.rs:144-196 (plain L1, mean over H*W*3) and the RGB slice at trainer.rs:278
  const int64_t np = W * H;
  const float scale = weight / (float)(3 * np);
  double sum = 0.0;
  for (int64_t i = 0; i < np; ++i) {
    for (int c = 0; c < 3; ++c) {
      const float d = pred[i * 4 + c] - target[i * 3 + c];
      sum += std::fabs((double)d);
      if (v_out) v_out[i * 4 + c] = d > 0.0f ? scale : (d < 0.0f ? -scale : 0.0f);
    }
    if (v_out) v_out[i * 4 + 3] = 0.0f;
  }
  return (double)weight * sum / (double)(3 * np);
}

float orc_powi(float a, int32_t b) {
  // compiler-rt __powisf2 (what Rust's f32::powi lowers to for a runtime exponent)
  const bool recip = b < 0;
  float r = 1.0f;
  while (true) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0f / r : r;
}

void orc_adam_group(int64_t n_rows, int64_t row_len, float* param, const float* grad, float* m, float* v,
                    const float* lr_row, float beta1, float beta2, float eps, int32_t step, float weight_decay) {
  // engine/optimizer.rs:122-162 (weight decay: :128-130)
  const float omb1 = 1.0f - beta1, omb2 = 1.0f - beta2;
  const float bc1 = 1.0f - orc_powi(beta1, step);
  const float bc2 = 1.0f - orc_powi(beta2, step);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_rows; ++r)
    for (int64_t c = 0; c < row_len; ++c) {
      const int64_t i = r * row_len + c;
      float g = grad[i];
      if (weight_decay != 0.0f) g = g + param[i] * weight_decay;
      const float gsq = g * g;
      const float m1 = m[i] * beta1 + g * omb1;
      const float m2 = v[i] * beta2 + gsq * omb2;
      const float mh = m1 / bc1;
      const float vh = m2 / bc2;
      const float upd = mh / (rgs_sqrt(vh) + eps);
      param[i] = param[i] - upd * lr_row[c];
      m[i] = m1;
      v[i] = m2;
    }
}

static Pipe<double>* make_f64(int64_t n, int deg, const double* tf, const double* sh, const double* op,
                              const orc_camera* cam, const double* bg, double blur, int tw) {
  Pipe<double>* p = new Pipe<double>(*cam);
  p->n = n;
  p->deg = deg;
  p->nk = (deg + 1) * (deg + 1);
  p->tf = tf;
  p->sh = sh;
  p->op = op;
  for (int i = 0; i < 3; ++i) p->bg[i] = bg ? bg[i] : 0.0;
  p->blur = blur > 0.0 ? blur : 0.0;
  p->tw = tw;
  return p;
}

double orc_forward_loss_f64(int64_t n, int deg, const double* tf, const double* sh, const double* op,
                            const orc_camera* cam, const double* bg, double blur, int tw, const float* target,
                            int flags, const double* fvd, const double* fcomp, double* out, int64_t frozen_g,
                            const uint8_t* frozen_pix, const double* frozen_proj) {
  Pipe<double>* p = make_f64(n, deg, tf, sh, op, cam, bg, blur, tw);
  if (flags & 1) p->frozen_vd = fvd;
  if (flags & 2) p->frozen_comp = fcomp;
  p->bwd_termination = (flags & 4) != 0;
  if (frozen_pix && frozen_proj && frozen_g >= 0) {
    p->frozen_g = frozen_g;
    p->frozen_pix = frozen_pix;
    const double* f = frozen_proj;
    p->frozen_proj = Proj<double>{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8]};
  }
  p->forward();
  const int64_t np = (int64_t)p->v.W * p->v.H;
  double sum = 0.0;
  for (int64_t i = 0; i < np; ++i)
    for (int c = 0; c < 3; ++c) sum += std::fabs(p->img[i * 4 + c] - (double)target[i * 3 + c]);
  if (out) std::copy(p->img.begin(), p->img.end(), out);
  delete p;
  return sum / (double)(3 * np);
}

// The whole pipeline in FP64 with its own decisions: forward, compositing
// backward and projection backward.  The gradient of the FP64 function whose
// finite differences tests take, free of FP32-versus-FP64 threshold
// mismatches (a splat that the FP32 forward stops at T = 1e-4 and the FP64
// forward does not).
int orc_grad_f64(int64_t n, int deg, const double* tf, const double* sh, const double* op, const orc_camera* cam,
                 const double* bg, double blur, int tw, const double* v_out, int mode, double* v_tf, double* v_sh,
                 double* v_op) {
  if (mode != 0 && mode != 1) return 1;
  Pipe<double>* p = make_f64(n, deg, tf, sh, op, cam, bg, blur, tw);
  p->forward();
  std::vector<double> vs, ss;
  p->raster_bwd(v_out, mode, vs, ss);
  p->project_bwd(vs, ss, mode, v_tf, v_sh, v_op, nullptr);
  delete p;
  return 0;
}

// The unperturbed FP64 forward's crossing map for finite differences: per
// pixel the global id of the splat that takes T below 1e-4 (-1: none), and
// that splat's projection (x, y, a, b, c, r, g, b, opacity) per global id
// (zeros for culled Gaussians).
void orc_crossing_map_f64(int64_t n, int deg, const double* tf, const double* sh, const double* op,
                          const orc_camera* cam, const double* bg, double blur, int tw, int flags, const double* fvd,
                          const double* fcomp, int64_t* pix_gid, double* proj9) {
  Pipe<double>* p = make_f64(n, deg, tf, sh, op, cam, bg, blur, tw);
  if (flags & 1) p->frozen_vd = fvd;
  if (flags & 2) p->frozen_comp = fcomp;
  p->forward();
  std::copy(p->cross_gid.begin(), p->cross_gid.end(), pix_gid);
  std::fill(proj9, proj9 + n * 9, 0.0);
  for (size_t c = 0; c < p->proj.size(); ++c) {
    const Proj<double>& q = p->proj[c];
    double* o = proj9 + (size_t)p->order[c] * 9;
    o[0] = q.x; o[1] = q.y; o[2] = q.a; o[3] = q.b; o[4] = q.c; o[5] = q.r; o[6] = q.g; o[7] = q.bl; o[8] = q.o;
  }
  delete p;
}

void orc_projection_terms_f64(int64_t n, const double* tf, const double* op, const orc_camera* cam, double blur,
                              double* viewdir, double* comp) {
  Pipe<double>* p = make_f64(n, 0, tf, nullptr, op, cam, nullptr, blur, 16);
  for (int64_t g = 0; g < n; ++g) {
    const double* q = tf + g * 10;
    double mc[3];
    p->mean_c(g, mc);
    const double d[3] = {q[0] - p->v.campos[0], q[1] - p->v.campos[1], q[2] - p->v.campos[2]};
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double s = 1.0 / std::sqrt(std::max(dd, 1e-12));
    for (int i = 0; i < 3; ++i) viewdir[g * 3 + i] = d[i] * s;
    double qq[4] = {q[3], q[4], q[5], q[6]};
    const double qn = qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2] + qq[3] * qq[3];
    if (qn < 1e-6 || mc[2] < 0.01) {
      comp[g] = std::nan("");
      continue;
    }
    for (int i = 0; i < 4; ++i) qq[i] /= std::sqrt(qn);
    const double sc[3] = {std::exp(q[7]), std::exp(q[8]), std::exp(q[9])};
    M2<double> cov2d = calc_cov2d(calc_cov3d(sc, qq), mc, p->v);
    comp[g] = compensate_cov2d(cov2d, p->blur);
  }
  delete p;
}

// sh.rs:35-112 evaluate_sh_rgb with its f32 literals (degrees 0-4; missing
// coefficients read as zero, sh.rs:114-121).
static void tr_sh_rgb(const float* c, int nk, int degree, const float vd[3], float col[3]) {
  auto co = [&](int k, int ch) { return k < nk ? c[k * 3 + ch] : 0.0f; };
  for (int ch = 0; ch < 3; ++ch) col[ch] = 0.2820948f * co(0, ch);
  auto add = [&](int k, float s) {
    for (int ch = 0; ch < 3; ++ch) col[ch] += co(k, ch) * s;
  };
  const float x = vd[0], y = vd[1], z = vd[2];
  if (degree >= 1) {
    add(1, 0.48860252f * -y);
    add(2, 0.48860252f * z);
    add(3, 0.48860252f * -x);
  }
  if (degree >= 2) {
    const float z2 = z * z, f0b = -1.0925485f * z, f1a = 0.54627424f, fc1 = x * x - y * y, fs1 = 2.0f * x * y;
    add(4, f1a * fs1);
    add(5, f0b * y);
    add(6, 0.9461747f * z2 - 0.31539157f);
    add(7, f0b * x);
    add(8, f1a * fc1);
    if (degree >= 3) {
      const float f0c = -2.285229f * z2 + 0.4570458f, f1b = 1.4453057f * z, f2a = -0.5900436f;
      const float fc2 = x * fc1 - y * fs1, fs2 = x * fs1 + y * fc1;
      add(9, f2a * fs2);
      add(10, f1b * fs1);
      add(11, f0c * y);
      add(12, z * (1.8658817f * z2 - 1.119529f));
      add(13, f0c * x);
      add(14, f1b * fc1);
      add(15, f2a * fc2);
      if (degree >= 4) {
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
  for (int ch = 0; ch < 3; ++ch) col[ch] += 0.5f;
}

void orc_tr_sh_rgb(int64_t n, int degree, const float* sh, const float* dirs, float* rgb) {
  const int nk = (degree + 1) * (degree + 1);
  for (int64_t i = 0; i < n; ++i) tr_sh_rgb(sh + i * nk * 3, nk, degree, dirs + i * 3, rgb + i * 3);
}

// Rust's `f as usize` for the non-negative clamp results of
// compute_tile_bounds: truncation, NaN and negatives to 0, saturating above.
static uint64_t tr_as_usize(float f) {
  if (!(f > 0.0f)) return 0;
  if (f >= 18446744073709551616.0f) return UINT64_MAX;
  return (uint64_t)f;
}

void orc_tiled_renderer(int64_t n, int sh_degree, const float* positions, const float* log_scales,
                        const float* rotations, const float* opacity_logits, const float* sh, const orc_camera* cam,
                        const float* background, float cov_blur, int gpu_math, float* color, float* depth_out) {
  // render/tiled_renderer.rs:174-492 (render_camera_splats, project_projected_gaussian,
  // compute_tile_bounds, render_projected_with_background), single-threaded.
  // gpu_math = 0: libm expf, as the reference; 1: rgs_exp, the library's
  // exponential (the two differ by a few ulp at most).
  // Rust's f32::max / f32::min return the non-NaN operand: fmaxf / fminf.
  auto EXP = [gpu_math](float v) { return gpu_math ? rgs_exp(v) : std::exp(v); };
  struct PG {
    float x, y, depth, cxx, cxy, cyy, opacity, col[3];
  };
  const uint32_t W = cam->width, H = cam->height;
  const uint64_t tw = 16;
  const uint64_t ntx = (W + tw - 1) / tw, nty = (H + tw - 1) / tw;
  const float blur = std::fmax(cov_blur, 0.0f);  // with_raster_cov_blur (tiled_renderer.rs:68-71)
  float r[3][3], tr3[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) r[i][j] = cam->viewmat[j * 4 + i];  // rows
    tr3[i] = cam->viewmat[12 + i];
  }
  const int nk = (sh_degree + 1) * (sh_degree + 1);
  std::vector<PG> pr;
  pr.reserve(n);
  for (int64_t idx = 0; idx < n; ++idx) {
    const float* pos = positions + idx * 3;
    float vd[3] = {pos[0] - cam->cam_pos[0], pos[1] - cam->cam_pos[1], pos[2] - cam->cam_pos[2]};
    const float len = std::sqrt(vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2]);
    if (len <= 1e-12f) {  // normalize3, tiled_renderer.rs:507-513
      vd[0] = 0.0f; vd[1] = 0.0f; vd[2] = 1.0f;
    } else {
      for (int i = 0; i < 3; ++i) vd[i] /= len;
    }
    float col[3];
    tr_sh_rgb(sh + idx * nk * 3, nk, sh_degree, vd, col);
    // project_projected_gaussian (tiled_renderer.rs:389-492)
    const float wx = pos[0], wy = pos[1], wz = pos[2];
    const float ccx = r[0][0] * wx + r[0][1] * wy + r[0][2] * wz + tr3[0];
    const float ccy = r[1][0] * wx + r[1][1] * wy + r[1][2] * wz + tr3[1];
    const float ccz = r[2][0] * wx + r[2][1] * wy + r[2][2] * wz + tr3[2];
    if (ccz <= 0.0f) continue;
    // A NaN depth passes the test above, but such a splat's centre is NaN, so
    // it never contributes; kept in a tile list it would make the depth
    // comparison a non-total order, under which the reference's sort_by
    // leaves the order unspecified (and may panic).  It is left out.
    if (std::isnan(ccz)) continue;
    const float ppx = cam->fx * ccx / ccz + cam->cx, ppy = cam->fy * ccy / ccz + cam->cy;
    const float* q = rotations + idx * 4;
    const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    const float r00 = 1 - 2 * (qy * qy + qz * qz), r01 = 2 * (qx * qy - qw * qz), r02 = 2 * (qx * qz + qw * qy);
    const float r10 = 2 * (qx * qy + qw * qz), r11 = 1 - 2 * (qx * qx + qz * qz), r12 = 2 * (qy * qz - qw * qx);
    const float r20 = 2 * (qx * qz - qw * qy), r21 = 2 * (qy * qz + qw * qx), r22 = 1 - 2 * (qx * qx + qy * qy);
    const float sx = std::fabs(std::fabs(EXP(log_scales[idx * 3]))),
                sy = std::fabs(std::fabs(EXP(log_scales[idx * 3 + 1]))),
                sz = std::fabs(std::fabs(EXP(log_scales[idx * 3 + 2])));
    const float sxx = sx * sx, syy = sy * sy, szz = sz * sz;
    const float cw00 = r00 * r00 * sxx + r01 * r01 * syy + r02 * r02 * szz;
    const float cw01 = r00 * r10 * sxx + r01 * r11 * syy + r02 * r12 * szz;
    const float cw02 = r00 * r20 * sxx + r01 * r21 * syy + r02 * r22 * szz;
    const float cw11 = r10 * r10 * sxx + r11 * r11 * syy + r12 * r12 * szz;
    const float cw12 = r10 * r20 * sxx + r11 * r21 * syy + r12 * r22 * szz;
    const float cw22 = r20 * r20 * sxx + r21 * r21 * syy + r22 * r22 * szz;
    const float m00 = r[0][0] * cw00 + r[0][1] * cw01 + r[0][2] * cw02;
    const float m01 = r[0][0] * cw01 + r[0][1] * cw11 + r[0][2] * cw12;
    const float m02 = r[0][0] * cw02 + r[0][1] * cw12 + r[0][2] * cw22;
    const float m10 = r[1][0] * cw00 + r[1][1] * cw01 + r[1][2] * cw02;
    const float m11 = r[1][0] * cw01 + r[1][1] * cw11 + r[1][2] * cw12;
    const float m12 = r[1][0] * cw02 + r[1][1] * cw12 + r[1][2] * cw22;
    const float m20 = r[2][0] * cw00 + r[2][1] * cw01 + r[2][2] * cw02;
    const float m21 = r[2][0] * cw01 + r[2][1] * cw11 + r[2][2] * cw12;
    const float m22 = r[2][0] * cw02 + r[2][1] * cw12 + r[2][2] * cw22;
    const float cc00 = m00 * r[0][0] + m01 * r[0][1] + m02 * r[0][2];
    const float cc01 = m00 * r[1][0] + m01 * r[1][1] + m02 * r[1][2];
    const float cc02 = m00 * r[2][0] + m01 * r[2][1] + m02 * r[2][2];
    const float cc11 = m10 * r[1][0] + m11 * r[1][1] + m12 * r[1][2];
    const float cc12 = m10 * r[2][0] + m11 * r[2][1] + m12 * r[2][2];
    const float cc22 = m20 * r[2][0] + m21 * r[2][1] + m22 * r[2][2];
    const float iz = 1.0f / ccz, iz2 = iz * iz;
    const float jx0 = cam->fx * iz, jx2 = -cam->fx * ccx * iz2, jy1 = cam->fy * iz, jy2 = -cam->fy * ccy * iz2;
    float cxx = jx0 * (cc00 * jx0 + cc02 * jx2) + jx2 * (cc02 * jx0 + cc22 * jx2);
    const float cxy = jx0 * (cc01 * jy1 + cc02 * jy2) + jx2 * (cc12 * jy1 + cc22 * jy2);
    float cyy = jy1 * (cc11 * jy1 + cc12 * jy2) + jy2 * (cc12 * jy1 + cc22 * jy2);
    const float det_raw = cxx * cyy - cxy * cxy;
    cxx += blur;
    cyy += blur;
    const float det_bl = std::fmax(cxx * cyy - cxy * cxy, 1e-12f);
    if (det_raw <= 0.0f) continue;
    const float opacity = 1.0f / (1.0f + EXP(-opacity_logits[idx])) * std::sqrt(det_raw / det_bl);
    if (opacity < 1.0f / 255.0f) continue;
    pr.push_back({ppx, ppy, ccz, cxx, cxy, cyy, opacity, {col[0], col[1], col[2]}});
  }
  std::vector<float> alpha_buf((size_t)W * H, 0.0f);
  std::fill(color, color + (size_t)W * H * 3, 0.0f);
  std::fill(depth_out, depth_out + (size_t)W * H, 3.4028235e38f);
  std::vector<std::vector<uint32_t>> lists((size_t)ntx * nty);
  for (uint32_t idx = 0; idx < pr.size(); ++idx) {  // compute_tile_bounds with tile_alpha = 4
    const PG& g = pr[idx];
    const float sgx = std::fmax(std::sqrt(g.cxx * 4.0f), 1.0f), sgy = std::fmax(std::sqrt(g.cyy * 4.0f), 1.0f);
    const uint64_t xmin = tr_as_usize(std::fmax(g.x - 3.0f * sgx, 0.0f));
    const uint64_t xmax = tr_as_usize(std::fmin(g.x + 3.0f * sgx, (float)W - 1.0f));
    const uint64_t ymin = tr_as_usize(std::fmax(g.y - 3.0f * sgy, 0.0f));
    const uint64_t ymax = tr_as_usize(std::fmin(g.y + 3.0f * sgy, (float)H - 1.0f));
    const uint64_t txmin = xmin / tw, txmax = (xmax + tw - 1) / tw, tymin = ymin / tw, tymax = (ymax + tw - 1) / tw;
    for (uint64_t ty = tymin; ty < tymax; ++ty)
      for (uint64_t tx = txmin; tx < txmax; ++tx) lists[ty * ntx + tx].push_back(idx);
  }
  for (uint64_t ty = 0; ty < nty; ++ty)
    for (uint64_t tx = 0; tx < ntx; ++tx) {
      auto& L = lists[ty * ntx + tx];
      if (L.empty()) continue;
      // stable sort by depth; a NaN depth compares equal (partial_cmp fallback)
      std::stable_sort(L.begin(), L.end(), [&](uint32_t a, uint32_t b) { return pr[a].depth < pr[b].depth; });
      const uint32_t x0 = (uint32_t)(tx * tw), y0 = (uint32_t)(ty * tw);
      const uint32_t x1 = std::min<uint32_t>(x0 + (uint32_t)tw, W), y1 = std::min<uint32_t>(y0 + (uint32_t)tw, H);
      for (uint32_t gi : L) {
        const PG& g = pr[gi];
        const float det = g.cxx * g.cyy - g.cxy * g.cxy;
        if (det < 1e-10f) continue;
        const float inv = 1.0f / det;
        for (uint32_t py = y0; py < y1; ++py)
          for (uint32_t px = x0; px < x1; ++px) {
            const size_t id = (size_t)py * W + px;
            if (alpha_buf[id] >= 0.999f) continue;
            const float dx = g.x - ((float)px + 0.5f), dy = g.y - ((float)py + 0.5f);
            const float d2 = (g.cyy * dx * dx - 2.0f * g.cxy * dx * dy + g.cxx * dy * dy) * inv;
            if (d2 < 9.0f) {
              const float w = EXP(-0.5f * d2) * g.opacity;
              if (w > 0.001f) {
                const float a = w * (1.0f - alpha_buf[id]);
                if (a > 0.0f) {
                  for (int ch = 0; ch < 3; ++ch) color[id * 3 + ch] += g.col[ch] * a;
                  alpha_buf[id] += a;
                  if (depth_out[id] == 3.4028235e38f && a > 0.01f) depth_out[id] = g.depth;
                }
              }
            }
          }
      }
    }
  const float* bg = background;
  for (size_t id = 0; id < (size_t)W * H; ++id) {
    const float t = 1.0f - alpha_buf[id];
    for (int ch = 0; ch < 3; ++ch) color[id * 3 + ch] += (bg ? bg[ch] : 0.0f) * t;
  }
  // f32::clamp(0, 1): NaN stays NaN
  for (size_t i = 0; i < (size_t)W * H * 3; ++i) {
    const float v = color[i];
    color[i] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  }
}

void orc_math_exp(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = rgs_exp(x[i]);
}
void orc_math_exp_core(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = rgs_exp_core(x[i]);
}
void orc_math_exp_core_fast(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = rgs_exp_core_fast(x[i]);
}
// rgs_expf_libm against this C library's expf over every float input: the
// number whose result bits differ (two NaNs count as equal); the first
// max_out of those inputs (as bits) go to out.
int64_t orc_expf_libm_check(uint32_t* out, int max_out) {
  int64_t bad = 0;
  int stored = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (int64_t u = 0; u <= 0xffffffffLL; ++u) {
    const uint32_t w = (uint32_t)u;
    const float x = rgs_u2f(w);
    const float a = std::exp(x), b = rgs_expf_libm(x);
    if (rgs_f2u(a) != rgs_f2u(b) && !(a != a && b != b)) {
      ++bad;
#pragma omp critical
      if (stored < max_out) out[stored++] = w;
    }
  }
  return bad;
}

void orc_math_log(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = rgs_log(x[i]);
}
void orc_sh_color(int degree, const float* coeffs, const float* viewdir, float* rgb) {
  float c[3];
  sh_color_generic<float, float>(degree, viewdir, coeffs, c);
  for (int i = 0; i < 3; ++i) rgb[i] = c[i] + 0.5f;
}
void orc_sort_by_key(int64_t n, const uint32_t* keys, const uint32_t* values, uint32_t* ok, uint32_t* ov) {
  std::vector<int64_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
  for (int64_t i = 0; i < n; ++i) {
    ok[i] = keys[idx[i]];
    ov[i] = values[idx[i]];
  }
}
void orc_inclusive_scan(int64_t n, const uint32_t* in, uint32_t* out) {
  uint32_t s = 0;
  for (int64_t i = 0; i < n; ++i) {
    s += in[i];
    out[i] = s;
  }
}
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
