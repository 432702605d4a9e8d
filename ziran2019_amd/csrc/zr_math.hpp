// zr_math.hpp -- per-particle math for the gfx950 kernels: SPGrid page codes,
// quadratic B-spline weights, the 3x3 implicit-QR SVD and the fixed-corotated
// constitutive model.  All index arguments are compile-time so the 3x3
// matrices live in VGPRs (no scratch).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define ZR_HD __host__ __device__ __forceinline__

namespace zr {

// ---------------------------------------------------------------------------
// SPGrid addressing for GridState<double,3> (128 B nodes, 2x4x4 pages).
// Lib/SPGrid/Core/SPGrid_Mask.h:22-174 with data_bits = 7, block bits (1,2,2):
// a node offset is (page_code << 12) | (elem << 7), elem = x0<<4 | y<<2 | z
// (low x bit, two low y bits, two low z bits) and the page code interleaves
// page coordinates (px = x>>1, pz = z>>2, py = y>>2) as bits x,z,y,x,z,y,...
// ---------------------------------------------------------------------------
// Page coordinates have up to 11 bits (px of a 4096-node extent: 2048 pages);
// the interleave is done in 64 bits and the code fits 31 bits.
ZR_HD uint32_t spread3(uint32_t v)
{
  uint64_t x = v & 0x7ffu;
  x = (x | (x << 32)) & 0x001f00000000ffffull;
  x = (x | (x << 16)) & 0x001f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return (uint32_t)x;
}
ZR_HD uint32_t compact3(uint32_t v)
{
  uint64_t x = v & 0x49249249u;
  x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
  x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
  x = (x ^ (x >> 8)) & 0x001f0000ff0000ffull;
  x = (x ^ (x >> 16)) & 0x001f00000000ffffull;
  x = (x ^ (x >> 32)) & 0x00000000001fffffull;
  return (uint32_t)x;
}
ZR_HD uint32_t page_encode(uint32_t px, uint32_t py, uint32_t pz)
{
  return spread3(px) | (spread3(pz) << 1) | (spread3(py) << 2);
}
ZR_HD void page_decode(uint32_t code, uint32_t& px, uint32_t& py, uint32_t& pz)
{
  px = compact3(code);
  pz = compact3(code >> 1);
  py = compact3(code >> 2);
}
ZR_HD uint32_t elem_of(int x, int y, int z)
{
  return (uint32_t)(((x & 1) << 4) | ((y & 3) << 2) | (z & 3));
}

// Workgroups are dispatched round-robin over the 8 XCDs of the device
// (block b runs on XCD b mod 8).  Groups are numbered in Morton page order,
// so neighbouring groups share tile nodes; this bijection gives each XCD a
// contiguous range of groups so that the shared nodes stay in one L2.
__device__ __forceinline__ uint32_t xcd_group(uint32_t b, uint32_t ngroups)
{
  const uint32_t x = b & 7u, l = b >> 3;
  const uint32_t q = ngroups >> 3, r = ngroups & 7u;
  return x * q + (x < r ? x : r) + l;
}

// MathTools.h:21-25
ZR_HD int int_floor(double x)
{
  int i = (int)x;
  return i - (i > x);
}
ZR_HD int int_floor(float x)
{
  int i = (int)x;
  return i - (i > x);
}

// Quadratic B-spline (BSplines.h:16-20,55-81); x is in index space.
template <class T>
ZR_HD int bspline_w(T x, T w[3])
{
  int b = int_floor(x - (T)0.5);
  T d0 = x - (T)b;
  T z = (T)1.5 - d0;
  w[0] = (T)0.5 * z * z;
  T d1 = d0 - (T)1;
  w[1] = (T)0.75 - d1 * d1;
  T d2 = (T)1 - d1;
  T zz = (T)1.5 - d2;
  w[2] = (T)0.5 * zz * zz;
  return b;
}
template <class T>
ZR_HD int bspline_wdw(T x, T w[3], T dw[3])
{
  int b = int_floor(x - (T)0.5);
  T d0 = x - (T)b;
  T z = (T)1.5 - d0;
  w[0] = (T)0.5 * z * z;
  T d1 = d0 - (T)1;
  w[1] = (T)0.75 - d1 * d1;
  T d2 = (T)1 - d1;
  T zz = (T)1.5 - d2;
  w[2] = (T)0.5 * zz * zz;
  dw[0] = -z;
  dw[1] = (T)-2 * d1;
  dw[2] = zz;
  return b;
}

// ---------------------------------------------------------------------------
// 3x3 matrix in registers, row-major a[r][c].
// ---------------------------------------------------------------------------
template <class T>
struct M3 {
  T a[3][3];
};

template <class T>
ZR_HD void set_identity(M3<T>& m)
{
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) m.a[r][c] = (r == c) ? (T)1 : (T)0;
}

template <class T>
ZR_HD T det3(const M3<T>& f)
{
  return f.a[0][0] * (f.a[1][1] * f.a[2][2] - f.a[1][2] * f.a[2][1]) -
         f.a[0][1] * (f.a[1][0] * f.a[2][2] - f.a[1][2] * f.a[2][0]) +
         f.a[0][2] * (f.a[1][0] * f.a[2][1] - f.a[1][1] * f.a[2][0]);
}

// ---------------------------------------------------------------------------
// Implicit-shifted QR SVD of a 3x3 (Gast, Fu, Jiang, Teran 2016), in the form
// of Lib/Ziran/Math/Linear/{Givens.h:32-271, ImplicitQRSVD.h:45-533}.
// Output: F = U diag(s) V^T, U and V rotations, |s0|>=|s1|>=|s2|, only s2 < 0.
// ---------------------------------------------------------------------------
template <class T>
struct Givens {
  T c, s;
};

template <class T>
ZR_HD Givens<T> gv_compute(T a, T b)
{
  Givens<T> g{ (T)1, (T)0 };
  T sq = sqrt(a * a + b * b);
  if (sq != (T)0) {
    T t = (T)1 / sq;
    g.c = a * t;
    g.s = -b * t;
  }
  return g;
}
template <class T>
ZR_HD Givens<T> gv_compute_unconv(T a, T b)
{
  Givens<T> g{ (T)0, (T)1 };
  T sq = sqrt(a * a + b * b);
  if (sq != (T)0) {
    T t = (T)1 / sq;
    g.s = a * t;
    g.c = b * t;
  }
  return g;
}
template <int I, int K, class T>
ZR_HD void gv_row(const Givens<T>& g, M3<T>& m)
{
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    T t1 = m.a[I][j], t2 = m.a[K][j];
    m.a[I][j] = g.c * t1 - g.s * t2;
    m.a[K][j] = g.s * t1 + g.c * t2;
  }
}
template <int I, int K, class T>
ZR_HD void gv_col(const Givens<T>& g, M3<T>& m)
{
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    T t1 = m.a[j][I], t2 = m.a[j][K];
    m.a[j][I] = g.c * t1 - g.s * t2;
    m.a[j][K] = g.s * t1 + g.c * t2;
  }
}

template <class T>
ZR_HD void zero_chase(M3<T>& h, M3<T>& u, M3<T>& v)
{
  Givens<T> r1 = gv_compute(h.a[0][0], h.a[1][0]);
  Givens<T> r2;
  if (h.a[1][0] != (T)0)
    r2 = gv_compute(h.a[0][0] * h.a[0][1] + h.a[1][0] * h.a[1][1], h.a[0][0] * h.a[0][2] + h.a[1][0] * h.a[1][2]);
  else
    r2 = gv_compute(h.a[0][1], h.a[0][2]);
  gv_row<0, 1>(r1, h);
  gv_col<1, 2>(r2, h);
  gv_col<1, 2>(r2, v);
  Givens<T> r3 = gv_compute(h.a[1][1], h.a[2][1]);
  gv_row<1, 2>(r3, h);
  gv_col<0, 1>(r1, u);
  gv_col<1, 2>(r3, u);
}

template <class T>
ZR_HD void svd2(T a00, T a01, T a10, T a11, Givens<T>& u, T& s0, T& s1, Givens<T>& v)
{
  T x0 = a00 + a11, x1 = a10 - a01;
  T den = sqrt(x0 * x0 + x1 * x1);
  u.c = (T)1;
  u.s = (T)0;
  if (den != (T)0) {
    u.c = x0 / den;
    u.s = -x1 / den;
  }
  T sx = u.c * a00 - u.s * a10;
  T sy = u.c * a01 - u.s * a11;
  T sz = u.s * a01 + u.c * a11;
  T cosine, sine;
  if (sy * sy == (T)0) {
    cosine = (T)1;
    sine = (T)0;
    s0 = sx;
    s1 = sz;
  } else {
    T tau = (T)0.5 * (sx - sz);
    T w = sqrt(tau * tau + sy * sy);
    T t = (tau > (T)0) ? sy / (tau + w) : sy / (tau - w);
    cosine = (T)1 / sqrt(t * t + (T)1);
    sine = -t * cosine;
    T c2 = cosine * cosine, csy = (T)2 * cosine * sine * sy, sn2 = sine * sine;
    s0 = c2 * sx - csy + sn2 * sz;
    s1 = sn2 * sx + csy + c2 * sz;
  }
  if (s0 < s1) {
    T t = s0;
    s0 = s1;
    s1 = t;
    v.c = -sine;
    v.s = cosine;
  } else {
    v.c = cosine;
    v.s = sine;
  }
  T nc = u.c * v.c - u.s * v.s;
  T ns = u.s * v.c + u.c * v.s;
  u.c = nc;
  u.s = ns;
}

template <class T>
ZR_HD T wilkinson_shift(T a1, T b1, T a2)
{
  T d = (T)0.5 * (a1 - a2);
  T bs = b1 * b1;
  return a2 - copysign(bs / (fabs(d) + sqrt(d * d + bs)), d);
}

template <int TT, class T>
ZR_HD void process_block(M3<T>& b, M3<T>& u, T s[3], M3<T>& v)
{
  constexpr int other = (TT == 1) ? 0 : 2;
  s[other] = b.a[other][other];
  Givens<T> gu, gvv;
  svd2(b.a[TT][TT], b.a[TT][TT + 1], b.a[TT + 1][TT], b.a[TT + 1][TT + 1], gu, s[TT], s[TT + 1], gvv);
  gv_col<TT, TT + 1>(gu, u);
  gv_col<TT, TT + 1>(gvv, v);
}

template <int I, class T>
ZR_HD void neg_col(M3<T>& m)
{
#pragma unroll
  for (int r = 0; r < 3; ++r) m.a[r][I] = -m.a[r][I];
}
template <int I, int K, class T>
ZR_HD void swap_col(M3<T>& m)
{
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    T t = m.a[r][I];
    m.a[r][I] = m.a[r][K];
    m.a[r][K] = t;
  }
}
template <int I, class T>
ZR_HD void flip_sign(M3<T>& u, T s[3])
{
  s[I] = -s[I];
  neg_col<I>(u);
}
template <class T>
ZR_HD void swp(T& a, T& b)
{
  T t = a;
  a = b;
  b = t;
}

template <class T>
ZR_HD void sort0(M3<T>& u, T s[3], M3<T>& v)
{
  if (fabs(s[1]) >= fabs(s[2])) {
    if (s[1] < (T)0) {
      flip_sign<1>(u, s);
      flip_sign<2>(u, s);
    }
    return;
  }
  if (s[2] < (T)0) {
    flip_sign<1>(u, s);
    flip_sign<2>(u, s);
  }
  swp(s[1], s[2]);
  swap_col<1, 2>(u);
  swap_col<1, 2>(v);
  if (s[1] > s[0]) {
    swp(s[0], s[1]);
    swap_col<0, 1>(u);
    swap_col<0, 1>(v);
  } else {
    neg_col<2>(u);
    neg_col<2>(v);
  }
}

template <class T>
ZR_HD void sort1(M3<T>& u, T s[3], M3<T>& v)
{
  if (fabs(s[0]) >= s[1]) {
    if (s[0] < (T)0) {
      flip_sign<0>(u, s);
      flip_sign<2>(u, s);
    }
    return;
  }
  swp(s[0], s[1]);
  swap_col<0, 1>(u);
  swap_col<0, 1>(v);
  if (fabs(s[1]) < fabs(s[2])) {
    swp(s[1], s[2]);
    swap_col<1, 2>(u);
    swap_col<1, 2>(v);
  } else {
    neg_col<1>(u);
    neg_col<1>(v);
  }
  if (s[1] < (T)0) {
    flip_sign<1>(u, s);
    flip_sign<2>(u, s);
  }
}

template <class T>
struct SvdTol;
template <>
struct SvdTol<double> {
  static constexpr double value = 1024.0 * 2.220446049250313e-16;  // ImplicitQRSVD.h:526-533
};
template <>
struct SvdTol<float> {
  static constexpr float value = 128.0f * 1.1920928955078125e-07f;  // ImplicitQRSVD.h:518-524
};

constexpr int kSvdMaxSweeps = 64;

template <class T>
ZR_HD int svd3(const M3<T>& a, M3<T>& u, T s[3], M3<T>& v)
{
  T tol = SvdTol<T>::value;
  M3<T> b = a;
  set_identity(u);
  set_identity(v);
  {
    Givens<T> r = gv_compute(b.a[1][0], b.a[2][0]);
    gv_row<1, 2>(r, b);
    gv_col<1, 2>(r, u);
    zero_chase(b, u, v);
  }
  int count = 0;
  T a1 = b.a[0][0], b1 = b.a[0][1], a2 = b.a[1][1], a3 = b.a[2][2], b2 = b.a[1][2];
  T g1 = a1 * b1, g2 = a2 * b2;
  T scale = (T)0.5 * sqrt(a1 * a1 + a2 * a2 + a3 * a3 + b1 * b1 + b2 * b2);
  tol *= (scale > (T)1 ? scale : (T)1);
  while (fabs(b2) > tol && fabs(b1) > tol && fabs(a1) > tol && fabs(a2) > tol && fabs(a3) > tol &&
         count < kSvdMaxSweeps) {
    T mu = wilkinson_shift(a2 * a2 + b1 * b1, g2, a3 * a3 + b2 * b2);
    Givens<T> r = gv_compute(a1 * a1 - mu, g1);
    gv_col<0, 1>(r, b);
    gv_col<0, 1>(r, v);
    zero_chase(b, u, v);
    a1 = b.a[0][0];
    b1 = b.a[0][1];
    a2 = b.a[1][1];
    a3 = b.a[2][2];
    b2 = b.a[1][2];
    g1 = a1 * b1;
    g2 = a2 * b2;
    count++;
  }
  if (fabs(b2) <= tol) {
    process_block<0>(b, u, s, v);
    sort0(u, s, v);
  } else if (fabs(b1) <= tol) {
    process_block<1>(b, u, s, v);
    sort1(u, s, v);
  } else if (fabs(a2) <= tol) {
    Givens<T> r1 = gv_compute_unconv(b.a[1][2], b.a[2][2]);
    gv_row<1, 2>(r1, b);
    gv_col<1, 2>(r1, u);
    process_block<0>(b, u, s, v);
    sort0(u, s, v);
  } else if (fabs(a3) <= tol) {
    Givens<T> r1 = gv_compute(b.a[1][1], b.a[1][2]);
    gv_col<1, 2>(r1, b);
    gv_col<1, 2>(r1, v);
    Givens<T> r2 = gv_compute(b.a[0][0], b.a[0][2]);
    gv_col<0, 2>(r2, b);
    gv_col<0, 2>(r2, v);
    process_block<0>(b, u, s, v);
    sort0(u, s, v);
  } else if (fabs(a1) <= tol) {
    Givens<T> r1 = gv_compute_unconv(b.a[0][1], b.a[1][1]);
    gv_row<0, 1>(r1, b);
    gv_col<0, 1>(r1, u);
    Givens<T> r2 = gv_compute_unconv(b.a[0][2], b.a[2][2]);
    gv_row<0, 2>(r2, b);
    gv_col<0, 2>(r2, u);
    process_block<1>(b, u, s, v);
    sort1(u, s, v);
  } else {
    process_block<0>(b, u, s, v);
    sort0(u, s, v);
  }
  return count;
}

// ---------------------------------------------------------------------------
// Fixed-corotated elasticity (Stomakhin et al. 2012): Kirchhoff stress
//   tau = P F^T = 2 mu (F - R) F^T + lam (J - 1) J I,   R = U V^T.
// ---------------------------------------------------------------------------
template <class T>
ZR_HD void fcr_kirchhoff(T mu, T lam, const M3<T>& f, M3<T>& tau)
{
  M3<T> u, v;
  T s[3];
  svd3(f, u, s, v);
  M3<T> r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = u.a[i][0] * v.a[j][0] + u.a[i][1] * v.a[j][1] + u.a[i][2] * v.a[j][2];
  T J = det3(f);
  T vol_term = lam * (J - (T)1) * J;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T acc = (T)0;
#pragma unroll
      for (int k = 0; k < 3; ++k) acc += (f.a[i][k] - r.a[i][k]) * f.a[j][k];
      tau.a[i][j] = (T)2 * mu * acc + (i == j ? vol_term : (T)0);
    }
}

// ---------------------------------------------------------------------------
// StvkWithHencky (StvkWithHencky.cpp:114-143): with F = U S V^T,
//   tau = P F^T = U diag(2 mu log|s_i| + lam sum_k log|s_k|) U^T.
// ---------------------------------------------------------------------------
template <class T>
ZR_HD void hencky_kirchhoff(T mu, T lam, const M3<T>& f, M3<T>& tau)
{
  M3<T> u, v;
  T s[3];
  svd3(f, u, s, v);
  T l[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) l[i] = log(fabs(s[i]));
  const T tr = l[0] + l[1] + l[2];
  T d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) d[i] = (T)2 * mu * l[i] + lam * tr;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      tau.a[i][j] = u.a[i][0] * d[0] * u.a[j][0] + u.a[i][1] * d[1] * u.a[j][1] + u.a[i][2] * d[2] * u.a[j][2];
}

// ---------------------------------------------------------------------------
// NeoHookeanBorden with degradation g (NeoHookeanBorden.cpp:16-86):
//   kappa = 0.666666 mu + lambda, tau_dev = mu J^(-2/3) dev(F F^T),
//   tau_vol = J kappa/2 (J - 1/J) I; g scales tau_dev (and tau_vol if J >= 1).
// ---------------------------------------------------------------------------
template <class T>
ZR_HD void nhb_kirchhoff(T mu, T lam, T g, const M3<T>& f, M3<T>& tau)
{
  const T J = det3(f);
  M3<T> b;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) b.a[i][j] = f.a[i][0] * f.a[j][0] + f.a[i][1] * f.a[j][1] + f.a[i][2] * f.a[j][2];
  const T trb = b.a[0][0] + b.a[1][1] + b.a[2][2];
  const T sdev = mu * pow(J, (T)(-2.0 / 3.0));
  const T kappa = (T)0.666666 * mu + lam;
  const T vol = J * (kappa / (T)2 * (J - (T)1 / J));
  const bool tension = J >= (T)1;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const T dev = sdev * (b.a[i][j] - (i == j ? trb / (T)3 : (T)0));
      const T v = i == j ? vol : (T)0;
      tau.a[i][j] = tension ? g * (dev + v) : g * dev + v;
    }
}

template <class T>
ZR_HD void usv_t(const M3<T>& u, const T s[3], const M3<T>& v, M3<T>& f)
{
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      f.a[i][j] = u.a[i][0] * s[0] * v.a[j][0] + u.a[i][1] * s[1] * v.a[j][1] + u.a[i][2] * s[2] * v.a[j][2];
}

// VonMisesStvkHencky::projectStrain (PlasticityApplier.cpp:464-501).
// Returns true when f was projected; yield_stress hardens by xi * dgamma;
// broken is set when the projected first principal stress reaches fail_stress.
template <class T>
ZR_HD bool von_mises_project(T mu, T lam, T xi, T fail_stress, T& yield_stress, M3<T>& f, bool& broken)
{
  M3<T> u, v;
  T s[3], e[3], h[3];
  svd3(f, u, s, v);
#pragma unroll
  for (int i = 0; i < 3; ++i) e[i] = log(s[i] > (T)1e-4 ? s[i] : (T)1e-4);
  const T tr = e[0] + e[1] + e[2];
  T n2 = (T)0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    h[i] = e[i] - tr / (T)3;
    n2 += h[i] * h[i];
  }
  const T nrm = sqrt(n2);
  const T dg = nrm - yield_stress / ((T)2 * mu);
  if (dg <= (T)0) return false;
  yield_stress -= xi * dg;
  T H[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) H[i] = e[i] - (dg / nrm) * h[i];
  const T tau0 = (T)2 * mu * H[0] + lam * (H[0] + H[1] + H[2]);
#pragma unroll
  for (int i = 0; i < 3; ++i) H[i] = exp(H[i]);
  usv_t(u, H, v, f);
  if (tau0 >= fail_stress) broken = true;
  return true;
}

// DruckerPragerStvkHencky::projectStrain (PlasticityApplier.cpp:275-316).
template <class T>
ZR_HD void drucker_prager_project(T mu, T lam, T alpha, T beta, T cohesion, bool volume_correction, T& logJp,
                                  M3<T>& f)
{
  M3<T> u, v;
  T s[3], e[3], h[3];
  svd3(f, u, s, v);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const T a = fabs(s[i]);
    e[i] = log(a > (T)1e-4 ? a : (T)1e-4) - cohesion;
  }
  const T sum = e[0] + e[1] + e[2];
  const T tr = sum + logJp;
  T n2 = (T)0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    h[i] = e[i] - tr / (T)3;
    n2 += h[i] * h[i];
  }
  T H[3];
  if (tr >= (T)0) {  // case II: the tip
    const T c = exp(cohesion);
    H[0] = H[1] = H[2] = c;
    usv_t(u, H, v, f);
    if (volume_correction) logJp = beta * sum + logJp;
    return;
  }
  if (mu == (T)0) return;
  logJp = (T)0;
  const T nrm = sqrt(n2);
  const T dg = nrm + ((T)3 * lam + (T)2 * mu) / ((T)2 * mu) * tr * alpha;
  if (dg <= (T)0) {  // case I: inside
#pragma unroll
    for (int i = 0; i < 3; ++i) H[i] = e[i] + cohesion;
  } else {  // case III: onto the cone
#pragma unroll
    for (int i = 0; i < 3; ++i) H[i] = e[i] - (dg / nrm) * h[i] + cohesion;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) H[i] = exp(H[i]);
  usv_t(u, H, v, f);
}

// NonAssociativeCamClay::projectStrain (PlasticityApplier.cpp:113-135) on a
// NeoHookeanBorden model (kappa = 0.666666 mu + lambda,
// NeoHookeanBorden.cpp:20): the live branch is Compare_With_Physbam
// (PlasticityApplier.cpp:36-109) on the singular values, with
// p0 = kappa (1e-5 + sinh(xi max(-logJp, 0))): the tips of the yield surface
// (p beyond p0 or below -beta p0) set Se to a volumetric state, a trial state
// outside the surface (y >= 1e-4) keeps its mean of Se^2 and scales its
// deviatoric part onto the surface, with logJp hardened through the
// intersection of the line to the surface's centre (the reference's "fake"
// hardening).  M comes from the friction angle (PlasticityApplier.cpp:10-19).
template <class T>
ZR_HD void cam_clay_project(T mu, T lam, T M, T beta, T xi, T& logJp, M3<T>& f)
{
  M3<T> u, v;
  T Se[3];
  svd3(f, u, Se, v);
  const T kappa = (T)0.666666 * mu + lam;
  const T p0 = kappa * ((T)0.00001 + sinh(xi * (-logJp > (T)0 ? -logJp : (T)0)));
  const T a = ((T)1 + (T)2 * beta) * ((T)6 - (T)3) / (T)2;
  const T b = beta * p0;
  const T c = p0;
  const T M2 = M * M;
  T Je = (T)1;
#pragma unroll
  for (int i = 0; i < 3; ++i) Je *= Se[i];
  T Se2[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) Se2[i] = Se[i] * Se[i];
  T mean = (T)0;
#pragma unroll
  for (int i = 0; i < 3; ++i) mean += Se2[i];
  mean /= (T)3;
  const T sc = mu * pow(Je, -(T)2 / (T)3);
  T sh[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) sh[i] = sc * (Se2[i] - mean);
  const T Uprime = kappa / (T)2 * (Je - (T)1 / Je);
  const T p_trial = -Uprime * Je;
  if (p_trial > c || p_trial < -b) {  // the tips
    const T Je_new = p_trial > c ? sqrt(-(T)2 * c / kappa + (T)1) : sqrt((T)2 * b / kappa + (T)1);
    const T se = pow(Je_new, (T)1 / (T)3);
    Se[0] = Se[1] = Se[2] = se;
    logJp += log(Je / Je_new);
  } else {
    const T k = sqrt(-M2 * (p_trial + b) * (p_trial - c) / a);
    const T sn2 = (sh[0] * sh[0] + sh[1] * sh[1]) + sh[2] * sh[2];
    const T sn = sqrt(sn2);
    const T y = a * (sn * sn) + M2 * (p_trial + b) * (p_trial - c);
    // inside the yield surface Se stays; F is rebuilt from the SVD either way
    if (y >= (T)1e-4) {
      const T pc = ((T)1 - beta) * p0 / (T)2;
      if (p0 > (T)1e-4 && p_trial < p0 - (T)1e-4 && p_trial > -beta * p0 + (T)1e-4) {
        const T aa = M2 * ((p_trial - pc) * (p_trial - pc)) / (a * (sn * sn));
        const T dd = (T)1 + aa;
        const T ff = aa * beta * p0 - aa * p0 - (T)2 * pc;
        const T gg = pc * pc - aa * beta * (p0 * p0);
        const T zz = sqrt(fabs(ff * ff - (T)4 * dd * gg));
        const T p1 = (-ff + zz) / ((T)2 * dd);
        const T p2 = (-ff - zz) / ((T)2 * dd);
        const T p_fake = (p_trial - pc) * (p1 - pc) > (T)0 ? p1 : p2;
        const T Je_new_fake = sqrt(fabs(-(T)2 * p_fake / kappa + (T)1));
        if (Je_new_fake > (T)1e-4) logJp += log(Je / Je_new_fake);
      }
      const T ssum = (Se2[0] + Se2[1]) + Se2[2];
      const T scale = k / mu * pow(Je, (T)2 / (T)3);
#pragma unroll
      for (int i = 0; i < 3; ++i) Se[i] = sqrt(scale * sh[i] / sn + (T)1 / (T)3 * ssum);
    }
  }
  usv_t(u, Se, v, f);
}

}  // namespace zr
