/*
 * zro_linalg.c -- oracle (TEST INFRASTRUCTURE ONLY): 3x3 SVD by implicit-
 * shifted QR and the fixed-corotated constitutive model, FP64.
 *
 * The SVD restates the method of Gast, Fu, Jiang, Teran, "Implicit-shifted
 * Symmetric QR Singular Value Decomposition of 3x3 Matrices" (UCLA tech report
 * 2016), in the form the reference uses:
 *   Lib/Ziran/Math/Linear/Givens.h:32-271          (Givens, zeroChase, makeUpperBidiag)
 *   Lib/Ziran/Math/Linear/ImplicitQRSVD.h:45-163   (2x2 polar / SVD)
 *   Lib/Ziran/Math/Linear/ImplicitQRSVD.h:220-345  (wilkinsonShift, process, flipSign, sort)
 *   Lib/Ziran/Math/Linear/ImplicitQRSVD.h:355-533  (3x3 driver, tolerances)
 * Output convention (same as the reference): F = U diag(sigma) V^T with U, V
 * rotations (det +1), |sigma0| >= |sigma1| >= |sigma2|, sigma0, sigma1 >= 0,
 * only sigma2 may be negative.
 * Deviation: the QR sweep loop is capped at 64 iterations (the reference loop
 * is unbounded); convergence takes far fewer for finite input.
 *
 * Fixed-corotated (Stomakhin et al. 2012; Jiang et al. 2016 MPM course notes):
 *   psi = mu |F - R|^2 + lam/2 (J-1)^2
 *   P   = 2 mu (F - R) + lam (J-1) J F^-T
 *   dP  = 2 mu (dF - dR) + lam (JF^-T : dF) JF^-T + lam (J-1) d(JF^-T)
 *   with U^T dR V = skew part W, W_ij = (M_ij - M_ji)/(s_i + s_j), M = U^T dF V.
 * The reference has no fixed-corotated model in Lib/ (only a commented-out
 * CorotatedElasticity in Projects/admm/AdmmFormula.h:57-693); Lame parameters
 * follow NeoHookeanBorden::setLameParameters (NeoHookeanBorden.cpp:16-21).
 */
#include <math.h>
#include <float.h>
#include <string.h>
#include "zr_oracle.h"

typedef double M3[3][3];

/* column-major 9-array <-> row-major [3][3] */
static void from_cm(const double* a, M3 m)
{
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m[r][c] = a[c * 3 + r];
}
static void to_cm(M3 m, double* a)
{
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[c * 3 + r] = m[r][c];
}
static void m_ident(M3 m)
{
  memset(m, 0, sizeof(M3));
  m[0][0] = m[1][1] = m[2][2] = 1.0;
}

/* ---------------- Givens rotations (Givens.h:32-180) ---------------- */
typedef struct { int i, k; double c, s; } givens;

static givens gv_make(int i, int k) { givens g = { i, k, 1.0, 0.0 }; return g; }
/* (c -s; s c)(a; b) = (*; 0)  -- Givens.h:73-87 */
static void gv_compute(givens* g, double a, double b)
{
  double d = a * a + b * b;
  g->c = 1.0; g->s = 0.0;
  double sq = sqrt(d);
  if (sq != 0.0) { double t = 1.0 / sq; g->c = a * t; g->s = -b * t; }
}
/* (c -s; s c)(a; b) = (0; *)  -- Givens.h:94-107 */
static void gv_compute_unconv(givens* g, double a, double b)
{
  double d = a * a + b * b;
  g->c = 0.0; g->s = 1.0;
  double sq = sqrt(d);
  if (sq != 0.0) { double t = 1.0 / sq; g->s = a * t; g->c = b * t; }
}
/* G^T A on rows i,k -- Givens.h:129-139 */
static void gv_row(const givens* g, M3 a)
{
  for (int j = 0; j < 3; ++j) {
    double t1 = a[g->i][j], t2 = a[g->k][j];
    a[g->i][j] = g->c * t1 - g->s * t2;
    a[g->k][j] = g->s * t1 + g->c * t2;
  }
}
/* A G on columns i,k -- Givens.h:148-158 */
static void gv_col(const givens* g, M3 a)
{
  for (int j = 0; j < 3; ++j) {
    double t1 = a[j][g->i], t2 = a[j][g->k];
    a[j][g->i] = g->c * t1 - g->s * t2;
    a[j][g->k] = g->s * t1 + g->c * t2;
  }
}

/* zero chasing of a 3x3 with one bulge at (1,0): Givens.h:196-239 */
static void zero_chase(M3 h, M3 u, M3 v)
{
  givens r1 = gv_make(0, 1);
  gv_compute(&r1, h[0][0], h[1][0]);
  givens r2 = gv_make(1, 2);
  if (h[1][0] != 0.0)
    gv_compute(&r2, h[0][0] * h[0][1] + h[1][0] * h[1][1], h[0][0] * h[0][2] + h[1][0] * h[1][2]);
  else
    gv_compute(&r2, h[0][1], h[0][2]);
  gv_row(&r1, h);
  gv_col(&r2, h);
  gv_col(&r2, v);
  givens r3 = gv_make(1, 2);
  gv_compute(&r3, h[1][1], h[2][1]);
  gv_row(&r3, h);
  gv_col(&r1, u);
  gv_col(&r3, u);
}

/* Givens.h:251-271 */
static void make_upper_bidiag(M3 h, M3 u, M3 v)
{
  m_ident(u);
  m_ident(v);
  givens r = gv_make(1, 2);
  gv_compute(&r, h[1][0], h[2][0]);
  gv_row(&r, h);
  gv_col(&r, u);
  zero_chase(h, u, v);
}

/* ---------------- 2x2 polar / SVD (ImplicitQRSVD.h:45-163) ----------- */
/* a = [a00 a01; a10 a11]; returns rotation (as c,s) and symmetric s2 */
static void polar2(double a00, double a01, double a10, double a11, givens* r, double s2[2][2])
{
  double x0 = a00 + a11, x1 = a10 - a01;
  double den = sqrt(x0 * x0 + x1 * x1);
  r->c = 1.0; r->s = 0.0;
  if (den != 0.0) { r->c = x0 / den; r->s = -x1 / den; }
  /* S = G^T A on the 2x2 */
  s2[0][0] = r->c * a00 - r->s * a10;
  s2[0][1] = r->c * a01 - r->s * a11;
  s2[1][0] = r->s * a00 + r->c * a10;
  s2[1][1] = r->s * a01 + r->c * a11;
}

static void svd2(double a00, double a01, double a10, double a11, givens* u, double* sig, givens* v)
{
  double s2[2][2];
  polar2(a00, a01, a10, a11, u, s2);
  double x = s2[0][0], y = s2[0][1], z = s2[1][1];
  double y2 = y * y, cosine, sine;
  if (y2 == 0.0) {
    cosine = 1.0; sine = 0.0;
    sig[0] = x; sig[1] = z;
  } else {
    double tau = 0.5 * (x - z);
    double w = sqrt(tau * tau + y2);
    double t = (tau > 0.0) ? y / (tau + w) : y / (tau - w);
    cosine = 1.0 / sqrt(t * t + 1.0);
    sine = -t * cosine;
    double c2 = cosine * cosine, csy = 2.0 * cosine * sine * y, sn2 = sine * sine;
    sig[0] = c2 * x - csy + sn2 * z;
    sig[1] = sn2 * x + csy + c2 * z;
  }
  if (sig[0] < sig[1]) {
    double t = sig[0]; sig[0] = sig[1]; sig[1] = t;
    v->c = -sine; v->s = cosine;
  } else {
    v->c = cosine; v->s = sine;
  }
  /* U *= V (compose two rotations of the same plane): Givens.h:164-170 */
  double nc = u->c * v->c - u->s * v->s;
  double ns = u->s * v->c + u->c * v->s;
  u->c = nc; u->s = ns;
}

/* ImplicitQRSVD.h:220-232 */
static double wilkinson_shift(double a1, double b1, double a2)
{
  double d = 0.5 * (a1 - a2);
  double bs = b1 * b1;
  return a2 - copysign(bs / (fabs(d) + sqrt(d * d + bs)), d);
}

/* process<t>: 2x2 SVD of the block at (t,t): ImplicitQRSVD.h:237-252 */
static void process_block(int t, M3 b, M3 u, double* sig, M3 v)
{
  int other = (t == 1) ? 0 : 2;
  sig[other] = b[other][other];
  givens gu = gv_make(t, t + 1), gvv = gv_make(t, t + 1);
  svd2(b[t][t], b[t][t + 1], b[t + 1][t], b[t + 1][t + 1], &gu, sig + t, &gvv);
  gv_col(&gu, u);
  gv_col(&gvv, v);
}

static void flip_sign(int i, M3 u, double* sig)
{
  sig[i] = -sig[i];
  for (int r = 0; r < 3; ++r) u[r][i] = -u[r][i];
}
static void swap_cols(M3 a, int i, int k)
{
  for (int r = 0; r < 3; ++r) { double t = a[r][i]; a[r][i] = a[r][k]; a[r][k] = t; }
}
static void neg_col(M3 a, int i)
{
  for (int r = 0; r < 3; ++r) a[r][i] = -a[r][i];
}
static void swap_d(double* a, double* b) { double t = *a; *a = *b; *b = t; }

/* sort<0>: ImplicitQRSVD.h:257-294 */
static void sort0(M3 u, double* sig, M3 v)
{
  if (fabs(sig[1]) >= fabs(sig[2])) {
    if (sig[1] < 0) { flip_sign(1, u, sig); flip_sign(2, u, sig); }
    return;
  }
  if (sig[2] < 0) { flip_sign(1, u, sig); flip_sign(2, u, sig); }
  swap_d(&sig[1], &sig[2]); swap_cols(u, 1, 2); swap_cols(v, 1, 2);
  if (sig[1] > sig[0]) {
    swap_d(&sig[0], &sig[1]); swap_cols(u, 0, 1); swap_cols(v, 0, 1);
  } else {
    neg_col(u, 2); neg_col(v, 2);
  }
}

/* sort<1>: ImplicitQRSVD.h:299-334 */
static void sort1(M3 u, double* sig, M3 v)
{
  if (fabs(sig[0]) >= sig[1]) {
    if (sig[0] < 0) { flip_sign(0, u, sig); flip_sign(2, u, sig); }
    return;
  }
  swap_d(&sig[0], &sig[1]); swap_cols(u, 0, 1); swap_cols(v, 0, 1);
  if (fabs(sig[1]) < fabs(sig[2])) {
    swap_d(&sig[1], &sig[2]); swap_cols(u, 1, 2); swap_cols(v, 1, 2);
  } else {
    neg_col(u, 1); neg_col(v, 1);
  }
  if (sig[1] < 0) { flip_sign(1, u, sig); flip_sign(2, u, sig); }
}

#define ZRO_SVD_MAX_SWEEPS 64

/* ImplicitQRSVD.h:355-517 with tol = 1024 eps (ImplicitQRSVD.h:526-533) */
static int svd3_rm(M3 a, M3 u, double* sig, M3 v)
{
  double tol = 1024.0 * DBL_EPSILON;
  M3 b;
  memcpy(b, a, sizeof(M3));
  make_upper_bidiag(b, u, v);
  int count = 0;
  double a1 = b[0][0], b1 = b[0][1], a2 = b[1][1], a3 = b[2][2], b2 = b[1][2];
  double g1 = a1 * b1, g2 = a2 * b2;
  double scale = 0.5 * sqrt(a1 * a1 + a2 * a2 + a3 * a3 + b1 * b1 + b2 * b2);
  tol *= (scale > 1.0 ? scale : 1.0);
  while (fabs(b2) > tol && fabs(b1) > tol && fabs(a1) > tol && fabs(a2) > tol &&
         fabs(a3) > tol && count < ZRO_SVD_MAX_SWEEPS) {
    double mu = wilkinson_shift(a2 * a2 + b1 * b1, g2, a3 * a3 + b2 * b2);
    givens r = gv_make(0, 1);
    gv_compute(&r, a1 * a1 - mu, g1);
    gv_col(&r, b);
    gv_col(&r, v);
    zero_chase(b, u, v);
    a1 = b[0][0]; b1 = b[0][1]; a2 = b[1][1]; a3 = b[2][2]; b2 = b[1][2];
    g1 = a1 * b1; g2 = a2 * b2;
    count++;
  }
  if (fabs(b2) <= tol) {
    process_block(0, b, u, sig, v);
    sort0(u, sig, v);
  } else if (fabs(b1) <= tol) {
    process_block(1, b, u, sig, v);
    sort1(u, sig, v);
  } else if (fabs(a2) <= tol) {
    givens r1 = gv_make(1, 2);
    gv_compute_unconv(&r1, b[1][2], b[2][2]);
    gv_row(&r1, b);
    gv_col(&r1, u);
    process_block(0, b, u, sig, v);
    sort0(u, sig, v);
  } else if (fabs(a3) <= tol) {
    givens r1 = gv_make(1, 2);
    gv_compute(&r1, b[1][1], b[1][2]);
    gv_col(&r1, b);
    gv_col(&r1, v);
    givens r2 = gv_make(0, 2);
    gv_compute(&r2, b[0][0], b[0][2]);
    gv_col(&r2, b);
    gv_col(&r2, v);
    process_block(0, b, u, sig, v);
    sort0(u, sig, v);
  } else if (fabs(a1) <= tol) {
    givens r1 = gv_make(0, 1);
    gv_compute_unconv(&r1, b[0][1], b[1][1]);
    gv_row(&r1, b);
    gv_col(&r1, u);
    givens r2 = gv_make(0, 2);
    gv_compute_unconv(&r2, b[0][2], b[2][2]);
    gv_row(&r2, b);
    gv_col(&r2, u);
    process_block(1, b, u, sig, v);
    sort1(u, sig, v);
  } else {
    /* only reachable when the sweep cap was hit: treat as converged */
    process_block(0, b, u, sig, v);
    sort0(u, sig, v);
  }
  return count;
}

int zro_svd3(const double* F, double* U, double* sigma, double* V)
{
  M3 a, u, v;
  from_cm(F, a);
  int c = svd3_rm(a, u, sigma, v);
  to_cm(u, U);
  to_cm(v, V);
  return c;
}

/* ImplicitQRSVD.h:546-558 */
void zro_polar3(const double* F, double* R, double* S)
{
  M3 a, u, v, r, s;
  double sig[3];
  from_cm(F, a);
  svd3_rm(a, u, sig, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0, accs = 0.0;
      for (int k = 0; k < 3; ++k) {
        acc += u[i][k] * v[j][k];
        accs += v[i][k] * sig[k] * v[j][k];
      }
      r[i][j] = acc;
      s[i][j] = accs;
    }
  to_cm(r, R);
  to_cm(s, S);
}

/* ---------------- fixed-corotated ---------------- */
void zro_lame(double E, double nu, double* mu, double* lam)
{
  *lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *mu = E / (2.0 * (1.0 + nu));
}

static double det3(M3 f)
{
  return f[0][0] * (f[1][1] * f[2][2] - f[1][2] * f[2][1])
       - f[0][1] * (f[1][0] * f[2][2] - f[1][2] * f[2][0])
       + f[0][2] * (f[1][0] * f[2][1] - f[1][1] * f[2][0]);
}

/* cofactor matrix J F^-T (DenseExt.h cofactorMatrix) */
static void cofactor3(M3 f, M3 c)
{
  c[0][0] = f[1][1] * f[2][2] - f[1][2] * f[2][1];
  c[0][1] = f[1][2] * f[2][0] - f[1][0] * f[2][2];
  c[0][2] = f[1][0] * f[2][1] - f[1][1] * f[2][0];
  c[1][0] = f[0][2] * f[2][1] - f[0][1] * f[2][2];
  c[1][1] = f[0][0] * f[2][2] - f[0][2] * f[2][0];
  c[1][2] = f[0][1] * f[2][0] - f[0][0] * f[2][1];
  c[2][0] = f[0][1] * f[1][2] - f[0][2] * f[1][1];
  c[2][1] = f[0][2] * f[1][0] - f[0][0] * f[1][2];
  c[2][2] = f[0][0] * f[1][1] - f[0][1] * f[1][0];
}

/* differential of the cofactor matrix along d (bilinear in f, d) */
static void cofactor3_diff(M3 f, M3 d, M3 c)
{
#define CD(a, b, e, g) (d[a][b] * f[e][g] + f[a][b] * d[e][g])
  c[0][0] = CD(1, 1, 2, 2) - CD(1, 2, 2, 1);
  c[0][1] = CD(1, 2, 2, 0) - CD(1, 0, 2, 2);
  c[0][2] = CD(1, 0, 2, 1) - CD(1, 1, 2, 0);
  c[1][0] = CD(0, 2, 2, 1) - CD(0, 1, 2, 2);
  c[1][1] = CD(0, 0, 2, 2) - CD(0, 2, 2, 0);
  c[1][2] = CD(0, 1, 2, 0) - CD(0, 0, 2, 1);
  c[2][0] = CD(0, 1, 1, 2) - CD(0, 2, 1, 1);
  c[2][1] = CD(0, 2, 1, 0) - CD(0, 0, 1, 2);
  c[2][2] = CD(0, 0, 1, 1) - CD(0, 1, 1, 0);
#undef CD
}

static void rotation_of(M3 f, M3 r, M3 u, double* sig, M3 v)
{
  svd3_rm(f, u, sig, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r[i][j] = u[i][0] * v[j][0] + u[i][1] * v[j][1] + u[i][2] * v[j][2];
}

double zro_fcr_psi(double mu, double lam, const double* F)
{
  M3 f, u, v;
  double sig[3];
  from_cm(F, f);
  svd3_rm(f, u, sig, v);
  double J = det3(f);
  double a = (sig[0] - 1.0) * (sig[0] - 1.0) + (sig[1] - 1.0) * (sig[1] - 1.0) + (sig[2] - 1.0) * (sig[2] - 1.0);
  return mu * a + 0.5 * lam * (J - 1.0) * (J - 1.0);
}

void zro_fcr_first_piola(double mu, double lam, const double* F, double* P)
{
  M3 f, r, u, v, c, p;
  double sig[3];
  from_cm(F, f);
  rotation_of(f, r, u, sig, v);
  cofactor3(f, c);
  double J = det3(f);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      p[i][j] = 2.0 * mu * (f[i][j] - r[i][j]) + lam * (J - 1.0) * c[i][j];
  to_cm(p, P);
}

void zro_fcr_kirchhoff(double mu, double lam, const double* F, double* tau)
{
  M3 f, r, u, v, t;
  double sig[3];
  from_cm(F, f);
  rotation_of(f, r, u, sig, v);
  double J = det3(f);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k) acc += (f[i][k] - r[i][k]) * f[j][k];
      t[i][j] = 2.0 * mu * acc + (i == j ? lam * (J - 1.0) * J : 0.0);
    }
  to_cm(t, tau);
}

void zro_fcr_first_piola_differential(double mu, double lam, const double* F,
                                      const double* dF, double* dP)
{
  M3 f, d, r, u, v, c, dc, m, w, dr, p;
  double sig[3];
  from_cm(F, f);
  from_cm(dF, d);
  rotation_of(f, r, u, sig, v);
  cofactor3(f, c);
  cofactor3_diff(f, d, dc);
  double J = det3(f);
  /* M = U^T dF V */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) acc += u[k][i] * d[k][l] * v[l][j];
      m[i][j] = acc;
    }
  /* skew W */
  for (int i = 0; i < 3; ++i) w[i][i] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j) {
      double den = sig[i] + sig[j];
      if (fabs(den) < 1e-12) den = (den < 0 ? -1e-12 : 1e-12);
      double x = (m[i][j] - m[j][i]) / den;
      w[i][j] = x;
      w[j][i] = -x;
    }
  /* dR = U W V^T */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) acc += u[i][k] * w[k][l] * v[j][l];
      dr[i][j] = acc;
    }
  double cd = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) cd += c[i][j] * d[i][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      p[i][j] = 2.0 * mu * (d[i][j] - dr[i][j]) + lam * cd * c[i][j] + lam * (J - 1.0) * dc[i][j];
  to_cm(p, dP);
}

/* ---- StvkWithHencky (Lib/Ziran/Physics/ConstitutiveModel/StvkWithHencky.cpp) ----
   psi = mu tr((log S)^2) + lambda/2 (tr log S)^2            (:121-128)
   P   = U (2 mu S^-1 log S + lambda tr(log S) S^-1) V^T      (:130-143)
   tau = P F^T (HyperelasticConstitutiveModel.h:75-80) */
double zro_stvk_hencky_psi(double mu, double lam, const double* F)
{
  M3 f, u, v;
  double sig[3];
  from_cm(F, f);
  svd3_rm(f, u, sig, v);
  double l2 = 0.0, tr = 0.0;
  for (int i = 0; i < 3; ++i) {
    double l = log(fabs(sig[i]));
    l2 += l * l;
    tr += l;
  }
  return mu * l2 + 0.5 * lam * tr * tr;
}

void zro_stvk_hencky_first_piola(double mu, double lam, const double* F, double* P)
{
  M3 f, u, v, p;
  double sig[3], ls[3], ph[3];
  from_cm(F, f);
  svd3_rm(f, u, sig, v);
  double tr = 0.0;
  for (int i = 0; i < 3; ++i) {
    ls[i] = log(fabs(sig[i]));
    tr += ls[i];
  }
  for (int i = 0; i < 3; ++i) {
    double inv = 1.0 / sig[i];
    ph[i] = 2.0 * mu * inv * ls[i] + lam * tr * inv;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) p[i][j] = u[i][0] * ph[0] * v[j][0] + u[i][1] * ph[1] * v[j][1] + u[i][2] * ph[2] * v[j][2];
  to_cm(p, P);
}

void zro_stvk_hencky_kirchhoff(double mu, double lam, const double* F, double* tau)
{
  double P[9];
  M3 p, f, t;
  zro_stvk_hencky_first_piola(mu, lam, F, P);
  from_cm(P, p);
  from_cm(F, f);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = p[i][0] * f[j][0] + p[i][1] * f[j][1] + p[i][2] * f[j][2];
  to_cm(t, tau);
}

/* ---- NeoHookeanBorden (Lib/Ziran/Physics/ConstitutiveModel/NeoHookeanBorden.cpp) ----
   kappa = 0.666666 mu + lambda (:16-21), degradation g on psi+ (:38-52) */
static double nhb_kappa(double mu, double lam) { return 0.666666 * mu + lam; }

static void nhb_parts(double mu, double lam, const double* F, double* dev, double* vol, double* J)
{
  M3 f;
  from_cm(F, f);
  double j = det3(f);
  double a = pow(j, -1.0 / 3.0);
  double tr = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) tr += (a * f[r][c]) * (a * f[r][c]);
  *dev = mu / 2.0 * (tr - 3.0);
  *vol = nhb_kappa(mu, lam) / 2.0 * ((j * j - 1.0) / 2.0 - log(j));
  *J = j;
}

double zro_nhb_psi(double mu, double lam, double g, const double* F)
{
  double dev, vol, J;
  nhb_parts(mu, lam, F, &dev, &vol, &J);
  double pos = J >= 1.0 ? vol + dev : dev;
  double neg = J >= 1.0 ? 0.0 : vol;
  return g * pos + neg;
}

/* get_psi_pos (:54-67) */
double zro_nhb_psi_pos(double mu, double lam, const double* F)
{
  double dev, vol, J;
  nhb_parts(mu, lam, F, &dev, &vol, &J);
  return J >= 1.0 ? vol + dev : dev;
}

/* kirchhoff (:69-86) */
void zro_nhb_kirchhoff(double mu, double lam, double g, const double* F, double* tau)
{
  M3 f, b, t;
  from_cm(F, f);
  double J = det3(f);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) b[i][j] = f[i][0] * f[j][0] + f[i][1] * f[j][1] + f[i][2] * f[j][2];
  double trb = b[0][0] + b[1][1] + b[2][2];
  double s = mu * pow(J, -2.0 / 3.0);
  double prime = nhb_kappa(mu, lam) / 2.0 * (J - 1.0 / J);
  double vol = J * prime;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double dev = s * (b[i][j] - (i == j ? trb / 3.0 : 0.0));
      double v = i == j ? vol : 0.0;
      t[i][j] = J >= 1.0 ? g * (dev + v) : g * dev + v;
    }
  to_cm(t, tau);
}

/* ---- return mapping (Lib/Ziran/Physics/PlasticityApplier.cpp) ---- */
static void rebuild(M3 u, const double* s, M3 v, double* F)
{
  M3 f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) f[i][j] = u[i][0] * s[0] * v[j][0] + u[i][1] * s[1] * v[j][1] + u[i][2] * s[2] * v[j][2];
  to_cm(f, F);
}

/* VonMisesStvkHencky::projectStrain (:464-501).  Returns 1 when F changed.
   *yield_stress is hardened by xi * delta_gamma; *broken is set when the
   projected first principal Kirchhoff stress reaches fail_stress. */
int zro_von_mises_project(double mu, double lam, double xi, double fail_stress, double* yield_stress, double* F,
                          int* broken)
{
  M3 f, u, v;
  double sig[3], eps[3], H[3], e[3];
  from_cm(F, f);
  svd3_rm(f, u, sig, v);
  double tr = 0.0;
  for (int i = 0; i < 3; ++i) {
    eps[i] = log(sig[i] > 1e-4 ? sig[i] : 1e-4);
    tr += eps[i];
  }
  double n2 = 0.0;
  double hat[3];
  for (int i = 0; i < 3; ++i) {
    hat[i] = eps[i] - tr / 3.0;
    n2 += hat[i] * hat[i];
  }
  double nrm = sqrt(n2);
  double dg = nrm - *yield_stress / (2.0 * mu);
  if (dg <= 0.0) return 0;
  *yield_stress -= xi * dg;
  for (int i = 0; i < 3; ++i) H[i] = eps[i] - (dg / nrm) * hat[i];
  double tau0 = 2.0 * mu * H[0] + lam * (H[0] + H[1] + H[2]);
  for (int i = 0; i < 3; ++i) e[i] = exp(H[i]);
  rebuild(u, e, v, F);
  if (broken && tau0 >= fail_stress) *broken = 1;
  return 1;
}

/* DruckerPragerStvkHencky::projectStrain (:275-316); friction angle enters
   through alpha (:249-260, zro_drucker_prager_alpha). */
double zro_drucker_prager_alpha(double friction_angle_deg)
{
  double s = sin(friction_angle_deg / 180.0 * 3.141592653);
  return sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s);
}

int zro_drucker_prager_project(double mu, double lam, double alpha, double beta, double cohesion,
                               int volume_correction, double* logJp, double* F)
{
  M3 f, u, v;
  double sig[3], eps[3], hat[3], H[3], e[3];
  from_cm(F, f);
  svd3_rm(f, u, sig, v);
  double sum = 0.0;
  for (int i = 0; i < 3; ++i) {
    double a = fabs(sig[i]);
    eps[i] = log(a > 1e-4 ? a : 1e-4) - cohesion;
    sum += eps[i];
  }
  double tr = sum + *logJp;
  double n2 = 0.0;
  for (int i = 0; i < 3; ++i) {
    hat[i] = eps[i] - tr / 3.0;
    n2 += hat[i] * hat[i];
  }
  if (tr >= 0.0) { /* case II: project to the tip */
    double c = exp(cohesion);
    double cs[3] = { c, c, c };
    rebuild(u, cs, v, F);
    if (volume_correction) *logJp = beta * sum + *logJp;
    return 1;
  }
  if (mu == 0.0) return 0;
  *logJp = 0.0;
  double nrm = sqrt(n2);
  double dg = nrm + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * alpha;
  if (dg <= 0.0) { /* case I: inside the yield surface */
    for (int i = 0; i < 3; ++i) H[i] = eps[i] + cohesion;
  } else { /* case III: project onto the cone */
    for (int i = 0; i < 3; ++i) H[i] = eps[i] - (dg / nrm) * hat[i] + cohesion;
  }
  for (int i = 0; i < 3; ++i) e[i] = exp(H[i]);
  rebuild(u, e, v, F);
  return 1;
}

/* ---- FElasticNonequilibratedBasedMpmForceHelper::evolveStrain
   (FElasticNonequilibratedBasedMpmForceHelper.h:62-87) ---- */
/* Fneq = (I + dt gradV) Fneq (the form of the engine's F update), then the
   Hencky strain relaxation with alpha = 2 mu / eta_d and
   beta = 2 (2 mu + 3 lambda) / (9 eta_v) - 2 mu / (3 eta_d):
   eps = (eps_trial - dt beta / (1 + dt (alpha + 3 beta)) t 1) / (1 + dt alpha),
   where t is Eigen's trace() of the 3x1 vector eps_trial, i.e. its first
   entry (the reference calls trace() on a vector). */
void zro_visco_relax(double mu, double lam, double visc_d, double visc_v, double dt, const double* gradV_cm,
                     double* Fneq_cm)
{
  double Fo[3][3], G[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      Fo[r][c] = Fneq_cm[c * 3 + r];
      G[r][c] = gradV_cm[c * 3 + r];
    }
  M3 f, u, v;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double acc = Fo[r][c];
      for (int k = 0; k < 3; ++k) acc += dt * G[r][k] * Fo[k][c];
      f[r][c] = acc;
    }
  double sig[3], eps[3], e[3];
  svd3_rm(f, u, sig, v);
  const double alpha = 2.0 * mu / visc_d;
  const double beta = 2.0 * (2.0 * mu + lam * 3.0) / (9.0 * visc_v) - 2.0 * mu / (visc_d * 3.0);
  for (int i = 0; i < 3; ++i) eps[i] = log(fabs(sig[i]));
  const double t = eps[0];
  const double k = dt * beta / (1.0 + dt * (alpha + 3.0 * beta)) * t;
  for (int i = 0; i < 3; ++i) {
    const double ei = 1.0 / (1.0 + dt * alpha) * (eps[i] - k);
    e[i] = exp(ei);
  }
  double out[9];
  rebuild(u, e, v, out);
  memcpy(Fneq_cm, out, sizeof(out));
}

/* ---- NonAssociativeCamClay (PlasticityApplier.cpp:10-19, 36-135) ---- */
/* M from the friction angle in degrees (the constructor, :10-19) */
double zro_cam_clay_M(double friction_angle_deg, int dim)
{
  double sin_phi = sin(friction_angle_deg / 180.0 * 3.141592653);
  double mohr_columb_friction = sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi);
  return mohr_columb_friction * (double)dim / sqrt(2.0 / (6.0 - dim));
}

/* projectStrain on a NeoHookeanBorden model (kappa = 0.666666 mu + lambda,
   NeoHookeanBorden.cpp:20): the live branch is Compare_With_Physbam on the
   singular values with p0 = kappa (1e-5 + sinh(xi max(-logJp, 0))), then
   F = U diag(Se) V^T.  Returns 1 when the state was outside the surface or
   at a tip. */
int zro_cam_clay_project(double mu, double lam, double M, double beta, double xi, double* logJp, double* F)
{
  M3 f, u, v;
  double Se[3];
  from_cm(F, f);
  svd3_rm(f, u, Se, v);
  const double kappa = 0.666666 * mu + lam;
  const double mlj = -*logJp;
  const double p0 = kappa * (0.00001 + sinh(xi * (mlj > 0.0 ? mlj : 0.0)));
  const double a = (1.0 + 2.0 * beta) * (6.0 - 3.0) / 2.0;
  const double b = beta * p0, c = p0, M2 = M * M;
  double Je = 1.0;
  for (int i = 0; i < 3; ++i) Je *= Se[i];
  double Se2[3], sum = 0.0;
  for (int i = 0; i < 3; ++i) {
    Se2[i] = Se[i] * Se[i];
    sum += Se2[i];
  }
  sum /= 3.0;
  double s_hat[3];
  const double scal = mu * pow(Je, -2.0 / 3.0);
  for (int i = 0; i < 3; ++i) s_hat[i] = scal * (Se2[i] - sum);
  const double Uprime = kappa / 2.0 * (Je - 1.0 / Je);
  const double p_trial = -Uprime * Je;
  int moved = 0;
  if (p_trial > c) { /* tips */
    double Je_new = sqrt(-2.0 * c / kappa + 1.0);
    for (int i = 0; i < 3; ++i) Se[i] = pow(Je_new, 1.0 / 3.0);
    *logJp += log(Je / Je_new);
    moved = 1;
  } else if (p_trial < -b) {
    double Je_new = sqrt(2.0 * b / kappa + 1.0);
    for (int i = 0; i < 3; ++i) Se[i] = pow(Je_new, 1.0 / 3.0);
    *logJp += log(Je / Je_new);
    moved = 1;
  } else {
    double k = sqrt(-M2 * (p_trial + b) * (p_trial - c) / a);
    double s_norm = sqrt((s_hat[0] * s_hat[0] + s_hat[1] * s_hat[1]) + s_hat[2] * s_hat[2]);
    double y = a * (s_norm * s_norm) + M2 * (p_trial + b) * (p_trial - c);
    if (y >= 1e-4) { /* outside: onto the surface, with the "fake" hardening */
      double pc = (1.0 - beta) * p0 / 2.0;
      if (p0 > 1e-4 && p_trial < p0 - 1e-4 && p_trial > -beta * p0 + 1e-4) {
        double aa = M2 * ((p_trial - pc) * (p_trial - pc)) / (a * (s_norm * s_norm));
        double dd = 1.0 + aa;
        double ff = aa * beta * p0 - aa * p0 - 2.0 * pc;
        double gg = pc * pc - aa * beta * (p0 * p0);
        double zz = sqrt(fabs(ff * ff - 4.0 * dd * gg));
        double p1 = (-ff + zz) / (2.0 * dd);
        double p2 = (-ff - zz) / (2.0 * dd);
        double p_fake = (p_trial - pc) * (p1 - pc) > 0.0 ? p1 : p2;
        double Je_new_fake = sqrt(fabs(-2.0 * p_fake / kappa + 1.0));
        if (Je_new_fake > 1e-4) *logJp += log(Je / Je_new_fake);
      }
      double ssum = (Se2[0] + Se2[1]) + Se2[2];
      double scale = k / mu * pow(Je, 2.0 / 3.0);
      for (int i = 0; i < 3; ++i) Se[i] = sqrt(scale * s_hat[i] / s_norm + 1.0 / 3.0 * ssum);
      moved = 1;
    }
  }
  rebuild(u, Se, v, F);
  return moved;
}

/* ---- adjoint of the fixed-corotated Kirchhoff stress (no reference counterpart) ---- */
/* reverse of tau = 2 mu (F - R) F^T + lam (J - 1) J I at F (column-major
   9-vectors): adds to Fbar, returns mubar / lambar */
void zro_fcr_kirchhoff_adjoint(double mu, double lam, const double* Fcm, const double* taubar_cm, double* Fbar_cm,
                               double* mubar, double* lambar)
{
  M3 f, u, v, tb, r, fb, A, Ah, Mb, c;
  double sig[3];
  from_cm(Fcm, f);
  from_cm(taubar_cm, tb);
  from_cm(Fbar_cm, fb);
  svd3_rm(f, u, sig, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = u[i][0] * v[j][0] + u[i][1] * v[j][1] + u[i][2] * v[j][2];
  double J = f[0][0] * (f[1][1] * f[2][2] - f[2][1] * f[1][2]) - f[0][1] * (f[1][0] * f[2][2] - f[2][0] * f[1][2]) +
             f[0][2] * (f[1][0] * f[2][1] - f[2][0] * f[1][1]);
  c[0][0] = f[1][1] * f[2][2] - f[1][2] * f[2][1];
  c[0][1] = f[1][2] * f[2][0] - f[1][0] * f[2][2];
  c[0][2] = f[1][0] * f[2][1] - f[1][1] * f[2][0];
  c[1][0] = f[0][2] * f[2][1] - f[0][1] * f[2][2];
  c[1][1] = f[0][0] * f[2][2] - f[0][2] * f[2][0];
  c[1][2] = f[0][1] * f[2][0] - f[0][0] * f[2][1];
  c[2][0] = f[0][1] * f[1][2] - f[0][2] * f[1][1];
  c[2][1] = f[0][2] * f[1][0] - f[0][0] * f[1][2];
  c[2][2] = f[0][0] * f[1][1] - f[0][1] * f[1][0];
  double trtb = tb[0][0] + tb[1][1] + tb[2][2];
  /* A = taubar F  (for the dR term) */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = tb[i][0] * f[0][j] + tb[i][1] * f[1][j] + tb[i][2] * f[2][j];
  /* Ah = U^T A V */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) acc += u[k][i] * A[k][l] * v[l][j];
      Ah[i][j] = acc;
    }
  for (int i = 0; i < 3; ++i) Mb[i][i] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j) {
      double den = sig[i] + sig[j];
      if (fabs(den) < 1e-12) den = (den < 0 ? -1e-12 : 1e-12);
      double x = (Ah[i][j] - Ah[j][i]) / den;
      Mb[i][j] = x;
      Mb[j][i] = -x;
    }
  double mb = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      /* 2 mu (tb + tb^T) F */
      double t1 = 0.0, t2 = 0.0, t3 = 0.0;
      for (int k = 0; k < 3; ++k) {
        t1 += (tb[i][k] + tb[k][i]) * f[k][j];
        t2 += tb[k][i] * r[k][j];             /* tb^T R */
        t3 += u[i][k] * (Mb[k][0] * v[j][0] + Mb[k][1] * v[j][1] + Mb[k][2] * v[j][2]);  /* U Mb V^T */
      }
      fb[i][j] += 2.0 * mu * (t1 - t2 - t3) + lam * (2.0 * J - 1.0) * trtb * c[i][j];
      double fr = 0.0;
      for (int k = 0; k < 3; ++k) fr += (f[i][k] - r[i][k]) * f[j][k];
      mb += tb[i][j] * 2.0 * fr;
    }
  to_cm(fb, Fbar_cm);
  *mubar = mb;
  *lambar = trtb * (J - 1.0) * J;
}

