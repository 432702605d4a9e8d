// zr_admm.hpp -- the ADMM step of Projects/admm/AdmmSimulationFoam.h
// (StvkWithHencky elasticity with no / von Mises / Drucker-Prager plasticity,
// or elastic with the viscoelastic branch, enable_visco) for the gfx950
// engine.  Included by zr_engine.hip inside its anonymous namespace after
// zr_adjoint.hpp.
//
// Per step: APIC P2G; initStep (warm-started dv, collision projection,
// F = (I + dt grad(vn + dv)) Fn, omega from the PD-clamped Hencky Hessians,
// y rescaled by omega_old / omega); then admm_iterations x { collision
// update, local F step (SVD + Newton on the singular values with the return
// map's derivative), global dv step (MINRES on M + sum_p rho omega^2 dt^2
// (Fn^T grad w)(Fn^T grad w)^T, applied matrix-free, with the collision
// projection and the diagonal preconditioner), dual update y }.  Particle
// passes use the group's LDS tile of a node vector; scatters use per-wave
// LDS tiles staged per group and summed per node in a fixed order.


struct AdmmParticles {
  double* F;       // [9][n] ADMM F
  double* y;       // [9][n] scaled dual
  double* om;      // [n] omega
  double* last;    // [3][n] last dv per particle
  // the viscoelastic branch (the reference's entries i + particles.count):
  // its ADMM F, dual and omega, the particle's Fe_Nonequilibrated (FFn) and
  // the branch's StvkWithHencky model and viscosities
  double* F2;      // [9][n]
  double* y2;      // [9][n]
  double* om2;     // [n]
  const double* fneq;
  const double* mu2;
  const double* lam2;
  double visc_d, visc_v;
};

struct AdmmNodes {
  double* vn;      // [3][cap]
  double* dv;      // [3][cap]
  double* proj;    // [9][cap] collision projection
  int32_t* coll;   // [cap]
  double* diag;    // [cap]
  int64_t stride;
};

// ---- per-particle math (restating StvkWithHencky.cpp:34-204, PlasticityApplier.cpp) ----
__device__ __forceinline__ void hencky_dE(double mu, double lam, const double s[3], double dE[3])
{
  const double l0 = log(fabs(s[0])), l1 = log(fabs(s[1])), l2 = log(fabs(s[2]));
  const double sl = l0 + l1 + l2;
  dE[0] = (2.0 * mu * l0 + lam * sl) / s[0];
  dE[1] = (2.0 * mu * l1 + lam * sl) / s[1];
  dE[2] = (2.0 * mu * l2 + lam * sl) / s[2];
}
__device__ __forceinline__ void hencky_ddE(double mu, double lam, const double s[3], double d[3][3])
{
  const double l0 = log(fabs(s[0])), l1 = log(fabs(s[1])), l2 = log(fabs(s[2]));
  const double g = 2.0 * mu + lam;
  d[0][0] = (g * (1.0 - l0) - lam * (l1 + l2)) / (s[0] * s[0]);
  d[1][1] = (g * (1.0 - l1) - lam * (l0 + l2)) / (s[1] * s[1]);
  d[2][2] = (g * (1.0 - l2) - lam * (l0 + l1)) / (s[2] * s[2]);
  d[0][1] = d[1][0] = lam / (s[0] * s[1]);
  d[0][2] = d[2][0] = lam / (s[0] * s[2]);
  d[1][2] = d[2][1] = lam / (s[1] * s[2]);
}
template <int I, int J>
__device__ __forceinline__ double hencky_Bij_pd2(double mu, double lam, const double s[3], double clamp)
{
  double dE[3];
  hencky_dE(mu, lam, s, dE);
  const double sum = s[I] + s[J];
  const double mag = fabs(sum), sign = sum > 0.0 ? 1.0 : -1.0;
  const double cl = (mag > clamp ? mag : clamp) * sign;
  const double BP = 0.5 * (dE[I] + dE[J]) / cl;
  const double sl = log(fabs(s[0])) + log(fabs(s[1])) + log(fabs(s[2]));
  const double p = s[I] / s[J] - 1.0;
  const double l1p = fabs(p) < 1e-8 ? 1.0 : log1p(p) / p;
  const double interlock = log(fabs(s[J])) - s[J] * (l1p / s[J]);
  const double BM = -0.5 * (lam * sl + 2.0 * mu * interlock) / (s[I] * s[J]);
  // |makePD2D([[BP+BM, BM-BP],[BM-BP, BP+BM]])|_F^2 from its eigenvalues
  const double a = BP + BM, b = BM - BP;
  const double e0 = a - b, e1 = a + b;
  return (e0 > 0.0 ? e0 * e0 : 0.0) + (e1 > 0.0 ? e1 * e1 : 0.0);
}

// eigenvalues of a symmetric 3x3 matrix (trigonometric closed form of the
// characteristic cubic, O. K. Smith 1961)
__device__ __forceinline__ void sym3_eigs(const double A[3][3], double l[3])
{
  const double a01 = 0.5 * (A[0][1] + A[1][0]), a02 = 0.5 * (A[0][2] + A[2][0]), a12 = 0.5 * (A[1][2] + A[2][1]);
  const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  const double q = (A[0][0] + A[1][1] + A[2][2]) / 3.0;
  const double b00 = A[0][0] - q, b11 = A[1][1] - q, b22 = A[2][2] - q;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  if (p2 <= 0.0) {
    l[0] = l[1] = l[2] = q;
    return;
  }
  const double p = sqrt(p2 / 6.0), ip = 1.0 / p;
  const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip, c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  const double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
  double r = 0.5 * det;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  const double phi = acos(r) / 3.0;
  const double e0 = q + 2.0 * p * cos(phi);
  const double e2 = q + 2.0 * p * cos(phi + 2.0943951023931957);
  l[0] = e2;
  l[1] = 3.0 * q - e0 - e2;
  l[2] = e0;
}

struct PlasticParams {
  int type;
  double a, b, c;   // VM: yield stress; DP: alpha, -, cohesion
};

__device__ __forceinline__ PlasticParams plastic_params(const ParticlesView& P, const int32_t* mat, const double* pparam,
                                                        const double* pstate, uint32_t i)
{
  PlasticParams q{ ZR_PLASTIC_NONE, 0.0, 0.0, 0.0 };
  if (!mat) return q;
  q.type = mat[i] >> 4;
  if (q.type == ZR_PLASTIC_VON_MISES) {
    q.a = pstate[i];
  } else if (q.type == ZR_PLASTIC_DRUCKER_PRAGER) {
    q.a = pparam[i];
    q.c = pparam[2 * P.stride + i];
  }
  return q;
}

// projectSigma (PlasticityApplier.cpp:321-346, 528-547; DummyPlasticity: sigma)
__device__ __forceinline__ void admm_project_sigma(const PlasticParams& q, double mu, double lam, const double s[3],
                                                   double out[3])
{
  out[0] = s[0];
  out[1] = s[1];
  out[2] = s[2];
  if (q.type == ZR_PLASTIC_NONE) return;
  double e[3], h[3], tr = 0.0, n2 = 0.0;
  if (q.type == ZR_PLASTIC_VON_MISES) {
    for (int d = 0; d < 3; ++d) {
      e[d] = log(s[d] > 1e-4 ? s[d] : 1e-4);
      tr += e[d];
    }
    for (int d = 0; d < 3; ++d) {
      h[d] = e[d] - tr / 3.0;
      n2 += h[d] * h[d];
    }
    const double nrm = sqrt(n2), dg = nrm - q.a / (2.0 * mu);
    if (dg <= 0.0) return;
    for (int d = 0; d < 3; ++d) out[d] = exp(e[d] - (dg / nrm) * h[d]);
  } else {
    for (int d = 0; d < 3; ++d) {
      const double a = fabs(s[d]);
      e[d] = log(a > 1e-4 ? a : 1e-4) - q.c;
      tr += e[d];
    }
    for (int d = 0; d < 3; ++d) {
      h[d] = e[d] - tr / 3.0;
      n2 += h[d] * h[d];
    }
    if (tr >= 0.0) {
      for (int d = 0; d < 3; ++d) out[d] = exp(q.c);
      return;
    }
    const double nrm = sqrt(n2);
    const double dg = nrm + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * q.a;
    for (int d = 0; d < 3; ++d) out[d] = exp(dg <= 0.0 ? e[d] + q.c : e[d] - (dg / nrm) * h[d] + q.c);
  }
}

// projectSigmaAndDerivative (PlasticityApplier.cpp:389-423, 580-606; Dummy: identity)
__device__ __forceinline__ void admm_project_sigma_d(const PlasticParams& q, double mu, double lam, const double s[3],
                                                     double zp[3], double D[3][3])
{
  for (int a = 0; a < 3; ++a) {
    zp[a] = s[a];
    for (int b = 0; b < 3; ++b) D[a][b] = a == b ? 1.0 : 0.0;
  }
  if (q.type == ZR_PLASTIC_NONE) return;
  double e[3], h[3], tr = 0.0, n2 = 0.0;
  if (q.type == ZR_PLASTIC_VON_MISES) {
    for (int d = 0; d < 3; ++d) {
      e[d] = log(s[d] > 1e-4 ? s[d] : 1e-4);
      tr += e[d];
    }
    for (int d = 0; d < 3; ++d) {
      h[d] = e[d] - tr / 3.0;
      n2 += h[d] * h[d];
    }
    const double nrm = sqrt(n2), dg = nrm - q.a / (2.0 * mu);
    if (dg <= 0.0) return;
    double w[3], sh[3], sn = 0.0;
    for (int d = 0; d < 3; ++d) {
      w[d] = 1.0 / s[d];
      sn += h[d] * h[d];
    }
    sn = sqrt(sn);
    for (int d = 0; d < 3; ++d) {
      sh[d] = h[d] / sn;
      zp[d] = exp(e[d] - (dg / nrm) * h[d]);
    }
    const double a1 = 1.0 - q.a / (2.0 * mu * sn), a2 = q.a / (2.0 * mu * sn);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const double Iw = r == c ? w[c] : 0.0;
        D[r][c] = zp[r] * (Iw - a1 * (Iw - w[c] / 3.0) - a2 * sh[r] * sh[c] * w[c]);
      }
  } else {
    for (int d = 0; d < 3; ++d) {
      const double a = fabs(s[d]);
      e[d] = log(a > 1e-4 ? a : 1e-4) - q.c;
      tr += e[d];
    }
    for (int d = 0; d < 3; ++d) {
      h[d] = e[d] - tr / 3.0;
      n2 += h[d] * h[d];
    }
    const double nrm = sqrt(n2);
    const double dg = nrm + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * q.a;
    if (dg <= 0.0) {
      for (int d = 0; d < 3; ++d) zp[d] = exp(e[d] + q.c);
      return;
    }
    if (tr > 0.0 || nrm == 0.0) {
      for (int d = 0; d < 3; ++d) {
        zp[d] = exp(q.c);
        for (int c = 0; c < 3; ++c) D[d][c] = 0.0;
      }
      return;
    }
    double w[3], sh[3], sn = 0.0;
    const double k = tr;
    for (int d = 0; d < 3; ++d) {
      w[d] = 1.0 / s[d];
      sn += h[d] * h[d];
    }
    sn = sqrt(sn);
    for (int d = 0; d < 3; ++d) sh[d] = h[d] / sn;
    const double pp = q.a * k * (3.0 * lam + 2.0 * mu) / (2.0 * mu * sn);
    for (int d = 0; d < 3; ++d) zp[d] = exp(e[d] - (dg / nrm) * h[d] + q.c);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        const double t1 = ((1.0 + 2.0 * pp) / 3.0 - pp / k * e[r]) * w[c];
        const double t2 = pp * ((r == c ? 1.0 : 0.0) - sh[r] * sh[c]) * w[c];
        D[r][c] = zp[r] * (t1 - t2);
      }
  }
}

// initOmega (AdmmSimulationFoam.h:222-274)
__device__ __forceinline__ double admm_omega(const PlasticParams& q, double mu, double lam, double vol,
                                             const double s[3])
{
  double A[3][3], l[3], ps[3];
  hencky_ddE(mu, lam, s, A);
  sym3_eigs(A, l);
  double total = 0.0;
  for (int d = 0; d < 3; ++d) total += l[d] > 0.0 ? l[d] * l[d] : 0.0;
  admm_project_sigma(q, mu, lam, s, ps);
  total += hencky_Bij_pd2<0, 1>(mu, lam, ps, 1e-6);
  total += hencky_Bij_pd2<0, 2>(mu, lam, ps, 1e-6);
  total += hencky_Bij_pd2<1, 2>(mu, lam, ps, 1e-6);
  return sqrt(vol * sqrt(total));
}

// initOmega for the viscoelastic branch (:254-270): its StvkWithHencky model,
// no projection
__device__ __forceinline__ double admm_visco_omega(double mu, double lam, double vol, const double s[3])
{
  const PlasticParams none{ ZR_PLASTIC_NONE, 0.0, 0.0, 0.0 };
  return admm_omega(none, mu, lam, vol, s);
}

// the viscoelastic branch's return map in FStepSolve (:583-603): relaxed
// singular values zp = exp((log|z| - c t 1) / (1 + dt alpha)),
// c = dt beta / (1 + dt (alpha + 3 beta)), t = log|z_0| (Eigen's trace() of
// the 3x1 log vector is its first entry), and the derivative the reference
// forms, D = diag(zp) (diag(1/z) - c 1 (1/z)^T) / (1 + dt alpha)
__device__ __forceinline__ void admm_visco_map(double mu, double lam, double visc_d, double visc_v, double dt,
                                               const double z[3], double zp[3], double D[3][3])
{
  const double alpha = 2.0 * mu / visc_d;
  const double beta = 2.0 * (2.0 * mu + lam * 3.0) / (9.0 * visc_v) - 2.0 * mu / (visc_d * 3.0);
  const double c = dt * beta / (1.0 + dt * (alpha + 3.0 * beta));
  const double h = 1.0 / (1.0 + dt * alpha);
  double e[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) e[d] = log(fabs(z[d]));
  const double t = e[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) zp[d] = exp(h * (e[d] - c * t));
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) D[a][b] = zp[a] * h * ((a == b ? 1.0 / z[b] : 0.0) - c / z[b]);
}

__device__ __forceinline__ void inv3(const double H[3][3], double Hi[3][3])
{
  const double det = H[0][0] * (H[1][1] * H[2][2] - H[2][1] * H[1][2]) -
                     H[0][1] * (H[1][0] * H[2][2] - H[2][0] * H[1][2]) +
                     H[0][2] * (H[1][0] * H[2][1] - H[2][0] * H[1][1]);
  Hi[0][0] = (H[1][1] * H[2][2] - H[1][2] * H[2][1]) / det;
  Hi[0][1] = (H[0][2] * H[2][1] - H[0][1] * H[2][2]) / det;
  Hi[0][2] = (H[0][1] * H[1][2] - H[0][2] * H[1][1]) / det;
  Hi[1][0] = (H[1][2] * H[2][0] - H[1][0] * H[2][2]) / det;
  Hi[1][1] = (H[0][0] * H[2][2] - H[0][2] * H[2][0]) / det;
  Hi[1][2] = (H[0][2] * H[1][0] - H[0][0] * H[1][2]) / det;
  Hi[2][0] = (H[1][0] * H[2][1] - H[1][1] * H[2][0]) / det;
  Hi[2][1] = (H[0][1] * H[2][0] - H[0][0] * H[2][1]) / det;
  Hi[2][2] = (H[0][0] * H[1][1] - H[0][1] * H[1][0]) / det;
}

// multiObjectCollisionWithFakeSeparate (CollisionObject.cpp:108-160) in the
// frame at the step's start.
//
// A separating object releases the node when its relative normal velocity
// (v - v_obj).n is positive.  The ADMM loop repeats that test on nodes the
// previous iteration projected, whose relative normal velocity is then
// exactly zero: the projection removed it, and the global step only moves dv
// inside the projected space because its preconditioner is one scalar per
// node (AdmmSimulationFoam.h:476-503).  Tested against zero in floating point,
// those nodes are released or kept by rounding noise, and a 1e-15 change of
// the input changes the step's result by percent.  The test here is against
// a rounding bound, (v - v_obj).n > kSeparateRoundoff sum_d (|v_d| + |v_obj,d|),
// which gives the exact-arithmetic outcome: zero relative normal velocity
// stays in contact.
constexpr double kSeparateRoundoff = 1e-12;
__device__ bool admm_collision(const ColliderSet& cs, const double x[3], double v[3], double P[3][3])
{
  bool any = false;
  int nslip = 0;
  double K[3][3];
  for (int k = 0; k < cs.n; ++k) {
    const DevCollider& c = cs.c[k];
    double xb[3], phi, n[3], vo[3];
    if (!collider_query(c, 0, x, phi, n, xb)) continue;
    collider_velocity(c, 0, xb, vo);
    const double old[3] = { v[0], v[1], v[2] };
    bool col = detect_and_resolve(c, 0, x, v);
    if (col && c.type == ZR_SEPARATE) {
      double dot = 0.0, scale = 0.0;
      for (int d = 0; d < 3; ++d) {
        dot += (old[d] - vo[d]) * n[d];
        scale += fabs(old[d]) + fabs(vo[d]);
      }
      if (dot > kSeparateRoundoff * scale) {
        v[0] = old[0];
        v[1] = old[1];
        v[2] = old[2];
        col = false;
      }
    }
    any = any || col;
    if (!col) continue;
    if (c.type == ZR_STICKY) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = 0.0;
      return true;
    }
    for (int q = 0; q < nslip; ++q) {
      const double dot = K[q][0] * n[0] + K[q][1] * n[1] + K[q][2] * n[2];
      for (int d = 0; d < 3; ++d) n[d] -= dot * K[q][d];
    }
    const double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0.0) {
      for (int d = 0; d < 3; ++d) K[nslip][d] = n[d] / l;
      if (++nslip == 3) break;
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      P[i][j] = i == j ? 1.0 : 0.0;
      for (int q = 0; q < nslip; ++q) P[i][j] -= K[q][i] * K[q][j];
    }
  return any;
}

// ---- node passes ----

// initStep (:296-319): vn = v, dv = 0 or the warm start from last_dv (staged
// sums of last_dv w (3) and w (1))
__global__ __launch_bounds__(256) void k_admm_init_nodes(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                         const double* __restrict__ stage, const double* __restrict__ gm,
                                                         const double* __restrict__ gv, AdmmNodes N, int warm,
                                                         double dt)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const int64_t S = N.stride;
  double vn[3] = { 0.0, 0.0, 0.0 }, dv[3] = { 0.0, 0.0, 0.0 };
  if (gm[k] != 0.0) {
    for (int d = 0; d < 3; ++d) vn[d] = gv[d * S + k];
    if (warm) {
      const uint32_t e = k & 31u;
      double f[4];
      sum_tiles_src<4>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
      for (int d = 0; d < 3; ++d) dv[d] = f[d] * dt / f[3];
    }
  }
  for (int d = 0; d < 3; ++d) {
    N.vn[d * S + k] = vn[d];
    N.dv[d * S + k] = dv[d];
  }
}

// updateCollisionNodes (:276-294)
__global__ __launch_bounds__(256) void k_admm_collide(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                                      const double* __restrict__ gm, AdmmNodes N, double dx,
                                                      ColliderSet cs)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const int64_t S = N.stride;
  int32_t hit = 0;
  if (gm[k] != 0.0 && cs.n > 0) {
    double xi[3], vi[3], P[3][3];
    node_pos(node_coord(slot_page, k), dx, xi);
    for (int d = 0; d < 3; ++d) vi[d] = N.dv[d * S + k] + N.vn[d * S + k];
    if (admm_collision(cs, xi, vi, P)) {
      hit = 1;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) N.proj[(3 * i + j) * S + k] = P[i][j];
      for (int d = 0; d < 3; ++d) N.dv[d * S + k] = vi[d] - N.vn[d * S + k];
    }
  }
  N.coll[k] = hit;
}

// ---- particle passes (one workgroup per group) ----

// tile of a node 3-vector (a + b when b is given), zero at inactive nodes
__device__ __forceinline__ void admm_tile3(const int32_t* gslot, const double* gm, const double* a, const double* b,
                                           int64_t S, double* tv, int32_t* nslot)
{
  const int tid = threadIdx.x;
  if (tid < 8) nslot[tid] = gslot[tid];
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += blockDim.x) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    const bool on = sl >= 0 && gm[node] != 0.0;
    for (int d = 0; d < 3; ++d)
      tv[d * kTileNodes + t] = on ? a[d * S + node] + (b ? b[d * S + node] : 0.0) : 0.0;
  }
  __syncthreads();
}

// G = sum_i v_i grad w_i^T over the tile
__device__ __forceinline__ void admm_grad_tile(const double* tv, const double w[3][3], const double dw[3][3],
                                               double inv_dx, int lx, int ly, int lz, double G[3][3])
{
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) G[r][c] = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double gw[3];
        wgrad(w, dw, inv_dx, a, b, c, gw);
        const int t = ((lx + a) * kTileY + (ly + b)) * kTileZ + (lz + c);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const double v = tv[r * kTileNodes + t];
          G[r][0] += v * gw[0];
          G[r][1] += v * gw[1];
          G[r][2] += v * gw[2];
        }
      }
}

__device__ __forceinline__ void load_F(const double* F, int64_t S, uint32_t i, double M[3][3])
{
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) M[r][c] = F[(c * 3 + r) * S + i];
}

// FStepSolve's viscoelastic branch (:566-611): F2 = U diag(z) V^T with z the
// Newton solution (at most 20 iterations) of
// vol dE(zp(z)) + rho omega2^2 (z - sigma) = 0, sigma the singular values of
// FFn - y2 / (rho omega2) + dt G FFn, G = grad(vn + dv)
__device__ __forceinline__ void admm_visco_fstep(const AdmmParticles& A, int64_t S, uint32_t i, double vol,
                                                 const double G[3][3], double dt, double rho, double local_tol)
{
  double Fv[3][3];
  load_F(A.fneq, S, i, Fv);
  const double om = A.om2[i];
  M3<double> fd, u, v;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      fd.a[r][c] = Fv[r][c] - A.y2[(c * 3 + r) * S + i] / (om * rho) +
                   dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
  double sg[3];
  svd3(fd, u, sg, v);
  double z[3];
  {
    double Fa[3][3];
    load_F(A.F2, S, i, Fa);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) acc += u.a[a][d] * Fa[a][bb] * v.a[bb][d];
      z[d] = acc;
    }
  }
  const double mu = A.mu2[i], lam = A.lam2[i], coeff = rho * om * om;
  for (int it = 0; it < 20; ++it) {
    double zp[3], D[3][3], dE[3], ddE[3][3], gg[3], H[3][3], Hi[3][3];
    admm_visco_map(mu, lam, A.visc_d, A.visc_v, dt, z, zp, D);
    hencky_dE(mu, lam, zp, dE);
    hencky_ddE(mu, lam, zp, ddE);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      gg[a] = vol * dE[a] + coeff * (z[a] - sg[a]);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
        H[a][bb] = vol * (ddE[a][0] * D[0][bb] + ddE[a][1] * D[1][bb] + ddE[a][2] * D[2][bb]) + (a == bb ? coeff : 0.0);
    }
    inv3(H, Hi);
    double sn = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double st = -(Hi[a][0] * gg[0] + Hi[a][1] * gg[1] + Hi[a][2] * gg[2]);
      z[a] += st;
      sn += st * st;
    }
    if (sqrt(sn) < local_tol) break;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      A.F2[(c * 3 + r) * S + i] = u.a[r][0] * z[0] * v.a[c][0] + u.a[r][1] * z[1] * v.a[c][1] + u.a[r][2] * z[2] * v.a[c][2];
}


// initStep F and omega (:359-377), the y rescale, and the preconditioner
// scatter c_p |Fn^T grad w|^2 (:480-500), one staged field; VISCO: the same
// for the viscoelastic branch's F2 from FFn (:372-400, 488-497)
template <bool VISCO>
__global__ __launch_bounds__(256) void k_admm_init_particles(ParticlesView P, AdmmParticles A, AdmmNodes N,
                                                             const int32_t* __restrict__ mat,
                                                             const double* __restrict__ pparam,
                                                             const double* __restrict__ pstate,
                                                             const uint32_t* __restrict__ order,
                                                             const uint32_t* __restrict__ group_start,
                                                             const uint32_t* __restrict__ group_page,
                                                             const int32_t* __restrict__ group_nslot,
                                                             const double* __restrict__ gm, double* stage, int first,
                                                             double dt, double rho, double inv_dx)
{
  __shared__ double tv[3 * kTileNodes];
  __shared__ double tile[4 * kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * kTileNodes; k += 256) tile[k] = 0.0;
  admm_tile3(group_nslot + 8 * (size_t)g, gm, N.vn, N.dv, N.stride, tv, nslot);
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  double* wt = tile + (tid >> 6) * kTileNodes;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double G[3][3], Fn[3][3];
    admm_grad_tile(tv, w, dw, inv_dx, lx, ly, lz, G);
    load_F(P.F, S, i, Fn);
    M3<double> f, u, v;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        f.a[r][c] = Fn[r][c] + dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
        A.F[(c * 3 + r) * S + i] = f.a[r][c];
      }
    double s[3];
    svd3(f, u, s, v);
    const PlasticParams q = plastic_params(P, mat, pparam, pstate, i);
    const double om = admm_omega(q, P.mu[i], P.lam[i], P.vol[i], s);
    const double old = first ? 0.0 : A.om[i];
#pragma unroll
    for (int c = 0; c < 9; ++c) A.y[c * S + i] = first ? 0.0 : A.y[c * S + i] * old / om;
    A.om[i] = om;
    const double cp = rho * om * om * dt * dt;
    double Fv[3][3], cp2 = 0.0;
    if constexpr (VISCO) {
      load_F(A.fneq, S, i, Fv);
      M3<double> f2, u2, v2;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          f2.a[r][c] = Fv[r][c] + dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
          A.F2[(c * 3 + r) * S + i] = f2.a[r][c];
        }
      double s2[3];
      svd3(f2, u2, s2, v2);
      const double om2 = admm_visco_omega(A.mu2[i], A.lam2[i], P.vol[i], s2);
      const double old2 = first ? 0.0 : A.om2[i];
#pragma unroll
      for (int c = 0; c < 9; ++c) A.y2[c * S + i] = first ? 0.0 : A.y2[c * S + i] * old2 / om2;
      A.om2[i] = om2;
      cp2 = rho * om2 * om2 * dt * dt;
    }
    opaque_weights(w, dw);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3], fg[3];
          wgrad(w, dw, inv_dx, a, bb, c, gw);
#pragma unroll
          for (int d = 0; d < 3; ++d) fg[d] = Fn[0][d] * gw[0] + Fn[1][d] * gw[1] + Fn[2][d] * gw[2];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          double val = cp * (fg[0] * fg[0] + fg[1] * fg[1] + fg[2] * fg[2]);
          if constexpr (VISCO) {
#pragma unroll
            for (int d = 0; d < 3; ++d) fg[d] = Fv[0][d] * gw[0] + Fv[1][d] * gw[1] + Fv[2][d] * gw[2];
            val += cp2 * (fg[0] * fg[0] + fg[1] * fg[1] + fg[2] * fg[2]);
          }
          atomicAdd(&wt[t], val);
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * kTileNodes;
  for (int k = tid; k < kTileNodes; k += 256)
    out[k] = tile[k] + tile[kTileNodes + k] + tile[2 * kTileNodes + k] + tile[3 * kTileNodes + k];
}

// diag = m + staged sums
__global__ __launch_bounds__(256) void k_admm_diag(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                   const double* __restrict__ stage, const double* __restrict__ gm,
                                                   double* diag)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  double d = 0.0;
  if (gm[k] != 0.0) {
    const uint32_t e = k & 31u;
    double f[1];
    sum_tiles_src<1>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
    d = gm[k] + f[0];
  }
  diag[k] = d;
}

// the local step of particle i (FStepSolve, AdmmSimulationFoam.h:507-615)
// given G = grad(vn + dv) at the particle
template <bool VISCO>
__device__ __forceinline__ void admm_local_step(const ParticlesView& P, const AdmmParticles& A,
                                                const int32_t* __restrict__ mat, const double* __restrict__ pparam,
                                                const double* __restrict__ pstate, uint32_t i, const double G[3][3],
                                                double dt, double rho, double local_tol)
{
  const int64_t S = P.stride;
  double Fn[3][3];
  load_F(P.F, S, i, Fn);
  const double om = A.om[i];
  M3<double> fd, u, v;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      fd.a[r][c] = Fn[r][c] - A.y[(c * 3 + r) * S + i] / (om * rho) +
                   dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
  double sg[3];
  svd3(fd, u, sg, v);
  double z[3];
  {
    double Fa[3][3];
    load_F(A.F, S, i, Fa);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int bb = 0; bb < 3; ++bb) acc += u.a[a][d] * Fa[a][bb] * v.a[bb][d];
      z[d] = acc;
    }
  }
  const PlasticParams q = plastic_params(P, mat, pparam, pstate, i);
  const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i], coeff = rho * om * om;
  for (int it = 0; it < 5; ++it) {
    double zp[3], D[3][3], dE[3], ddE[3][3], gg[3], H[3][3];
    admm_project_sigma_d(q, mu, lam, z, zp, D);
    hencky_dE(mu, lam, zp, dE);
    hencky_ddE(mu, lam, zp, ddE);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      gg[a] = vol * dE[a] + coeff * (z[a] - sg[a]);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
        H[a][bb] = vol * (ddE[a][0] * D[0][bb] + ddE[a][1] * D[1][bb] + ddE[a][2] * D[2][bb]) +
                   (a == bb ? coeff : 0.0);
    }
    const double det = H[0][0] * (H[1][1] * H[2][2] - H[2][1] * H[1][2]) -
                       H[0][1] * (H[1][0] * H[2][2] - H[2][0] * H[1][2]) +
                       H[0][2] * (H[1][0] * H[2][1] - H[2][0] * H[1][1]);
    double Hi[3][3];
    Hi[0][0] = (H[1][1] * H[2][2] - H[1][2] * H[2][1]) / det;
    Hi[0][1] = (H[0][2] * H[2][1] - H[0][1] * H[2][2]) / det;
    Hi[0][2] = (H[0][1] * H[1][2] - H[0][2] * H[1][1]) / det;
    Hi[1][0] = (H[1][2] * H[2][0] - H[1][0] * H[2][2]) / det;
    Hi[1][1] = (H[0][0] * H[2][2] - H[0][2] * H[2][0]) / det;
    Hi[1][2] = (H[0][2] * H[1][0] - H[0][0] * H[1][2]) / det;
    Hi[2][0] = (H[1][0] * H[2][1] - H[1][1] * H[2][0]) / det;
    Hi[2][1] = (H[0][1] * H[2][0] - H[0][0] * H[2][1]) / det;
    Hi[2][2] = (H[0][0] * H[1][1] - H[0][1] * H[1][0]) / det;
    double sn = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double st = -(Hi[a][0] * gg[0] + Hi[a][1] * gg[1] + Hi[a][2] * gg[2]);
      z[a] += st;
      sn += st * st;
    }
    if (sqrt(sn) < local_tol) break;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      A.F[(c * 3 + r) * S + i] = u.a[r][0] * z[0] * v.a[c][0] + u.a[r][1] * z[1] * v.a[c][1] + u.a[r][2] * z[2] * v.a[c][2];
  if constexpr (VISCO) admm_visco_fstep(A, S, i, vol, G, dt, rho, local_tol);
}

// FStepSolve (:507-615): F = U diag(z) V^T with z the Newton solution of
// vol dE(proj(z)) + rho omega^2 (z - sigma) = 0, sigma the singular values
// of Fn - y/(rho omega) + dt grad(vn + dv) Fn; VISCO: then the viscoelastic
// branch's F2.  In two passes, so that the SVD and the Newton iterations run
// in a kernel without the tile (134 VGPRs, 3 waves/SIMD, against one kernel
// of 256 VGPRs with spills): k_admm_fgrad writes grad(vn + dv) of every
// particle (9 rows of n, particle order), k_admm_flocal runs the local step
// per particle from it
__global__ __launch_bounds__(256) void k_admm_fgrad(ParticlesView P, AdmmNodes N, const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ group_start,
                                                    const uint32_t* __restrict__ group_page,
                                                    const int32_t* __restrict__ group_nslot,
                                                    const double* __restrict__ gm, double inv_dx,
                                                    double* __restrict__ gs)
{
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  const uint32_t g = blockIdx.x;
  admm_tile3(group_nslot + 8 * (size_t)g, gm, N.vn, N.dv, N.stride, tv, nslot);
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    const uint32_t i = order[start + j];
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double G[3][3];
    admm_grad_tile(tv, w, dw, inv_dx, lx, ly, lz, G);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) gs[(c * 3 + r) * S + i] = G[r][c];
  }
}

template <bool VISCO>
__global__ __launch_bounds__(256) void k_admm_flocal(ParticlesView P, AdmmParticles A,
                                                     const int32_t* __restrict__ mat,
                                                     const double* __restrict__ pparam,
                                                     const double* __restrict__ pstate, uint32_t n,
                                                     const double* __restrict__ gs, double dt, double rho,
                                                     double local_tol)
{
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const int64_t S = P.stride;
  double G[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) G[r][c] = gs[(c * 3 + r) * S + i];
  admm_local_step<VISCO>(P, A, mat, pparam, pstate, i, G, dt, rho, local_tol);
}

// dvStep rhs scatter (:167-205): dt rho omega^2 (F - Fn + y/(rho omega) - dt grad(vn) Fn) Fn^T grad w
// MODE 0; the operator's particle term c_p (grad(x) Fn) Fn^T grad w: MODE 1;
// VISCO adds the viscoelastic branch's terms with F2, y2, omega2 and FFn
template <int MODE, bool VISCO>
__global__ __launch_bounds__(256) void k_admm_scatter(ParticlesView P, AdmmParticles A, AdmmNodes N,
                                                      const double* __restrict__ xvec,
                                                      const uint32_t* __restrict__ order,
                                                      const uint32_t* __restrict__ group_start,
                                                      const uint32_t* __restrict__ group_page,
                                                      const int32_t* __restrict__ group_nslot,
                                                      const double* __restrict__ gm, double* stage, double dt,
                                                      double rho, double inv_dx, const int32_t* ai)
{
  if (MODE == 1 && ai[AI_DONE]) return;
  constexpr int SS = 3 * kTileNodes;
  __shared__ double tv[3 * kTileNodes];
  __shared__ double tile[4 * SS];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * SS; k += 256) tile[k] = 0.0;
  admm_tile3(group_nslot + 8 * (size_t)g, gm, MODE == 0 ? N.vn : xvec, nullptr, N.stride, tv, nslot);
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  double* wt = tile + (tid >> 6) * SS;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double G[3][3], Fn[3][3], Y[3][3];
    admm_grad_tile(tv, w, dw, inv_dx, lx, ly, lz, G);
    load_F(P.F, S, i, Fn);
    const double om = A.om[i];
    const double c2 = MODE == 0 ? dt * rho * om * om : rho * om * om * dt * dt;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double gf = G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c];
        if (MODE == 0)
          Y[r][c] = c2 * (A.F[(c * 3 + r) * S + i] - Fn[r][c] + A.y[(c * 3 + r) * S + i] / (om * rho) - dt * gf);
        else
          Y[r][c] = c2 * gf;
      }
    double Fv[3][3], Y2[3][3];
    if constexpr (VISCO) {
      load_F(A.fneq, S, i, Fv);
      const double om2 = A.om2[i];
      const double c22 = MODE == 0 ? dt * rho * om2 * om2 : rho * om2 * om2 * dt * dt;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double gf = G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c];
          if (MODE == 0)
            Y2[r][c] = c22 * (A.F2[(c * 3 + r) * S + i] - Fv[r][c] + A.y2[(c * 3 + r) * S + i] / (om2 * rho) - dt * gf);
          else
            Y2[r][c] = c22 * gf;
        }
    }
    opaque_weights(w, dw);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3], fg[3];
          wgrad(w, dw, inv_dx, a, bb, c, gw);
#pragma unroll
          for (int d = 0; d < 3; ++d) fg[d] = Fn[0][d] * gw[0] + Fn[1][d] * gw[1] + Fn[2][d] * gw[2];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          if constexpr (VISCO) {
            double fg2[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) fg2[d] = Fv[0][d] * gw[0] + Fv[1][d] * gw[1] + Fv[2][d] * gw[2];
#pragma unroll
            for (int r = 0; r < 3; ++r)
              atomicAdd(&wt[r * kTileNodes + t], (Y[r][0] * fg[0] + Y[r][1] * fg[1] + Y[r][2] * fg[2]) +
                                                     (Y2[r][0] * fg2[0] + Y2[r][1] * fg2[1] + Y2[r][2] * fg2[2]));
          } else {
#pragma unroll
            for (int r = 0; r < 3; ++r)
              atomicAdd(&wt[r * kTileNodes + t], Y[r][0] * fg[0] + Y[r][1] * fg[1] + Y[r][2] * fg[2]);
          }
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * SS;
  for (int k = tid; k < SS; k += 256) out[k] = tile[k] + tile[SS + k] + tile[2 * SS + k] + tile[3 * SS + k];
}

__device__ __forceinline__ void admm_project_node(const AdmmNodes& N, uint32_t k, double r[3])
{
  if (!N.coll[k]) return;
  const int64_t S = N.stride;
  double t[3];
  for (int i = 0; i < 3; ++i)
    t[i] = N.proj[(3 * i) * S + k] * r[0] + N.proj[(3 * i + 1) * S + k] * r[1] + N.proj[(3 * i + 2) * S + k] * r[2];
  r[0] = t[0];
  r[1] = t[1];
  r[2] = t[2];
}

// rhs = dt m g + staged sums (:167-173,206-209)
__global__ __launch_bounds__(256) void k_admm_rhs(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                  const double* __restrict__ stage, const double* __restrict__ gm,
                                                  double* rhs, int64_t S, double dt, double gx, double gy, double gz)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  double f[3] = { 0.0, 0.0, 0.0 };
  const double m = gm[k];
  if (m != 0.0) {
    const uint32_t e = k & 31u;
    sum_tiles_src<3>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
    f[0] += dt * m * gx;
    f[1] += dt * m * gy;
    f[2] += dt * m * gz;
  }
  for (int d = 0; d < 3; ++d) rhs[d * S + k] = f[d];
}

// MINRES start (Minres.h:40-101): q = P(b - A x), z = q / diag, partial z.q;
// m_k, m_{k-1} and q_k start at zero (Minres::reinitialize)
// (T: the precision the Krylov vectors are stored in; arithmetic in FP64)
template <class T>
__global__ __launch_bounds__(256) void k_minres_init(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                     const double* __restrict__ stage, const double* __restrict__ gm,
                                                     AdmmNodes N, const double* __restrict__ b,
                                                     const double* __restrict__ x, T* q, T* z, T* zero0, T* zero1,
                                                     T* zero2, double* partial)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const int64_t S = N.stride;
  double acc = 0.0;
  if (k < nn) {
    double r[3] = { 0.0, 0.0, 0.0 }, zz[3] = { 0.0, 0.0, 0.0 };
    const double m = gm[k];
    if (m != 0.0) {
      const uint32_t e = k & 31u;
      double f[3];
      sum_tiles_src<3>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
      for (int d = 0; d < 3; ++d) r[d] = b[d * S + k] - (m * x[d * S + k] + f[d]);
      admm_project_node(N, k, r);
      const double dg = N.diag[k];
      for (int d = 0; d < 3; ++d) {
        zz[d] = r[d] / dg;
        acc += zz[d] * r[d];
      }
    }
    for (int d = 0; d < 3; ++d) {
      q[d * S + k] = (T)r[d];
      z[d * S + k] = (T)zz[d];
      zero0[d * S + k] = (T)0;
      zero1[d * S + k] = (T)0;
      zero2[d * S + k] = (T)0;
    }
  }
  block_partial(acc, partial);
}

// q_{k+1} = P(A z) (z is the new m_k), partial of alpha = z . q
template <class T>
__global__ __launch_bounds__(256) void k_minres_apply(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                      const double* __restrict__ stage,
                                                      const double* __restrict__ gm, AdmmNodes N,
                                                      const T* __restrict__ z, T* q, const int32_t* ai,
                                                      double* partial)
{
  if (ai[AI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const int64_t S = N.stride;
  double acc = 0.0;
  if (k < nn) {
    double r[3] = { 0.0, 0.0, 0.0 };
    const double m = gm[k];
    if (m != 0.0) {
      const uint32_t e = k & 31u;
      double f[3];
      sum_tiles_src<3>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
      for (int d = 0; d < 3; ++d) r[d] = m * (double)z[d * S + k] + f[d];
      admm_project_node(N, k, r);
      for (int d = 0; d < 3; ++d) acc += (double)z[d * S + k] * r[d];
    }
    for (int d = 0; d < 3; ++d) q[d * S + k] = (T)r[d];
  }
  block_partial(acc, partial);
}

// q -= alpha q_k + beta_k q_{k-1}; z = q / diag; partial of z . q
template <class T>
__global__ __launch_bounds__(256) void k_minres_lanczos(uint32_t nn, const double* __restrict__ gm, AdmmNodes N,
                                                        T* q, const T* __restrict__ qk, const T* __restrict__ qkm1,
                                                        T* z, const double* sc, const int32_t* ai, double* partial)
{
  if (ai[AI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const int64_t S = N.stride;
  double acc = 0.0;
  if (k < nn && gm[k] != 0.0) {
    const double alpha = sc[AD_ALPHA], beta = sc[AD_BETA_K];
    const double dg = N.diag[k];
    for (int d = 0; d < 3; ++d) {
      double v = (double)q[d * S + k];
      v -= alpha * (double)qk[d * S + k];
      v -= beta * (double)qkm1[d * S + k];
      const T vt = (T)v, zt = (T)(v / dg);
      q[d * S + k] = vt;
      z[d * S + k] = zt;
      acc += (double)zt * (double)vt;
    }
  }
  block_partial(acc, partial);
}

// normalise q, z by beta_{k+1}; m_k = (z_old - delta m_{k-1} - eps m_{k-2}) / gamma; x += t_k m_k
template <class T>
__global__ __launch_bounds__(256) void k_minres_update(uint32_t nn, const double* __restrict__ gm, AdmmNodes N,
                                                       T* q, T* z, T* mk, const T* __restrict__ mkm1,
                                                       const T* __restrict__ mkm2, double* x, const double* sc,
                                                       const int32_t* ai)
{
  if (ai[AI_DONE + 2]) return;  // finished before this iteration's Givens step
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn || gm[k] == 0.0) return;
  const int64_t S = N.stride;
  const double bk = sc[AD_BETA_KP1], gamma = sc[AD_GAMMA], delta = sc[AD_DELTA], eps = sc[AD_EPS], tk = sc[AD_TK];
  for (int d = 0; d < 3; ++d) {
    if (bk > 0.0) {
      q[d * S + k] = (T)((double)q[d * S + k] / bk);
      z[d * S + k] = (T)((double)z[d * S + k] / bk);
    }
    const T m = (T)(((double)mk[d * S + k] - delta * (double)mkm1[d * S + k] - eps * (double)mkm2[d * S + k]) / gamma);
    mk[d * S + k] = m;
    x[d * S + k] += tk * (double)m;
  }
}

// the start vectors divided by beta_1 = |r0|_P (Minres.h:95-98)
template <class T>
__global__ __launch_bounds__(256) void k_minres_scale(uint32_t nn, const double* __restrict__ gm, int64_t S, T* q,
                                                      T* z, const double* sc, const int32_t* ai)
{
  if (ai[AI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const double b = sc[AD_BETA_KP1];
  if (k >= nn || gm[k] == 0.0 || !(b > 0.0)) return;
  for (int d = 0; d < 3; ++d) {
    q[d * S + k] = (T)((double)q[d * S + k] / b);
    z[d * S + k] = (T)((double)z[d * S + k] / b);
  }
}

// ---- the FP32-bulk global step (configuration 3: FP32 bulk, FP64 SVD and
// projection): the operator's per-particle data in FP32, in sorted order,
// formed once per step after initStep; the operator apply reads it
// coalesced, gathers and forms its terms in FP32, and adds them into the
// FP64 LDS tiles (FP32 LDS adds are far slower on this part); the Krylov
// vectors are FP32, their inner products and the solution dv FP64.
struct AdmmCache32 {
  float* frac;     // [3][n] x / dx - base node
  float* F;        // [9][n] Fn, column-major
  float* cp;       // [n] rho omega^2 dt^2
  float* F2;       // [9][n] FFn of the viscoelastic branch (VISCO)
  float* cp2;      // [n]
  uint8_t* cell;   // [n] element code of the base node in its page
  int64_t stride;  // n
};

template <bool VISCO>
__global__ __launch_bounds__(256) void k_admm_cache32(ParticlesView P, AdmmParticles A,
                                                      const uint32_t* __restrict__ order, uint32_t n, AdmmCache32 C,
                                                      double inv_dx, double rho, double dt)
{
  const uint32_t ks = blockIdx.x * 256u + threadIdx.x;
  if (ks >= n) return;
  const uint32_t i = order[ks];
  const int64_t S = P.stride, CS = C.stride;
  int b[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double xi = P.x[d * S + i] * inv_dx;
    b[d] = int_floor(xi - 0.5);
    C.frac[d * CS + ks] = (float)(xi - (double)b[d]);
  }
  C.cell[ks] = (uint8_t)elem_of(b[0], b[1], b[2]);
#pragma unroll
  for (int k = 0; k < 9; ++k) C.F[k * CS + ks] = (float)P.F[k * S + i];
  const double om = A.om[i];
  C.cp[ks] = (float)(rho * om * om * dt * dt);
  if constexpr (VISCO) {
#pragma unroll
    for (int k = 0; k < 9; ++k) C.F2[k * CS + ks] = (float)A.fneq[k * S + i];
    const double om2 = A.om2[i];
    C.cp2[ks] = (float)(rho * om2 * om2 * dt * dt);
  }
}

// the global operator's particle term in FP32 (k_admm_scatter<1>'s sums):
// c_p (grad(x) Fn) Fn^T grad w (+ the viscoelastic branch's term with FFn),
// x given in precision TX
template <bool VISCO, class TX>
__global__ __launch_bounds__(256) void k_admm_op32(AdmmCache32 C, const uint32_t* __restrict__ group_start,
                                                   const int32_t* __restrict__ group_nslot,
                                                   const double* __restrict__ gm, const TX* __restrict__ xvec,
                                                   int64_t NS, double* stage, float inv_dx, const int32_t* ai)
{
  if (ai[AI_DONE]) return;
  constexpr int SS = 3 * kTileNodes;
  __shared__ float tv[3 * kTileNodes];
  __shared__ double tile[4 * SS];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * SS; k += 256) tile[k] = 0.0;
  if (tid < 8) nslot[tid] = group_nslot[8 * (size_t)g + tid];
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += 256) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    const bool on = sl >= 0 && gm[node] != 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) tv[d * kTileNodes + t] = on ? (float)xvec[d * NS + node] : 0.0f;
  }
  __syncthreads();
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  const int64_t CS = C.stride;
  double* wt = tile + (tid >> 6) * SS;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t ks = start + j;
    const uint32_t e = C.cell[ks];
    const int lx = (int)(e >> 4), ly = (int)((e >> 2) & 3u), lz = (int)(e & 3u);
    float w[3][3], dw[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) bspline_frac(C.frac[d * CS + ks], w[d], dw[d]);
    float G[3][3] = { { 0.f } };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float gw[3];
          wgradT(w, dw, inv_dx, a, bb, c, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const float v = tv[r * kTileNodes + t];
            G[r][0] += v * gw[0];
            G[r][1] += v * gw[1];
            G[r][2] += v * gw[2];
          }
        }
    float Fn[3][3], Y[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) Fn[r][c] = C.F[(c * 3 + r) * CS + ks];
    const float cp = C.cp[ks];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) Y[r][c] = cp * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
    float Fv[3][3], Y2[3][3];
    if constexpr (VISCO) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) Fv[r][c] = C.F2[(c * 3 + r) * CS + ks];
      const float cp2 = C.cp2[ks];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Y2[r][c] = cp2 * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
    }
    opaque_weights(w, dw);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          float gw[3], fg[3];
          wgradT(w, dw, inv_dx, a, bb, c, gw);
#pragma unroll
          for (int d = 0; d < 3; ++d) fg[d] = Fn[0][d] * gw[0] + Fn[1][d] * gw[1] + Fn[2][d] * gw[2];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            float val = Y[r][0] * fg[0] + Y[r][1] * fg[1] + Y[r][2] * fg[2];
            if constexpr (VISCO) {
              float fg2[3];
#pragma unroll
              for (int d = 0; d < 3; ++d) fg2[d] = Fv[0][d] * gw[0] + Fv[1][d] * gw[1] + Fv[2][d] * gw[2];
              val += Y2[r][0] * fg2[0] + Y2[r][1] * fg2[1] + Y2[r][2] * fg2[2];
            }
            atomicAdd(&wt[r * kTileNodes + t], (double)val);
          }
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * SS;
  for (int k = tid; k < SS; k += 256) out[k] = tile[k] + tile[SS + k] + tile[2 * SS + k] + tile[3 * SS + k];
}


// uStep (:638-671): y += rho omega (F - Fn - dt grad(vn + dv) Fn); VISCO
// also y2 += rho omega2 (F2 - FFn - dt grad(vn + dv) FFn)
template <bool VISCO>
__global__ __launch_bounds__(256) void k_admm_ustep(ParticlesView P, AdmmParticles A, AdmmNodes N,
                                                    const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ group_start,
                                                    const uint32_t* __restrict__ group_page,
                                                    const int32_t* __restrict__ group_nslot,
                                                    const double* __restrict__ gm, double dt, double rho,
                                                    double inv_dx)
{
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  const uint32_t g = blockIdx.x;
  admm_tile3(group_nslot + 8 * (size_t)g, gm, N.vn, N.dv, N.stride, tv, nslot);
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    const uint32_t i = order[start + j];
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double G[3][3], Fn[3][3];
    admm_grad_tile(tv, w, dw, inv_dx, lx, ly, lz, G);
    load_F(P.F, S, i, Fn);
    const double om = A.om[i];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double t = A.F[(c * 3 + r) * S + i] - Fn[r][c] -
                         dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
        A.y[(c * 3 + r) * S + i] += rho * om * t;
      }
    if constexpr (VISCO) {
      double Fv[3][3];
      load_F(A.fneq, S, i, Fv);
      const double om2 = A.om2[i];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double t = A.F2[(c * 3 + r) * S + i] - Fv[r][c] -
                           dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
          A.y2[(c * 3 + r) * S + i] += rho * om2 * t;
        }
    }
  }
}

// warm-start scatter of last dv (P2G of last_dv w and w, 4 staged fields)
__global__ __launch_bounds__(256) void k_admm_lastdv_p2g(ParticlesView P, AdmmParticles A,
                                                         const uint32_t* __restrict__ order,
                                                         const uint32_t* __restrict__ group_start,
                                                         const uint32_t* __restrict__ group_page, double* stage,
                                                         double inv_dx)
{
  constexpr int SS = 4 * kTileNodes;
  __shared__ double tile[4 * SS];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * SS; k += 256) tile[k] = 0.0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  __syncthreads();
  double* wt = tile + (tid >> 6) * SS;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const int64_t S = P.stride;
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    double w[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_w(P.x[d * S + i] * inv_dx, w[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    const double ld[3] = { A.last[i], A.last[S + i], A.last[2 * S + i] };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double wt_ = w[0][a] * w[1][bb] * w[2][c];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          atomicAdd(&wt[t], ld[0] * wt_);
          atomicAdd(&wt[kTileNodes + t], ld[1] * wt_);
          atomicAdd(&wt[2 * kTileNodes + t], ld[2] * wt_);
          atomicAdd(&wt[3 * kTileNodes + t], wt_);
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * SS;
  for (int k = tid; k < SS; k += 256) out[k] = tile[k] + tile[SS + k] + tile[2 * SS + k] + tile[3 * SS + k];
}

// end of admmStep (:789-800): new v = vn + dv on the grid, last_dv per particle
__global__ __launch_bounds__(256) void k_admm_finish_nodes(uint32_t nn, const double* __restrict__ gm, AdmmNodes N,
                                                           double* gvs)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const int64_t S = N.stride;
  const bool on = gm[k] != 0.0;
  for (int d = 0; d < 3; ++d) gvs[d * S + k] = on ? N.vn[d * S + k] + N.dv[d * S + k] : 0.0;
}

__global__ __launch_bounds__(256) void k_admm_lastdv(ParticlesView P, AdmmParticles A, AdmmNodes N,
                                                     const uint32_t* __restrict__ order,
                                                     const uint32_t* __restrict__ group_start,
                                                     const uint32_t* __restrict__ group_page,
                                                     const int32_t* __restrict__ group_nslot,
                                                     const double* __restrict__ gm, double dt, double inv_dx)
{
  __shared__ double tv[3 * kTileNodes];
  __shared__ double tw[kTileNodes];
  __shared__ int32_t nslot[8];
  const uint32_t g = blockIdx.x;
  admm_tile3(group_nslot + 8 * (size_t)g, gm, N.dv, nullptr, N.stride, tv, nslot);
  for (int t = threadIdx.x; t < kTileNodes; t += 256) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    tw[t] = (sl >= 0 && gm[(int64_t)sl * 32 + elem_of(tx, ty, tz)] != 0.0) ? 1.0 : 0.0;
  }
  __syncthreads();
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    const uint32_t i = order[start + j];
    double w[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_w(P.x[d * S + i] * inv_dx, w[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double acc[3] = { 0.0, 0.0, 0.0 }, ws = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          const double wt_ = w[0][a] * w[1][bb] * w[2][c] * tw[t];
          acc[0] += tv[t] * wt_;
          acc[1] += tv[kTileNodes + t] * wt_;
          acc[2] += tv[2 * kTileNodes + t] * wt_;
          ws += wt_;
        }
    for (int d = 0; d < 3; ++d) A.last[d * S + i] = acc[d] / (ws * dt);
  }
}

__global__ void k_admm_reset_flags(int32_t* ai)
{
  if (threadIdx.x < AI_N) ai[threadIdx.x] = 0;
}
