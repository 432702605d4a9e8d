// zr_adjoint.hpp -- reverse of the explicit MLS step (the differentiable
// form of the engine; no counterpart in the reference).  Included by
// zr_engine.hip inside its anonymous namespace after zr_damage.hpp.
//
// The forward step from particle state (x, v, C, F) with fixed-corotated
// (mu, lambda) is  tau = tau(F), P2G (m, p, f), v* = collide(p/m + dt g +
// dt f/m), G2P (v', C' = D^-1 sum w v* d^T, x' = x + dt v', F' = (I + dt C') F).
// Given the adjoint of (x', v', C', F') the three kernels below produce the
// adjoint of (x, v, C, F) and accumulate those of mu and lambda, holding the
// grid topology of the step fixed (weights are differentiated in x):
//   k_adj_g2p : F and C' terms, v*-adjoint scattered to per-wave LDS tiles
//               and staged per group, x terms of the G2P weights;
//   k_adj_grid: v*-adjoint summed per node, through the collider Jacobian,
//               to the adjoints of m, p, f (7 values per node);
//   k_adj_p2g : gathers them per particle: x, v, C and the stress adjoint,
//               then tau's reverse into F, mu, lambda.
// Node sums are the fixed-order staged-tile sums, so results are
// deterministic.

constexpr int kAdjGridFields = 7;  // m, p[3], f[3]

struct AdjointView {
  double* x;     // [3][n] adjoint of x (output on entry, input on exit)
  double* v;     // [3][n]
  double* C;     // [9][n]
  double* F;     // [9][n]
  double* mu;    // [n] accumulated
  double* lam;   // [n]
  double* xg;    // [3][n] x-adjoint of the G2P part (scratch)
  double* Fg;    // [9][n] F-adjoint of the F update (scratch)
  double* vt;    // [9][n] V0 tau of the step's input (scratch)
  double* tb;    // [9][n] adjoint of V0 tau (scratch)
};

// d v_after / d v_before of the node's collider handling: multi_object_collision
// in the frame at t, then trial_collision in the frame at t + dt when nothing
// collided, for objects moving as x = R s X + b.  The object's velocity at
// the node does not depend on v, so each collider acts on v - v_obj through
// its world normal n = R N: sticky gives 0, a slipping or approaching
// contact P = I - n n^T; Coulomb friction on the sliding branch (|t| >
// -mu dot, t the tangential part) scales t by k = 1 + mu dot / |t|, with
// d t' / d rel = k P + t (mu n / |t| - mu dot t / |t|^3)^T, and the sticking
// branch gives 0.  The trial collision v -= (phi / dt) n at x + dt v has
// I - n n^T - phi grad n, grad n = P / (phi + s r) for a sphere of radius
// s r.  Returns false for a node at a sphere's centre in the trial collision.
__device__ bool collide_jacobian(const ColliderSet& cs, double dt, const double x[3], const double vpre[3],
                                 double J[3][3])
{
  double v[3] = { vpre[0], vpre[1], vpre[2] };
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[i][j] = i == j ? 1.0 : 0.0;
  bool any = false;
  for (int k = 0; k < cs.n; ++k) {
    const DevCollider& c = cs.c[k];
    double xb[3], phi, N[3], vo[3];
    if (!collider_query(c, 0, x, phi, N, xb)) continue;
    any = true;
    if (c.type == ZR_STICKY) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] = 0.0;
      return true;
    }
    collider_velocity(c, 0, xb, vo);
    double rel[3], dot = 0.0;
    for (int d = 0; d < 3; ++d) rel[d] = v[d] - vo[d];
    for (int d = 0; d < 3; ++d) dot += rel[d] * N[d];
    if (c.type == ZR_SLIP || dot < 0.0) {
      double P[3][3], Jr[3][3], T[3][3], t[3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          P[i][j] = (i == j ? 1.0 : 0.0) - N[i] * N[j];
          Jr[i][j] = P[i][j];
        }
      for (int d = 0; d < 3; ++d) t[d] = rel[d] - N[d] * dot;
      if (c.friction != 0.0 && dot < 0.0) {
        const double mu = c.friction;
        const double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (-dot * mu < tn) {
          const double kf = 1.0 + mu * dot / tn;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Jr[i][j] = kf * P[i][j] + t[i] * (mu * N[j] / tn - mu * dot * t[j] / (tn * tn * tn));
          for (int d = 0; d < 3; ++d) t[d] *= kf;
        } else {
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Jr[i][j] = 0.0;
          t[0] = t[1] = t[2] = 0.0;
        }
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i][j] = Jr[i][0] * J[0][j] + Jr[i][1] * J[1][j] + Jr[i][2] * J[2][j];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] = T[i][j];
      for (int d = 0; d < 3; ++d) rel[d] = t[d];
    }
    for (int d = 0; d < 3; ++d) v[d] = rel[d] + vo[d];
  }
  if (any) return true;
  for (int k = 0; k < cs.n; ++k) {
    const DevCollider& c = cs.c[k];
    double xt[3], xb[3], phi, n[3];
    for (int d = 0; d < 3; ++d) xt[d] = x[d] + dt * v[d];
    if (!collider_query(c, 1, xt, phi, n, xb)) continue;
    double A[3][3], T[3][3];
    if (c.shape == ZR_HALF_SPACE) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = (i == j ? 1.0 : 0.0) - n[i] * n[j];
    } else {
      const double dist = phi + c.f[1].s * c.radius;
      if (dist < 1e-7 * c.f[1].s) return false;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double pr = (i == j ? 1.0 : 0.0) - n[i] * n[j];
          A[i][j] = pr - phi / dist * pr;
        }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[i][j] = A[i][0] * J[0][j] + A[i][1] * J[1][j] + A[i][2] * J[2][j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] = T[i][j];
    for (int d = 0; d < 3; ++d) v[d] -= (phi / dt) * n[d];
  }
  return true;
}

// reverse of tau = 2 mu (F - R) F^T + lam (J - 1) J I from the SVD of F:
// adds (dtau/dF)^T taubar to Fb, returns mubar and lambar
__device__ __forceinline__ void fcr_kirchhoff_adjoint(double mu, double lam, const M3<double>& f,
                                                      const M3<double>& u, const double s[3], const M3<double>& v,
                                                      const double tb[3][3], double Fb[3][3], double& mub,
                                                      double& lab)
{
  double r[3][3], A[3][3], Ah[3][3], Mb[3][3], c[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r[i][j] = u.a[i][0] * v.a[j][0] + u.a[i][1] * v.a[j][1] + u.a[i][2] * v.a[j][2];
  const double J = det3(f);
  c[0][0] = f.a[1][1] * f.a[2][2] - f.a[1][2] * f.a[2][1];
  c[0][1] = f.a[1][2] * f.a[2][0] - f.a[1][0] * f.a[2][2];
  c[0][2] = f.a[1][0] * f.a[2][1] - f.a[1][1] * f.a[2][0];
  c[1][0] = f.a[0][2] * f.a[2][1] - f.a[0][1] * f.a[2][2];
  c[1][1] = f.a[0][0] * f.a[2][2] - f.a[0][2] * f.a[2][0];
  c[1][2] = f.a[0][1] * f.a[2][0] - f.a[0][0] * f.a[2][1];
  c[2][0] = f.a[0][1] * f.a[1][2] - f.a[0][2] * f.a[1][1];
  c[2][1] = f.a[0][2] * f.a[1][0] - f.a[0][0] * f.a[1][2];
  c[2][2] = f.a[0][0] * f.a[1][1] - f.a[0][1] * f.a[1][0];
  const double trtb = tb[0][0] + tb[1][1] + tb[2][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = tb[i][0] * f.a[0][j] + tb[i][1] * f.a[1][j] + tb[i][2] * f.a[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) acc += u.a[k][i] * A[k][l] * v.a[l][j];
      Ah[i][j] = acc;
    }
#pragma unroll
  for (int i = 0; i < 3; ++i) Mb[i][i] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i + 1; j < 3; ++j) {
      double den = s[i] + s[j];
      if (fabs(den) < 1e-12) den = den < 0.0 ? -1e-12 : 1e-12;
      const double x = (Ah[i][j] - Ah[j][i]) / den;
      Mb[i][j] = x;
      Mb[j][i] = -x;
    }
  double mb = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double t1 = 0.0, t2 = 0.0, t3 = 0.0, fr = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        t1 += (tb[i][k] + tb[k][i]) * f.a[k][j];
        t2 += tb[k][i] * r[k][j];
        t3 += u.a[i][k] * (Mb[k][0] * v.a[j][0] + Mb[k][1] * v.a[j][1] + Mb[k][2] * v.a[j][2]);
        fr += (f.a[i][k] - r[i][k]) * f.a[j][k];
      }
      Fb[i][j] += 2.0 * mu * (t1 - t2 - t3) + lam * (2.0 * J - 1.0) * trtb * c[i][j];
      mb += tb[i][j] * 2.0 * fr;
    }
  mub = mb;
  lab = trtb * (J - 1.0) * J;
}

// G2P + F update reverse (per group)
__global__ __launch_bounds__(256) void k_adj_g2p(ParticlesView P, AdjointView A, const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ group_start,
                                                 const uint32_t* __restrict__ group_page,
                                                 const int32_t* __restrict__ group_nslot,
                                                 const double* __restrict__ gvs, int64_t gstride, double* stage,
                                                 double dx, double inv_dx, double Dinv, double dt)
{
  constexpr int SS = 3 * kTileNodes;
  __shared__ double tv[3 * kTileNodes];
  __shared__ double tile[4 * SS];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  if (tid < 8) nslot[tid] = group_nslot[8 * (size_t)g + tid];
  for (int k = tid; k < 4 * SS; k += 256) tile[k] = 0.0;
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += 256) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    for (int d = 0; d < 3; ++d) tv[d * kTileNodes + t] = sl >= 0 ? gvs[d * gstride + node] : 0.0;
  }
  __syncthreads();
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
    // forward C' = D^-1 sum w v* d^T
    double B[3][3] = { { 0.0 } };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double d0 = (double)(b[0] + a) * dx - xp[0];
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const double wab = w[0][a] * w[1][bb];
        const double d1 = (double)(b[1] + bb) * dx - xp[1];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double wt_ = wab * w[2][c];
          const double d2 = (double)(b[2] + c) * dx - xp[2];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double gv = tv[r * kTileNodes + t];
            B[r][0] += wt_ * gv * d0;
            B[r][1] += wt_ * gv * d1;
            B[r][2] += wt_ * gv * d2;
          }
        }
      }
    }
    // F' = F + dt C' F reverse; C'bar += dt F'bar F^T
    double Bb[3][3];
    {
      double Fo[3][3], Fb[3][3], Cn[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          Fo[r][c] = P.F[(c * 3 + r) * S + i];
          Fb[r][c] = A.F[(c * 3 + r) * S + i];
          Cn[r][c] = B[r][c] * Dinv;
        }
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double acc = Fb[r][c];
#pragma unroll
          for (int k = 0; k < 3; ++k) acc += dt * Cn[k][r] * Fb[k][c];
          A.Fg[(c * 3 + r) * S + i] = acc;
          double cb = 0.0;
#pragma unroll
          for (int k = 0; k < 3; ++k) cb += Fb[r][k] * Fo[c][k];
          Bb[r][c] = Dinv * (A.C[(c * 3 + r) * S + i] + dt * cb);
        }
    }
    double vb[3], xb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xb[d] = A.x[d * S + i];
      vb[d] = A.v[d * S + i] + dt * xb[d];
    }
    opaque_weights(w, dw);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double d0 = (double)(b[0] + a) * dx - xp[0];
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const double d1 = (double)(b[1] + bb) * dx - xp[1];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double d2 = (double)(b[2] + c) * dx - xp[2];
          double gw[3];
          const double wt_ = wgrad(w, dw, inv_dx, a, bb, c, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          double sc = 0.0, btv[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double vs = tv[r * kTileNodes + t];
            const double e = vb[r] + Bb[r][0] * d0 + Bb[r][1] * d1 + Bb[r][2] * d2;
            atomicAdd(&wt[r * kTileNodes + t], wt_ * e);
            sc += vs * e;
            btv[0] += Bb[r][0] * vs;
            btv[1] += Bb[r][1] * vs;
            btv[2] += Bb[r][2] * vs;
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) xb[d] += gw[d] * sc - wt_ * btv[d];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) A.xg[d * S + i] = xb[d];
  }
  __syncthreads();
  double* out = stage + (size_t)g * SS;
  for (int k = tid; k < SS; k += 256) out[k] = tile[k] + tile[SS + k] + tile[2 * SS + k] + tile[3 * SS + k];
}

// grid update reverse (per node): adjoints of m, p, f; counts nodes whose
// collider branch has no derivative
__global__ __launch_bounds__(256) void k_adj_grid(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                                  const int32_t* __restrict__ slot_src,
                                                  const double* __restrict__ stage, GridView G, double* ag,
                                                  double dx, double dt, double gx, double gy, double gz,
                                                  ColliderSet cs, uint32_t* bad)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const int64_t S = G.stride;
  const double m = G.m[k];
  double out[kAdjGridFields] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
  if (m != 0.0) {
    const NodeCoord nc = node_coord(slot_page, k);
    double avs[3];
    sum_tiles_src<3>(stage, slot_src + 8 * (size_t)(k >> 5), nc.ex, nc.ey, nc.ez, avs);
    double xi[3];
    node_pos(nc, dx, xi);
    const double v[3] = { G.v[k], G.v[S + k], G.v[2 * S + k] };
    const double f[3] = { G.f[k], G.f[S + k], G.f[2 * S + k] };
    const double dt_over_mi = dt / m;
    const double vpre[3] = { v[0] + gx * dt + f[0] * dt_over_mi, v[1] + gy * dt + f[1] * dt_over_mi,
                             v[2] + gz * dt + f[2] * dt_over_mi };
    double J[3][3];
    if (!collide_jacobian(cs, dt, xi, vpre, J)) atomicAdd(bad, 1u);
    double mb = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double vb = J[0][d] * avs[0] + J[1][d] * avs[1] + J[2][d] * avs[2];
      out[1 + d] = vb / m;
      out[4 + d] = dt * vb / m;
      mb -= (v[d] + f[d] * dt_over_mi) * vb / m;
    }
    out[0] = mb;
  }
#pragma unroll
  for (int q = 0; q < kAdjGridFields; ++q) ag[q * S + k] = out[q];
}

// P2G + stress reverse in three passes, so that the pass holding the node
// tile holds no SVD (one kernel of 256 VGPRs + 134 AGPRs ran at one wave
// per SIMD): k_adj_stress writes V0 tau per particle, k_adj_p2g reverses
// the momentum and stress transfer over the tile, k_adj_stress_rev reverses
// the stress (the SVD again) into the F, mu and lambda adjoints.

// forward V0 tau = V0 (2 mu (F - R) F^T + lam (J - 1) J I), per particle
__global__ __launch_bounds__(256) void k_adj_stress(ParticlesView P, AdjointView A, uint32_t n)
{
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const int64_t S = P.stride;
  M3<double> f, u, v;
  double s[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) f.a[r][c] = P.F[(c * 3 + r) * S + i];
  svd3(f, u, s, v);
  const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i];
  const double J = det3(f);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double rk = u.a[r][0] * v.a[k][0] + u.a[r][1] * v.a[k][1] + u.a[r][2] * v.a[k][2];
        acc += (f.a[r][k] - rk) * f.a[c][k];
      }
      A.vt[(c * 3 + r) * S + i] = vol * (2.0 * mu * acc + (r == c ? lam * (J - 1.0) * J : 0.0));
    }
}

// the transfer's reverse over the group's node tile (per group): x, v and C
// adjoints, and the adjoint of V0 tau into A.tb
__global__ __launch_bounds__(256) void k_adj_p2g(ParticlesView P, AdjointView A, const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ group_start,
                                                 const uint32_t* __restrict__ group_page,
                                                 const int32_t* __restrict__ group_nslot,
                                                 const double* __restrict__ ag, int64_t gstride, double dx,
                                                 double inv_dx, double Dinv)
{
  __shared__ double ta[kAdjGridFields * kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  if (tid < 8) nslot[tid] = group_nslot[8 * (size_t)g + tid];
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += 256) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    for (int q = 0; q < kAdjGridFields; ++q) ta[q * kTileNodes + t] = sl >= 0 ? ag[q * gstride + node] : 0.0;
  }
  __syncthreads();
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  for (uint32_t j = tid; j < n; j += 256) {
    const uint32_t i = order[start + j];
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    const double mass = P.m[i];
    double vt[3][3], Cm[3][3], vp[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        vt[r][c] = A.vt[(c * 3 + r) * S + i];
        Cm[r][c] = P.C[(c * 3 + r) * S + i];
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) vp[d] = P.v[d * S + i];
    double Cb[3][3] = { { 0.0 } }, tb[3][3] = { { 0.0 } }, vbar[3] = { 0.0, 0.0, 0.0 }, xb[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double d0 = (double)(b[0] + a) * dx - xp[0];
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const double d1 = (double)(b[1] + bb) * dx - xp[1];
        // the row's tile address, opaque so that the row's reads are not
        // hoisted ahead of the previous rows' sums (register budget)
        int row = ((lx + a) * kTileY + (ly + bb)) * kTileZ + lz;
        asm volatile("" : "+v"(row));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double d2 = (double)(b[2] + c) * dx - xp[2];
          double gw[3];
          const double wt_ = wgrad(w, dw, inv_dx, a, bb, c, gw);
          const int t = row + c;
          const double am = ta[t];
          const double ap[3] = { ta[kTileNodes + t], ta[2 * kTileNodes + t], ta[3 * kTileNodes + t] };
          const double af[3] = { ta[4 * kTileNodes + t], ta[5 * kTileNodes + t], ta[6 * kTileNodes + t] };
          double spp = 0.0, sf = 0.0;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double mom = mass * vp[r] + mass * (Cm[r][0] * d0 + Cm[r][1] * d1 + Cm[r][2] * d2);
            const double td = vt[r][0] * d0 + vt[r][1] * d1 + vt[r][2] * d2;
            spp += ap[r] * mom;
            sf += af[r] * td;
            vbar[r] += wt_ * mass * ap[r];
            Cb[r][0] += wt_ * mass * ap[r] * d0;
            Cb[r][1] += wt_ * mass * ap[r] * d1;
            Cb[r][2] += wt_ * mass * ap[r] * d2;
            tb[r][0] -= wt_ * Dinv * af[r] * d0;
            tb[r][1] -= wt_ * Dinv * af[r] * d1;
            tb[r][2] -= wt_ * Dinv * af[r] * d2;
          }
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double ctp = Cm[0][d] * ap[0] + Cm[1][d] * ap[1] + Cm[2][d] * ap[2];
            const double vtf = vt[0][d] * af[0] + vt[1][d] * af[1] + vt[2][d] * af[2];
            xb[d] += gw[d] * (am * mass + spp - Dinv * sf) - wt_ * mass * ctp + wt_ * Dinv * vtf;
          }
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      A.x[d * S + i] = A.xg[d * S + i] + xb[d];
      A.v[d * S + i] = vbar[d];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        A.C[(c * 3 + r) * S + i] = Cb[r][c];
        A.tb[(c * 3 + r) * S + i] = tb[r][c];
      }
  }
}

// the stress's reverse per particle: F adjoint = that of the F update plus
// (d V0 tau / dF)^T tb, and the mu and lambda adjoints accumulated
__global__ __launch_bounds__(256) void k_adj_stress_rev(ParticlesView P, AdjointView A, uint32_t n)
{
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const int64_t S = P.stride;
  M3<double> f, u, v;
  double s[3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) f.a[r][c] = P.F[(c * 3 + r) * S + i];
  svd3(f, u, s, v);
  const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i];
  double Fb[3][3], tb[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      Fb[r][c] = A.Fg[(c * 3 + r) * S + i];
      tb[r][c] = A.tb[(c * 3 + r) * S + i] * vol;
    }
  double mub, lab;
  fcr_kirchhoff_adjoint(mu, lam, f, u, s, v, tb, Fb, mub, lab);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) A.F[(c * 3 + r) * S + i] = Fb[r][c];
  A.mu[i] += mub;
  A.lam[i] += lab;
}
