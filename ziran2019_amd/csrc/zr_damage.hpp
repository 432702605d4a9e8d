// zr_damage.hpp -- CD-MPM phase-field damage and the naive Rankine damage
// (Projects/fracture/FractureSimulation.h:209-554, PhaseField.h:70-76) for
// the gfx950 engine.  Included by zr_engine.hip inside its anonymous
// namespace, after zr_implicit.hpp.
//
// The phase-field system is solved on the active nodes (m != 0) of the
// current sort with the fracture project's preconditioned CG
// (Projects/fracture/ConjugateGradient.h:50-97), lumped form:
//   A x = D x + Mg sum_p kappa_p (grad x_p) . grad w_ip,
//   D_i = sum_p vol_p (M pf_Fp_p + 1/dt) w_ip   (M > 0; vol_p pf_Fp_p w_ip otherwise),
//   kappa_p = 4 l0_p^2 vol_p, Mg = M (M > 0; 1 otherwise), vol_p = V0_p det F_p,
//   rhs_i = (M + phi_i/dt) sum_p vol_p w_ip  (sum_p vol_p w_ip when M <= 0),
//   phi_i = sum_p c_p w_ip / sum_p w_ip.
// The lumped term of the operator is D_i x_i, so only the gradient term needs
// a particle pass.  Every scalar of CG lives on the device; all node sums are
// the fixed-order staged-tile sums of P2G and every dot product is a
// deterministic two-level reduction.

enum {
  PF_RHO = 8,      // mr . r of the current iteration
  PF_SQ = 9,       // s . q
  PF_ALPHA = 10,
  PF_BETA = 11,    // rho / rho_old for the next direction
  PF_NORM = 12,    // |r|
  PF_REL = 13,     // relative tolerance, cg_relative |r0|
};
enum { CI_PF_IT = 2 };

constexpr int kPfFields = 4;  // sum w, sum c w, sum vol w, sum D_p w

struct PfParticles {
  double* c;
  double* H;
  const double* l0;
  const double* osc;   // one_over_sigma_c
  double* Fp;          // pf_Fp
  const double* Hmax;
  const double* k;     // residual_phase
  const int32_t* allow;
  double* vol;         // V0 det F of this solve
};

struct PfScalars {
  double M, dt, tol;
  int32_t max_it;
};

__device__ __forceinline__ double det_soa(const double* F, int64_t S, uint32_t i)
{
  const double f00 = F[i], f10 = F[S + i], f20 = F[2 * S + i];
  const double f01 = F[3 * S + i], f11 = F[4 * S + i], f21 = F[5 * S + i];
  const double f02 = F[6 * S + i], f12 = F[7 * S + i], f22 = F[8 * S + i];
  return f00 * (f11 * f22 - f21 * f12) - f01 * (f10 * f22 - f20 * f12) + f02 * (f10 * f21 - f20 * f11);
}

// phaseP2G (:209-257) with the rhs and diagonal sums (:303-325, :405-437):
// per-wave LDS tiles, staged per group like P2G
__global__ __launch_bounds__(256) void k_pf_p2g(ParticlesView P, PfParticles pf, PfScalars ps,
                                                const uint32_t* __restrict__ order,
                                                const uint32_t* __restrict__ group_start,
                                                const uint32_t* __restrict__ group_page, double* __restrict__ stage,
                                                double inv_dx)
{
  constexpr int SS = kPfFields * kTileNodes;
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
  const double M = ps.M;
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    const double vol = P.vol[i] * det_soa(P.F, S, i);
    pf.vol[i] = vol;
    const double dg = M > 0.0 ? vol * (M * pf.Fp[i] + 1.0 / ps.dt) : vol * pf.Fp[i];
    const double c = pf.c[i];
    double w[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_w(xp[d] * inv_dx, w[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const double wab = w[0][a] * w[1][bb];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          const double wt_ = wab * w[2][cc];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + cc);
          if (M > 0.0) {
            atomicAdd(&wt[t], wt_);
            atomicAdd(&wt[kTileNodes + t], c * wt_);
          }
          atomicAdd(&wt[2 * kTileNodes + t], vol * wt_);
          atomicAdd(&wt[3 * kTileNodes + t], dg * wt_);
        }
      }
  }
  __syncthreads();
  double* out = stage + (size_t)g * SS;
  for (int k = tid; k < SS; k += 256) out[k] = tile[k] + tile[SS + k] + tile[2 * SS + k] + tile[3 * SS + k];
}

// node pass: phi, rhs, D, then the CG start (fracture/ConjugateGradient.h:58-65):
// x = 0, r = rhs, mr = r / D; partials of mr.r (rho) and r.r
__global__ __launch_bounds__(256) void k_pf_nodes(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                  const double* __restrict__ stage, const double* __restrict__ gm,
                                                  PfScalars ps, double* phase, double* rhs, double* diag, double* x,
                                                  double* r, double* mr, double* partial, uint32_t nb)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double rho = 0.0, rr = 0.0;
  if (k < nn) {
    double ph = 0.0, b = 0.0, D = 0.0, z = 0.0;
    if (gm[k] != 0.0) {
      const uint32_t e = k & 31u;
      double f[kPfFields];
      sum_tiles_src<kPfFields>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u),
                               (int)(e & 3u), f);
      if (ps.M > 0.0) {
        ph = f[1] / f[0];
        b = (ps.M + ph / ps.dt) * f[2];
      } else {
        b = f[2];
      }
      D = f[3];
      z = b / D;
      rho = z * b;
      rr = b * b;
    }
    phase[k] = ph;
    rhs[k] = b;
    diag[k] = D;
    x[k] = 0.0;
    r[k] = b;
    mr[k] = z;
  }
  __shared__ double smem[4];
  const double s1 = block_sum256(rho, smem);
  __syncthreads();
  const double s2 = block_sum256(rr, smem);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s1;
    partial[nb + blockIdx.x] = s2;
  }
}

// the gradient term of the operator on the direction s = mr + beta s
// (:338-403), built on the fly in the LDS tile; one scalar field staged per group
__global__ __launch_bounds__(256) void k_pf_apply(const double* __restrict__ X, int64_t S, PfParticles pf,
                                                  PfScalars ps, const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ group_start,
                                                  const uint32_t* __restrict__ group_page,
                                                  const int32_t* __restrict__ group_nslot,
                                                  const double* __restrict__ gm, const double* __restrict__ mr,
                                                  const double* __restrict__ s, const double* sc,
                                                  const int32_t* ci, double* __restrict__ stage, double inv_dx)
{
  if (ci[CI_DONE]) return;
  __shared__ double tile[4 * kTileNodes];
  __shared__ double tv[kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  const bool first = ci[CI_PF_IT] == 0;  // s = mr on the first iteration
  const double beta = first ? 0.0 : sc[PF_BETA];
  if (tid < 8) nslot[tid] = group_nslot[8 * (size_t)g + tid];
  for (int k = tid; k < 4 * kTileNodes; k += 256) tile[k] = 0.0;
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += 256) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    tv[t] = (sl >= 0 && gm[node] != 0.0) ? (first ? mr[node] : mr[node] + beta * s[node]) : 0.0;
  }
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  __syncthreads();
  double* wt = tile + (tid >> 6) * kTileNodes;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const double gscale = ps.M > 0.0 ? ps.M : 1.0;
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const double xp[3] = { X[i], X[S + i], X[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double gc[3] = { 0.0, 0.0, 0.0 };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          double gw[3];
          wgrad(w, dw, inv_dx, a, bb, cc, gw);
          const double sv = tv[((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + cc)];
          gc[0] += sv * gw[0];
          gc[1] += sv * gw[1];
          gc[2] += sv * gw[2];
        }
    const double l0 = pf.l0[i];
    const double kap = pf.vol[i] * 4.0 * l0 * l0 * gscale;
#pragma unroll
    for (int d = 0; d < 3; ++d) gc[d] *= kap;
    opaque_weights(w, dw);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          double gw[3];
          wgrad(w, dw, inv_dx, a, bb, cc, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + cc);
          atomicAdd(&wt[t], gc[0] * gw[0] + gc[1] * gw[1] + gc[2] * gw[2]);
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * kTileNodes;
  for (int k = tid; k < kTileNodes; k += 256)
    out[k] = tile[k] + tile[kTileNodes + k] + tile[2 * kTileNodes + k] + tile[3 * kTileNodes + k];
}

// s = mr + beta s, q = D s + (staged gradient term); partials of s.q
__global__ __launch_bounds__(256) void k_pf_ap(uint32_t nn, const int32_t* __restrict__ slot_src,
                                               const double* __restrict__ stage, const double* __restrict__ gm,
                                               const double* __restrict__ diag, const double* __restrict__ mr,
                                               double* s, double* q, const double* sc, const int32_t* ci,
                                               double* partial)
{
  if (ci[CI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const bool first = ci[CI_PF_IT] == 0;
  const double beta = first ? 0.0 : sc[PF_BETA];
  double sq = 0.0;
  if (k < nn) {
    double sv = 0.0, qv = 0.0;
    if (gm[k] != 0.0) {
      const uint32_t e = k & 31u;
      double f[1];
      sum_tiles_src<1>(stage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u),
                       (int)(e & 3u), f);
      sv = first ? mr[k] : mr[k] + beta * s[k];
      qv = diag[k] * sv + f[0];
      sq = sv * qv;
    }
    s[k] = sv;
    q[k] = qv;
  }
  block_partial(sq, partial);
}

// x += alpha s, r -= alpha q, mr = r / D; partials of r.r and mr.r
__global__ __launch_bounds__(256) void k_pf_update(uint32_t nn, const double* __restrict__ gm,
                                                   const double* __restrict__ diag, const double* __restrict__ s,
                                                   const double* __restrict__ q, double* x, double* r, double* mr,
                                                   const double* sc, const int32_t* ci, double* partial, uint32_t nb)
{
  if (ci[CI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const double alpha = sc[PF_ALPHA];
  double rr = 0.0, rho = 0.0;
  if (k < nn && gm[k] != 0.0) {
    x[k] = alpha * s[k] + x[k];
    const double rv = -alpha * q[k] + r[k];
    r[k] = rv;
    const double z = rv / diag[k];
    mr[k] = z;
    rr = rv * rv;
    rho = z * rv;
  }
  __shared__ double smem[4];
  const double s1 = block_sum256(rr, smem);
  __syncthreads();
  const double s2 = block_sum256(rho, smem);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = s1;
    partial[nb + blockIdx.x] = s2;
  }
}

enum { PFR_INIT = 0, PFR_ALPHA = 1, PFR_CHECK = 2 };

// one-block ordered sums of the partials and the CG scalar logic
__global__ __launch_bounds__(256) void k_pf_reduce(const double* __restrict__ partial, uint32_t nb, double* sc,
                                                   int32_t* ci, int mode, double cg_relative, PfScalars ps)
{
  __shared__ double smem[4];
  if (mode != PFR_INIT && ci[CI_DONE]) return;
  const double a0 = thread_partials256(partial, nb);
  const double a1 = mode != PFR_ALPHA ? thread_partials256(partial + nb, nb) : 0.0;
  const double s0 = block_sum256(a0, smem);
  __syncthreads();
  const double s1 = block_sum256(a1, smem);
  if (threadIdx.x != 0) return;
  if (mode == PFR_INIT) {  // s0 = mr.r, s1 = r.r
    sc[PF_RHO] = s0;
    sc[PF_NORM] = sqrt(s1);
    sc[PF_REL] = sqrt(s1) * cg_relative;
    ci[CI_DONE] = 0;
    ci[CI_PF_IT] = 0;
  } else if (mode == PFR_ALPHA) {  // s0 = s.q
    sc[PF_SQ] = s0;
    sc[PF_ALPHA] = s0 != 0.0 ? sc[PF_RHO] / s0 : 1.7976931348623157e308;
  } else {  // s0 = r.r, s1 = new mr.r
    const double norm = sqrt(s0);
    sc[PF_NORM] = norm;
    const int it = ci[CI_PF_IT];
    if (norm < ps.tol || norm < sc[PF_REL] || it == ps.max_it) {
      ci[CI_DONE] = 1;
      ci[CI_ITERS] = it;
    } else {
      sc[PF_BETA] = s1 / sc[PF_RHO];
      sc[PF_RHO] = s1;
      ci[CI_PF_IT] = it + 1;
    }
  }
}

// phaseG2P (:457-507): c from the solved increment, then the history update
// (PhaseField::Update_Phase_Field_Fp) and g = c^2 (1 - k) + k
__global__ __launch_bounds__(256) void k_pf_g2p(ParticlesView P, PfParticles pf, PfScalars ps,
                                                const uint32_t* __restrict__ order,
                                                const uint32_t* __restrict__ group_start,
                                                const uint32_t* __restrict__ group_page,
                                                const int32_t* __restrict__ group_nslot,
                                                const double* __restrict__ gm, const double* __restrict__ x,
                                                const double* __restrict__ phase, double* gdeg, double inv_dx)
{
  __shared__ double tv[kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  if (tid < 8) nslot[tid] = group_nslot[8 * (size_t)g + tid];
  __syncthreads();
  const bool parabolic = ps.M > 0.0;
  for (int t = tid; t < kTileNodes; t += 256) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    double v = 0.0;
    if (sl >= 0 && gm[node] != 0.0) v = parabolic ? x[node] - phase[node] : x[node];
    tv[t] = v;
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
    double c = pf.c[i];
    if (pf.allow[i]) {
      const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
      double w[3][3];
      int b[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) b[d] = bspline_w(xp[d] * inv_dx, w[d]);
      const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
      if ((unsigned)lx <= 1u && (unsigned)ly <= 3u && (unsigned)lz <= 3u) {
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) {
            const double wab = w[0][a] * w[1][bb];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
              acc += wab * w[2][cc] * tv[((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + cc)];
          }
        const double nc = parabolic ? c + acc : acc;
        const double cm = c < nc ? c : nc;
        c = cm > 0.0 ? cm : 0.0;
        pf.c[i] = c;
      }
    }
    // psi+ of NeoHookeanBorden::get_psi_pos (NeoHookeanBorden.cpp:54-67)
    M3<double> f;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc)
#pragma unroll
      for (int r = 0; r < 3; ++r) f.a[r][cc] = P.F[(cc * 3 + r) * S + i];
    const double mu = P.mu[i], lam = P.lam[i];
    const double J = det3(f);
    const double aJ = pow(J, -1.0 / 3.0);
    double tr = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) tr += (aJ * f.a[r][cc]) * (aJ * f.a[r][cc]);
    const double psi_dev = mu / 2.0 * (tr - 3.0);
    const double kappa = 0.666666 * mu + lam;
    const double psi_vol = kappa / 2.0 * ((J * J - 1.0) / 2.0 - log(J));
    const double psi = J >= 1.0 ? psi_vol + psi_dev : psi_dev;
    const double kres = pf.k[i];
    if (psi > pf.H[i]) {
      const double hm = pf.Hmax[i];
      const double H = psi < hm ? psi : hm;
      pf.H[i] = H;
      pf.Fp[i] = 4.0 * pf.l0[i] * (1.0 - kres) * H * pf.osc[i] + 1.0;
    }
    gdeg[i] = c * c * (1.0 - kres) + kres;
  }
}

// largest eigenvalue of a symmetric 3x3 matrix, trigonometric closed form of
// the characteristic cubic (O. K. Smith, CACM 4(4):168, 1961)
__device__ __forceinline__ double sym3_max_eig(const M3<double>& A)
{
  const double a00 = A.a[0][0], a11 = A.a[1][1], a22 = A.a[2][2];
  const double a01 = 0.5 * (A.a[0][1] + A.a[1][0]), a02 = 0.5 * (A.a[0][2] + A.a[2][0]),
               a12 = 0.5 * (A.a[1][2] + A.a[2][1]);
  const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  const double q = (a00 + a11 + a22) / 3.0;
  const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  if (p2 <= 0.0) return q;
  const double p = sqrt(p2 / 6.0);
  const double ip = 1.0 / p;
  const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip, c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  const double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
  double r = 0.5 * det;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  return q + 2.0 * p * cos(acos(r) / 3.0);
}

// FractureSimulation::deleteParticles (:162-207), run by reinitialize() at
// the start of every step (MpmSimulationBase.cpp:262): particles whose c is
// below the threshold are parked at (4, 4, 4) at rest with F = I and take no
// further damage
__global__ __launch_bounds__(256) void k_pf_delete(ParticlesView P, const double* __restrict__ c,
                                                   int32_t* __restrict__ allow, double threshold)
{
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= (uint32_t)P.stride) return;
  if (!(c[i] < threshold)) return;
  const int64_t S = P.stride;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    P.x[d * S + i] = 4.0;
    P.v[d * S + i] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) P.F[k * S + i] = (k % 4 == 0) ? 1.0 : 0.0;
  allow[i] = 0;
}

// solvePhaseFieldSystemNaive (:509-554): Rankine cut of c by the largest
// principal Cauchy stress of the undegraded model; the FP64 threshold test
__global__ __launch_bounds__(256) void k_pf_naive(ParticlesView P, PfParticles pf, double sigma_F, double* gdeg)
{
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= (uint32_t)P.stride) return;
  const int64_t S = P.stride;
  M3<double> f, tau;
#pragma unroll
  for (int cc = 0; cc < 3; ++cc)
#pragma unroll
    for (int r = 0; r < 3; ++r) f.a[r][cc] = P.F[(cc * 3 + r) * S + i];
  const double J = det3(f);
  nhb_kirchhoff(P.mu[i], P.lam[i], 1.0, f, tau);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) tau.a[r][cc] = tau.a[r][cc] / J;
  const double smax = sym3_max_eig(tau);
  double c = pf.c[i];
  if (smax > sigma_F) {
    const double cut = sigma_F / smax;
    c = c < cut ? c : cut;
  }
  pf.c[i] = c;
  const double kres = pf.k[i];
  gdeg[i] = c * c * (1.0 - kres) + kres;
}
