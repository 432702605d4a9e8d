// zr_minres.hpp -- Minres::solve (Lib/Ziran/Math/Linear/Minres.h:69-178) on
// the device: the scalar recurrences (Lanczos alpha / beta, the Givens QR of
// the tridiagonal, the residual estimate and the termination test) run in one
// thread after an ordered block reduction, so the host only polls a flag.
// Used by the ADMM global step and by the implicit Newton step.  Included by
// zr_engine.hip inside its anonymous namespace after zr_implicit.hpp.

// device scalars (double) and flags (int32) of a MINRES solve
enum {
  AD_RES = 0,      // MINRES residual (preconditioned norm)
  AD_LTOL = 1,     // local tolerance
  AD_ALPHA = 2,
  AD_BETA_K = 3,
  AD_BETA_KP1 = 4,
  AD_GAMMA = 5,
  AD_DELTA = 6,
  AD_EPS = 7,
  AD_TK = 8,
  AD_LT0 = 9,      // last two Givens-transformed rhs components
  AD_LT1 = 10,
  AD_GK_C = 11, AD_GK_S = 12, AD_GKM1_C = 13, AD_GKM1_S = 14,
};
enum { AI_DONE = 0, AI_ITERS = 1, AI_N = 4 };

__device__ __forceinline__ void g2_rot(double c, double s, double v[2])
{
  const double t1 = v[0], t2 = v[1];
  v[0] = c * t1 - s * t2;
  v[1] = s * t1 + c * t2;
}

enum { MR_INIT = 0, MR_ALPHA = 1, MR_GIVENS = 2 };

// the MINRES scalar logic (Minres.h:75-178) in one thread after an ordered sum
__global__ __launch_bounds__(256) void k_minres_reduce(const double* __restrict__ partial, uint32_t nb, double* sc,
                                                       int32_t* ai, int mode, double tol, double rel, int it,
                                                       int max_it)
{
  __shared__ double smem[4];
  if (mode != MR_INIT && ai[AI_DONE]) return;
  const double acc = thread_partials256(partial, nb);
  const double v = block_sum256(acc, smem);
  if (threadIdx.x != 0) return;
  if (mode == MR_INIT) {
    const double res = sqrt(v);
    sc[AD_RES] = res;
    sc[AD_BETA_KP1] = res;
    const double lt = rel * res < tol ? rel * res : tol;
    sc[AD_LTOL] = lt;
    sc[AD_LT0] = res;
    sc[AD_LT1] = 0.0;
    sc[AD_GK_C] = 1.0;
    sc[AD_GK_S] = 0.0;
    sc[AD_GKM1_C] = 1.0;
    sc[AD_GKM1_S] = 0.0;
    ai[AI_ITERS] = 0;
    ai[AI_DONE] = res < lt || max_it <= 0 ? 1 : 0;
    ai[AI_DONE + 2] = ai[AI_DONE];
  } else if (mode == MR_ALPHA) {
    sc[AD_ALPHA] = v;
  } else {
    const double bb = v > 0.0 ? v : 0.0;
    const double beta_kp1 = sqrt(bb);
    sc[AD_BETA_KP1] = beta_kp1;
    // applyAllPreviousGivensRotationsAndDetermineNewGivens
    const double gkm2_c = sc[AD_GKM1_C], gkm2_s = sc[AD_GKM1_S];
    const double gkm1_c = sc[AD_GK_C], gkm1_s = sc[AD_GK_S];
    double ek[2] = { 0.0, sc[AD_BETA_K] };
    g2_rot(gkm2_c, gkm2_s, ek);
    double dk[2] = { ek[1], sc[AD_ALPHA] };
    g2_rot(gkm1_c, gkm1_s, dk);
    double tv[2] = { dk[1], beta_kp1 };
    double c = 1.0, s = 0.0;  // GivensRotation::compute (Givens.h:56-62)
    const double d = sqrt(tv[0] * tv[0] + tv[1] * tv[1]);
    if (d) {
      const double t = 1.0 / d;
      c = tv[0] * t;
      s = -tv[1] * t;
    }
    g2_rot(c, s, tv);
    double lt[2] = { sc[AD_LT0], sc[AD_LT1] };
    g2_rot(c, s, lt);
    sc[AD_EPS] = ek[0];
    sc[AD_DELTA] = dk[0];
    sc[AD_GAMMA] = tv[0];
    sc[AD_TK] = lt[0];
    sc[AD_LT0] = lt[1];
    sc[AD_LT1] = 0.0;
    sc[AD_GKM1_C] = gkm1_c;
    sc[AD_GKM1_S] = gkm1_s;
    sc[AD_GK_C] = c;
    sc[AD_GK_S] = s;
    const double res = lt[1] < 0.0 ? -lt[1] : lt[1];
    sc[AD_RES] = res;
    ai[AI_ITERS] = it + 1;
    if (res < sc[AD_LTOL] || it + 1 >= max_it) ai[AI_DONE] = 1;
  }
}

// the start of each MINRES iteration: beta_k = beta_{k+1}; the update of this
// iteration still runs when the convergence flag is raised by its Givens step
__global__ void k_minres_begin(double* sc, int32_t* ai)
{
  if (threadIdx.x != 0) return;
  ai[AI_DONE + 2] = ai[AI_DONE];
  if (ai[AI_DONE]) return;
  sc[AD_BETA_K] = sc[AD_BETA_KP1];
}

// ---- the implicit Newton step's MINRES (BackwardEuler::computeStep with the
// reference's default linear solver, BackwardEuler.h:55-62,200-240): vectors
// in the operator's precision T (FP32 in the FP32-bulk form), inner products
// and scalars in FP64, the mass-lumped or Jacobi preconditioner ----

// start (Minres.h:69-101) with x = 0 as computeStep sets it: q = P(b - A 0) = b
// (the Newton residual is already projected), z = M^-1 q; q_k, m_k and m_{k-1}
// start at zero (Minres::reinitialize, :40-66); partials of z . q
template <class T>
__global__ __launch_bounds__(256) void k_imr_init(uint32_t nn, const double* __restrict__ gm, NodeVec3 b,
                                                  NodeVec<T> x, NodeVec<T> q, NodeVec<T> z, NodeVec<T> zero0,
                                                  NodeVec<T> zero1, NodeVec<T> zero2, NodeVec<T> pinv,
                                                  double* __restrict__ partial)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    for (int d = 0; d < 3; ++d) {
      const T qv = (T)b(d, k);
      const T zv = m != 0.0 ? (pinv.p ? qv * pinv(d, k) : qv / (T)m) : (T)0;
      x(d, k) = (T)0;
      q(d, k) = qv;
      z(d, k) = zv;
      zero0(d, k) = (T)0;
      zero1(d, k) = (T)0;
      zero2(d, k) = (T)0;
      contrib += (double)qv * (double)zv;
    }
  }
  block_partial(contrib, partial);
}

// the start vectors divided by beta_1 (Minres.h:95-98)
template <class T>
__global__ __launch_bounds__(256) void k_imr_scale(uint32_t nn, const double* __restrict__ gm, NodeVec<T> q,
                                                   NodeVec<T> z, const double* sc, const int32_t* ai)
{
  if (ai[AI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  const double b = sc[AD_BETA_KP1];
  if (k >= nn || gm[k] == 0.0 || !(b > 0.0)) return;
  for (int d = 0; d < 3; ++d) {
    q(d, k) /= (T)b;
    z(d, k) /= (T)b;
  }
}

// q -= alpha q_k; q -= beta_k q_{k-1}; z = M^-1 q; partials of z . q
// (Minres.h:123-130)
template <class T>
__global__ __launch_bounds__(256) void k_imr_lanczos(uint32_t nn, const double* __restrict__ gm, NodeVec<T> q,
                                                     NodeVec<T> qk, NodeVec<T> qkm1, NodeVec<T> z, NodeVec<T> pinv,
                                                     const double* sc, const int32_t* ai, double* partial)
{
  if (ai[AI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    if (m != 0.0) {
      const T alpha = (T)sc[AD_ALPHA], beta = (T)sc[AD_BETA_K];
      for (int d = 0; d < 3; ++d) {
        T v = q(d, k);
        v -= alpha * qk(d, k);
        v -= beta * qkm1(d, k);
        q(d, k) = v;
        const T zv = pinv.p ? v * pinv(d, k) : v / (T)m;
        z(d, k) = zv;
        contrib += (double)zv * (double)v;
      }
    }
  }
  block_partial(contrib, partial);
}

// q, z divided by beta_{k+1}; m_k = (m_k - delta m_{k-1} - eps m_{k-2}) / gamma;
// x += t_k m_k (Minres.h:131-143)
template <class T>
__global__ __launch_bounds__(256) void k_imr_update(uint32_t nn, const double* __restrict__ gm, NodeVec<T> q,
                                                    NodeVec<T> z, NodeVec<T> mk, NodeVec<T> mkm1, NodeVec<T> mkm2,
                                                    NodeVec<T> x, const double* sc, const int32_t* ai)
{
  if (ai[AI_DONE + 2]) return;  // finished before this iteration's Givens step
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn || gm[k] == 0.0) return;
  const double bk = sc[AD_BETA_KP1];
  const T gamma = (T)sc[AD_GAMMA], delta = (T)sc[AD_DELTA], eps = (T)sc[AD_EPS], tk = (T)sc[AD_TK];
  for (int d = 0; d < 3; ++d) {
    if (bk > 0.0) {
      q(d, k) /= (T)bk;
      z(d, k) /= (T)bk;
    }
    const T m = (mk(d, k) - delta * mkm1(d, k) - eps * mkm2(d, k)) / gamma;
    mk(d, k) = m;
    x(d, k) += tk * m;
  }
}
