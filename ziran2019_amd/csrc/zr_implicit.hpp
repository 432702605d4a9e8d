// zr_implicit.hpp -- kernels of the implicit backward-Euler step (APIC
// transfers, Newton with a matrix-free preconditioned CG), FP64 plain path.
// Included by zr_engine.hip inside its anonymous namespace.
//
// Node vectors are structure-of-arrays [3][stride] over page slots x 32
// elements (nodes with m == 0 hold zeros).  Reductions are deterministic:
// per-block partial sums in a fixed shuffle tree, then one block sums the
// partials in index order.
//
// Per-particle cache written by the state kernel and read by the operator:
// U, V (9 each, column-major) and Sigma (3) of the SVD of the Newton iterate
// F = (I + dt grad(vn + dv)) Fn.

enum {
  SC_NORM2 = 0,   // Newton residual sum |r|^2 / m
  SC_ZTR = 1,     // CG z^T r
  SC_PAP = 2,     // CG p^T A p
  SC_ZTR_NEW = 3,
  SC_BETA = 4,
  SC_LOCAL_TOL = 5,
  SC_N = 16
};
enum { CI_DONE = 0, CI_ITERS = 1, CI_N = 4 };

constexpr int kImpFields = 3;
constexpr int kImpStage = kImpFields * kTileNodes;

struct NodeVec3 {
  double* p;
  int64_t stride;
  __device__ __forceinline__ double& operator()(int d, int64_t k) const { return p[d * stride + k]; }
};

// weights and gradients exactly as MpmGrid::iterateKernel forms them
// (MpmGrid.h:231-268): dw scaled by 1/dx, products in i, j, k order
__device__ __forceinline__ double wgrad(const double w[3][3], const double dw[3][3], double inv_dx, int a, int b,
                                        int c, double gw[3])
{
  const double wi = w[0][a], wj = w[1][b], wk = w[2][c];
  const double dwidxi = inv_dx * dw[0][a];
  const double wij = wi * wj;
  const double dwijdxi = dwidxi * wj;
  const double dwijdxj = wi * inv_dx * dw[1][b];
  gw[0] = dwijdxi * wk;
  gw[1] = dwijdxj * wk;
  gw[2] = wij * inv_dx * dw[2][c];
  return wij * wk;
}

// deterministic block sum of doubles (blockDim 256); result valid on thread 0
__device__ __forceinline__ double block_sum256(double v, double* smem4)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) smem4[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = ((smem4[0] + smem4[1]) + smem4[2]) + smem4[3];
  return s;
}

// collision projection P = I - K K^T of the node (CollisionObject.h:17-40),
// with the normal basis of multiObjectCollision (CollisionObject.cpp:66-105)
__device__ __forceinline__ void project_node(const ColliderSet& cs, double t, const double xi[3], double r[3])
{
  double K[3][3];
  int nslip = 0;
  for (int k = 0; k < cs.n; ++k) {
    double X[3], phi, n[3];
    for (int d = 0; d < 3; ++d) X[d] = xi[d] - cs.c[k].velocity[d] * t;
    if (!ls_query_inside(cs.c[k], X, phi, n)) continue;
    if (cs.c[k].type == ZR_STICKY) {
      r[0] = r[1] = r[2] = 0.0;
      return;
    }
    for (int q = 0; q < nslip; ++q) {
      double dot = K[q][0] * n[0] + K[q][1] * n[1] + K[q][2] * n[2];
      for (int d = 0; d < 3; ++d) n[d] -= dot * K[q][d];
    }
    double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0.0) {
      for (int d = 0; d < 3; ++d) K[nslip][d] = n[d] / l;
      if (++nslip == 3) break;
    }
  }
  for (int q = 0; q < nslip; ++q) {
    double dot = K[q][0] * r[0] + K[q][1] * r[1] + K[q][2] * r[2];
    for (int d = 0; d < 3; ++d) r[d] -= dot * K[q][d];
  }
}

// sum of the <= 8 staged tiles overlapping node (slot page q, element e)
template <int NF>
__device__ __forceinline__ void sum_tiles(const double* stage, const int32_t* page_group, uint32_t qx, uint32_t qy,
                                          uint32_t qz, int ex, int ey, int ez, double acc[NF])
{
  constexpr int stride = NF * kTileNodes;
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = 0.0;
  for (int a = 0; a < 2; ++a) {
    if (a > (int)qx) break;
    const int tx = ex + 2 * a;
    for (int b = 0; b < 2; ++b) {
      const int ty = ey + 4 * b;
      if (ty >= kTileY || b > (int)qy) break;
      for (int c = 0; c < 2; ++c) {
        const int tz = ez + 4 * c;
        if (tz >= kTileZ || c > (int)qz) break;
        const int32_t src = page_group[page_encode(qx - a, qy - b, qz - c)];
        if (src < 0) continue;
        const double* base = stage + (size_t)src * stride + (tx * kTileY + ty) * kTileZ + tz;
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[f] += base[f * kTileNodes];
      }
    }
  }
}

struct NodeCoord {
  uint32_t qx, qy, qz;
  int ex, ey, ez;
};
__device__ __forceinline__ NodeCoord node_coord(const uint32_t* slot_page, uint32_t k)
{
  NodeCoord c;
  page_decode(slot_page[k >> 5], c.qx, c.qy, c.qz);
  const uint32_t e = k & 31u;
  c.ex = (int)(e >> 4);
  c.ey = (int)((e >> 2) & 3u);
  c.ez = (int)(e & 3u);
  return c;
}
__device__ __forceinline__ void node_pos(const NodeCoord& c, double dx, double xi[3])
{
  xi[0] = (double)(2 * (int)c.qx + c.ex) * dx;
  xi[1] = (double)(4 * (int)c.qy + c.ey) * dx;
  xi[2] = (double)(4 * (int)c.qz + c.ez) * dx;
}

// v* of the 4x6x6 footprint of group page (px,py,pz) -> LDS, summing up to two
// node vectors (vn + dv); the 8 neighbour slots are looked up once
__device__ __forceinline__ void load_tile3(uint32_t px, uint32_t py, uint32_t pz, const int32_t* page_slot,
                                           const double* a, const double* b, int64_t stride, double* tv,
                                           int32_t* nslot)
{
  const int tid = threadIdx.x;
  if (tid < 8) nslot[tid] = page_slot[page_encode(px + (tid >> 2), py + ((tid >> 1) & 1), pz + (tid & 1))];
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += blockDim.x) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    for (int d = 0; d < 3; ++d)
      tv[d * kTileNodes + t] = sl >= 0 ? a[d * stride + node] + (b ? b[d * stride + node] : 0.0) : 0.0;
  }
  __syncthreads();
}

struct ImpCache {
  double* U;  // [9][n]
  double* V;  // [9][n]
  double* S;  // [3][n]
};

// Newton state: F = (I + dt grad(vn + dv)) Fn, SVD cached, and the stress
// term vPFnT grad w staged per group (MpmForceBase.cpp:305-344,
// FBasedMpmForceHelper.cpp:68-91, rasterizeForceToTVStack<false>)
__global__ __launch_bounds__(256) void k_be_state(ParticlesView P, ImpCache cache, const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ group_start,
                                                  const uint32_t* __restrict__ group_page,
                                                  const int32_t* __restrict__ page_slot, const double* __restrict__ vn,
                                                  const double* __restrict__ dv, int64_t gstride, double dx,
                                                  double inv_dx, double dt, double* __restrict__ stage)
{
  __shared__ double tile[4 * kImpStage];
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * kImpStage; k += 256) tile[k] = 0.0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  load_tile3(px, py, pz, page_slot, vn, dv, gstride, tv, nslot);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  double* wt = tile + (tid >> 6) * kImpStage;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const int64_t S = P.stride;
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
    double G[3][3] = { { 0.0 } };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgrad(w, dw, inv_dx, a, bb, c, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double vv = tv[r * kTileNodes + t];
            G[r][0] += vv * gw[0];
            G[r][1] += vv * gw[1];
            G[r][2] += vv * gw[2];
          }
        }
    M3<double> Fn, F;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) Fn.a[r][c] = P.F[(c * 3 + r) * S + i];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double acc = Fn.a[r][c];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc += dt * G[r][k] * Fn.a[k][c];
        F.a[r][c] = acc;
      }
    M3<double> u, v;
    double s[3];
    svd3(F, u, s, v);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        cache.U[(c * 3 + r) * S + i] = u.a[r][c];
        cache.V[(c * 3 + r) * S + i] = v.a[r][c];
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) cache.S[d * S + i] = s[d];
    // P = U diag(2 mu (s - 1) + lam (J - 1) cof(s)) V^T
    const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i];
    const double J = s[0] * s[1] * s[2];
    const double cof[3] = { s[1] * s[2], s[0] * s[2], s[0] * s[1] };
    double ph[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ph[d] = 2.0 * mu * (s[d] - 1.0) + lam * (J - 1.0) * cof[d];
    double Pm[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) Pm[r][c] = u.a[r][0] * ph[0] * v.a[c][0] + u.a[r][1] * ph[1] * v.a[c][1] + u.a[r][2] * ph[2] * v.a[c][2];
    double vp[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        vp[r][c] = vol * (Pm[r][0] * Fn.a[c][0] + Pm[r][1] * Fn.a[c][1] + Pm[r][2] * Fn.a[c][2]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgrad(w, dw, inv_dx, a, bb, c, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r)
            atomicAdd(&wt[r * kTileNodes + t], vp[r][0] * gw[0] + vp[r][1] * gw[1] + vp[r][2] * gw[2]);
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * kImpStage;
  for (int k = tid; k < kImpStage; k += 256)
    out[k] = tile[k] + tile[kImpStage + k] + tile[2 * kImpStage + k] + tile[3 * kImpStage + k];
}

// Hessian-vector product, particle part: ds = V0 dP(grad p Fn) Fn^T staged
// as ds grad w (MpmForceBase.cpp:258-302, FBasedMpmForceHelper.cpp:132-155)
__global__ __launch_bounds__(256) void k_hvp(ParticlesView P, ImpCache cache, const uint32_t* __restrict__ order,
                                             const uint32_t* __restrict__ group_start,
                                             const uint32_t* __restrict__ group_page,
                                             const int32_t* __restrict__ page_slot, const double* __restrict__ pvec,
                                             int64_t gstride, double inv_dx, double* __restrict__ stage,
                                             const int32_t* __restrict__ ci)
{
  if (ci && ci[CI_DONE]) return;
  __shared__ double tile[4 * kImpStage];
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * kImpStage; k += 256) tile[k] = 0.0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  load_tile3(px, py, pz, page_slot, pvec, nullptr, gstride, tv, nslot);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  double* wt = tile + (tid >> 6) * kImpStage;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const int64_t S = P.stride;
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
    double G[3][3] = { { 0.0 } };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgrad(w, dw, inv_dx, a, bb, c, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double vv = tv[r * kTileNodes + t];
            G[r][0] += vv * gw[0];
            G[r][1] += vv * gw[1];
            G[r][2] += vv * gw[2];
          }
        }
    double Fn[3][3], u[3][3], v[3][3], s[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        Fn[r][c] = P.F[(c * 3 + r) * S + i];
        u[r][c] = cache.U[(c * 3 + r) * S + i];
        v[r][c] = cache.V[(c * 3 + r) * S + i];
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d] = cache.S[d * S + i];
    // dF = G Fn ; M = U^T dF V
    double dF[3][3], T1[3][3], M[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) dF[r][c] = G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) T1[r][c] = u[0][r] * dF[0][c] + u[1][r] * dF[1][c] + u[2][r] * dF[2][c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) M[r][c] = T1[r][0] * v[0][c] + T1[r][1] * v[1][c] + T1[r][2] * v[2][c];
    const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i];
    const double J = s[0] * s[1] * s[2];
    const double cof[3] = { s[1] * s[2], s[0] * s[2], s[0] * s[1] };
    const double cm = cof[0] * M[0][0] + cof[1] * M[1][1] + cof[2] * M[2][2];
    // skew part of U^T dR V
    double W01 = 0.0, W02 = 0.0, W12 = 0.0;
    {
      double d01 = s[0] + s[1], d02 = s[0] + s[2], d12 = s[1] + s[2];
      if (fabs(d01) < 1e-12) d01 = d01 < 0.0 ? -1e-12 : 1e-12;
      if (fabs(d02) < 1e-12) d02 = d02 < 0.0 ? -1e-12 : 1e-12;
      if (fabs(d12) < 1e-12) d12 = d12 < 0.0 ? -1e-12 : 1e-12;
      W01 = (M[0][1] - M[1][0]) / d01;
      W02 = (M[0][2] - M[2][0]) / d02;
      W12 = (M[1][2] - M[2][1]) / d12;
    }
    const double lj = lam * (J - 1.0);
    double H[3][3];
    H[0][0] = 2.0 * mu * M[0][0] + lam * cm * cof[0] + lj * (s[2] * M[1][1] + s[1] * M[2][2]);
    H[1][1] = 2.0 * mu * M[1][1] + lam * cm * cof[1] + lj * (s[2] * M[0][0] + s[0] * M[2][2]);
    H[2][2] = 2.0 * mu * M[2][2] + lam * cm * cof[2] + lj * (s[1] * M[0][0] + s[0] * M[1][1]);
    H[0][1] = 2.0 * mu * (M[0][1] - W01) - lj * s[2] * M[1][0];
    H[1][0] = 2.0 * mu * (M[1][0] + W01) - lj * s[2] * M[0][1];
    H[0][2] = 2.0 * mu * (M[0][2] - W02) - lj * s[1] * M[2][0];
    H[2][0] = 2.0 * mu * (M[2][0] + W02) - lj * s[1] * M[0][2];
    H[1][2] = 2.0 * mu * (M[1][2] - W12) - lj * s[0] * M[2][1];
    H[2][1] = 2.0 * mu * (M[2][1] + W12) - lj * s[0] * M[1][2];
    // dP = U H V^T ; ds = vol dP Fn^T
    double T2[3][3], dP[3][3], ds[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) T2[r][c] = u[r][0] * H[0][c] + u[r][1] * H[1][c] + u[r][2] * H[2][c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) dP[r][c] = T2[r][0] * v[c][0] + T2[r][1] * v[c][1] + T2[r][2] * v[c][2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) ds[r][c] = vol * (dP[r][0] * Fn[c][0] + dP[r][1] * Fn[c][1] + dP[r][2] * Fn[c][2]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgrad(w, dw, inv_dx, a, bb, c, gw);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
#pragma unroll
          for (int r = 0; r < 3; ++r)
            atomicAdd(&wt[r * kTileNodes + t], ds[r][0] * gw[0] + ds[r][1] * gw[1] + ds[r][2] * gw[2]);
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * kImpStage;
  for (int k = tid; k < kImpStage; k += 256)
    out[k] = tile[k] + tile[kImpStage + k] + tile[2 * kImpStage + k] + tile[3 * kImpStage + k];
}

// buildInitialDvAndVnForNewton (MpmSimulationBase.cpp:1105-1136)
__global__ __launch_bounds__(256) void k_be_init(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                                 const double* __restrict__ gm, const double* __restrict__ gv,
                                                 NodeVec3 vn, NodeVec3 dv, double dx, double dt, double gx, double gy,
                                                 double gz, double time, ColliderSet cs)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const double m = gm[k];
  if (m == 0.0) {
    for (int d = 0; d < 3; ++d) vn(d, k) = dv(d, k) = 0.0;
    return;
  }
  const int64_t S = vn.stride;
  const double v[3] = { gv[k], gv[S + k], gv[2 * S + k] };
  double vi[3] = { v[0], v[1], v[2] };
  double xi[3];
  node_pos(node_coord(slot_page, k), dx, xi);
  const bool hit = multi_object_collision(cs, time, xi, vi);
  const double g[3] = { gx, gy, gz };
  for (int d = 0; d < 3; ++d) {
    dv(d, k) = hit ? vi[d] - v[d] : g[d] * dt;
    vn(d, k) = v[d];
  }
}

// computeResidual (BackwardEuler.h:101-124): r = dt m g - dt sum vPFnT grad w - m dv,
// projected; partial sums of |r|^2 / m (computeNorm, BackwardEuler.h:127-140)
__global__ __launch_bounds__(256) void k_be_residual(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                                     const int32_t* __restrict__ page_group,
                                                     const double* __restrict__ stage, const double* __restrict__ gm,
                                                     NodeVec3 dv, NodeVec3 r, double dx, double dt, double gx,
                                                     double gy, double gz, double time, ColliderSet cs,
                                                     double* __restrict__ partial)
{
  __shared__ double smem[4];
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    double rv[3] = { 0.0, 0.0, 0.0 };
    if (m != 0.0) {
      const NodeCoord nc = node_coord(slot_page, k);
      double f[kImpFields];
      sum_tiles<kImpFields>(stage, page_group, nc.qx, nc.qy, nc.qz, nc.ex, nc.ey, nc.ez, f);
      const double g[3] = { gx, gy, gz };
      for (int d = 0; d < 3; ++d) rv[d] = dt * g[d] * m - dt * f[d] - m * dv(d, k);
      double xi[3];
      node_pos(nc, dx, xi);
      project_node(cs, time, xi, rv);
      contrib = (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]) / m;
    }
    for (int d = 0; d < 3; ++d) r(d, k) = rv[d];
  }
  const double s = block_sum256(contrib, smem);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// sums nb partials in index order into sc[slot]; skipped once CG is done
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ partial, uint32_t nb, double* sc, int slot,
                                                const int32_t* ci)
{
  __shared__ double smem[4];
  if (ci && ci[CI_DONE]) return;
  double acc = 0.0;
  for (uint32_t b = threadIdx.x; b < nb; b += 256) acc += partial[b];
  const double s = block_sum256(acc, smem);
  if (threadIdx.x == 0) sc[slot] = s;
}

// CG start (ConjugateGradient.h:105-116): x = 0, r = b, z = r / m, p = z
__global__ __launch_bounds__(256) void k_cg_init(uint32_t nn, const double* __restrict__ gm, NodeVec3 b, NodeVec3 x,
                                                 NodeVec3 r, NodeVec3 q, NodeVec3 p, double* __restrict__ partial)
{
  __shared__ double smem[4];
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    for (int d = 0; d < 3; ++d) {
      const double rv = b(d, k);
      const double qv = m != 0.0 ? rv / m : 0.0;
      x(d, k) = 0.0;
      r(d, k) = rv;
      q(d, k) = qv;
      p(d, k) = qv;
      contrib += rv * qv;
    }
  }
  const double s = block_sum256(contrib, smem);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// local_tolerance = min(rel |r0|_P, tol)  (ConjugateGradient.h:117-119)
__global__ void k_cg_setup(double* sc, int32_t* ci, double rel, double tol)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double norm = sqrt(sc[SC_ZTR]);
  const double lt = rel * norm < tol ? rel * norm : tol;
  sc[SC_LOCAL_TOL] = lt;
  ci[CI_DONE] = norm < lt ? 1 : 0;
}

// Ap = M p + dt^2 sum ds grad w, projected; partial p^T A p (BackwardEuler.h:322-351)
__global__ __launch_bounds__(256) void k_cg_ap(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                               const int32_t* __restrict__ page_group,
                                               const double* __restrict__ stage, const double* __restrict__ gm,
                                               NodeVec3 p, NodeVec3 ap, double dx, double dt, double time,
                                               ColliderSet cs, int project, double* __restrict__ partial,
                                               const int32_t* ci)
{
  if (ci && ci[CI_DONE]) return;
  __shared__ double smem[4];
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    double av[3] = { 0.0, 0.0, 0.0 };
    if (m != 0.0) {
      const NodeCoord nc = node_coord(slot_page, k);
      double f[kImpFields];
      sum_tiles<kImpFields>(stage, page_group, nc.qx, nc.qy, nc.qz, nc.ex, nc.ey, nc.ez, f);
      for (int d = 0; d < 3; ++d) av[d] = m * p(d, k) + dt * dt * f[d];
      if (project) {
        double xi[3];
        node_pos(nc, dx, xi);
        project_node(cs, time, xi, av);
      }
    }
    for (int d = 0; d < 3; ++d) {
      ap(d, k) = av[d];
      contrib += av[d] * p(d, k);
    }
  }
  const double s = block_sum256(contrib, smem);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x += alpha p, r -= alpha Ap, z = r / m ; partial z^T r
__global__ __launch_bounds__(256) void k_cg_update(uint32_t nn, const double* __restrict__ gm, NodeVec3 x, NodeVec3 r,
                                                   NodeVec3 q, NodeVec3 p, NodeVec3 ap, const double* __restrict__ sc,
                                                   double* __restrict__ partial, const int32_t* ci)
{
  if (ci[CI_DONE]) return;
  __shared__ double smem[4];
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double alpha = sc[SC_ZTR] / sc[SC_PAP];
    const double m = gm[k];
    for (int d = 0; d < 3; ++d) {
      x(d, k) += p(d, k) * alpha;
      const double rv = r(d, k) - ap(d, k) * alpha;
      r(d, k) = rv;
      const double qv = m != 0.0 ? rv / m : 0.0;
      q(d, k) = qv;
      contrib += qv * rv;
    }
  }
  const double s = block_sum256(contrib, smem);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_cg_post(double* sc, int32_t* ci)
{
  if (threadIdx.x != 0 || blockIdx.x != 0 || ci[CI_DONE]) return;
  const double znew = sc[SC_ZTR_NEW];
  sc[SC_BETA] = znew / sc[SC_ZTR];
  sc[SC_ZTR] = znew;
  ci[CI_ITERS] += 1;
  if (sqrt(znew) < sc[SC_LOCAL_TOL]) ci[CI_DONE] = 1;
}

__global__ __launch_bounds__(256) void k_cg_p(uint32_t nn, NodeVec3 q, NodeVec3 p, const double* __restrict__ sc,
                                              const int32_t* ci)
{
  if (ci[CI_DONE]) return;
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const double beta = sc[SC_BETA];
  for (int d = 0; d < 3; ++d) p(d, k) = q(d, k) + beta * p(d, k);
}

__global__ __launch_bounds__(256) void k_axpy3(uint32_t nn, NodeVec3 y, NodeVec3 x)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  for (int d = 0; d < 3; ++d) y(d, k) += x(d, k);
}

// constructNewVelocityFromNewtonResult (MpmSimulationBase.cpp:867-877)
__global__ __launch_bounds__(256) void k_be_finish(uint32_t nn, const double* __restrict__ gm,
                                                   const double* __restrict__ gv, NodeVec3 dv,
                                                   double* __restrict__ gvs)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const double m = gm[k];
  const int64_t S = dv.stride;
  for (int d = 0; d < 3; ++d) gvs[d * S + k] = m != 0.0 ? gv[d * S + k] + dv(d, k) : 0.0;
}
