// zr_implicit.hpp -- kernels of the implicit backward-Euler step (APIC
// transfers, Newton with a matrix-free preconditioned CG), FP64 plain path.
// Included by zr_engine.hip inside its anonymous namespace.
//
// Node vectors are structure-of-arrays [3][stride] over page slots x 32
// elements (nodes with m == 0 hold zeros).  Reductions are deterministic:
// per-block partial sums in a fixed shuffle tree, then one block sums the
// partials in index order.
//
// Per-particle operator data, written by the state kernel in SORTED order
// (position start(g) + j of particle j of group g) so that the operator reads
// it with unit stride and without the particle index, from the SVD
// F = U diag(s) V^T of the Newton iterate F = (I + dt grad(vn + dv)) Fn:
//   U (9), Q = Fn V (9), s (3), k = (2 mu V0, lam V0) (2),
//   frac = x/dx - base (3) and the base cell relative to the group page.
// With M = U^T dF V = U^T (G Fn) V = U^T G Q, the operator's particle term is
//   V0 dP Fn^T = U H(M) V^T Fn^T = U H(M) Q^T,
// H the fixed-corotated differential in the singular frame (scaled by V0).
// The operator runs in T = double (plain path) or T = float (FP32 bulk); the
// state, the Newton residual and every inner product stay FP64.

enum {
  SC_NORM2 = 0,   // Newton residual sum |r|^2 / m
  SC_ZTR = 1,     // CG z^T r of even iterations
  SC_PAP = 2,     // CG p^T A p
  SC_ZTR_NEW = 3, // CG z^T r of odd iterations
  SC_BETA = 4,
  SC_LOCAL_TOL = 5,
  SC_ALPHA = 6,
  SC_N = 16
};
// z_k^T r_k lives in alternating slots so that the blocks of one kernel can
// read the previous value while block 0 stores the new one
__device__ __forceinline__ int ztr_slot(int k) { return (k & 1) ? SC_ZTR_NEW : SC_ZTR; }
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

// this thread's share of an ordered sum of nb block partials (blockDim 256):
// eight independent strided accumulators so that the loads are all in flight
// at once, combined in a fixed order
__device__ __forceinline__ double thread_partials256(const double* __restrict__ partial, uint32_t nb)
{
  double a[8] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
  uint32_t b = threadIdx.x;
  for (; b + 7u * 256u < nb; b += 8u * 256u) {
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] += partial[b + 256u * j];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (b + 256u * j < nb) a[j] += partial[b + 256u * j];
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// the ordered sum of nb block partials (blockDim 256), on every thread
__device__ __forceinline__ double block_total256(const double* __restrict__ partial, uint32_t nb, double* smem5)
{
  const double s = block_sum256(thread_partials256(partial, nb), smem5);
  if (threadIdx.x == 0) smem5[4] = s;
  __syncthreads();
  return smem5[4];
}

// block partial of a grid-wide sum; the partials are added in index order
// by a one-block kernel (a device-scope fence per block would force L2
// write-backs across the XCDs of the device)
__device__ __forceinline__ void block_partial(double contrib, double* partial)
{
  __shared__ double smem[4];
  const double s = block_sum256(contrib, smem);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// collision projection P = I - K K^T of the node (CollisionObject.h:17-40),
// with the normal basis of multiObjectCollision (CollisionObject.cpp:66-105)
__device__ __forceinline__ void project_node(const ColliderSet& cs, const double xi[3], double r[3])
{
  double K[3][3];
  int nslip = 0;
  for (int k = 0; k < cs.n; ++k) {
    double xb[3], phi, n[3];
    if (!collider_query(cs.c[k], 0, xi, phi, n, xb)) continue;
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
template <int NF, class T>
__device__ __forceinline__ void sum_tiles(const T* stage, const int32_t* page_group, uint32_t qx, uint32_t qy,
                                          uint32_t qz, int ex, int ey, int ez, T acc[NF])
{
  constexpr int stride = NF * kTileNodes;
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = (T)0;
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
        const T* base = stage + (size_t)src * stride + (tx * kTileY + ty) * kTileZ + tz;
#pragma unroll
        for (int f = 0; f < NF; ++f) acc[f] += base[f * kTileNodes];
      }
    }
  }
}

// sum of the <= 8 staged tiles overlapping a node, from the slot's
// precomputed source groups (o = 4a + 2b + c, -1 when absent)
template <int NF, class T>
__device__ __forceinline__ void sum_tiles_src(const T* stage, const int32_t* src8, int ex, int ey, int ez,
                                              T acc[NF])
{
  constexpr int stride = NF * kTileNodes;
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = (T)0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int a = o >> 2, b = (o >> 1) & 1, c = o & 1;
    const int tx = ex + 2 * a, ty = ey + 4 * b, tz = ez + 4 * c;
    if (ty >= kTileY || tz >= kTileZ) continue;
    const int32_t src = src8[o];
    if (src < 0) continue;
    const T* base = stage + (size_t)src * stride + (tx * kTileY + ty) * kTileZ + tz;
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] += base[f * kTileNodes];
  }
}

// the same with the page's 8 source groups already in registers
template <int NF, class T>
__device__ __forceinline__ void sum_tiles_src8(const T* stage, const int32_t src8[8], int ex, int ey, int ez,
                                               T acc[NF])
{
  constexpr int stride = NF * kTileNodes;
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = (T)0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int a = o >> 2, b = (o >> 1) & 1, c = o & 1;
    const int tx = ex + 2 * a, ty = ey + 4 * b, tz = ez + 4 * c;
    if (ty >= kTileY || tz >= kTileZ) continue;
    if (src8[o] < 0) continue;
    const T* base = stage + (size_t)src8[o] * stride + (tx * kTileY + ty) * kTileZ + tz;
#pragma unroll
    for (int f = 0; f < NF; ++f) acc[f] += base[f * kTileNodes];
  }
}
__device__ __forceinline__ void load_src8(const int32_t* __restrict__ slot_src, uint32_t k, int32_t src8[8])
{
  const int4* v = reinterpret_cast<const int4*>(slot_src + 8 * (size_t)(k >> 5));
  const int4 lo = v[0], hi = v[1];
  src8[0] = lo.x;
  src8[1] = lo.y;
  src8[2] = lo.z;
  src8[3] = lo.w;
  src8[4] = hi.x;
  src8[5] = hi.y;
  src8[6] = hi.z;
  src8[7] = hi.w;
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
template <class T, class TS>
__device__ __forceinline__ void load_tile3(uint32_t px, uint32_t py, uint32_t pz, const int32_t* page_slot,
                                           const TS* a, const TS* b, int64_t stride, T* tv, int32_t* nslot)
{
  const int tid = threadIdx.x;
  if (tid < 8) nslot[tid] = page_slot[page_encode(px + (tid >> 2), py + ((tid >> 1) & 1), pz + (tid & 1))];
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += blockDim.x) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    for (int d = 0; d < 3; ++d)
      tv[d * kTileNodes + t] = sl >= 0 ? (T)(a[d * stride + node] + (b ? b[d * stride + node] : (TS)0)) : (T)0;
  }
  __syncthreads();
}

// tile load from the group's precomputed neighbour slots (8 entries)
template <class T, class TS>
__device__ __forceinline__ void load_tile3_slots(const int32_t* gslot, const TS* a, int64_t stride, T* tv,
                                                 int32_t* nslot)
{
  const int tid = threadIdx.x;
  if (tid < 8) nslot[tid] = gslot[tid];
  __syncthreads();
  for (int t = tid; t < kTileNodes; t += blockDim.x) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    for (int d = 0; d < 3; ++d) tv[d * kTileNodes + t] = sl >= 0 ? (T)a[d * stride + node] : (T)0;
  }
  __syncthreads();
}

// CG search direction on the fly: p_new = q + beta p_old for the tile nodes
// tile node t (< kTileNodes) of a group's 4x6x6 footprint: which of the 8
// neighbour page slots holds it (4a + 2b + c) and its element in that page
__device__ __forceinline__ void tile_node_source(int t, int& src, int& elem)
{
  const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
  src = ((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2);
  elem = (int)elem_of(tx, ty, tz);
}

template <class T>
struct HvpCache {
  T* U;           // [9][stride] column-major
  T* Q;           // [9][stride] Fn V, column-major
  T* S;           // [3][stride] singular values
  T* K;           // [2][stride] 2 mu V0, lam V0
  T* frac;        // [3][stride] x/dx - base node, in [0.5, 1.5)
  uint32_t* cell; // [stride] lx | ly << 1 | lz << 3 within the group page
  int64_t stride;
};

template <class T>
struct NodeVec {
  T* p;
  int64_t stride;
  __device__ __forceinline__ T& operator()(int d, int64_t k) const { return p[d * stride + k]; }
};

// quadratic B-spline weights from d0 = x - base (BSplines.h:55-81)
template <class T>
__device__ __forceinline__ void bspline_frac(T d0, T w[3], T dw[3])
{
  const T z = (T)1.5 - d0;
  w[0] = (T)0.5 * z * z;
  const T d1 = d0 - (T)1;
  w[1] = (T)0.75 - d1 * d1;
  const T d2 = (T)1 - d1;
  const T zz = (T)1.5 - d2;
  w[2] = (T)0.5 * zz * zz;
  dw[0] = -z;
  dw[1] = (T)-2 * d1;
  dw[2] = zz;
}

template <class T>
__device__ __forceinline__ T wgradT(const T w[3][3], const T dw[3][3], T inv_dx, int a, int b, int c, T gw[3])
{
  const T wi = w[0][a], wj = w[1][b], wk = w[2][c];
  const T dwidxi = inv_dx * dw[0][a];
  const T wij = wi * wj;
  const T dwijdxi = dwidxi * wj;
  const T dwijdxj = wi * inv_dx * dw[1][b];
  gw[0] = dwijdxi * wk;
  gw[1] = dwijdxj * wk;
  gw[2] = wij * inv_dx * dw[2][c];
  return wij * wk;
}

// keeps the compiler from carrying values computed before this point across
// it (the weights are recomputed after the per-particle math instead of
// holding all 27 gradients live)
__device__ __forceinline__ void opaque(double& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void opaque(float& v) { asm volatile("" : "+v"(v)); }
// a scatter after a gather recomputes the 27 weight gradients from the 1-D
// weights rather than holding them from the gather (register budget)
template <class T>
__device__ __forceinline__ void opaque_weights(T w[3][3], T dw[3][3])
{
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      opaque(w[d][a]);
      opaque(dw[d][a]);
    }
}

// quotient in the operator's precision: IEEE division in FP64; in FP32 the
// hardware reciprocal (1 ulp), well inside the FP32 operator's own rounding
template <class T>
__device__ __forceinline__ T op_div(T a, T b)
{
  return a / b;
}
template <>
__device__ __forceinline__ float op_div<float>(float a, float b)
{
  return a * __builtin_amdgcn_rcpf(b);
}

// fixed-corotated differential in the singular frame: for F = U diag(s) V^T
// and M = U^T dF V, H = U^T (V0 dP) V with k2mu = 2 mu V0, klam = lam V0
// (the rotation's part through W = (M - M^T) / (s_i + s_j), |s_i + s_j|
// kept >= 1e-12), from the pair sums d_ij = s_i + s_j
template <class T>
__device__ __forceinline__ void fcr_frame_differential_sums(const T s[3], T d01, T d02, T d12, T k2mu, T klam,
                                                            const T M[3][3], T H[3][3])
{
  const T J = s[0] * s[1] * s[2];
  const T cof[3] = { s[1] * s[2], s[0] * s[2], s[0] * s[1] };
  const T cm = cof[0] * M[0][0] + cof[1] * M[1][1] + cof[2] * M[2][2];
  const T eps = (T)1e-12;
  if (fabs(d01) < eps) d01 = d01 < (T)0 ? -eps : eps;
  if (fabs(d02) < eps) d02 = d02 < (T)0 ? -eps : eps;
  if (fabs(d12) < eps) d12 = d12 < (T)0 ? -eps : eps;
  const T W01 = op_div(M[0][1] - M[1][0], d01);
  const T W02 = op_div(M[0][2] - M[2][0], d02);
  const T W12 = op_div(M[1][2] - M[2][1], d12);
  const T lj = klam * (J - (T)1);
  H[0][0] = k2mu * M[0][0] + klam * cm * cof[0] + lj * (s[2] * M[1][1] + s[1] * M[2][2]);
  H[1][1] = k2mu * M[1][1] + klam * cm * cof[1] + lj * (s[2] * M[0][0] + s[0] * M[2][2]);
  H[2][2] = k2mu * M[2][2] + klam * cm * cof[2] + lj * (s[1] * M[0][0] + s[0] * M[1][1]);
  H[0][1] = k2mu * (M[0][1] - W01) - lj * s[2] * M[1][0];
  H[1][0] = k2mu * (M[1][0] + W01) - lj * s[2] * M[0][1];
  H[0][2] = k2mu * (M[0][2] - W02) - lj * s[1] * M[2][0];
  H[2][0] = k2mu * (M[2][0] + W02) - lj * s[1] * M[0][2];
  H[1][2] = k2mu * (M[1][2] - W12) - lj * s[0] * M[2][1];
  H[2][1] = k2mu * (M[2][1] + W12) - lj * s[0] * M[1][2];
}
template <class T>
__device__ __forceinline__ void fcr_frame_differential(const T s[3], T k2mu, T klam, const T M[3][3], T H[3][3])
{
  fcr_frame_differential_sums(s, s[0] + s[1], s[0] + s[2], s[1] + s[2], k2mu, klam, M, H);
}
// The FP32 operator data hold t = s_1 + s_2, summed in FP64, in place of
// s_2.  Only s_2 can be negative (the SVD's sign convention), so s_1 + s_2
// and s_0 + s_2 are the sums that cancel as an element inverts, and the
// rotation terms divide by them; from t both come out to FP32 relative
// accuracy (s_0 - s_1 is exact when s_1 <= s_0 <= 2 s_1 and far from
// cancelling otherwise), where the sum of two rounded singular values
// would carry an absolute error of their size.
__device__ __forceinline__ void fcr_frame_differential_t(const float st[3], float k2mu, float klam,
                                                         const float M[3][3], float H[3][3])
{
  const float t = st[2];
  const float s[3] = { st[0], st[1], t - st[1] };
  fcr_frame_differential_sums(s, s[0] + s[1], (s[0] - s[1]) + t, t, k2mu, klam, M, H);
}

constexpr int kCells = 32;  // 2 x 4 x 4 cells of a page

// Newton state: F = (I + dt grad(vn + dv)) Fn (FP64, with the SVD), the
// stress term vPFnT grad w staged per group (MpmForceBase.cpp:305-344,
// FBasedMpmForceHelper.cpp:68-91, rasterizeForceToTVStack<false>), and the
// operator data of every particle in sorted order, in precision T; unless
// cell_order is null, with the cells of each group's first 256 particles by
// particle count, largest first (ties by cell), in cell_order[32 g ..] for
// the operator apply's task deal.  JAC: also the operator's diagonal
// without the mass, sum_p (V0 dP(e_d h^T) h)_d with h = Fn^T grad w_ip, staged
// per group in dstage (the node kernel adds m and dt^2), for the Jacobi
// preconditioner; in the singular frame, with u = U^T e_d and g = Q^T grad w,
// it is u^T H(u g^T) g = sum_c g_c u^T H(u e_c^T) g (H linear in M).
template <class T, bool JAC>
__global__ __launch_bounds__(256) void k_be_state(ParticlesView P, HvpCache<T> cache,
                                                  const uint32_t* __restrict__ order,
                                                  const uint32_t* __restrict__ group_start,
                                                  const uint32_t* __restrict__ group_page,
                                                  const int32_t* __restrict__ page_slot, const double* __restrict__ vn,
                                                  const double* __restrict__ dv, int64_t gstride, double dx,
                                                  double inv_dx, double dt, double* __restrict__ stage,
                                                  uint8_t* __restrict__ cell_order, double* __restrict__ dstage)
{
  __shared__ double tile[4 * kImpStage];
  __shared__ double dtile[JAC ? 4 * kImpStage : 1];
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  __shared__ int ccount[kCells];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * kImpStage; k += 256) tile[k] = 0.0;
  if (JAC)
    for (int k = tid; k < 4 * kImpStage; k += 256) dtile[k] = 0.0;
  if (tid < kCells) ccount[tid] = 0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  load_tile3(px, py, pz, page_slot, vn, dv, gstride, tv, nslot);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  double* wt = tile + (tid >> 6) * kImpStage;
  double* wd = dtile + (JAC ? (tid >> 6) * kImpStage : 0);
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const int64_t S = P.stride, CS = cache.stride;
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const uint32_t ks = start + j;
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3], d0[3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double xi = xp[d] * inv_dx;
      b[d] = int_floor(xi - 0.5);
      d0[d] = xi - (double)b[d];
      bspline_frac(d0[d], w[d], dw[d]);
    }
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    if (cell_order && j0 == 0) atomicAdd(&ccount[lx | (ly << 1) | (lz << 3)], 1);
    const int t0 = (lx * kTileY + ly) * kTileZ + lz;
    double G[3][3] = { { 0.0 } };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgradT(w, dw, inv_dx, a, bb, c, gw);
          const int t = t0 + (a * kTileY + bb) * kTileZ + c;
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
    const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        cache.U[(c * 3 + r) * CS + ks] = (T)u.a[r][c];
        cache.Q[(c * 3 + r) * CS + ks] = (T)(Fn.a[r][0] * v.a[0][c] + Fn.a[r][1] * v.a[1][c] + Fn.a[r][2] * v.a[2][c]);
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cache.S[d * CS + ks] = (T)(d == 2 && sizeof(T) == sizeof(float) ? s[1] + s[2] : s[d]);
      cache.frac[d * CS + ks] = (T)d0[d];
    }
    cache.K[ks] = (T)(2.0 * mu * vol);
    cache.K[CS + ks] = (T)(lam * vol);
    cache.cell[ks] = (uint32_t)(lx | (ly << 1) | (lz << 3));
    // P = U diag(2 mu (s - 1) + lam (J - 1) cof(s)) V^T
    const double J = s[0] * s[1] * s[2];
    const double cof[3] = { s[1] * s[2], s[0] * s[2], s[0] * s[1] };
    double ph[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ph[d] = 2.0 * mu * (s[d] - 1.0) + lam * (J - 1.0) * cof[d];
    double Pm[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        Pm[r][c] = u.a[r][0] * ph[0] * v.a[c][0] + u.a[r][1] * ph[1] * v.a[c][1] + u.a[r][2] * ph[2] * v.a[c][2];
    double vp[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        vp[r][c] = vol * (Pm[r][0] * Fn.a[c][0] + Pm[r][1] * Fn.a[c][1] + Pm[r][2] * Fn.a[c][2]);
    // JAC: Kd[d][c][r] = (u^T H(u e_c^T))_r, u = row d of U; Qm = Fn V
    double Kd[JAC ? 3 : 1][3][3], Qm[3][3];
    if (JAC) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          Qm[r][c] = Fn.a[r][0] * v.a[0][c] + Fn.a[r][1] * v.a[1][c] + Fn.a[r][2] * v.a[2][c];
      const double k2mu = 2.0 * mu * vol, klam = lam * vol;
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double M[3][3], H[3][3];
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) M[r][q] = q == c ? u.a[d][r] : 0.0;
          fcr_frame_differential(s, k2mu, klam, M, H);
#pragma unroll
          for (int q = 0; q < 3; ++q)
            Kd[JAC ? d : 0][c][q] = u.a[d][0] * H[0][q] + u.a[d][1] * H[1][q] + u.a[d][2] * H[2][q];
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        opaque(w[a][d]);
        opaque(dw[a][d]);
      }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgradT(w, dw, inv_dx, a, bb, c, gw);
          const int t = t0 + (a * kTileY + bb) * kTileZ + c;
#pragma unroll
          for (int r = 0; r < 3; ++r)
            atomicAdd(&wt[r * kTileNodes + t], vp[r][0] * gw[0] + vp[r][1] * gw[1] + vp[r][2] * gw[2]);
          if (JAC) {
            double gq[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) gq[q] = Qm[0][q] * gw[0] + Qm[1][q] * gw[1] + Qm[2][q] * gw[2];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              double acc = 0.0;
#pragma unroll
              for (int c = 0; c < 3; ++c)
                acc += gq[c] * (Kd[JAC ? d : 0][c][0] * gq[0] + Kd[JAC ? d : 0][c][1] * gq[1] +
                                Kd[JAC ? d : 0][c][2] * gq[2]);
              atomicAdd(&wd[d * kTileNodes + t], acc);
            }
          }
        }
  }
  __syncthreads();
  double* out = stage + (size_t)g * kImpStage;
  for (int k = tid; k < kImpStage; k += 256)
    out[k] = tile[k] + tile[kImpStage + k] + tile[2 * kImpStage + k] + tile[3 * kImpStage + k];
  if (JAC) {
    double* dout = dstage + (size_t)g * kImpStage;
    for (int k = tid; k < kImpStage; k += 256)
      dout[k] = dtile[k] + dtile[kImpStage + k] + dtile[2 * kImpStage + k] + dtile[3 * kImpStage + k];
  }
  if (cell_order && tid < kCells) {
    const int cnt = ccount[tid];
    int rank = 0;
#pragma unroll
    for (int e2 = 0; e2 < kCells; ++e2) {
      const int c2 = ccount[e2];
      rank += (c2 > cnt || (c2 == cnt && e2 < tid)) ? 1 : 0;
    }
    cell_order[(size_t)g * kCells + rank] = (uint8_t)tid;
  }
}

// The Newton state of k_be_state<T, false> in three passes, so that the
// SVD runs in a kernel of its own at a higher occupancy than the fused
// kernel's one wave per SIMD:
//   k_be_grad    per group: grad(vn + dv) from the LDS tile, F = Fn + dt G Fn
//                into fb (9 rows of n, sorted order), the fractions, cells
//                and cell order of the operator data
//   k_be_svd     per particle (sorted order): the SVD of F, the operator's
//                U, Q, S, K, and vPFnT = V0 P(F) Fn^T back into fb
//   k_be_scatter per group: vPFnT grad w into the staged tiles
// The arithmetic and its order are those of k_be_state.
template <class T>
__global__ __launch_bounds__(256) void k_be_grad(ParticlesView P, HvpCache<T> cache,
                                                 const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ group_start,
                                                 const uint32_t* __restrict__ group_page,
                                                 const int32_t* __restrict__ page_slot,
                                                 const double* __restrict__ vn, const double* __restrict__ dv,
                                                 int64_t gstride, double inv_dx, double dt, double* __restrict__ fb,
                                                 uint8_t* __restrict__ cell_order)
{
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  __shared__ int ccount[kCells];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  if (tid < kCells) ccount[tid] = 0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  load_tile3(px, py, pz, page_slot, vn, dv, gstride, tv, nslot);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const int64_t S = P.stride, CS = cache.stride;
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const uint32_t ks = start + j;
    const double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3], d0[3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double xi = xp[d] * inv_dx;
      b[d] = int_floor(xi - 0.5);
      d0[d] = xi - (double)b[d];
      bspline_frac(d0[d], w[d], dw[d]);
    }
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    if (cell_order && j0 == 0) atomicAdd(&ccount[lx | (ly << 1) | (lz << 3)], 1);
    const int t0 = (lx * kTileY + ly) * kTileZ + lz;
    double G[3][3] = { { 0.0 } };
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgradT(w, dw, inv_dx, a, bb, c, gw);
          const int t = t0 + (a * kTileY + bb) * kTileZ + c;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double vv = tv[r * kTileNodes + t];
            G[r][0] += vv * gw[0];
            G[r][1] += vv * gw[1];
            G[r][2] += vv * gw[2];
          }
        }
    double Fn[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) Fn[r][c] = P.F[(c * 3 + r) * S + i];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double acc = Fn[r][c];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc += dt * G[r][k] * Fn[k][c];
        fb[(c * 3 + r) * CS + ks] = acc;
      }
#pragma unroll
    for (int d = 0; d < 3; ++d) cache.frac[d * CS + ks] = (T)d0[d];
    cache.cell[ks] = (uint32_t)(lx | (ly << 1) | (lz << 3));
  }
  __syncthreads();
  if (cell_order && tid < kCells) {
    const int cnt = ccount[tid];
    int rank = 0;
#pragma unroll
    for (int e2 = 0; e2 < kCells; ++e2) {
      const int c2 = ccount[e2];
      rank += (c2 > cnt || (c2 == cnt && e2 < tid)) ? 1 : 0;
    }
    cell_order[(size_t)g * kCells + rank] = (uint8_t)tid;
  }
}

template <class T>
__global__ __launch_bounds__(256) void k_be_svd(ParticlesView P, HvpCache<T> cache,
                                                const uint32_t* __restrict__ order, uint32_t n,
                                                double* __restrict__ fb)
{
  const uint32_t ks = blockIdx.x * 256u + threadIdx.x;
  if (ks >= n) return;
  const uint32_t i = order[ks];
  const int64_t S = P.stride, CS = cache.stride;
  M3<double> F, u, v;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) F.a[r][c] = fb[(c * 3 + r) * CS + ks];
  double s[3];
  svd3(F, u, s, v);
  double Fn[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) Fn[r][c] = P.F[(c * 3 + r) * S + i];
  const double mu = P.mu[i], lam = P.lam[i], vol = P.vol[i];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      cache.U[(c * 3 + r) * CS + ks] = (T)u.a[r][c];
      cache.Q[(c * 3 + r) * CS + ks] = (T)(Fn[r][0] * v.a[0][c] + Fn[r][1] * v.a[1][c] + Fn[r][2] * v.a[2][c]);
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) cache.S[d * CS + ks] = (T)(d == 2 && sizeof(T) == sizeof(float) ? s[1] + s[2] : s[d]);
  cache.K[ks] = (T)(2.0 * mu * vol);
  cache.K[CS + ks] = (T)(lam * vol);
  // P = U diag(2 mu (s - 1) + lam (J - 1) cof(s)) V^T
  const double J = s[0] * s[1] * s[2];
  const double cof[3] = { s[1] * s[2], s[0] * s[2], s[0] * s[1] };
  double ph[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) ph[d] = 2.0 * mu * (s[d] - 1.0) + lam * (J - 1.0) * cof[d];
  double Pm[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      Pm[r][c] = u.a[r][0] * ph[0] * v.a[c][0] + u.a[r][1] * ph[1] * v.a[c][1] + u.a[r][2] * ph[2] * v.a[c][2];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      fb[(c * 3 + r) * CS + ks] = vol * (Pm[r][0] * Fn[c][0] + Pm[r][1] * Fn[c][1] + Pm[r][2] * Fn[c][2]);
}

__global__ __launch_bounds__(256) void k_be_scatter(ParticlesView P, const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ group_start,
                                                    const uint32_t* __restrict__ group_page, int64_t cstride,
                                                    double inv_dx, const double* __restrict__ fb,
                                                    double* __restrict__ stage)
{
  __shared__ double tile[4 * kImpStage];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < 4 * kImpStage; k += 256) tile[k] = 0.0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  __syncthreads();
  double* wt = tile + (tid >> 6) * kImpStage;
  const uint32_t perm = (uint32_t)((tid & 31) * 8 + (tid >> 5));
  const int64_t S = P.stride;
  for (uint32_t j0 = 0; j0 < n; j0 += 256) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    const uint32_t ks = start + j;
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double xi = P.x[d * S + i] * inv_dx;
      b[d] = int_floor(xi - 0.5);
      bspline_frac(xi - (double)b[d], w[d], dw[d]);
    }
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    const int t0 = (lx * kTileY + ly) * kTileZ + lz;
    double vp[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) vp[r][c] = fb[(c * 3 + r) * cstride + ks];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int bb = 0; bb < 3; ++bb)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          wgradT(w, dw, inv_dx, a, bb, c, gw);
          const int t = t0 + (a * kTileY + bb) * kTileZ + c;
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

constexpr int kAB = 9;        // (a, b) stencil offsets; c is handled inside a task

// Node-major layout of the per-(cell, a, b, c) partial sums of the operator
// apply: tile node t = (tx, ty, tz) owns n(t) = na(tx) nb(ty) nb(tz)
// consecutive entries from off(t), one per stencil offset (a, b, c) whose
// source cell t - (a, b, c) lies in the page, ordered by a, then b, then c.
// Per axis: na = {1,2,2,1} (2 cells), nb = {1,2,3,3,2,1} (4 cells), with
// prefix sums {0,1,3,5} and {0,1,3,6,9,11}; 864 entries in all.
constexpr int hvp_na(int tx) { return (0x1221 >> (4 * tx)) & 15; }
constexpr int hvp_nb(int ty) { return (0x123321 >> (4 * ty)) & 15; }
constexpr int hvp_node_off(int tx, int ty, int tz)
{
  return ((0x5310 >> (4 * tx)) & 15) * 144 +
         hvp_na(tx) * (((0xB96310 >> (4 * ty)) & 15) * 12 + hvp_nb(ty) * ((0xB96310 >> (4 * tz)) & 15));
}
constexpr int hvp_slot(int tx, int ty, int tz, int a, int b, int c)
{
  return hvp_node_off(tx, ty, tz) +
         ((a - (tx > 1 ? tx - 1 : 0)) * hvp_nb(ty) + (b - (ty > 3 ? ty - 3 : 0))) * hvp_nb(tz) +
         (c - (tz > 3 ? tz - 3 : 0));
}

// slots of task (cell e, a, b)'s three c offsets, 10 bits each
struct HvpTaskSlots {
  uint32_t v[kCells * kAB];
};
constexpr HvpTaskSlots hvp_task_slots()
{
  HvpTaskSlots t{};
  for (int task = 0; task < kCells * kAB; ++task) {
    const int e = task / kAB, a = (task % kAB) / 3, b = task % 3;
    const int tx = (e & 1) + a, ty = ((e >> 1) & 3) + b, cz = (e >> 3) & 3;
    uint32_t v = 0;
    for (int c = 0; c < 3; ++c) v |= (uint32_t)hvp_slot(tx, ty, cz + c, a, b, c) << (10 * c);
    t.v[task] = v;
  }
  return t;
}
// every entry of the layout belongs to exactly one (task, c)
constexpr bool hvp_task_slots_valid()
{
  const HvpTaskSlots t = hvp_task_slots();
  int hits[kCells * kAB * 3] = {};
  for (int task = 0; task < kCells * kAB; ++task)
    for (int c = 0; c < 3; ++c) {
      const uint32_t k = (t.v[task] >> (10 * c)) & 1023u;
      if (k >= (uint32_t)(kCells * kAB * 3)) return false;
      ++hits[k];
    }
  for (int k = 0; k < kCells * kAB * 3; ++k)
    if (hits[k] != 1) return false;
  return true;
}
static_assert(hvp_task_slots_valid(), "node-major layout is a bijection");
static_assert(hvp_task_slots().v[0] == (0u | (2u << 10) | (5u << 20)), "slots of task (cell 0, a 0, b 0)");
__constant__ HvpTaskSlots kHvpTaskSlots = hvp_task_slots();

// particles per chunk of the operator apply, and the waves per SIMD its LDS
// allows (the register budget is held to it)
template <class T>
struct HvpChunk;
template <>
struct HvpChunk<float> {
  static constexpr int value = 256;
  static constexpr int threads = 256;
  static constexpr int waves = 4;
  static constexpr bool frame_early = true;  // singular frame requested before the gather
};
template <>
struct HvpChunk<double> {
  static constexpr int value = 256;
  static constexpr int threads = 256;
  static constexpr int waves = 2;
  static constexpr bool frame_early = false;  // after it: FP64 registers
};

// one particle's operator data with the cells of its neighbours in the
// chunk (~0 past the chunk's ends): particles are sorted by cell within a
// group, so a particle starts / ends its cell's run where they differ.
// The loads take a uniform base and a 32-bit byte offset (a field of the
// operator data spans less than 4 GiB: at most 2^29 - 1 particles).
template <class T>
struct HvpParticle {
  T fr[3], u[3][3], q[3][3], s[3], k2mu, klam;
  uint32_t cell, cprev, cnext;
};
template <class V>
__device__ __forceinline__ V ld_off(const V* base, uint32_t byte_off)
{
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ void hvp_load(const HvpCache<T>& cache, uint32_t ks, uint32_t j, uint32_t m,
                                         HvpParticle<T>& P)
{
  const int64_t CS = cache.stride;
  // the offsets are formed here, at the issue point, so that the compiler
  // does not hoist (and hold) the addresses of a prefetch
  uint32_t o = ks * (uint32_t)sizeof(T), oc = ks * 4u;
  asm volatile("" : "+v"(o), "+v"(oc));
  P.cell = ld_off(cache.cell, oc);
  P.cprev = j > 0 ? ld_off(cache.cell, oc - 4u) : ~0u;
  P.cnext = j + 1 < m ? ld_off(cache.cell, oc + 4u) : ~0u;
#pragma unroll
  for (int d = 0; d < 3; ++d) P.fr[d] = ld_off(cache.frac + d * CS, o);
}
template <class T>
__device__ __forceinline__ void hvp_load_frame(const HvpCache<T>& cache, uint32_t ks, HvpParticle<T>& P)
{
  const int64_t CS = cache.stride;
  uint32_t o = ks * (uint32_t)sizeof(T);
  asm volatile("" : "+v"(o));
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      P.u[r][c] = ld_off(cache.U + (c * 3 + r) * CS, o);
      P.q[r][c] = ld_off(cache.Q + (c * 3 + r) * CS, o);
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) P.s[d] = ld_off(cache.S + d * CS, o);
  P.k2mu = ld_off(cache.K, o);
  P.klam = ld_off(cache.K + CS, o);
}

// work counters of the persistent operator apply: one per XCD, each in its
// own 256-byte line
constexpr int kHvpCounterStride = 64;
constexpr int kHvpCounters = 8;

// predicted cost of a group in the operator apply, in particle units: a
// fixed part (tile, node sums, barriers), one per particle and a phase-2
// pass per further chunk of 256 particles (fitted to per-group timings of
// both precisions on MI355X)
constexpr uint32_t kHvpGroupCost = 384, kHvpChunkCost = 256;
__device__ __forceinline__ uint32_t hvp_group_cost(uint32_t n)
{
  return kHvpGroupCost + n + (n > 0 ? kHvpChunkCost * ((n - 1) / 256) : 0u);
}

// block b of a persistent launch runs on XCD x = b mod 8 as its l-th block;
// XCD x owns groups [xb[x], xb[x + 1]) and nbx blocks
struct XcdGroups {
  uint32_t lo, hi, nbx, l;
};
__device__ __forceinline__ XcdGroups xcd_groups(uint32_t b, uint32_t nblocks, const uint32_t* xb)
{
  const uint32_t x = b & 7u, q = nblocks >> 3, r = nblocks & 7u;
  return { xb[x], xb[x + 1], q + (x < r ? 1u : 0u), b >> 3 };
}

// XCD ranges of the persistent operator apply, of equal predicted cost in
// proportion to each XCD's blocks, for nblk32 (FP32) and nblk64 (FP64)
// blocks: bounds[9 p + x] is the first group of XCD x; and the order in
// which each range is handed out, largest groups first (by 16-particle
// bucket; the order within a bucket is arbitrary and does not affect the
// result, every group writing its own staged tile): perm[p A + k].  One
// block of 1024.
__global__ __launch_bounds__(1024) void k_xcd_split(const uint32_t* __restrict__ group_start, uint32_t A,
                                                    uint32_t nblk32, uint32_t nblk64, uint32_t* __restrict__ bounds,
                                                    uint32_t* __restrict__ perm)
{
  using u64 = unsigned long long;
  __shared__ u64 wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  u64 t = 0;
  for (uint32_t g = tid; g < A; g += 1024) t += hvp_group_cost(group_start[g + 1] - group_start[g]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if (lane == 0) wsum[wid] = t;
  if (tid < 18) bounds[tid] = (tid % 9 == 8) ? A : 0u;
  __syncthreads();
  u64 total = 0;
  for (int w = 0; w < 16; ++w) total += wsum[w];
  u64 T[2][7];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const uint32_t nb = p ? nblk64 : nblk32, q = nb >> 3, r = nb & 7u;
#pragma unroll
    for (int x = 1; x < 8; ++x) T[p][x - 1] = total * (u64)(x * q + ((uint32_t)x < r ? (uint32_t)x : r)) / nb;
  }
  __syncthreads();
  // the group whose cost interval (excl, incl] holds a target closes the range before it
  u64 carry = 0;
  for (uint32_t base = 0; base < A; base += 1024) {
    const uint32_t g = base + tid;
    const u64 c = g < A ? hvp_group_cost(group_start[g + 1] - group_start[g]) : 0;
    u64 v = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const u64 y = __shfl_up(v, off, 64);
      if (lane >= off) v += y;
    }
    if (lane == 63) wsum[wid] = v;
    __syncthreads();
    u64 wpre = 0, tile = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wid) wpre += wsum[w];
      tile += wsum[w];
    }
    const u64 incl = carry + wpre + v, excl = incl - c;
    if (g < A) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int x = 1; x < 8; ++x)
          if (excl < T[p][x - 1] && T[p][x - 1] <= incl) bounds[9 * p + x] = g + 1;
    }
    carry += tile;
    __syncthreads();
  }
  // largest-first order within each range: counting sort on (range, 31 - size bucket)
  __shared__ uint32_t cnt[256];
  __shared__ uint32_t xb[9];
  for (int p = 0; p < 2; ++p) {
    if (tid < 256) cnt[tid] = 0u;
    if (tid < 9) xb[tid] = bounds[9 * p + tid];
    __syncthreads();
    for (uint32_t g = tid; g < A; g += 1024) {
      uint32_t x = 0;
      while (x < 7u && g >= xb[x + 1]) ++x;
      const uint32_t nsz = group_start[g + 1] - group_start[g], bk = nsz / 16u < 31u ? nsz / 16u : 31u;
      atomicAdd(&cnt[x * 32u + (31u - bk)], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t run = 0;
      for (int k = 0; k < 256; ++k) {
        const uint32_t c = cnt[k];
        cnt[k] = run;
        run += c;
      }
    }
    __syncthreads();
    for (uint32_t g = tid; g < A; g += 1024) {
      uint32_t x = 0;
      while (x < 7u && g >= xb[x + 1]) ++x;
      const uint32_t nsz = group_start[g + 1] - group_start[g], bk = nsz / 16u < 31u ? nsz / 16u : 31u;
      perm[(size_t)p * A + atomicAdd(&cnt[x * 32u + (31u - bk)], 1u)] = g;
    }
    __syncthreads();
  }
}

// requests of the next group gn: bounds, this thread's tile node value(s),
// first-chunk particle (cells and fractions) and cell order entry
template <class T>
__device__ __forceinline__ void hvp_issue_next(const HvpCache<T>& cache, const uint32_t* __restrict__ group_start,
                                               const int32_t* __restrict__ group_nslot, const T* __restrict__ qvec,
                                               const T* __restrict__ pvec, int64_t gstride, uint32_t gn, bool tnode,
                                               const uint8_t* __restrict__ cell_order, uint32_t& nstart, uint32_t& nn,
                                               int32_t& nsl, int& nord, HvpParticle<T>& P1, T tq[3], T tp[3])
{
  constexpr int CH = HvpChunk<T>::value;
  const int tid = threadIdx.x;
  if (tid < kCells) nord = cell_order[(size_t)gn * kCells + tid];
  nstart = group_start[gn];
  nn = group_start[gn + 1] - nstart;
  const uint32_t m1 = nn < (uint32_t)CH ? nn : (uint32_t)CH;
  if ((uint32_t)tid < m1) hvp_load(cache, nstart + tid, tid, m1, P1);
  if (tnode) {
    int tsrc, telem;
    tile_node_source(tid, tsrc, telem);
    nsl = group_nslot[8 * (size_t)gn + tsrc];
    if (nsl >= 0) {
      const int64_t node = (int64_t)nsl * 32 + telem;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        tp[d] = pvec[d * gstride + node];
        if (qvec) tq[d] = qvec[d * gstride + node];
      }
    }
  }
}

// Hessian-vector product, particle part: ds = V0 dP(grad p Fn) Fn^T rasterised
// as ds grad w (MpmForceBase.cpp:258-302, FBasedMpmForceHelper.cpp:132-155),
// in precision T from the sorted operator data.  No atomics, fixed order:
//   phase 1  thread per particle: separable gather of grad p from the LDS
//            tile, ds from the singular-frame differential; ds / dx^2 and
//            the 1-D weights go to LDS, the cell runs' ends to cbeg / cend
//   phase 2  task per (cell, a, b): sums ds grad w over the cell's particles
//            in index order for the three c offsets (separable form
//            gw = (dw_a w_b w_c, w_a dw_b w_c, w_a w_b dw_c) / dx) into its
//            three entries of the node-major layout; the tasks go to the
//            lanes by the cell order of k_be_state (cells with most
//            particles first), so that a wave's lanes loop about as often
//   phase 3  thread per tile node: sums its <= 18 partials, a contiguous run of
//            the node-major layout, into the staged tile
// Persistent blocks, one wave of them resident, dealt to the XCDs as
// xcd_groups describes (ranges and hand-out order from k_xcd_split); further
// positions in the XCD's order come from its work counter in wctr (zero at
// launch; the kernel that follows resets them), so that the load balances
// within an XCD.  The counter is read three groups ahead, the order two and
// the next group's bounds, page slots, tile values and first-chunk cells and
// fractions one group ahead, so that they are in flight while this group's
// phases 2 and 3 run.  A particle's singular frame
// is requested at the start of phase 1 (FP32) or after the gather (FP64),
// within the register budget of the resident waves.
// qvec == nullptr: applies the operator to pvec itself; otherwise to the CG
// direction qvec + beta pvec (beta from sc)
template <class T>
__global__ __launch_bounds__(HvpChunk<T>::threads, HvpChunk<T>::waves) void k_hvp(HvpCache<T> cache, const uint32_t* __restrict__ group_start,
                                             const int32_t* __restrict__ group_nslot, const T* __restrict__ qvec,
                                             const T* __restrict__ pvec, const double* sc,
                                             int64_t gstride, T inv_dx, T* __restrict__ stage,
                                             int32_t* __restrict__ ci, const uint32_t* __restrict__ xb,
                                             const uint32_t* __restrict__ order, uint32_t* __restrict__ wctr,
                                             const double* __restrict__ upd_partial, uint32_t n_upd, int cg_iter,
                                             double* __restrict__ scw, const uint8_t* __restrict__ cell_order)
{
  constexpr int CH = HvpChunk<T>::value;
  constexpr int kTasks = kCells * kAB;
  constexpr int NT = HvpChunk<T>::threads;
  static_assert(kTasks <= 2 * NT && CH <= NT, "at most two task rounds, one particle per thread");
  static_assert(kTileNodes <= NT, "one tile node per thread");
  const XcdGroups xg = xcd_groups(blockIdx.x, gridDim.x, xb);
  if (xg.lo + xg.l >= xg.hi) return;
  uint32_t g = order[xg.lo + xg.l];
  uint32_t* const xctr = wctr + kHvpCounterStride * (blockIdx.x & 7u);
  __shared__ uint32_t next_group;
  __shared__ T tv[3 * kTileNodes];
  // ds (9, row-major), w (3x3), dw (3x3); rows padded by one element so that
  // the rows one particle's tasks read together fall in different banks
  __shared__ T pds[27][CH + 1];
  __shared__ T contrib[3][kTasks * 3];
  __shared__ int cbeg[kCells], cend[kCells];
  __shared__ int cell_rank[kCells];  // the group's cells in task-deal order
  const int tid = threadIdx.x;
  const T inv_dx2 = inv_dx * inv_dx;
  const bool tnode = tid < kTileNodes;
  // thread 0's pipeline over the XCD's hand-out order: the next group's
  // number (read a group ago) and the position after it (counted a group
  // ago); ~0u past the range
  uint32_t ahead_g = ~0u, ahead_k = ~0u;
  if (tid == 0) {
    const uint32_t k1 = xg.lo + xg.nbx + atomicAdd(xctr, 1u);
    ahead_g = k1 < xg.hi ? order[k1] : ~0u;
    if (k1 < xg.hi) ahead_k = xg.lo + xg.nbx + atomicAdd(xctr, 1u);
  }

  // prologue: the first group's tile and first-chunk particle data,
  // requested before beta is summed (the work counters this touches on a
  // finished solve are reset by the kernel that follows)
  uint32_t start = group_start[g];
  uint32_t n = group_start[g + 1] - start;
  // this thread's particle of the current chunk
  HvpParticle<T> P;
  int32_t sl0 = -1;
  T tq0[3], tp0[3];  // read only where loaded (sl0 >= 0)
  {
    const uint32_t m0 = n < (uint32_t)CH ? n : (uint32_t)CH;
    if ((uint32_t)tid < m0) {
      hvp_load(cache, start + tid, tid, m0, P);
    }
    if (tnode) {
      int tsrc, telem;
      tile_node_source(tid, tsrc, telem);
      sl0 = group_nslot[8 * (size_t)g + tsrc];
      if (sl0 >= 0) {
        const int64_t node = (int64_t)sl0 * 32 + telem;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          tp0[d] = pvec[d * gstride + node];
          if (qvec) tq0[d] = qvec[d * gstride + node];
        }
      }
    }
  }
  // a finished solve's launches stop here (its flag is read beside the
  // requests above rather than before them)
  if (ci && ci[CI_DONE]) return;
  T beta = (T)0;
  if (qvec && cg_iter > 0) {
    // ConjugateGradient.h:136-143 for the update of iteration cg_iter - 1:
    // z^T r from its block partials, beta, the count and the termination
    // test, alike in every block; block 0 records them
    __shared__ double red[5];
    const double ztr = block_total256(upd_partial, n_upd, red);
    const double b64 = ztr / sc[ztr_slot(cg_iter - 1)];
    const bool done = sqrt(ztr) < sc[SC_LOCAL_TOL];
    if (blockIdx.x == 0 && tid == 0) {
      scw[ztr_slot(cg_iter)] = ztr;
      scw[SC_BETA] = b64;
      ci[CI_ITERS] += 1;
      if (done) ci[CI_DONE] = 1;
    }
    if (done) return;
    beta = (T)b64;
  }
  if (tnode) {
#pragma unroll
    for (int d = 0; d < 3; ++d) tv[d * kTileNodes + tid] = sl0 >= 0 ? (qvec ? tq0[d] + beta * tp0[d] : tp0[d]) : (T)0;
  }
  if (tid < kCells) {
    cbeg[tid] = CH;
    cend[tid] = 0;
    cell_rank[tid] = cell_order[(size_t)g * kCells + tid];
  }
  __syncthreads();

  for (;;) {
    // the next group (read a group ago), visible to the block after this
    // group's first barrier; and the one after it
    if (tid == 0) {
      next_group = ahead_g;
      ahead_g = ahead_k < xg.hi ? order[ahead_k] : ~0u;
      if (ahead_k < xg.hi) ahead_k = xg.lo + xg.nbx + atomicAdd(xctr, 1u);
    }
    uint32_t gn = 0, nstart = 0, nn = 0;
    bool more = false;
    int32_t nsl = -1;
    int nord = 0;
    HvpParticle<T> P1;
    T tq[3], tp[3];  // read only where loaded (more && nsl >= 0)
    bool issued = false;
    if (n == 0) {  // an empty group stages zeros
      for (int k = tid; k < 3 * kTasks * 3; k += NT) (&contrib[0][0])[k] = (T)0;
      __syncthreads();
    }
    for (uint32_t c0 = 0; c0 < n; c0 += CH) {
      const uint32_t m = n - c0 < (uint32_t)CH ? n - c0 : (uint32_t)CH;
      if (c0 > 0) {
        if (tid < kCells) {
          cbeg[tid] = CH;
          cend[tid] = 0;
        }
        __syncthreads();
      }
      if ((uint32_t)tid < m) {
        if (HvpChunk<T>::frame_early) hvp_load_frame(cache, start + c0 + tid, P);
        const uint32_t cell = P.cell;
        const int t0 = ((int)(cell & 1u) * kTileY + (int)((cell >> 1) & 3u)) * kTileZ + (int)((cell >> 3) & 3u);
        T w[3][3], dw[3][3];
#pragma unroll
        for (int d = 0; d < 3; ++d) bspline_frac(P.fr[d], w[d], dw[d]);
        // separable gather: G[r][0] = sum dw_a w_b w_c p, G[r][1] = sum w_a dw_b w_c p,
        // G[r][2] = sum w_a w_b dw_c p (the 1/dx factor is applied to ds)
        T G[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T U[3], V[3], W[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            U[a] = (T)0;
            V[a] = (T)0;
            W[a] = (T)0;
#pragma unroll
            for (int bb = 0; bb < 3; ++bb) {
              const T* row = tv + r * kTileNodes + t0 + (a * kTileY + bb) * kTileZ;
              const T p0 = row[0], p1 = row[1], p2 = row[2];
              const T sc = p0 * w[2][0] + p1 * w[2][1] + p2 * w[2][2];
              const T tc = p0 * dw[2][0] + p1 * dw[2][1] + p2 * dw[2][2];
              U[a] += sc * w[1][bb];
              V[a] += sc * dw[1][bb];
              W[a] += tc * w[1][bb];
            }
          }
          G[r][0] = dw[0][0] * U[0] + dw[0][1] * U[1] + dw[0][2] * U[2];
          G[r][1] = w[0][0] * V[0] + w[0][1] * V[1] + w[0][2] * V[2];
          G[r][2] = w[0][0] * W[0] + w[0][1] * W[1] + w[0][2] * W[2];
        }
        if (!HvpChunk<T>::frame_early) hvp_load_frame(cache, start + c0 + tid, P);
        const T(&u)[3][3] = P.u;
        const T(&q)[3][3] = P.q;
        const T* s = P.s;
        const T k2mu = P.k2mu, klam = P.klam;
        // M = U^T (G Q)
        T T1[3][3], M[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) T1[r][c] = G[r][0] * q[0][c] + G[r][1] * q[1][c] + G[r][2] * q[2][c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) M[r][c] = u[0][r] * T1[0][c] + u[1][r] * T1[1][c] + u[2][r] * T1[2][c];
        T H[3][3];
        if constexpr (sizeof(T) == sizeof(float))
          fcr_frame_differential_t(s, k2mu, klam, M, H);
        else
          fcr_frame_differential(s, k2mu, klam, M, H);
        // ds = U H Q^T, stored times 1/dx^2 (the gather's and the scatter's 1/dx)
        T T2[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) T2[r][c] = u[r][0] * H[0][c] + u[r][1] * H[1][c] + u[r][2] * H[2][c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            pds[r * 3 + c][tid] = inv_dx2 * (T2[r][0] * q[c][0] + T2[r][1] * q[c][1] + T2[r][2] * q[c][2]);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            pds[9 + 3 * d + a][tid] = w[d][a];
            pds[18 + 3 * d + a][tid] = dw[d][a];
          }
        if (P.cprev != cell) cbeg[cell] = tid;
        if (P.cnext != cell) cend[cell] = tid + 1;
      }
      __syncthreads();
      if (c0 + CH >= n) {
        // last chunk: the next group's data are requested here and consumed
        // after phase 3
        gn = next_group;
        more = gn != ~0u;
        if (more) hvp_issue_next(cache, group_start, group_nslot, qvec, pvec, gstride, gn, tnode, cell_order, nstart,
                                 nn, nsl, nord, P1, tq, tp);
        issued = true;
      }
      // tasks past the first NT (the smallest cells) go to the last threads,
      // whose first tasks hold the smaller cells of the first round
      constexpr int kExtra = kTasks - NT;
      for (int k = tid; k < kTasks; k = (k < NT && tid >= NT - kExtra) ? tid + kExtra : kTasks) {
        const int rk = k / kAB, ab = k - rk * kAB;
        const int e = cell_rank[rk], task = e * kAB + ab;
        const int a = ab / 3, bb = ab - 3 * a;
        T acc[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) acc[c][r] = (T)0;
        const int jb = cbeg[e], je = cend[e];
        for (int jj = jb; jj < je; ++jj) {
          const T wa = pds[9 + a][jj], dwa = pds[18 + a][jj];
          const T wb = pds[12 + bb][jj], dwb = pds[21 + bb][jj];
          const T x0 = dwa * wb, x1 = wa * dwb, x2 = wa * wb;
          T A[3], B[3];
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            A[r] = pds[3 * r][jj] * x0 + pds[3 * r + 1][jj] * x1;
            B[r] = pds[3 * r + 2][jj] * x2;
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const T wc = pds[15 + c][jj], dwc = pds[24 + c][jj];
#pragma unroll
            for (int r = 0; r < 3; ++r) acc[c][r] = fma(B[r], dwc, fma(A[r], wc, acc[c][r]));
          }
        }
        // the first chunk stores every entry (zeros for an empty cell); later
        // chunks add to it
        const uint32_t ts = kHvpTaskSlots.v[task];
        if (c0 == 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int k = (int)((ts >> (10 * c)) & 1023u);
#pragma unroll
            for (int r = 0; r < 3; ++r) contrib[r][k] = acc[c][r];
          }
        } else if (jb < je) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int k = (int)((ts >> (10 * c)) & 1023u);
#pragma unroll
            for (int r = 0; r < 3; ++r) contrib[r][k] += acc[c][r];
          }
        }
      }
      if (c0 + CH < n) {  // the next chunk's particle
        const uint32_t c1 = c0 + CH;
        const uint32_t m2 = n - c1 < (uint32_t)CH ? n - c1 : (uint32_t)CH;
        if ((uint32_t)tid < m2) {
          hvp_load(cache, start + c1 + tid, tid, m2, P);
        }
      }
      __syncthreads();
    }
    if (!issued) {  // an empty group (no chunk ran; its zeroing barrier passed)
      gn = next_group;
      more = gn != ~0u;
      if (more) hvp_issue_next(cache, group_start, group_nslot, qvec, pvec, gstride, gn, tnode, cell_order, nstart,
                               nn, nsl, nord, P1, tq, tp);
    }
    // phase 3: each tile node sums its run of partials in (a, b, c) order
    if (tnode) {
      const int tx = tid / (kTileY * kTileZ), ty = (tid / kTileZ) % kTileY, tz = tid % kTileZ;
      const int k0 = hvp_node_off(tx, ty, tz), nk = hvp_na(tx) * hvp_nb(ty) * hvp_nb(tz);
      T acc0 = (T)0, acc1 = (T)0, acc2 = (T)0;
      for (int k = k0; k < k0 + nk; ++k) {
        acc0 += contrib[0][k];
        acc1 += contrib[1][k];
        acc2 += contrib[2][k];
      }
      T* out = stage + (size_t)g * kImpStage;
      out[tid] = acc0;
      out[kTileNodes + tid] = acc1;
      out[2 * kTileNodes + tid] = acc2;
    }
    if (!more) break;
    // the next group's tile (tv is free once this group's phase 1 is done)
    // and empty cell runs
    if (tnode) {
#pragma unroll
      for (int d = 0; d < 3; ++d) tv[d * kTileNodes + tid] = nsl >= 0 ? (qvec ? tq[d] + beta * tp[d] : tp[d]) : (T)0;
    }
    if (tid < kCells) {
      cbeg[tid] = CH;
      cend[tid] = 0;
      cell_rank[tid] = nord;
    }
    __syncthreads();
    g = gn;
    start = nstart;
    n = nn;
    P = P1;
  }
}

// buildInitialDvAndVnForNewton (MpmSimulationBase.cpp:1105-1136)
__global__ __launch_bounds__(256) void k_be_init(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                                 const double* __restrict__ gm, const double* __restrict__ gv,
                                                 NodeVec3 vn, NodeVec3 dv, double dx, double dt, double gx, double gy,
                                                 double gz, ColliderSet cs, uint8_t* __restrict__ collided)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const double m = gm[k];
  if (m == 0.0) {
    for (int d = 0; d < 3; ++d) vn(d, k) = dv(d, k) = 0.0;
    if (collided) collided[k] = 0;
    return;
  }
  const int64_t S = vn.stride;
  const double v[3] = { gv[k], gv[S + k], gv[2 * S + k] };
  double vi[3] = { v[0], v[1], v[2] };
  double xi[3];
  node_pos(node_coord(slot_page, k), dx, xi);
  const bool hit = multi_object_collision(cs, 0, xi, vi);
  if (collided) collided[k] = hit ? 1 : 0;
  const double g[3] = { gx, gy, gz };
  for (int d = 0; d < 3; ++d) {
    dv(d, k) = hit ? vi[d] - v[d] : g[d] * dt;
    vn(d, k) = v[d];
  }
}

// computeResidual (BackwardEuler.h:101-124): r = dt m g - dt sum vPFnT grad w - m dv,
// projected; sc[SC_NORM2] = sum |r|^2 / m (computeNorm, BackwardEuler.h:127-140)
__global__ __launch_bounds__(256) void k_be_residual(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                                     const int32_t* __restrict__ slot_src,
                                                     const double* __restrict__ stage, const double* __restrict__ gm,
                                                     NodeVec3 dv, NodeVec3 r, double dx, double dt, double gx,
                                                     double gy, double gz, ColliderSet cs,
                                                     double* __restrict__ partial)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    double rv[3] = { 0.0, 0.0, 0.0 };
    if (m != 0.0) {
      const NodeCoord nc = node_coord(slot_page, k);
      double f[kImpFields];
      sum_tiles_src<kImpFields>(stage, slot_src + 8 * (size_t)(k >> 5), nc.ex, nc.ey, nc.ez, f);
      const double g[3] = { gx, gy, gz };
      for (int d = 0; d < 3; ++d) rv[d] = dt * g[d] * m - dt * f[d] - m * dv(d, k);
      double xi[3];
      node_pos(nc, dx, xi);
      project_node(cs, xi, rv);
      contrib = (rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]) / m;
    }
    for (int d = 0; d < 3; ++d) r(d, k) = rv[d];
  }
  block_partial(contrib, partial);
}

// sums nb partials in index order into sc[slot]; skipped once CG is done
__global__ __launch_bounds__(256) void k_reduce(const double* __restrict__ partial, uint32_t nb, double* sc, int slot,
                                                const int32_t* ci)
{
  __shared__ double smem[4];
  if (ci && ci[CI_DONE]) return;
  const double acc = thread_partials256(partial, nb);
  const double s = block_sum256(acc, smem);
  if (threadIdx.x == 0) sc[slot] = s;
}

// JacobiPreconditioner (KrylovSolvers.h:32-57) on the operator's diagonal
// a_(k,d) = m_k + dt^2 (staged particle sums): pinv = 1 / max(16 eps, |a|),
// with the node's largest |a| on every axis of a node a collider touched at
// the step start, so that z = pinv r stays in the projected space; nodes
// without mass hold 0
template <class T>
__global__ __launch_bounds__(256) void k_jacobi(uint32_t nn, const int32_t* __restrict__ slot_src,
                                                const double* __restrict__ dstage, const double* __restrict__ gm,
                                                const uint8_t* __restrict__ collided, double dt, NodeVec<T> pinv)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  const double m = gm[k];
  if (m == 0.0) {
    for (int d = 0; d < 3; ++d) pinv(d, k) = (T)0;
    return;
  }
  const uint32_t e = k & 31u;
  double f[3];
  sum_tiles_src<3>(dstage, slot_src + 8 * (size_t)(k >> 5), (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
  double a[3], mx = 0.0;
  for (int d = 0; d < 3; ++d) {
    a[d] = fabs(m + dt * dt * f[d]);
    mx = a[d] > mx ? a[d] : mx;
  }
  const bool iso = collided[k] != 0;
  for (int d = 0; d < 3; ++d) pinv(d, k) = (T)(1.0 / fmax(16.0 * 2.220446049250313e-16, iso ? mx : a[d]));
}

// CG start (ConjugateGradient.h:105-119): x = 0, r = b, z = r / m, p = z,
// z^T r and the local tolerance min(rel |r0|_P, tol); beta = 0 so that the
// first direction built on the fly is p = z.  z = pinv r (Jacobi) when pinv
// is given, else r / m.
template <class T>
__global__ __launch_bounds__(256) void k_cg_init(uint32_t nn, const double* __restrict__ gm, NodeVec3 b, NodeVec<T> x,
                                                 NodeVec<T> r, NodeVec<T> q, NodeVec<T> p, NodeVec<T> pinv,
                                                 double* __restrict__ partial)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  double contrib = 0.0;
  if (k < nn) {
    const double m = gm[k];
    for (int d = 0; d < 3; ++d) {
      const T rv = (T)b(d, k);
      const T qv = m != 0.0 ? (pinv.p ? rv * pinv(d, k) : rv / (T)m) : (T)0;
      x(d, k) = (T)0;
      r(d, k) = rv;
      q(d, k) = qv;
      p(d, k) = qv;
      contrib += (double)rv * (double)qv;
    }
  }
  block_partial(contrib, partial);
}

// direction p = z + beta p, Ap = M p + dt^2 sum ds grad w (projected),
// alpha = z^T r / p^T A p  (ConjugateGradient.h:127-133, BackwardEuler.h:322-351)
template <class T>
__global__ __launch_bounds__(256) void k_cg_ap(uint32_t nn, const uint32_t* __restrict__ slot_page,
                                               const int32_t* __restrict__ slot_src, const T* __restrict__ stage,
                                               const double* __restrict__ gm, NodeVec<T> q, NodeVec<T> p,
                                               NodeVec<T> ap, double dx, double dt, ColliderSet cs,
                                               int project, double* __restrict__ partial, const double* sc,
                                               const int32_t* ci, uint32_t* __restrict__ hvp_ctr,
                                               const uint32_t* __restrict__ act, uint32_t nact)
{
  // the operator apply this kernel follows has drained its work counters
  if (blockIdx.x == 0 && threadIdx.x < kHvpCounters) hvp_ctr[kHvpCounterStride * threadIdx.x] = 0u;
  // act: the nodes with mass (ascending), when the caller has listed them;
  // otherwise every slot node
  // the first node's operands and its page's source groups are requested
  // before the flag and beta are known, so that the staged sums are one
  // dependent request away from the launch
  // A node without mass holds zeros in every CG vector (the start sets
  // them so and nothing adds to them), so its direction and A p are not
  // touched in a solve; the plain product (no q) still writes its zero A p.
  const uint32_t t0 = blockIdx.x * 256u + threadIdx.x, tstep = gridDim.x * 256u;
  const uint32_t count = act ? nact : nn;
  double m = 0.0;
  int32_t src8[8];
  T qv[3], pv0[3];
  auto fetch = [&](uint32_t k) {
    m = gm[k];
    load_src8(slot_src, k, src8);
    if (m != 0.0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        pv0[d] = p(d, k);
        if (q.p) qv[d] = q(d, k);
      }
    }
  };
  if (t0 < count) fetch(act ? act[t0] : t0);
  if (ci && ci[CI_DONE]) return;
  double contrib = 0.0;
  const T beta = q.p ? (T)sc[SC_BETA] : (T)0;
  for (uint32_t t = t0; t < count; t += tstep) {
    const uint32_t k = act ? act[t] : t;
    if (t != t0) fetch(k);
    if (m == 0.0) {
      if (!q.p)
        for (int d = 0; d < 3; ++d) ap(d, k) = (T)0;
      continue;
    }
    T pv[3];
    if (q.p) {
      for (int d = 0; d < 3; ++d) {
        pv[d] = qv[d] + beta * pv0[d];
        p(d, k) = pv[d];
      }
    } else {
      for (int d = 0; d < 3; ++d) pv[d] = pv0[d];
    }
    T av[3] = { (T)0, (T)0, (T)0 };
    {
      const uint32_t e = k & 31u;
      T f[kImpFields];
      sum_tiles_src8<kImpFields>(stage, src8, (int)(e >> 4), (int)((e >> 2) & 3u), (int)(e & 3u), f);
      const T dt2 = (T)(dt * dt);
      for (int d = 0; d < 3; ++d) av[d] = (T)m * pv[d] + dt2 * f[d];
      if (project) {
        double xi[3], a64[3] = { (double)av[0], (double)av[1], (double)av[2] };
        node_pos(node_coord(slot_page, k), dx, xi);
        project_node(cs, xi, a64);
        for (int d = 0; d < 3; ++d) av[d] = (T)a64[d];
      }
    }
    for (int d = 0; d < 3; ++d) {
      ap(d, k) = av[d];
      contrib += (double)av[d] * (double)pv[d];
    }
  }
  block_partial(contrib, partial);
}

// x += alpha p, r -= alpha Ap, z = r / m or pinv r (ConjugateGradient.h:127-135) with
// alpha = z^T r / p^T A p, p^T A p summed from its block partials in every
// block; this kernel's z^T r partials go to upd_partial (grid-stride, one
// partial per block)
template <class T>
__global__ __launch_bounds__(256) void k_cg_update(uint32_t nn, const double* __restrict__ gm, NodeVec<T> x,
                                                   NodeVec<T> r, NodeVec<T> q, NodeVec<T> p, NodeVec<T> ap,
                                                   NodeVec<T> pinv, double* __restrict__ sc,
                                                   const double* __restrict__ pap_partial,
                                                   uint32_t n_pap, double* __restrict__ upd_partial,
                                                   const int32_t* ci, int cg_iter)
{
  // this thread's first node's operands are requested before alpha is
  // known (they do not depend on it), so that the two waits overlap
  const uint32_t k0 = blockIdx.x * 256u + threadIdx.x, kstep = gridDim.x * 256u;
  double m = 0.0;
  T xv[3], pv[3], rv[3], av[3], jv[3];
  // a node without mass keeps its zeros (see k_cg_ap)
  auto fetch = [&](uint32_t k) {
    m = gm[k];
    if (m != 0.0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        xv[d] = x(d, k);
        pv[d] = p(d, k);
        rv[d] = r(d, k);
        av[d] = ap(d, k);
        jv[d] = pinv.p ? pinv(d, k) : (T)0;
      }
    }
  };
  if (k0 < nn) fetch(k0);
  if (ci[CI_DONE]) return;
  __shared__ double red[5];
  const double pap = block_total256(pap_partial, n_pap, red);
  const double a64 = sc[ztr_slot(cg_iter)] / pap;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[SC_PAP] = pap;
    sc[SC_ALPHA] = a64;
  }
  const T alpha = (T)a64;
  double contrib = 0.0;
  for (uint32_t k = k0; k < nn; k += kstep) {
    if (k != k0) fetch(k);
    if (m == 0.0) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x(d, k) = xv[d] + pv[d] * alpha;
      const T rn = rv[d] - av[d] * alpha;
      r(d, k) = rn;
      const T qv = pinv.p ? rn * jv[d] : rn / (T)m;
      q(d, k) = qv;
      contrib += (double)qv * (double)rn;
    }
  }
  block_partial(contrib, upd_partial);
}

// one-block reductions of the partials with the CG scalar updates
enum { RED_STORE = 0, RED_CG_INIT = 1, RED_FINISH = 2 };
__global__ __launch_bounds__(256) void k_reduce_cg(const double* __restrict__ partial, uint32_t nb, double* sc,
                                                   int32_t* ci, int mode, int slot, double rel, double tol)
{
  __shared__ double smem[4];
  if (mode != RED_STORE && mode != RED_CG_INIT && ci[CI_DONE]) return;
  const double acc = thread_partials256(partial, nb);
  const double v = block_sum256(acc, smem);
  if (threadIdx.x != 0) return;
  if (mode == RED_STORE) {
    sc[slot] = v;
  } else if (mode == RED_CG_INIT) {  // ConjugateGradient.h:115-119
    const double norm = sqrt(v);
    const double lt = rel * norm < tol ? rel * norm : tol;
    sc[SC_ZTR] = v;
    sc[SC_BETA] = 0.0;
    sc[SC_LOCAL_TOL] = lt;
    ci[CI_DONE] = norm < lt ? 1 : 0;
  } else {  // ConjugateGradient.h:136-143 for the last update of a solve (slot = its iteration)
    sc[ztr_slot(slot + 1)] = v;
    ci[CI_ITERS] += 1;
    if (sqrt(v) < sc[SC_LOCAL_TOL]) ci[CI_DONE] = 1;
  }
}

// Newton update dv += step (the FP64 outer iteration refines the FP32 solve)
template <class T>
__global__ __launch_bounds__(256) void k_axpy3(uint32_t nn, NodeVec3 y, NodeVec<T> x)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nn) return;
  for (int d = 0; d < 3; ++d) y(d, k) += (double)x(d, k);
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
