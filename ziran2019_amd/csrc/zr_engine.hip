// zr_engine.hip -- MI355X (gfx950) MPM time-step engine behind the C-ABI of
// include/zr_mpm.h.
//
// Step structure (one HIP stream, FP64 plain path):
//   bin      : per particle base node -> SPGrid page code; per-page counts via
//              atomics on a dense page table; occupancy bitmap
//   groups   : bitmap popcount scan -> sorted list of particle pages (groups)
//   scatter  : particles into their group segment, then an in-LDS bitonic sort
//              of each segment by (cell, index): the reference's sorted keys
//   pages    : 8 neighbour pages per group, first-touch order by atomicMin on
//              the touch index -> page slots (the reference's block list)
//   P2G      : one workgroup per group scatters its particles into per-wave
//              LDS tiles covering the 4x6x6 stencil footprint of the page,
//              then writes the tile to a staging buffer (no global atomics)
//   grid     : one thread per node of each slot sums the <= 8 overlapping
//              tiles in a fixed order, normalises, applies gravity/forces and
//              colliders (+ trial collision)
//   G2P      : one workgroup per group loads v* of the footprint into LDS and
//              gathers per particle: v, C, x, F update, CFL flags
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <string>
#include <vector>
#include <array>
#include <mutex>
#include <type_traits>

#include "../../include/zr_mpm.h"
#include "zr_math.hpp"
#include "zr_scan.hpp"

using namespace zr;

namespace {

constexpr int64_t kMaxParticles = (int64_t(1) << 29) - 1;
// blocks of the grid-stride node kernels of the CG loop (one partial each)
constexpr uint32_t kCgBlocks = 1024;
constexpr int kTileX = 4, kTileY = 6, kTileZ = 6;
constexpr int kTileNodes = kTileX * kTileY * kTileZ;  // 144
constexpr int kTransferThreads = 256;                  // block of the P2G transfers
// the APIC G2P (FP64 gradients of the weights, 164 registers) keeps more
// groups in flight per CU with blocks of two waves (measured: 194 -> 145 us
// on ~1M particles); the MLS G2P is fastest at four
template <bool MLS>
constexpr int g2p_threads() { return MLS ? 256 : 128; }
constexpr int kP2GFields = 7;                          // m, mv[3], f[3]
constexpr int kStageStride = kP2GFields * kTileNodes;  // doubles per group
constexpr int kMaxColliders = 8;
constexpr int kSortLds = 2048;  // particles per page sorted in LDS
constexpr uint32_t kNone = 0xffffffffu;

// a collider as the kernels read it: its rates and its frame x = R s X + b
// at two times, f[0] at the step's start t and f[1] at t + dt (the state
// updateState(t + dt) sets before the trial collision and G2P,
// MpmSimulationBase.cpp:273-278), each with R as a row-major matrix and 1/s
struct ColliderFrame {
  double b[3], R[9], s, inv_s;
};
struct DevCollider {
  int32_t shape, type;
  double friction, origin[3], normal[3], radius, velocity[3], omega[3], dsdt;
  ColliderFrame f[2];
};
struct ColliderSet {
  int32_t n;
  DevCollider c[kMaxColliders];
};

// device counters: the first CNT_SORT are zeroed by every sort; CNT_ADJ_BAD
// and CNT_CFL_RUN (the CFL flags of every step since it was last cleared)
// are cleared by the calls that read them
enum {
  CNT_GROUPS = 0,
  CNT_SLOTS = 1,
  CNT_DOMAIN = 2,
  CNT_CFL = 3,
  CNT_TOTAL = 4,
  CNT_SORT = 5,
  CNT_ADJ_BAD = 5,
  CNT_CFL_RUN = 6,
  CNT_N = 8
};

// ---------------------------------------------------------------------------
// collisions: AnalyticCollisionObject moving as x = R s X + b(t)
// (CollisionObject.cpp:48-105,232-246,279-296,340-405; AnalyticLevelSet.cpp:
// 111-119,259-287,435-450).  With R = I, s = 1 and omega = ds/dt = 0 every
// expression below reduces exactly to the translation-only one.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool ls_query_inside(const DevCollider& c, const double X[3], double& phi, double n[3])
{
  if (c.shape == ZR_HALF_SPACE) {
    double sd = c.normal[0] * (X[0] - c.origin[0]) + c.normal[1] * (X[1] - c.origin[1]) +
                c.normal[2] * (X[2] - c.origin[2]);
    phi = sd;
    if (sd <= 0.0) {
      n[0] = c.normal[0];
      n[1] = c.normal[1];
      n[2] = c.normal[2];
      return true;
    }
    return false;
  }
  double tc[3] = { X[0] - c.origin[0], X[1] - c.origin[1], X[2] - c.origin[2] };
  double d2 = tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2];
  if (d2 < c.radius * c.radius) {
    double dist = sqrt(d2);
    phi = dist - c.radius;
    if (dist < 1e-7) {
      n[0] = 1.0;
      n[1] = 0.0;
      n[2] = 0.0;
    } else {
      n[0] = tc[0] / dist;
      n[1] = tc[1] / dist;
      n[2] = tc[2] / dist;
    }
    return true;
  }
  return false;
}

// queryInside in world space (CollisionObject.cpp:279-296) in frame fr:
// xb = x - b, X = R^T xb / s; inside: n = R N and phi scaled by s
__device__ __forceinline__ bool collider_query(const DevCollider& c, int fr, const double x[3], double& phi,
                                               double n[3], double xb[3])
{
  const ColliderFrame& f = c.f[fr];
  double X[3], N[3];
  for (int d = 0; d < 3; ++d) xb[d] = x[d] - f.b[d];
  for (int d = 0; d < 3; ++d) X[d] = (f.R[d] * xb[0] + f.R[3 + d] * xb[1] + f.R[6 + d] * xb[2]) * f.inv_s;
  if (!ls_query_inside(c, X, phi, N)) return false;
  for (int d = 0; d < 3; ++d) n[d] = f.R[3 * d] * N[0] + f.R[3 * d + 1] * N[1] + f.R[3 * d + 2] * N[2];
  phi *= f.s;
  return true;
}

// the object's velocity at x: omega x (x - b) + (ds/dt / s)(x - b) + db/dt
__device__ __forceinline__ void collider_velocity(const DevCollider& c, int fr, const double xb[3], double vo[3])
{
  const double g = c.dsdt * c.f[fr].inv_s;
  vo[0] = (c.omega[1] * xb[2] - c.omega[2] * xb[1]) + g * xb[0] + c.velocity[0];
  vo[1] = (c.omega[2] * xb[0] - c.omega[0] * xb[2]) + g * xb[1] + c.velocity[1];
  vo[2] = (c.omega[0] * xb[1] - c.omega[1] * xb[0]) + g * xb[2] + c.velocity[2];
}

__device__ __forceinline__ bool detect_and_resolve(const DevCollider& c, int fr, const double x[3], double v[3])
{
  double xb[3], phi, N[3], vo[3];
  if (!collider_query(c, fr, x, phi, N, xb)) return false;
  collider_velocity(c, fr, xb, vo);
  for (int d = 0; d < 3; ++d) v[d] -= vo[d];
  if (c.type == ZR_STICKY) {
    v[0] = v[1] = v[2] = 0.0;
  } else {
    double dot = v[0] * N[0] + v[1] * N[1] + v[2] * N[2];
    if (c.type == ZR_SLIP || dot < 0.0) {
      for (int d = 0; d < 3; ++d) v[d] -= N[d] * dot;
      if (c.friction != 0.0 && dot < 0.0) {
        double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (-dot * c.friction < vn) {
          for (int d = 0; d < 3; ++d) v[d] += v[d] / vn * dot * c.friction;
        } else {
          v[0] = v[1] = v[2] = 0.0;
        }
      }
    }
  }
  for (int d = 0; d < 3; ++d) v[d] += vo[d];
  return true;
}

__device__ __forceinline__ bool multi_object_collision(const ColliderSet& cs, int fr, const double x[3], double v[3])
{
  bool any = false;
  for (int k = 0; k < cs.n; ++k) {
    bool hit = detect_and_resolve(cs.c[k], fr, x, v);
    any = any || hit;
    if (hit && cs.c[k].type == ZR_STICKY) break;
  }
  return any;
}

__device__ __forceinline__ void trial_collision(const ColliderSet& cs, int fr, const double x[3], double dt, double v[3])
{
  for (int k = 0; k < cs.n; ++k) {
    double xt[3], xb[3], phi, n[3];
    for (int d = 0; d < 3; ++d) xt[d] = x[d] + dt * v[d];
    if (collider_query(cs.c[k], fr, xt, phi, n, xb))
      for (int d = 0; d < 3; ++d) v[d] -= (phi / dt) * n[d];
  }
}

// ---------------------------------------------------------------------------
// binning and sort (MpmSimulationBase.cpp:1032-1103)
// ---------------------------------------------------------------------------
// (and zeroes the sort's counters)
__global__ void k_reset_tables(const uint32_t* group_page, uint32_t ngroups, const uint32_t* slot_page, uint32_t nslots,
                               uint32_t* page_count, int32_t* page_group, uint32_t* page_bits, uint32_t* page_sum,
                               uint32_t* page_touch, int32_t* page_slot, uint32_t* counters)
{
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (counters && t < CNT_SORT) counters[t] = 0u;
  if (t < ngroups) {
    uint32_t p = group_page[t];
    page_count[p] = 0u;
    page_group[p] = -1;
    page_bits[p >> 5] = 0u;
    page_sum[p >> 10] = 0u;
  }
  if (t < nslots) {
    uint32_t p = slot_page[t];
    page_touch[p] = kNone;
    page_slot[p] = -1;
  }
}

// page_bits flags the touched pages, page_sum the page_bits words that are
// not zero (bit w & 31 of word w >> 5), so that listing the pages reads the
// summary (a 4096^3 extent has 2^26 page_bits words and 2^21 summary words)
__global__ void k_bin(const double* __restrict__ x, int64_t stride, uint32_t n, double inv_dx, int lim,
                      uint32_t* page_count, uint32_t* page_bits, uint32_t* page_sum, uint32_t* p_page, uint32_t* p_re,
                      uint32_t* counters)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const int lane = threadIdx.x & 63;
  int b[3] = { 0, 0, 0 };
  bool bad = false;
  if (valid) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      b[d] = int_floor(x[d * stride + i] * inv_dx - 0.5);
      if (b[d] < 0 || b[d] > lim) {
        bad = true;
        b[d] = b[d] < 0 ? 0 : lim;
      }
    }
  }
  if (bad) atomicOr(&counters[CNT_DOMAIN], 1u);
  const uint32_t pc = page_encode((uint32_t)b[0] >> 1, (uint32_t)b[1] >> 2, (uint32_t)b[2] >> 2);
  const uint32_t e = elem_of(b[0], b[1], b[2]);
  // wave-aggregated counting: one atomic per distinct page in the wavefront;
  // lanes of a page get consecutive ranks in lane order
  uint64_t active = __ballot(valid);
  uint32_t r = 0;
  while (active) {
    const int leader = __ffsll((long long)active) - 1;
    const uint32_t lpc = __shfl(pc, leader, 64);
    const uint64_t m = __ballot(pc == lpc) & active;
    uint32_t base = 0;
    if (lane == leader) {
      base = atomicAdd(&page_count[lpc], (uint32_t)__popcll(m));
      if (base == 0u) {
        const uint32_t old = atomicOr(&page_bits[lpc >> 5], 1u << (lpc & 31u));
        if (old == 0u) atomicOr(&page_sum[lpc >> 10], 1u << ((lpc >> 5) & 31u));
      }
    }
    base = __shfl(base, leader, 64);
    if ((m >> lane) & 1ull) r = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    active &= ~m;
  }
  if (!valid) return;
  p_page[i] = pc;
  p_re[i] = (r << 5) | e;
}

// The touched pages in ascending page code (the reference's sorted key
// order): sum_base[s] is the first group of summary word s; 32 lanes per
// summary word, grid-stride, each lane one page_bits word, the lanes' page
// counts prefix-summed across the 32 lanes.
// the touched pages under each summary word (32 lanes per word, each
// counting one flagged page_bits word)
__global__ __launch_bounds__(256) void k_sum_counts(const uint32_t* __restrict__ page_sum, uint32_t nsum,
                                                    const uint32_t* __restrict__ page_bits,
                                                    uint32_t* __restrict__ sum_count)
{
  const uint32_t lane = threadIdx.x & 31u;
  for (uint32_t s = blockIdx.x * 8u + (threadIdx.x >> 5); s < nsum; s += gridDim.x * 8u) {
    const uint32_t sb = page_sum[s];
    uint32_t c = 0;
    if (sb) {  // alike in the 32 lanes of s
      c = ((sb >> lane) & 1u) ? (uint32_t)__popc(page_bits[s * 32u + lane]) : 0u;
#pragma unroll
      for (int off = 16; off > 0; off >>= 1) c += __shfl_xor(c, off, 32);
    }
    if (lane == 0) sum_count[s] = c;
  }
}

__global__ __launch_bounds__(256) void k_list_groups(const uint32_t* page_sum, uint32_t nsum,
                                                     const uint32_t* page_bits, const uint32_t* sum_base,
                                                     const uint32_t* page_count, uint32_t* group_page,
                                                     int32_t* page_group, uint32_t* group_count)
{
  const uint32_t lane = threadIdx.x & 31u;
  for (uint32_t s = blockIdx.x * 8u + (threadIdx.x >> 5); s < nsum; s += gridDim.x * 8u) {
    const uint32_t sb = page_sum[s];
    if (!sb) continue;  // alike in the 32 lanes of s
    const uint32_t w = s * 32u + lane;
    uint32_t bits = ((sb >> lane) & 1u) ? page_bits[w] : 0u;
    const uint32_t c = (uint32_t)__popc(bits);
    uint32_t incl = c;
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
      const uint32_t t = __shfl_up(incl, off, 32);
      if (lane >= (uint32_t)off) incl += t;
    }
    uint32_t g = sum_base[s] + incl - c;
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctz(bits);
      bits &= bits - 1u;
      const uint32_t pc = w * 32u + b;
      group_page[g] = pc;
      page_group[pc] = (int32_t)g;
      group_count[g] = page_count[pc];
      ++g;
    }
  }
}

// the group count the kernels after the page listing stop at: the
// device-side count, or the launch's hint when that is smaller (the host
// then repeats those kernels at the exact count)
struct GroupBound {
  const uint32_t* count;
  uint32_t hint;
  __device__ uint32_t get() const
  {
    const uint32_t a = *count;
    return a < hint ? a : hint;
  }
};
struct LdGroupCount {
  const uint32_t* group_count;
  GroupBound gb;
  __device__ uint32_t operator()(uint32_t i) const { return i < gb.get() ? group_count[i] : 0u; }
};
// first touch of a page: the minimum touch index stored for the page is this one
struct LdFirstTouch {
  const uint32_t* touch_page;
  const uint32_t* page_touch;
  GroupBound gb;
  __device__ uint32_t operator()(uint32_t i) const
  {
    return i < 8u * gb.get() && page_touch[touch_page[i]] == i ? 1u : 0u;
  }
};

// (thread 0 also closes group_start with the particle total)
__global__ void k_scatter(uint32_t n, const uint32_t* p_page, const uint32_t* p_re, const int32_t* page_group,
                          uint32_t* group_start, uint64_t* keys, GroupBound gb, const uint32_t* total)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t ng = gb.get();
  if (i == 0) group_start[ng] = *total;
  if (i >= n) return;
  uint32_t re = p_re[i];
  int32_t g = page_group[p_page[i]];
  if ((uint32_t)g >= ng) return;
  uint32_t dst = group_start[g] + (re >> 5);
  keys[dst] = ((uint64_t)(re & 31u) << 32) | (uint64_t)i;
}

// sorts each group's keys (cell << 32 | particle index) ascending: the
// particles are bucketed by cell (LDS counters), and each particle's place in
// its cell's bucket is the number of the bucket's indices below its own
__global__ __launch_bounds__(256) void k_sort_groups(const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ group_start,
                                                     uint32_t* __restrict__ order, GroupBound gb)
{
  __shared__ uint32_t bucket[kSortLds];  // particle indices by cell, unordered within a cell
  __shared__ uint32_t pidx[kSortLds];    // particle index of each key position
  __shared__ uint16_t code[kSortLds];    // cell | slot << 5 of each key position
  __shared__ uint32_t cnt[32], off[32];
  static_assert(kSortLds <= 2048, "slot << 5 fits 16 bits");
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  if (g >= gb.get()) return;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  if (n > (uint32_t)kSortLds) {
    // a page holding more than kSortLds particles: rank sort over its keys
    // in memory (the keys are unique)
    for (uint32_t k = tid; k < n; k += 256) {
      const uint64_t key = keys[start + k];
      uint32_t rank = 0;
      for (uint32_t j = 0; j < n; ++j) rank += keys[start + j] < key ? 1u : 0u;
      order[start + rank] = (uint32_t)(key & 0xffffffffull);
    }
    return;
  }
  if (tid < 32) cnt[tid] = 0u;
  __syncthreads();
  for (uint32_t k = tid; k < n; k += 256) {
    const uint64_t key = keys[start + k];
    const uint32_t c = (uint32_t)(key >> 32) & 31u;
    const uint32_t slot = atomicAdd(&cnt[c], 1u);
    code[k] = (uint16_t)(c | (slot << 5));
    pidx[k] = (uint32_t)key;
  }
  __syncthreads();
  if (tid < 64) {  // exclusive scan of the 32 cell counts
    const uint32_t v = tid < 32 ? cnt[tid] : 0u;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o, 64);
      if (tid >= o) incl += y;
    }
    if (tid < 32) off[tid] = incl - v;
  }
  __syncthreads();
  for (uint32_t k = tid; k < n; k += 256) {
    const uint32_t cs = code[k];
    bucket[off[cs & 31u] + (cs >> 5)] = pidx[k];
  }
  __syncthreads();
  for (uint32_t k = tid; k < n; k += 256) {
    const uint32_t c = code[k] & 31u, i = pidx[k];
    const uint32_t b0 = off[c], b1 = b0 + cnt[c];
    uint32_t rank = 0;
    for (uint32_t j = b0; j < b1; ++j) rank += bucket[j] < i ? 1u : 0u;
    order[start + b0 + rank] = i;
  }
}

// SPGrid_Page_Map::Set_Page over the 2x2x2 page neighbourhood of each group in
// the reference loop order i(x), j(y), k(z) (MpmSimulationBase.cpp:1081-1090)
__global__ void k_touch(GroupBound gb, const uint32_t* group_page, uint32_t* touch_page, uint32_t* page_touch)
{
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8u * gb.get()) return;
  uint32_t g = t >> 3, o = t & 7u;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  uint32_t q = page_encode(px + (o >> 2), py + ((o >> 1) & 1u), pz + (o & 1u));
  touch_page[t] = q;
  atomicMin(&page_touch[q], t);
}

__global__ void k_assign_slots(GroupBound gb, const uint32_t* touch_page, const uint32_t* page_touch,
                               const uint32_t* touch_slot, uint32_t* slot_page, int32_t* page_slot)
{
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8u * gb.get()) return;
  uint32_t q = touch_page[t];
  if (page_touch[q] != t) return;
  uint32_t s = touch_slot[t];
  slot_page[s] = q;
  page_slot[q] = (int32_t)s;
}

// the 8 neighbour page slots of every group, in touch order (o = 4a + 2b + c)
__global__ void k_group_slots(GroupBound gb, const uint32_t* touch_page, const int32_t* page_slot,
                              int32_t* group_nslot)
{
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8u * gb.get()) return;
  group_nslot[t] = page_slot[touch_page[t]];
}

// the <= 8 groups whose staged tiles overlap each slot page (o = 4a + 2b + c
// for the source page q - (a, b, c); -1 when absent)
__global__ void k_slot_sources(uint32_t nslots, const uint32_t* slot_page, const int32_t* page_group,
                               int32_t* slot_src)
{
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8u * nslots) return;
  const uint32_t s = t >> 3, o = t & 7u;
  const uint32_t a = o >> 2, b = (o >> 1) & 1u, c = o & 1u;
  uint32_t qx, qy, qz;
  page_decode(slot_page[s], qx, qy, qz);
  int32_t src = -1;
  if (a <= qx && b <= qy && c <= qz) src = page_group[page_encode(qx - a, qy - b, qz - c)];
  slot_src[t] = src;
}

// ---------------------------------------------------------------------------
// P2G, MLS with force (MpmSimulationBase.cpp:488-561) with the fixed-corotated
// Kirchhoff stress of FBasedMpmForceHelper::updateState fused in
// (FBasedMpmForceHelper.cpp:44-64).
// ---------------------------------------------------------------------------
struct ParticlesView {
  double* x;
  double* v;
  double* C;
  double* F;
  const double* m;
  const double* vol;
  const double* mu;
  const double* lam;
  int64_t stride;
  const int32_t* mat;   // model | plasticity << 4 per particle; null: all fixed-corotated
  const double* gdeg;   // NeoHookeanBorden degradation g
  const double* vtau_add;  // the viscoelastic branch's V0 tau, added to the stress (null: none)
  double* gradv;           // the APIC G2P writes grad v here when set (9 rows of n)
};

// constitutive_model.kirchhoff of the particle's model (FBasedMpmForceHelper.cpp:44-64).
// MODEL >= 0: every particle has that model (the kernel instance holds only
// its stress, so a NeoHookeanBorden scene does not pay the SVD's registers);
// kAnyModel: the model code of each particle.
constexpr int kAnyModel = -1;
template <int MODEL>
__device__ __forceinline__ void particle_kirchhoff(const ParticlesView& P, uint32_t i, const M3<double>& f,
                                                   M3<double>& tau)
{
  const int model = MODEL != kAnyModel ? MODEL : (P.mat ? (P.mat[i] & 15) : ZR_MODEL_FIXED_COROTATED);
  if (model == ZR_MODEL_STVK_HENCKY)
    hencky_kirchhoff(P.mu[i], P.lam[i], f, tau);
  else if (model == ZR_MODEL_NEO_HOOKEAN_BORDEN)
    nhb_kirchhoff(P.mu[i], P.lam[i], P.gdeg[i], f, tau);
  else
    fcr_kirchhoff(P.mu[i], P.lam[i], f, tau);
}

// the stress in precision T: a NeoHookeanBorden instance evaluates it in T,
// every other model in FP64 (its SVD), rounded to T
template <int MODEL, class T>
__device__ __forceinline__ void particle_kirchhoff_t(const ParticlesView& P, uint32_t i, M3<T>& tau)
{
  if constexpr (std::is_same<T, float>::value && MODEL == ZR_MODEL_NEO_HOOKEAN_BORDEN) {
    M3<float> f;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) f.a[r][c] = (float)P.F[(c * 3 + r) * P.stride + i];
    nhb_kirchhoff((float)P.mu[i], (float)P.lam[i], (float)P.gdeg[i], f, tau);
  } else {
    M3<double> f, t64;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) f.a[r][c] = P.F[(c * 3 + r) * P.stride + i];
    particle_kirchhoff<MODEL>(P, i, f, t64);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) tau.a[r][c] = (T)t64.a[r][c];
  }
}

// FORCE = true: MLS with force (explicit).  FORCE = false: APIC mass and
// momentum only, particlesToGridHelper<APIC> (MpmSimulationBase.cpp:563-608).
// T: the precision of the per-particle terms (stress, weights, products);
// the node sums are FP64 either way.  FUSED (with FORCE, inside a whole
// explicit step of size fdt): the force term goes into the momentum sum as
// mv + fdt f, so that the grid update's v + dt (f / m + g) becomes
// (mv + dt f) / m + dt g, with three fields fewer to add per node
// waves per SIMD the register budget is held to: three for the FP64 fused
// NeoHookeanBorden instance, which would otherwise miss that occupancy by
// one register
template <int MODEL, class T, bool FUSED>
constexpr int p2g_min_waves()
{
  return MODEL == ZR_MODEL_NEO_HOOKEAN_BORDEN && sizeof(T) == sizeof(double) && FUSED ? 3 : 1;
}
template <bool FORCE, int MODEL = kAnyModel, class T = double, bool FUSED = false>
__global__ __launch_bounds__(kTransferThreads, (p2g_min_waves<MODEL, T, FUSED>())) void k_p2g(ParticlesView P, const uint32_t* __restrict__ order,
                                             const uint32_t* __restrict__ group_start,
                                             const uint32_t* __restrict__ group_page, double* __restrict__ stage,
                                             double dx, double inv_dx, double Dinv, double fdt)
{
  static_assert(!FUSED || FORCE, "the fused form needs the force");
  constexpr int NF = FORCE && !FUSED ? kP2GFields : 4;
  constexpr int SS = NF * kTileNodes;
  constexpr int NT = kTransferThreads, NW = NT / 64;
  __shared__ double tile[NW * SS];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  for (int k = tid; k < NW * SS; k += NT) tile[k] = 0.0;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  __syncthreads();
  double* wt = tile + (tid >> 6) * SS;
  // lanes of a wave take particles NT/32 apart so that they mostly sit in
  // different cells and their LDS adds rarely collide
  const uint32_t perm = (uint32_t)((tid & 31) * (NT / 32) + (tid >> 5));
  const int64_t S = P.stride;
  for (uint32_t j0 = 0; j0 < n; j0 += NT) {
    const uint32_t j = j0 + perm;
    if (j >= n) continue;
    const uint32_t i = order[start + j];
    double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    const T mass = (T)P.m[i];
    T mom[3] = { mass * (T)P.v[i], mass * (T)P.v[S + i], mass * (T)P.v[2 * S + i] };
    T mC[3][3], st[3][3];
    if constexpr (FORCE) {
      M3<T> tau;
      particle_kirchhoff_t<MODEL, T>(P, i, tau);
      const T vol = (T)P.vol[i];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) st[r][c] = vol * tau.a[r][c];
      if (P.vtau_add)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) st[r][c] += (T)P.vtau_add[(c * 3 + r) * S + i];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) mC[r][c] = mass * (T)P.C[(c * 3 + r) * S + i];
    // base node and the offsets x_i - x_p in FP64, the weights and the
    // products in T
    T w[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double w64[3];
      b[d] = bspline_w(xp[d] * inv_dx, w64);
#pragma unroll
      for (int a = 0; a < 3; ++a) w[d][a] = (T)w64[a];
    }
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    const T Dinv_t = (T)Dinv;
    if constexpr (FUSED) {  // dt D^-1 V0 tau, so that a node takes (vd - st dx) w
      const T sc = (T)(fdt * Dinv);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) st[r][c] *= sc;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T dx0 = (T)((double)(b[0] + a) * dx - xp[0]);
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const T wab = w[0][a] * w[1][bb];
        const T dx1 = (T)((double)(b[1] + bb) * dx - xp[1]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T wt_ = wab * w[2][c];
          const T dx2 = (T)((double)(b[2] + c) * dx - xp[2]);
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          const T fw = wt_ * Dinv_t;
          atomicAdd(&wt[t], (double)(mass * wt_));
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            T vd = mC[r][0] * dx0 + mC[r][1] * dx1 + mC[r][2] * dx2 + mom[r];
            if constexpr (FUSED) {
              T fd = st[r][0] * dx0 + st[r][1] * dx1 + st[r][2] * dx2;
              atomicAdd(&wt[(1 + r) * kTileNodes + t], (double)((vd - fd) * wt_));
            } else {
              atomicAdd(&wt[(1 + r) * kTileNodes + t], (double)(vd * wt_));
              if constexpr (FORCE) {
                T fd = st[r][0] * dx0 + st[r][1] * dx1 + st[r][2] * dx2;
                atomicAdd(&wt[(4 + r) * kTileNodes + t], (double)(-(fd * fw)));
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  double* out = stage + (size_t)g * SS;
  for (int k = tid; k < SS; k += NT) {
    double acc = tile[k];
#pragma unroll
    for (int w = 1; w < NW; ++w) acc += tile[w * SS + k];
    out[k] = acc;
  }
}

// ---------------------------------------------------------------------------
// grid: sum of the overlapping staged tiles, normalisation, explicit update
// (MpmSimulationBase.cpp:474-485,818-865) and trial collision (:298-312)
// ---------------------------------------------------------------------------
struct GridView {
  double* m;
  double* v;   // [3][stride] normalised velocity
  double* f;   // [3][stride] rasterised force
  double* vs;  // [3][stride] new_v after the update
  int64_t stride;
};

enum { GRID_COMBINE = 0, GRID_COMBINE_UPDATE = 1, GRID_UPDATE = 2 };

template <int NF>
__global__ __launch_bounds__(256) void k_grid(int mode, uint32_t nslots, const uint32_t* __restrict__ slot_page,
                                              const int32_t* __restrict__ slot_src,
                                              const double* __restrict__ stage, GridView G, double dx, double dt,
                                              double gx, double gy, double gz, ColliderSet cs)
{
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= nslots * 32u) return;
  const uint32_t s = k >> 5, e = k & 31u;
  uint32_t qx, qy, qz;
  page_decode(slot_page[s], qx, qy, qz);
  const int ex = (int)(e >> 4), ey = (int)((e >> 2) & 3u), ez = (int)(e & 3u);
  double mass, v[3], f[3];
  const int64_t S = G.stride;
  if (mode != GRID_UPDATE) {
    double acc[kP2GFields];
#pragma unroll
    for (int q = 0; q < kP2GFields; ++q) acc[q] = 0.0;
    constexpr int SS = NF * kTileNodes;
    const int32_t* src8 = slot_src + 8 * (size_t)s;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const int a = o >> 2, b = (o >> 1) & 1, c = o & 1;
      const int tx = ex + 2 * a, ty = ey + 4 * b, tz = ez + 4 * c;
      if (ty >= kTileY || tz >= kTileZ) continue;
      const int32_t src = src8[o];
      if (src < 0) continue;
      const double* base = stage + (size_t)src * SS + (tx * kTileY + ty) * kTileZ + tz;
#pragma unroll
      for (int q = 0; q < NF; ++q) acc[q] += base[q * kTileNodes];
    }
    mass = acc[0];
    for (int d = 0; d < 3; ++d) {
      v[d] = mass != 0.0 ? acc[1 + d] / mass : 0.0;
      f[d] = acc[4 + d];
    }
    G.m[k] = mass;
    for (int d = 0; d < 3; ++d) {
      G.v[d * S + k] = v[d];
      G.f[d * S + k] = f[d];
    }
    if (mode == GRID_COMBINE) return;
  } else {
    mass = G.m[k];
    for (int d = 0; d < 3; ++d) {
      v[d] = G.v[d * S + k];
      f[d] = G.f[d * S + k];
    }
  }
  double vi[3] = { 0.0, 0.0, 0.0 };
  if (mass != 0.0) {
    const double xi[3] = { (double)(2 * (int)qx + ex) * dx, (double)(4 * (int)qy + ey) * dx,
                           (double)(4 * (int)qz + ez) * dx };
    const double dt_over_mi = dt / mass;
    vi[0] = v[0] + gx * dt + f[0] * dt_over_mi;
    vi[1] = v[1] + gy * dt + f[1] * dt_over_mi;
    vi[2] = v[2] + gz * dt + f[2] * dt_over_mi;
    bool collided = multi_object_collision(cs, 0, xi, vi);
    if (!collided) trial_collision(cs, 1, xi, dt, vi);
  }
  for (int d = 0; d < 3; ++d) G.vs[d * S + k] = vi[d];
}

#include "zr_implicit.hpp"
#include "zr_minres.hpp"
#include "zr_damage.hpp"
#include "zr_adjoint.hpp"
#include "zr_admm.hpp"

// ---------------------------------------------------------------------------
// G2P, MLS/APIC (MpmSimulationBase.cpp:906-1016) + evolveStrain
// (FBasedMpmForceHelper.cpp:95-110)
// ---------------------------------------------------------------------------
// MLS: grad v = C.  APIC (implicit path): grad v = sum v* grad w^T
// (MpmSimulationBase.cpp:941-956 with !USE_MLS_MPM).
template <bool MLS>
__global__ __launch_bounds__(g2p_threads<MLS>()) void k_g2p(ParticlesView P, const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ group_start,
                                                 const uint32_t* __restrict__ group_page,
                                                 const int32_t* __restrict__ page_slot, const double* __restrict__ gvs,
                                                 int64_t gstride, double dx, double inv_dx, double Dinv, double dt,
                                                 double cfl, uint32_t* counters)
{
  __shared__ double tv[3 * kTileNodes];
  __shared__ int32_t nslot[8];
  const int tid = threadIdx.x;
  const uint32_t g = blockIdx.x;
  const uint32_t start = group_start[g];
  const uint32_t n = group_start[g + 1] - start;
  uint32_t px, py, pz;
  page_decode(group_page[g], px, py, pz);
  if (tid < 8) nslot[tid] = page_slot[page_encode(px + (tid >> 2), py + ((tid >> 1) & 1), pz + (tid & 1))];
  __syncthreads();
  constexpr int NT = g2p_threads<MLS>();
  for (int t = tid; t < kTileNodes; t += NT) {
    const int tx = t / (kTileY * kTileZ), ty = (t / kTileZ) % kTileY, tz = t % kTileZ;
    const int32_t sl = nslot[((tx >> 1) << 2) | ((ty >> 2) << 1) | (tz >> 2)];
    const int64_t node = (int64_t)sl * 32 + elem_of(tx, ty, tz);
    for (int d = 0; d < 3; ++d) tv[d * kTileNodes + t] = sl >= 0 ? gvs[d * gstride + node] : 0.0;
  }
  __syncthreads();
  const int ox = 2 * (int)px, oy = 4 * (int)py, oz = 4 * (int)pz;
  const int64_t S = P.stride;
  const double dx2 = dx * dx;
  uint32_t flags = 0;
  for (uint32_t j = tid; j < n; j += NT) {
    const uint32_t i = order[start + j];
    double xp[3] = { P.x[i], P.x[S + i], P.x[2 * S + i] };
    double w[3][3], dw[3][3];
    int b[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) b[d] = bspline_wdw(xp[d] * inv_dx, w[d], dw[d]);
    const int lx = b[0] - ox, ly = b[1] - oy, lz = b[2] - oz;
    if ((unsigned)lx > 1u || (unsigned)ly > 3u || (unsigned)lz > 3u) continue;
    double pv[3] = { 0.0, 0.0, 0.0 }, B[3][3] = { { 0.0 } }, Gv[3][3] = { { 0.0 } };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double dx0 = (double)(b[0] + a) * dx - xp[0];
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const double wab = w[0][a] * w[1][bb];
        const double dx1 = (double)(b[1] + bb) * dx - xp[1];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double wt_ = wab * w[2][c];
          const double dx2c = (double)(b[2] + c) * dx - xp[2];
          const int t = ((lx + a) * kTileY + (ly + bb)) * kTileZ + (lz + c);
          double gw[3] = { 0.0, 0.0, 0.0 };
          if constexpr (!MLS) wgrad(w, dw, inv_dx, a, bb, c, gw);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double gv = tv[r * kTileNodes + t];
            pv[r] += wt_ * gv;
            B[r][0] += wt_ * gv * dx0;
            B[r][1] += wt_ * gv * dx1;
            B[r][2] += wt_ * gv * dx2c;
            if constexpr (!MLS) {
              Gv[r][0] += gv * gw[0];
              Gv[r][1] += gv * gw[1];
              Gv[r][2] += gv * gw[2];
            }
          }
        }
      }
    }
    double Cn[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cn[r][c] = B[r][c] * Dinv;
        P.C[(c * 3 + r) * S + i] = Cn[r][c];
      }
    double inc2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      P.v[d * S + i] = pv[d];
      const double inc = dt * pv[d];
      P.x[d * S + i] = xp[d] + inc;
      inc2 += inc * inc;
    }
    if constexpr (!MLS)
      if (P.gradv)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) P.gradv[(c * 3 + r) * S + i] = Gv[r][c];
    if (inc2 > dx2) flags |= 1u;
    if (inc2 > dx2 * 0.25 * (cfl * cfl)) flags |= 2u;
    double Fo[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) Fo[r][c] = P.F[(c * 3 + r) * S + i];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double acc = Fo[r][c];
#pragma unroll
        for (int k = 0; k < 3; ++k) acc += dt * (MLS ? Cn[r][k] : Gv[r][k]) * Fo[k][c];
        P.F[(c * 3 + r) * S + i] = acc;
      }
  }
  if (flags) {
    atomicOr(&counters[CNT_CFL], flags);
    atomicOr(&counters[CNT_CFL_RUN], flags);
  }
}

// ---------------------------------------------------------------------------
// evalMaxParticleSpeed (MpmSimulationBase.cpp:1143-1170): the component-wise
// maximum of (|v|, x, -x) over the particles, from (0, lowest, lowest) with
// lowest = numeric_limits<float>::lowest().  |v| is Eigen's norm(),
// sqrt((v0 v0 + v1 v1) + v2 v2), rounded after every operation; a maximum
// is exact in any order, so the result does not depend on the launch shape.
// ---------------------------------------------------------------------------
constexpr int kSpeedFields = 7;
constexpr uint32_t kSpeedBlocks = 512;
__device__ __forceinline__ double speed_of(double v0, double v1, double v2)
{
#pragma clang fp contract(off)
  return sqrt((v0 * v0 + v1 * v1) + v2 * v2);
}
__device__ __forceinline__ void speed_max(double& a, double b) { a = b > a ? b : a; }
__device__ __forceinline__ void speed_block_max(double m[kSpeedFields], double (*smem)[4])
{
#pragma unroll
  for (int j = 0; j < kSpeedFields; ++j)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) speed_max(m[j], __shfl_xor(m[j], off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < kSpeedFields; ++j) smem[j][wid] = m[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSpeedFields; ++j)
    for (int w = 0; w < 4; ++w) speed_max(m[j], smem[j][w]);
}
// component d of particle i at i * ps + d * cs: structure of arrays (ps = 1,
// cs = stride) or the reference's particle-major layout (ps = 3, cs = 1)
__global__ __launch_bounds__(256) void k_max_speed(const double* __restrict__ x, const double* __restrict__ v,
                                                   int64_t ps, int64_t cs, uint32_t n, double* __restrict__ partial)
{
  __shared__ double smem[kSpeedFields][4];
  double m[kSpeedFields];
  m[0] = 0.0;
  for (int j = 1; j < kSpeedFields; ++j) m[j] = -(double)__FLT_MAX__;
  // 64-bit index: with n up to 2^32 - 1 a 32-bit i + stride would wrap
  for (uint64_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const int64_t b = (int64_t)i * ps;
    speed_max(m[0], speed_of(v[b], v[b + cs], v[b + 2 * cs]));
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double xd = x[b + d * cs];
      speed_max(m[1 + d], xd);
      speed_max(m[4 + d], -xd);
    }
  }
  speed_block_max(m, smem);
  if (threadIdx.x < kSpeedFields) partial[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = m[threadIdx.x];
}
// one block: the block maxima of k_max_speed -> out[7] with the min corner
// negated back
__global__ __launch_bounds__(256) void k_max_speed_final(const double* __restrict__ partial, uint32_t nb,
                                                         double* __restrict__ out)
{
  __shared__ double smem[kSpeedFields][4];
  double m[kSpeedFields];
  m[0] = 0.0;
  for (int j = 1; j < kSpeedFields; ++j) m[j] = -(double)__FLT_MAX__;
  for (uint32_t b = threadIdx.x; b < nb; b += 256u)
#pragma unroll
    for (int j = 0; j < kSpeedFields; ++j) speed_max(m[j], partial[(size_t)j * nb + b]);
  speed_block_max(m, smem);
  if (threadIdx.x < kSpeedFields) out[threadIdx.x] = threadIdx.x >= 4 ? -m[threadIdx.x] : m[threadIdx.x];
}

// ---------------------------------------------------------------------------
// applyPlasticity after the F update (MpmSimulationBase.cpp:1015-1030): the
// return map of each particle's plasticity model, in FP64.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plasticity(uint32_t n, const int32_t* __restrict__ mat, double* F,
                                                    int64_t S, const double* __restrict__ mu,
                                                    const double* __restrict__ lam,
                                                    const double* __restrict__ pparam, double* pstate,
                                                    int32_t* broken)
{
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int plastic = mat[i] >> 4;
  if (plastic == ZR_PLASTIC_NONE) return;
  M3<double> f;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) f.a[r][c] = F[(c * 3 + r) * S + i];
  double st = pstate[i];
  if (plastic == ZR_PLASTIC_VON_MISES) {
    bool br = false;
    if (!von_mises_project(mu[i], lam[i], pparam[i], pparam[S + i], st, f, br)) return;
    if (br) broken[i] = 1;
  } else if (plastic == ZR_PLASTIC_DRUCKER_PRAGER) {
    drucker_prager_project(mu[i], lam[i], pparam[i], pparam[S + i], pparam[2 * S + i], pparam[3 * S + i] != 0.0,
                           st, f);
  } else {
    cam_clay_project(mu[i], lam[i], pparam[i], pparam[S + i], pparam[2 * S + i], st, f);
  }
  pstate[i] = st;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) F[(c * 3 + r) * S + i] = f.a[r][c];
}

// ---------------------------------------------------------------------------
// The viscoelastic branch, FElasticNonequilibratedBasedMpmForceHelper
// (FElasticNonequilibratedBasedMpmForceHelper.h:27-87): a second elastic
// strain Fe_Nonequilibrated per particle with a StvkWithHencky model of its
// own, whose V0 tau adds to the stress (updateState, :36-60) and which moves
// with the flow and relaxes viscously in the Hencky strain (evolveStrain,
// :62-87).  Explicit MLS step only (gradV = C).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_visco_stress(uint32_t n, const double* __restrict__ fneq,
                                                      const double* __restrict__ mu, const double* __restrict__ lam,
                                                      const double* __restrict__ vol, int64_t S,
                                                      double* __restrict__ vtau)
{
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  M3<double> f, tau;
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) f.a[r][c] = fneq[(c * 3 + r) * S + i];
  hencky_kirchhoff(mu[i], lam[i], f, tau);
  const double v0 = vol[i];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) vtau[(c * 3 + r) * S + i] = v0 * tau.a[r][c];
}

// Fneq <- (I + dt C) Fneq in the form of the F update, then the Hencky
// strain relaxation with alpha = 2 mu / eta_d and
// beta = 2 (2 mu + 3 lambda) / (9 eta_v) - 2 mu / (3 eta_d):
// eps = (eps_trial - dt beta / (1 + dt (alpha + 3 beta)) t 1) / (1 + dt alpha)
// where t is the reference's eps_trial.trace() of a 3x1 vector: its first
// entry (the log of the largest singular value).
__global__ __launch_bounds__(256) void k_visco_strain(uint32_t n, const double* __restrict__ C, double* fneq,
                                                      const double* __restrict__ mu, const double* __restrict__ lam,
                                                      int64_t S, double dt, double visc_d, double visc_v)
{
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double Fo[3][3], G[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      Fo[r][c] = fneq[(c * 3 + r) * S + i];
      G[r][c] = C[(c * 3 + r) * S + i];
    }
  M3<double> f, u, v;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double acc = Fo[r][c];
#pragma unroll
      for (int k = 0; k < 3; ++k) acc += dt * G[r][k] * Fo[k][c];
      f.a[r][c] = acc;
    }
  double sig[3], e[3];
  svd3(f, u, sig, v);
  const double m = mu[i], l = lam[i];
  const double alpha = 2.0 * m / visc_d;
  const double beta = 2.0 * (2.0 * m + l * 3.0) / (9.0 * visc_v) - 2.0 * m / (visc_d * 3.0);
  double eps[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) eps[d] = log(fabs(sig[d]));
  const double k = dt * beta / (1.0 + dt * (alpha + 3.0 * beta)) * eps[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) e[d] = exp(1.0 / (1.0 + dt * alpha) * (eps[d] - k));
  usv_t(u, e, v, f);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) fneq[(c * 3 + r) * S + i] = f.a[r][c];
}

// ---------------------------------------------------------------------------
// layout conversion and introspection
// ---------------------------------------------------------------------------
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, uint32_t n, int ncomp, int64_t stride)
{
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * ncomp) return;
  uint32_t i = (uint32_t)(t / ncomp);
  int c = (int)(t % ncomp);
  soa[c * stride + i] = aos[t];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, uint32_t n, int ncomp, int64_t stride)
{
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * ncomp) return;
  uint32_t i = (uint32_t)(t / ncomp);
  int c = (int)(t % ncomp);
  aos[t] = soa[c * stride + i];
}
__global__ void k_dump_sorter(uint32_t n, const uint32_t* order, const uint32_t* p_page, const uint32_t* p_re, uint64_t* out)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t i = order[k];
  uint64_t cell = ((uint64_t)p_page[i] << 5) | (p_re[i] & 31u);
  out[k] = (cell << 27) + i;
}
__global__ void k_dump_base_offset(uint32_t n, const uint32_t* p_page, const uint32_t* p_re, uint64_t* out)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t cell = ((uint64_t)p_page[i] << 5) | (p_re[i] & 31u);
  out[i] = cell << 7;
}
// the nodes with mass, ascending: act[scan[k]] = k
__global__ void k_active_list(uint32_t nn, const double* __restrict__ m, const uint32_t* __restrict__ scan,
                              uint32_t* __restrict__ act)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nn && m[k] != 0.0) act[scan[k]] = k;
}

__global__ void k_node_idx(uint32_t nn, const double* m, const uint32_t* scan, int64_t* idx)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  idx[k] = m[k] != 0.0 ? (int64_t)scan[k] : -1;
}

inline uint32_t blocks_for(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

// ===========================================================================
// engine
// ===========================================================================
struct zr_engine {
  zr_config cfg{};
  hipStream_t stream = nullptr;
  // blocks of the persistent operator apply that fit on the device at once
  // (FP32, FP64): resident blocks per CU times the CU count
  uint32_t hvp_resident[2] = { 0, 0 };
  std::string err;
  double time = 0.0;
  std::vector<zr_collider> colliders;

  uint32_t L = 0, lim = 0, ndense = 0, nwords = 0, nsum = 0;
  uint32_t* page_count = nullptr;
  int32_t* page_group = nullptr;
  uint32_t* page_bits = nullptr;
  uint32_t* page_sum = nullptr;
  uint32_t* page_sumcnt = nullptr;
  uint32_t* page_touch = nullptr;
  int32_t* page_slot = nullptr;

  int64_t n = 0;
  double *x = nullptr, *v = nullptr, *C = nullptr, *F = nullptr;
  double *m = nullptr, *vol = nullptr, *mu = nullptr, *lam = nullptr;
  // materials (null until zr_set_materials): model | plasticity << 4,
  // plasticity parameters [4][n], plastic state, degradation, broken flags
  int32_t* mat = nullptr;
  double *pparam = nullptr, *pstate = nullptr, *gdeg = nullptr;
  int32_t* broken = nullptr;
  bool has_models = false, has_plastic = false, all_hencky = false;
  int uniform_model = ZR_MODEL_FIXED_COROTATED;  // the model of every particle, or kAnyModel
  // viscoelastic branch: Fe_Nonequilibrated [9][n], its StvkWithHencky mu /
  // lambda, V0 tau of the branch [9][n], the two viscosities
  double *fneq = nullptr, *mu_neq = nullptr, *lam_neq = nullptr, *vtau_neq = nullptr;
  double visc_d = 0.0, visc_v = 0.0;
  double* g2p_gradv = nullptr;  // set around the APIC G2P whose grad v a step needs
  bool visco = false;
  // phase field: per particle (c, H, l0, 1/sigma_c, pf_Fp, H_max, k, vol, allow)
  // and per node (phase, rhs, D, x, r, mr, s, q) of the damage solve
  bool pf_on = false;
  zr_phase_field_opts pf_opts{};
  zr_phase_field_stats pf_stats{};
  double *pf_c = nullptr, *pf_H = nullptr, *pf_l0 = nullptr, *pf_osc = nullptr, *pf_Fp = nullptr,
         *pf_Hmax = nullptr, *pf_k = nullptr, *pf_vol = nullptr;
  int32_t* pf_allow = nullptr;
  double* pf_node = nullptr;  // 8 x grid capacity
  size_t pf_node_cap = 0;
  double* pf_partial = nullptr;
  size_t pf_partial_cap = 0;
  // differentiable form: particle adjoints [38][n] (x 3, v 3, C 9, F 9, mu,
  // lambda, G2P scratch x 3 and F 9), node adjoints [7][grid capacity],
  // device checkpoints of (x, v, C, F) with their times
  double* adj = nullptr;
  int64_t adj_n = 0;
  double* adj_grid = nullptr;
  size_t adj_grid_cap = 0;
  std::vector<double*> ckpt;
  std::vector<double> ckpt_time;
  // a backward rollout's recomputed steps: the sort tables and grid of each
  // (StepRecord), so that its adjoint step need not sort and rasterise again
  struct StepRecord {
    void* buf = nullptr;
    size_t cap = 0;
    uint32_t A = 0, S = 0;
  };
  std::vector<StepRecord> step_rec;
  // ADMM: per particle F 9, y 9, omega 1, last dv 3 (= 22 rows of n); per
  // node vn 3, dv 3, proj 9, diag 1, rhs 3 and MINRES q 3x3, z 3, m 3x3
  double* admm_p = nullptr;  // F 9, y 9, omega, last dv 3 (+ F2 9, y2 9, omega2 with the viscoelastic branch)
  int64_t admm_n = 0;
  int admm_steps = 0;
  bool admm_visco = false;
  // the FP32-bulk global step: sorted per-particle operator data (13 or 23
  // rows of n floats), element codes, and seven FP32 Krylov vectors
  float* admm_c32 = nullptr;
  uint8_t* admm_cell32 = nullptr;
  size_t admm_c32_cap = 0, admm_c32_rows = 0;
  float* admm_node32 = nullptr;
  size_t admm_node32_cap = 0;
  double* admm_node = nullptr;
  int32_t* admm_coll = nullptr;
  size_t admm_cap = 0;
  uint32_t *p_page = nullptr, *p_re = nullptr, *order = nullptr;
  uint64_t* keys = nullptr;
  uint32_t *group_page = nullptr, *group_count = nullptr, *group_start = nullptr;
  uint32_t *touch_page = nullptr, *touch_slot = nullptr;
  uint32_t* slot_page = nullptr;
  int32_t* group_nslot = nullptr;  // [8 * groups]
  int32_t* slot_src = nullptr;     // [8 * slots], capacity grid_cap / 32
  uint32_t* ticket = nullptr;      // grid-reduction ticket, kept at zero between kernels
  uint32_t *sum_base = nullptr;

  uint32_t A = 0, S = 0;  // groups, slots of the current sort
  bool sorted = false, rasterised = false, updated = false;

  double* stage = nullptr;
  size_t stage_cap = 0;  // groups
  double *gm = nullptr, *gv = nullptr, *gf = nullptr, *gvs = nullptr;
  size_t grid_cap = 0;  // nodes

  uint32_t* scan_tmp = nullptr;
  size_t scan_cap = 0;
  uint32_t* dcount = nullptr;
  // evalMaxParticleSpeed: 7 block maxima per block, then the 7 results
  double* speed = nullptr;
  uint32_t* hcount = nullptr;

  // implicit backward Euler: node vectors [3][grid_cap], reduction partials,
  // device scalars/flags (+ pinned mirrors), per-particle SVD cache
  bool mls_rasterised = true;  // kind of the last P2G (MLS with force or APIC)
  int grid_nf = 7;              // fields of its staged tiles (7: m, mv, f; 4: m, mv)
  double *ivn = nullptr, *idv = nullptr, *ir = nullptr, *ix = nullptr, *irc = nullptr, *iq = nullptr,
         *ip = nullptr, *iap = nullptr;
  size_t imp_cap = 0;
  double* partial = nullptr;
  size_t partial_cap = 0;
  double* partial2 = nullptr;  // second set of block partials (z^T r of the CG update)
  double* sc = nullptr;
  int32_t* ci = nullptr;
  uint32_t* hvp_ctr = nullptr;     // work counters of the persistent operator apply
  // the slot nodes with mass in ascending order, listed once per implicit
  // step for the CG node kernels, and their count and block count
  uint32_t* act = nullptr;
  uint32_t* act_scan = nullptr;
  size_t act_cap = 0;
  uint32_t nact = 0, ncg_act = 1;
  // per-particle 3x3 scratch (9 rows of n) between the passes of the Newton
  // state (F, then V0 P Fn^T, sorted order) and of the ADMM local step
  // (grad(vn + dv), particle order)
  double* be_fb = nullptr;
  size_t be_fb_cap = 0;
  // launch size of the sort's group kernels: the last sort's group count
  // with headroom (0: read the count first)
  uint32_t a_hint = 0;
  // the CFL flag counter is zero (the sort clears every counter; only G2P sets it)
  bool cfl_clear = false;
  // inside a rollout: a step does not wait for its CFL flags; the next
  // step's sort reads them (CNT_CFL_RUN) and stops the rollout there
  bool defer_cfl = false;
  uint32_t* hvp_bounds = nullptr;  // its XCD group ranges, FP32 then FP64 (2 x 9)
  uint32_t* hvp_order = nullptr;   // their hand-out orders (2 x hvp_order_cap)
  uint8_t* cell_order = nullptr;   // each group's cells by particle count (32 x hvp_order_cap)
  bool cell_order_set = false;     // written for this sort's groups
  // Jacobi preconditioner: staged diagonal sums per group (kImpStage each),
  // the inverse diagonal [3][grid_cap] in the CG vectors' precision (double
  // storage), the nodes a collider touched at the step start
  double* jstage = nullptr;
  size_t jstage_cap = 0;
  double* pinv = nullptr;
  uint8_t* collided = nullptr;
  size_t jac_cap = 0, collided_cap = 0;
  bool jacobi = false;             // this step's CG uses it
  size_t hvp_order_cap = 0;
  double* hsc = nullptr;
  int32_t* hci = nullptr;
  // FP32-bulk CG vectors and the operator data in both precisions
  float *fx = nullptr, *frc = nullptr, *fq = nullptr, *fp = nullptr, *fap = nullptr;
  size_t imp32_cap = 0;
  // the three Krylov vectors MINRES needs beyond CG's, per precision
  double* mr64 = nullptr;
  float* mr32 = nullptr;
  size_t mr64_cap = 0, mr32_cap = 0;
  double *c64 = nullptr;  // U9 Q9 S3 K2 frac3 = 26 rows of n
  float* c32 = nullptr;
  uint32_t* ccell = nullptr;
  int64_t c64_n = 0, c32_n = 0, ccell_n = 0;
  int32_t be_prec = ZR_FP64;
  bool be_ready = false;
  double be_dt = 0.0;

  // kernel timing: pairs of events per launch, per kernel class
  bool prof = false;
  std::vector<hipEvent_t> prof_ev[ZR_PROF_COUNT];
  size_t prof_used[ZR_PROF_COUNT] = { 0, 0, 0, 0, 0, 0 };
  uint64_t prof_seen[ZR_PROF_COUNT] = { 0, 0, 0, 0, 0, 0 };
  bool prof_phase[ZR_PROF_COUNT] = { false, false, false, false, false, false };
  bool prof_take[ZR_PROF_COUNT] = { false, false, false, false, false, false };
  int32_t prof_period = 1;
};

namespace {

int fail(zr_engine* e, int code, const std::string& msg)
{
  if (e) e->err = msg;
  return code;
}

#define ZR_HIP(e, call)                                                                     \
  do {                                                                                      \
    hipError_t _st = (call);                                                                \
    if (_st != hipSuccess)                                                                  \
      return fail((e), ZR_ERR_HIP, std::string(#call ": ") + hipGetErrorString(_st));       \
  } while (0)

#define ZR_LAUNCH_CHECK(e) ZR_HIP(e, hipGetLastError())

template <class T>
void dfree(T*& p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

int ensure_scan(zr_engine* e, uint32_t n)
{
  size_t nt = (n + kScanTile - 1) / kScanTile + 1;
  if (nt <= e->scan_cap) return ZR_OK;
  dfree(e->scan_tmp);
  ZR_HIP(e, hipMalloc(&e->scan_tmp, nt * sizeof(uint32_t)));
  e->scan_cap = nt;
  return ZR_OK;
}

template <class Ld>
int run_scan(zr_engine* e, Ld ld, uint32_t n, uint32_t* out, uint32_t* total_dev)
{
  int rc = ensure_scan(e, n);
  if (rc) return rc;
  if (n == 0) {
    ZR_HIP(e, hipMemsetAsync(total_dev, 0, sizeof(uint32_t), e->stream));
    return ZR_OK;
  }
  if (n <= kScanSingleMax) {
    k_scan_single<Ld><<<1, 1024, 0, e->stream>>>(ld, n, out, total_dev);
    ZR_LAUNCH_CHECK(e);
    return ZR_OK;
  }
  uint32_t nt = (n + kScanTile - 1) / kScanTile;
  k_scan_reduce<Ld><<<nt, kScanBlock, 0, e->stream>>>(ld, n, e->scan_tmp);
  k_scan_tile_sums<<<1, 1024, 0, e->stream>>>(e->scan_tmp, nt, total_dev);
  k_scan_down<Ld><<<nt, kScanBlock, 0, e->stream>>>(ld, n, e->scan_tmp, out);
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

// records one event of a start/stop pair for kernel class k when timing is
// on; every prof_period-th launch of a class is timed (calls alternate start
// and stop)
int prof_mark(zr_engine* e, int k)
{
  if (!e->prof) return ZR_OK;
  const bool start = !e->prof_phase[k];
  e->prof_phase[k] = start;
  if (start) e->prof_take[k] = (e->prof_seen[k]++ % (uint64_t)e->prof_period) == 0;
  if (!e->prof_take[k]) return ZR_OK;
  auto& ev = e->prof_ev[k];
  if (e->prof_used[k] == ev.size()) {
    hipEvent_t h;
    ZR_HIP(e, hipEventCreate(&h));
    ev.push_back(h);
  }
  ZR_HIP(e, hipEventRecord(ev[e->prof_used[k]++], e->stream));
  return ZR_OK;
}

int sync_counts(zr_engine* e)
{
  ZR_HIP(e, hipMemcpyAsync(e->hcount, e->dcount, CNT_N * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

void free_particles(zr_engine* e)
{
  dfree(e->x); dfree(e->v); dfree(e->C); dfree(e->F); dfree(e->m); dfree(e->vol); dfree(e->mu); dfree(e->lam);
  dfree(e->p_page); dfree(e->p_re); dfree(e->order); dfree(e->keys);
  dfree(e->group_page); dfree(e->group_count); dfree(e->group_start);
  dfree(e->touch_page); dfree(e->touch_slot); dfree(e->slot_page); dfree(e->group_nslot);
  dfree(e->mat); dfree(e->pparam); dfree(e->pstate); dfree(e->gdeg); dfree(e->broken);
  e->has_models = e->has_plastic = false;
  e->uniform_model = ZR_MODEL_FIXED_COROTATED;
  dfree(e->fneq); dfree(e->mu_neq); dfree(e->lam_neq); dfree(e->vtau_neq);
  e->visco = false;
  dfree(e->pf_c); dfree(e->pf_H); dfree(e->pf_l0); dfree(e->pf_osc); dfree(e->pf_Fp); dfree(e->pf_Hmax);
  dfree(e->pf_k); dfree(e->pf_vol); dfree(e->pf_allow);
  e->pf_on = false;
  dfree(e->adj);
  e->adj_n = 0;
  dfree(e->admm_p);
  e->admm_n = 0;
  e->admm_steps = 0;
  for (auto& c : e->ckpt) dfree(c);
  e->ckpt.clear();
  e->ckpt_time.clear();
  for (auto& r : e->step_rec) dfree(r.buf);
  e->step_rec.clear();
  e->n = 0;
}

// R from the quaternion <w,x,y,z> as Eigen::Quaternion::toRotationMatrix
// forms it (Rotation.h:53-57), row-major
void quat_to_matrix(const double q[4], double R[9])
{
#pragma clang fp contract(off)
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
  const double twx = tx * w, twy = ty * w, twz = tz * w, txx = tx * x, txy = ty * x, txz = tz * x, tyy = ty * y,
               tyz = tz * y, tzz = tz * z;
  R[0] = 1 - (tyy + tzz);
  R[1] = txy - twz;
  R[2] = txz + twy;
  R[3] = txy + twz;
  R[4] = 1 - (txx + tzz);
  R[5] = tyz - twx;
  R[6] = txz - twy;
  R[7] = tyz + twx;
  R[8] = 1 - (txx + tyy);
}
// The state updateState(t) sets for an object moving at constant rates from
// its time-0 state (CollisionObject.h:81-172; the reference's scene callbacks
// write this form, e.g. AdmmInit3D.h:298-313: rotation
// (cos(theta t/2), 0, sin(theta t/2), 0) with omega (0, theta, 0)):
// b(t) = translation + velocity t, s(t) = s0 + ds/dt t and
// R(t) = exp(t [omega]x) R0, as the quaternion product
// (cos(|omega| t/2), sin(|omega| t/2) omega/|omega|) q0.  A zero quaternion
// means the identity, scale 0 means 1.
void collider_frame(const zr_collider& c, double t, ColliderFrame& f)
{
#pragma clang fp contract(off)
  double q0[4] = { c.rotation[0], c.rotation[1], c.rotation[2], c.rotation[3] };
  if (q0[0] == 0.0 && q0[1] == 0.0 && q0[2] == 0.0 && q0[3] == 0.0) q0[0] = 1.0;
  const double* w = c.omega;
  const double wn = std::sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  double q[4] = { q0[0], q0[1], q0[2], q0[3] };
  if (wn != 0.0 && t != 0.0) {
    const double h = 0.5 * wn * t, sh = std::sin(h) / wn;
    const double a[4] = { std::cos(h), sh * w[0], sh * w[1], sh * w[2] };
    q[0] = a[0] * q0[0] - a[1] * q0[1] - a[2] * q0[2] - a[3] * q0[3];
    q[1] = a[0] * q0[1] + a[1] * q0[0] + a[2] * q0[3] - a[3] * q0[2];
    q[2] = a[0] * q0[2] - a[1] * q0[3] + a[2] * q0[0] + a[3] * q0[1];
    q[3] = a[0] * q0[3] + a[1] * q0[2] - a[2] * q0[1] + a[3] * q0[0];
  }
  quat_to_matrix(q, f.R);
  for (int d = 0; d < 3; ++d) f.b[d] = c.translation[d] + c.velocity[d] * t;
  f.s = (c.scale == 0.0 ? 1.0 : c.scale) + c.dscale_dt * t;
  f.inv_s = 1.0 / f.s;
}

// the colliders with their frames at the engine time t and at t + dt
ColliderSet collider_set(const zr_engine* e, double dt = 0.0)
{
  ColliderSet cs;
  memset(&cs, 0, sizeof(cs));
  cs.n = (int32_t)e->colliders.size();
  for (int k = 0; k < cs.n; ++k) {
    const zr_collider& a = e->colliders[k];
    DevCollider& c = cs.c[k];
    c.shape = a.shape;
    c.type = a.type;
    c.friction = a.friction;
    c.radius = a.radius;
    for (int d = 0; d < 3; ++d) {
      c.origin[d] = a.origin[d];
      c.normal[d] = a.normal[d];
      c.velocity[d] = a.velocity[d];
      c.omega[d] = a.omega[d];
    }
    c.dsdt = a.dscale_dt;
    collider_frame(a, e->time, c.f[0]);
    collider_frame(a, e->time + dt, c.f[1]);
  }
  return cs;
}

ParticlesView pview(zr_engine* e)
{
  return ParticlesView{ e->x, e->v, e->C, e->F, e->m, e->vol, e->mu, e->lam, e->n,
                        e->has_models ? e->mat : nullptr, e->gdeg, e->visco ? e->vtau_neq : nullptr, nullptr };
}

int ensure_grid(zr_engine* e)
{
  size_t nodes = (size_t)e->S * 32;
  if (nodes > e->grid_cap) {
    dfree(e->gm); dfree(e->gv); dfree(e->gf); dfree(e->gvs);
    size_t cap = nodes + nodes / 4 + 1024;
    ZR_HIP(e, hipMalloc(&e->gm, cap * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->gv, 3 * cap * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->gf, 3 * cap * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->gvs, 3 * cap * sizeof(double)));
    dfree(e->slot_src);
    ZR_HIP(e, hipMalloc(&e->slot_src, (cap / 32 + 1) * 8 * sizeof(int32_t)));
    e->grid_cap = cap;
  }
  if ((size_t)e->A > e->stage_cap) {
    dfree(e->stage);
    size_t cap = (size_t)e->A + e->A / 4 + 64;
    ZR_HIP(e, hipMalloc(&e->stage, cap * kStageStride * sizeof(double)));
    e->stage_cap = cap;
  }
  return ZR_OK;
}

// the sort after the page listing, for at most `hint` groups (GroupBound)
int sort_groups(zr_engine* e, uint32_t hint)
{
  hipStream_t s = e->stream;
  const uint32_t n = (uint32_t)e->n;
  const GroupBound gb{ e->dcount + CNT_GROUPS, hint };
  int rc = run_scan(e, LdGroupCount{ e->group_count, gb }, hint, e->group_start, e->dcount + CNT_TOTAL);
  if (rc) return rc;
  k_scatter<<<blocks_for(n, 256), 256, 0, s>>>(n, e->p_page, e->p_re, e->page_group, e->group_start, e->keys, gb,
                                               e->dcount + CNT_TOTAL);
  k_sort_groups<<<hint, 256, 0, s>>>(e->keys, e->group_start, e->order, gb);
  k_touch<<<blocks_for(8ull * hint, 256), 256, 0, s>>>(gb, e->group_page, e->touch_page, e->page_touch);
  ZR_LAUNCH_CHECK(e);
  rc = run_scan(e, LdFirstTouch{ e->touch_page, e->page_touch, gb }, 8u * hint, e->touch_slot,
                e->dcount + CNT_SLOTS);
  if (rc) return rc;
  k_assign_slots<<<blocks_for(8ull * hint, 256), 256, 0, s>>>(gb, e->touch_page, e->page_touch, e->touch_slot,
                                                             e->slot_page, e->page_slot);
  k_group_slots<<<blocks_for(8ull * hint, 256), 256, 0, s>>>(gb, e->touch_page, e->page_slot, e->group_nslot);
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

int do_sort(zr_engine* e, zr_step_stats* st)
{
  if (e->n <= 0) return fail(e, ZR_ERR_STATE, "no particles uploaded");
  const uint32_t n = (uint32_t)e->n;
  hipStream_t s = e->stream;
  // clear the dense tables touched by the previous sort
  uint32_t rst = e->A > e->S ? e->A : e->S;
  const uint32_t nrst = rst > (uint32_t)CNT_N ? rst : (uint32_t)CNT_N;
  e->sorted = e->rasterised = e->updated = false;
  e->cfl_clear = true;
  prof_mark(e, ZR_PROF_SORT);
  k_reset_tables<<<blocks_for(nrst, 256), 256, 0, s>>>(e->group_page, e->A, e->slot_page, e->S, e->page_count,
                                                      e->page_group, e->page_bits, e->page_sum, e->page_touch,
                                                      e->page_slot, e->dcount);
  k_bin<<<blocks_for(n, 256), 256, 0, s>>>(e->x, e->n, n, 1.0 / e->cfg.dx, (int)e->lim, e->page_count, e->page_bits,
                                           e->page_sum, e->p_page, e->p_re, e->dcount);
  ZR_LAUNCH_CHECK(e);
  uint32_t nbs = (e->nsum + 7u) / 8u;
  if (nbs > 2048u) nbs = 2048u;
  k_sum_counts<<<nbs, 256, 0, s>>>(e->page_sum, e->nsum, e->page_bits, e->page_sumcnt);
  int rc = run_scan(e, LdArray{ e->page_sumcnt }, e->nsum, e->sum_base, e->dcount + CNT_GROUPS);
  if (rc) return rc;
  {
    const uint32_t nb = nbs;
    k_list_groups<<<nb, 256, 0, s>>>(e->page_sum, e->nsum, e->page_bits, e->sum_base, e->page_count, e->group_page,
                                     e->page_group, e->group_count);
  }
  ZR_LAUNCH_CHECK(e);
  // group sizes, in-page order and slots: launched for a hint of the group
  // count (the last sort's with headroom), so that the host reads the counts
  // once, after them; the kernels stop at the device-side count, and a count
  // above the hint is redone at its exact size
  e->cell_order_set = false;
  uint32_t hint = e->a_hint;
  if (hint == 0) {  // no hint yet: the count first
    rc = sync_counts(e);
    if (rc) return rc;
    hint = e->hcount[CNT_GROUPS];
    if (e->hcount[CNT_DOMAIN]) {
      // tables hold clamped entries; keep A so the next sort clears them
      e->A = hint;
      e->S = 0;
      return fail(e, ZR_ERR_DOMAIN, "particle base node outside [0, 2^grid_log2 - 8]");
    }
  }
  if (hint > n) hint = n;
  if (hint == 0) hint = 1;
  rc = sort_groups(e, hint);
  if (rc) return rc;
  prof_mark(e, ZR_PROF_SORT);
  rc = sync_counts(e);
  if (rc) return rc;
  const uint32_t A = e->hcount[CNT_GROUPS];
  e->A = A;
  e->S = e->hcount[CNT_SLOTS];
  if (e->hcount[CNT_DOMAIN]) {
    // tables hold clamped entries; A and S are kept so the next sort clears them
    return fail(e, ZR_ERR_DOMAIN, "particle base node outside [0, 2^grid_log2 - 8]");
  }
  if (A > hint) {
    rc = sort_groups(e, A);
    if (rc) return rc;
    rc = sync_counts(e);
    if (rc) return rc;
    e->S = e->hcount[CNT_SLOTS];
  }
  e->a_hint = A + A / 4 + 64 < n ? A + A / 4 + 64 : n;
  // a rollout's earlier step asked for a restart: the particles are those
  // after that step (a sort moves none)
  if (e->defer_cfl && (e->hcount[CNT_CFL_RUN] & 1u)) return ZR_CFL_RESTART;
  rc = ensure_grid(e);
  if (rc) return rc;
  if (e->S) {
    k_slot_sources<<<blocks_for(8ull * e->S, 256), 256, 0, s>>>(e->S, e->slot_page, e->page_group, e->slot_src);
    ZR_LAUNCH_CHECK(e);
  }
  e->sorted = true;
  if (st) {
    st->num_groups = A;
    st->num_pages = e->S;
  }
  return ZR_OK;
}

// mls: the MLS P2G with force (explicit) or the APIC one (implicit, ADMM);
// fuse_dt > 0: a whole explicit step of that size follows, and the force
// goes into the momentum sum (k_p2g's FUSED form)
int do_p2g(zr_engine* e, bool mls, double fuse_dt = 0.0)
{
  if (!e->sorted) return fail(e, ZR_ERR_STATE, "particlesToGrid before sortParticlesAndPolluteGrid");
  const double dx = e->cfg.dx;
  prof_mark(e, ZR_PROF_P2G);
  if (mls && e->visco) {
    k_visco_stress<<<blocks_for((uint64_t)e->n, 256), 256, 0, e->stream>>>((uint32_t)e->n, e->fneq, e->mu_neq,
                                                                            e->lam_neq, e->vol, e->n, e->vtau_neq);
    ZR_LAUNCH_CHECK(e);
  }
  if (mls)
  {
    const ParticlesView pv = pview(e);
    const double Dinv = 4.0 / (dx * dx);
    const int model = e->has_models ? e->uniform_model : ZR_MODEL_FIXED_COROTATED;
    const bool fp32 = e->cfg.precision == ZR_FP32_REFINED, fused = fuse_dt > 0.0;
#define ZR_P2G_LAUNCH_T(M, T, F)                                                                        \
  k_p2g<true, M, T, F><<<e->A, kTransferThreads, 0, e->stream>>>(pv, e->order, e->group_start, e->group_page, \
                                                                 e->stage, dx, 1.0 / dx, Dinv, fuse_dt)
#define ZR_P2G_LAUNCH(M)                       \
  do {                                         \
    if (fp32 && fused)                         \
      ZR_P2G_LAUNCH_T(M, float, true);         \
    else if (fp32)                             \
      ZR_P2G_LAUNCH_T(M, float, false);        \
    else if (fused)                            \
      ZR_P2G_LAUNCH_T(M, double, true);        \
    else                                       \
      ZR_P2G_LAUNCH_T(M, double, false);       \
  } while (0)
    if (model == ZR_MODEL_NEO_HOOKEAN_BORDEN)
      ZR_P2G_LAUNCH(ZR_MODEL_NEO_HOOKEAN_BORDEN);
    else if (model == ZR_MODEL_STVK_HENCKY)
      ZR_P2G_LAUNCH(ZR_MODEL_STVK_HENCKY);
    else if (model == ZR_MODEL_FIXED_COROTATED)
      ZR_P2G_LAUNCH(ZR_MODEL_FIXED_COROTATED);
    else
      ZR_P2G_LAUNCH(kAnyModel);
#undef ZR_P2G_LAUNCH
#undef ZR_P2G_LAUNCH_T
    e->grid_nf = fused ? 4 : kP2GFields;
  }
  else
  {
    k_p2g<false><<<e->A, kTransferThreads, 0, e->stream>>>(pview(e), e->order, e->group_start, e->group_page, e->stage, dx,
                                              1.0 / dx, 4.0 / (dx * dx), 0.0);
    e->grid_nf = 4;
  }
  ZR_LAUNCH_CHECK(e);
  prof_mark(e, ZR_PROF_P2G);
  e->rasterised = true;
  e->mls_rasterised = mls;
  e->be_ready = false;
  return ZR_OK;
}

int do_grid(zr_engine* e, int mode, double dt)
{
  GridView G{ e->gm, e->gv, e->gf, e->gvs, (int64_t)e->grid_cap };
  ColliderSet cs = collider_set(e, dt);
  prof_mark(e, ZR_PROF_GRID);
  if (e->S) {
    const uint32_t nb = blocks_for((uint64_t)e->S * 32, 256);
    if (e->grid_nf == kP2GFields)
      k_grid<kP2GFields><<<nb, 256, 0, e->stream>>>(mode, e->S, e->slot_page, e->slot_src, e->stage, G, e->cfg.dx,
                                                    dt, e->cfg.gravity[0], e->cfg.gravity[1], e->cfg.gravity[2], cs);
    else
      k_grid<4><<<nb, 256, 0, e->stream>>>(mode, e->S, e->slot_page, e->slot_src, e->stage, G, e->cfg.dx, dt,
                                           e->cfg.gravity[0], e->cfg.gravity[1], e->cfg.gravity[2], cs);
  }
  ZR_LAUNCH_CHECK(e);
  prof_mark(e, ZR_PROF_GRID);
  if (mode != GRID_COMBINE) e->updated = true;
  return ZR_OK;
}

int do_g2p(zr_engine* e, double dt, zr_step_stats* st, bool mls)
{
  if (!e->updated) return fail(e, ZR_ERR_STATE, "gridToParticles before the grid update");
  const double dx = e->cfg.dx;
  if (!e->cfl_clear) ZR_HIP(e, hipMemsetAsync(e->dcount + CNT_CFL, 0, sizeof(uint32_t), e->stream));
  e->cfl_clear = false;
  ParticlesView pv = pview(e);
  pv.gradv = e->g2p_gradv;
  prof_mark(e, ZR_PROF_G2P);
  if (mls)
    k_g2p<true><<<e->A, g2p_threads<true>(), 0, e->stream>>>(pview(e), e->order, e->group_start, e->group_page, e->page_slot, e->gvs,
                                             (int64_t)e->grid_cap, dx, 1.0 / dx, 4.0 / (dx * dx), dt, e->cfg.cfl,
                                             e->dcount);
  else
    k_g2p<false><<<e->A, g2p_threads<false>(), 0, e->stream>>>(pv, e->order, e->group_start, e->group_page, e->page_slot,
                                              e->gvs, (int64_t)e->grid_cap, dx, 1.0 / dx, 4.0 / (dx * dx), dt,
                                              e->cfg.cfl, e->dcount);
  ZR_LAUNCH_CHECK(e);
  if (mls && e->visco) {  // the viscoelastic helper's evolveStrain with gradV = C
    k_visco_strain<<<blocks_for((uint64_t)e->n, 256), 256, 0, e->stream>>>(
        (uint32_t)e->n, e->C, e->fneq, e->mu_neq, e->lam_neq, e->n, dt, e->visc_d, e->visc_v);
    ZR_LAUNCH_CHECK(e);
  }
  if (e->has_plastic) {
    k_plasticity<<<blocks_for((uint64_t)e->n, 256), 256, 0, e->stream>>>(
        (uint32_t)e->n, e->mat, e->F, e->n, e->mu, e->lam, e->pparam, e->pstate, e->broken);
    ZR_LAUNCH_CHECK(e);
  }
  prof_mark(e, ZR_PROF_G2P);
  // particles moved: the sort is stale until the next one
  e->sorted = e->rasterised = e->updated = e->be_ready = false;
  if (e->defer_cfl) return ZR_OK;  // the flags are read by the next sort
  int rc = sync_counts(e);
  if (rc) return rc;
  uint32_t flags = e->hcount[CNT_CFL];
  if (st) st->cfl_flags = (int32_t)flags;
  return (flags & 1u) ? ZR_CFL_RESTART : ZR_OK;
}

// device scalars and flags of the Krylov solvers, with pinned host mirrors
int ensure_scalars(zr_engine* e)
{
  if (!e->sc) {
    ZR_HIP(e, hipMalloc(&e->sc, SC_N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->ci, CI_N * sizeof(int32_t)));
    ZR_HIP(e, hipMalloc(&e->hvp_ctr, kHvpCounters * kHvpCounterStride * sizeof(uint32_t)));
    ZR_HIP(e, hipMemset(e->hvp_ctr, 0, kHvpCounters * kHvpCounterStride * sizeof(uint32_t)));
    ZR_HIP(e, hipMalloc(&e->hvp_bounds, 18 * sizeof(uint32_t)));
    ZR_HIP(e, hipHostMalloc(&e->hsc, SC_N * sizeof(double)));
    ZR_HIP(e, hipHostMalloc(&e->hci, CI_N * sizeof(int32_t)));
  }
  return ZR_OK;
}

// ----------------------------------------------------------- phase field
PfParticles pf_view(zr_engine* e)
{
  return PfParticles{ e->pf_c, e->pf_H, e->pf_l0, e->pf_osc, e->pf_Fp, e->pf_Hmax, e->pf_k, e->pf_allow, e->pf_vol };
}

int do_pf(zr_engine* e, double dt)
{
  e->pf_stats = zr_phase_field_stats{};
  if (!e->pf_on || e->n == 0) return ZR_OK;
  const zr_phase_field_opts& o = e->pf_opts;
  hipStream_t s = e->stream;
  if (o.naive == ZR_DAMAGE_RANKINE) {
    k_pf_naive<<<blocks_for((uint64_t)e->n, 256), 256, 0, s>>>(pview(e), pf_view(e), o.sigma_F, e->gdeg);
    ZR_LAUNCH_CHECK(e);
    return ZR_OK;
  }
  if (!e->rasterised || !e->mls_rasterised)
    return fail(e, ZR_ERR_STATE, "phase-field solve before the MLS particlesToGrid of the step");
  int rc0 = ensure_scalars(e);
  if (rc0) return rc0;
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256);
  if (e->pf_node_cap < e->grid_cap) {
    dfree(e->pf_node);
    ZR_HIP(e, hipMalloc(&e->pf_node, 8 * e->grid_cap * sizeof(double)));
    e->pf_node_cap = e->grid_cap;
  }
  if (e->pf_partial_cap < 2 * (size_t)nb) {
    dfree(e->pf_partial);
    ZR_HIP(e, hipMalloc(&e->pf_partial, 2 * (size_t)nb * sizeof(double)));
    e->pf_partial_cap = 2 * (size_t)nb;
  }
  const size_t C = e->pf_node_cap;
  double *phase = e->pf_node, *rhs = phase + C, *diag = rhs + C, *x = diag + C, *r = x + C, *mr = r + C,
         *sv = mr + C, *q = sv + C;
  const double dx = e->cfg.dx;
  PfScalars ps{ o.parabolic_M, dt, o.cg_tolerance, o.cg_max_iterations };
  PfParticles pf = pf_view(e);
  k_pf_p2g<<<e->A, 256, 0, s>>>(pview(e), pf, ps, e->order, e->group_start, e->group_page, e->stage, 1.0 / dx);
  k_pf_nodes<<<nb, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, ps, phase, rhs, diag, x, r, mr, e->pf_partial,
                                nb);
  k_pf_reduce<<<1, 256, 0, s>>>(e->pf_partial, nb, e->sc, e->ci, PFR_INIT, o.cg_relative, ps);
  ZR_LAUNCH_CHECK(e);
  // the fracture CG runs iterations 0..max (it checks it == max after the update)
  const int total = o.cg_max_iterations + 1;
  const int check = 8;
  for (int it = 0; it < total; ++it) {
    k_pf_apply<<<e->A, 256, 0, s>>>(e->x, e->n, pf, ps, e->order, e->group_start, e->group_page, e->group_nslot,
                                    e->gm, mr, sv, e->sc, e->ci, e->stage, 1.0 / dx);
    k_pf_ap<<<nb, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, diag, mr, sv, q, e->sc, e->ci, e->pf_partial);
    k_pf_reduce<<<1, 256, 0, s>>>(e->pf_partial, nb, e->sc, e->ci, PFR_ALPHA, o.cg_relative, ps);
    k_pf_update<<<nb, 256, 0, s>>>(nn, e->gm, diag, sv, q, x, r, mr, e->sc, e->ci, e->pf_partial, nb);
    k_pf_reduce<<<1, 256, 0, s>>>(e->pf_partial, nb, e->sc, e->ci, PFR_CHECK, o.cg_relative, ps);
    ZR_LAUNCH_CHECK(e);
    if ((it + 1) % check == 0 || it + 1 == total) {
      ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      ZR_HIP(e, hipStreamSynchronize(s));
      if (e->hci[CI_DONE]) break;
    }
  }
  k_pf_g2p<<<e->A, 256, 0, s>>>(pview(e), pf, ps, e->order, e->group_start, e->group_page, e->group_nslot, e->gm, x,
                                phase, e->gdeg, 1.0 / dx);
  if (o.naive == ZR_DAMAGE_PHASE_FIELD_RANKINE)
    k_pf_naive<<<blocks_for((uint64_t)e->n, 256), 256, 0, s>>>(pview(e), pf, o.sigma_F, e->gdeg);
  ZR_LAUNCH_CHECK(e);
  ZR_HIP(e, hipMemcpyAsync(e->hsc, e->sc, SC_N * sizeof(double), hipMemcpyDeviceToHost, s));
  ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ZR_HIP(e, hipStreamSynchronize(s));
  e->pf_stats.cg_iterations = e->hci[CI_ITERS];
  e->pf_stats.residual_norm = e->hsc[PF_NORM];
  return ZR_OK;
}

// ------------------------------------------------------ differentiable form
constexpr int kAdjRows = 56;

AdjointView adj_view(zr_engine* e)
{
  const int64_t n = e->n;
  double* a = e->adj;
  return AdjointView{ a,          a + 3 * n,  a + 6 * n,  a + 15 * n, a + 24 * n,
                      a + 25 * n, a + 26 * n, a + 29 * n, a + 38 * n, a + 47 * n };
}

int ensure_adjoint(zr_engine* e)
{
  if (e->adj_n != e->n) {
    dfree(e->adj);
    ZR_HIP(e, hipMalloc(&e->adj, kAdjRows * (size_t)e->n * sizeof(double)));
    ZR_HIP(e, hipMemset(e->adj, 0, kAdjRows * (size_t)e->n * sizeof(double)));
    e->adj_n = e->n;
  }
  return ZR_OK;
}

bool adjoint_supported(zr_engine* e)
{
  if (e->has_models || e->has_plastic || e->pf_on || e->visco) {
    fail(e, ZR_ERR_ARG,
         "the adjoint supports fixed-corotated particles without plasticity, damage or a viscoelastic branch only");
    return false;
  }
  return true;
}

// reverse of one explicit step from the current (input) particle state
// check: wait for the step and report a collider branch without a
// derivative (CNT_ADJ_BAD, which the caller cleared); a rollout reads it once
// at its end instead
int adjoint_flags(zr_engine* e)
{
  int rc = sync_counts(e);
  if (rc) return rc;
  if (e->hcount[CNT_ADJ_BAD])
    return fail(e, ZR_ERR_ARG, "adjoint: a grid node at a sphere collider's centre in the trial collision (no derivative)");
  return ZR_OK;
}

// what the adjoint kernels read of a step's sort and grid: the engine's own
// (after a sort and the seven-field rasterisation of the step's input) or a
// StepRecord's copy; grid rows of `stride` nodes
struct AdjTables {
  uint32_t A, S;
  const uint32_t* order;
  const uint32_t* group_start;
  const uint32_t* group_page;
  const int32_t* group_nslot;
  const uint32_t* slot_page;
  const int32_t* slot_src;
  double *gm, *gv, *gf, *gvs;
  int64_t stride;
};

AdjTables engine_tables(zr_engine* e)
{
  return AdjTables{ e->A,        e->S,         e->order, e->group_start, e->group_page, e->group_nslot,
                    e->slot_page, e->slot_src, e->gm,    e->gv,          e->gf,         e->gvs,
                    (int64_t)e->grid_cap };
}

// the reverse of the step whose input the particles hold, on its tables
int adjoint_kernels(zr_engine* e, double dt, const AdjTables& t)
{
  const uint32_t nn = t.S * 32;
  if (e->adj_grid_cap < (size_t)t.stride) {
    dfree(e->adj_grid);
    e->adj_grid_cap = 0;
    ZR_HIP(e, hipMalloc(&e->adj_grid, kAdjGridFields * (size_t)t.stride * sizeof(double)));
    e->adj_grid_cap = (size_t)t.stride;
  }
  if ((size_t)t.A > e->stage_cap) return fail(e, ZR_ERR_STATE, "adjoint: staging buffer below the step's groups");
  const double dx = e->cfg.dx;
  hipStream_t s = e->stream;
  AdjointView A = adj_view(e);
  const uint32_t n = (uint32_t)e->n;
  k_adj_stress<<<blocks_for(n, 256), 256, 0, s>>>(pview(e), A, n);
  ZR_LAUNCH_CHECK(e);
  k_adj_g2p<<<t.A, 256, 0, s>>>(pview(e), A, t.order, t.group_start, t.group_page, t.group_nslot, t.gvs, t.stride,
                                e->stage, dx, 1.0 / dx, 4.0 / (dx * dx), dt);
  ZR_LAUNCH_CHECK(e);
  GridView G{ t.gm, t.gv, t.gf, t.gvs, t.stride };
  // the grid kernel reads the adjoint stride-wise like the grid fields
  if (nn) {
    k_adj_grid<<<blocks_for(nn, 256), 256, 0, s>>>(nn, t.slot_page, t.slot_src, e->stage, G, e->adj_grid, dx, dt,
                                                   e->cfg.gravity[0], e->cfg.gravity[1], e->cfg.gravity[2],
                                                   collider_set(e, dt), e->dcount + CNT_ADJ_BAD);
    ZR_LAUNCH_CHECK(e);
  }
  k_adj_p2g<<<t.A, 256, 0, s>>>(pview(e), A, t.order, t.group_start, t.group_page, t.group_nslot, e->adj_grid,
                                t.stride, dx, 1.0 / dx, 4.0 / (dx * dx));
  ZR_LAUNCH_CHECK(e);
  k_adj_stress_rev<<<blocks_for(n, 256), 256, 0, s>>>(pview(e), A, n);
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

int do_adjoint_step(zr_engine* e, double dt, bool check)
{
  int rc = do_sort(e, nullptr);
  if (rc) return rc;
  rc = do_p2g(e, true);
  if (rc) return rc;
  rc = do_grid(e, GRID_COMBINE_UPDATE, dt);
  if (rc) return rc;
  rc = adjoint_kernels(e, dt, engine_tables(e));
  if (rc) return rc;
  // the particles were not moved: the sort stays valid, the grid does too
  return check ? adjoint_flags(e) : ZR_OK;
}

// up to kCopySegs device copies in one launch (sizes in 4-byte words)
constexpr int kCopySegs = 16;
struct CopyList {
  const uint32_t* src[kCopySegs];
  uint32_t* dst[kCopySegs];
  uint64_t words[kCopySegs];
  int count;
};

__global__ __launch_bounds__(256) void k_multi_copy(CopyList L)
{
  const uint64_t stride = (uint64_t)gridDim.x * 256u;
  for (int q = 0; q < L.count; ++q) {
    const uint32_t* __restrict__ src = L.src[q];
    uint32_t* __restrict__ dst = L.dst[q];
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < L.words[q]; i += stride) dst[i] = src[i];
  }
}

// a StepRecord's layout: order [n], group_start [A + 1], group_page [A],
// group_nslot [8A], slot_page [S], slot_src [8S] (4-byte words), then m,
// v 3, f 3, v* 3 over the S * 32 nodes (doubles, 8-byte aligned)
struct RecordLayout {
  size_t order, gstart, gpage, gnslot, spage, ssrc, grid, bytes;
};

RecordLayout record_layout(size_t n, size_t A, size_t S)
{
  RecordLayout L;
  size_t w = 0;
  L.order = w; w += n;
  L.gstart = w; w += A + 1;
  L.gpage = w; w += A;
  L.gnslot = w; w += 8 * A;
  L.spage = w; w += S;
  L.ssrc = w; w += 8 * S;
  w += w & 1u;
  L.grid = w;
  w += 2 * 10 * 32 * S;
  L.bytes = 4 * w;
  return L;
}

// copies the engine's sort tables and grid (after the seven-field
// rasterisation and the grid update of a step) into record r
int record_step(zr_engine* e, zr_engine::StepRecord& r)
{
  const size_t n = (size_t)e->n, A = e->A, S = e->S, nodes = 32 * S;
  const RecordLayout L = record_layout(n, A, S);
  if (r.cap < L.bytes) {
    dfree(r.buf);
    r.cap = 0;
    const size_t cap = L.bytes + L.bytes / 4 + 4096;
    ZR_HIP(e, hipMalloc(&r.buf, cap));
    r.cap = cap;
  }
  r.A = e->A;
  r.S = e->S;
  uint32_t* w = static_cast<uint32_t*>(r.buf);
  CopyList c{};
  // 16 segments: six tables, the mass and three rows each of v, f and v*
  static_assert(kCopySegs >= 6 + 1 + 3 * 3, "record segments");
  auto add = [&](const void* src, size_t off, size_t words) {
    c.src[c.count] = static_cast<const uint32_t*>(src);
    c.dst[c.count] = w + off;
    c.words[c.count] = words;
    ++c.count;
  };
  add(e->order, L.order, n);
  add(e->group_start, L.gstart, A + 1);
  add(e->group_page, L.gpage, A);
  add(e->group_nslot, L.gnslot, 8 * A);
  add(e->slot_page, L.spage, S);
  add(e->slot_src, L.ssrc, 8 * S);
  add(e->gm, L.grid, 2 * nodes);
  for (int d = 0; d < 3; ++d) {
    add(e->gv + d * e->grid_cap, L.grid + 2 * nodes * (1 + d), 2 * nodes);
    add(e->gf + d * e->grid_cap, L.grid + 2 * nodes * (4 + d), 2 * nodes);
    add(e->gvs + d * e->grid_cap, L.grid + 2 * nodes * (7 + d), 2 * nodes);
  }
  const uint64_t b = blocks_for(L.bytes / 4, 256);
  k_multi_copy<<<b < 2048 ? (uint32_t)b : 2048u, 256, 0, e->stream>>>(c);
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

AdjTables record_tables(zr_engine* e, const zr_engine::StepRecord& r)
{
  const RecordLayout L = record_layout((size_t)e->n, r.A, r.S);
  uint32_t* w = static_cast<uint32_t*>(r.buf);
  double* g = reinterpret_cast<double*>(w + L.grid);
  const size_t nodes = 32 * (size_t)r.S;
  // grid rows of the record are `nodes` long; at least one node so that the
  // stride is positive for an empty grid
  const int64_t st = nodes ? (int64_t)nodes : 1;
  return AdjTables{ r.A,
                    r.S,
                    w + L.order,
                    w + L.gstart,
                    w + L.gpage,
                    reinterpret_cast<const int32_t*>(w + L.gnslot),
                    w + L.spage,
                    reinterpret_cast<const int32_t*>(w + L.ssrc),
                    g,
                    g + nodes,
                    g + 4 * nodes,
                    g + 7 * nodes,
                    st };
}

// one explicit step in the rasterisation the adjoint reads (force and
// momentum in their own fields), its tables and grid copied into r
int advance_recorded(zr_engine* e, double dt, zr_engine::StepRecord& r)
{
  int rc = do_sort(e, nullptr);
  if (rc) return rc;
  rc = do_p2g(e, true);
  if (rc) return rc;
  rc = do_grid(e, GRID_COMBINE_UPDATE, dt);
  if (rc) return rc;
  rc = record_step(e, r);
  if (rc) return rc;
  rc = do_g2p(e, dt, nullptr, true);
  if (rc < 0) return rc;
  e->time += dt;
  return rc;
}

// ------------------------------------------------------------------- ADMM
constexpr int kAdmmNodeRows = 3 + 3 + 9 + 1 + 3 + 9 + 3 + 9;  // vn dv proj diag rhs q(3) z m(3)

// the global operator's particle term of x into the staged tiles: the FP64
// scatter from the particle arrays, or the FP32 apply over the sorted cache
struct AdmmOp {
  bool fp32, visco;
  ParticlesView P;
  AdmmParticles Ap;
  AdmmNodes Nd;
  AdmmCache32 C;
  double dt, rho, inv_dx;
};

template <class TX>
void admm_apply(zr_engine* e, const AdmmOp& op, const TX* x)
{
  hipStream_t s = e->stream;
  if (op.fp32) {
    if (op.visco)
      k_admm_op32<true, TX><<<e->A, 256, 0, s>>>(op.C, e->group_start, e->group_nslot, e->gm, x, op.Nd.stride,
                                                 e->stage, (float)op.inv_dx, e->ci);
    else
      k_admm_op32<false, TX><<<e->A, 256, 0, s>>>(op.C, e->group_start, e->group_nslot, e->gm, x, op.Nd.stride,
                                                  e->stage, (float)op.inv_dx, e->ci);
  } else if constexpr (std::is_same<TX, double>::value) {
    if (op.visco)
      k_admm_scatter<1, true><<<e->A, 256, 0, s>>>(op.P, op.Ap, op.Nd, x, e->order, e->group_start, e->group_page,
                                                   e->group_nslot, e->gm, e->stage, op.dt, op.rho, op.inv_dx, e->ci);
    else
      k_admm_scatter<1, false><<<e->A, 256, 0, s>>>(op.P, op.Ap, op.Nd, x, e->order, e->group_start, e->group_page,
                                                    e->group_nslot, e->gm, e->stage, op.dt, op.rho, op.inv_dx, e->ci);
  }
}

// Minres::solve (Minres.h:69-149) on the global system for dv (FP64), the
// Krylov vectors in precision T (the seven vectors rotate by pointer)
template <class T>
int admm_minres(zr_engine* e, const AdmmOp& op, const double* rhs, T* const qb[3], T* zb, T* const mb[3],
                const zr_admm_opts* o, uint32_t nn, uint32_t nbn)
{
  hipStream_t s = e->stream;
  const AdmmNodes& Nd = op.Nd;
  const int64_t C = Nd.stride;
  k_admm_reset_flags<<<1, 64, 0, s>>>(e->ci);
  admm_apply<double>(e, op, Nd.dv);
  T *q = qb[2], *qk = qb[1], *qkm1 = qb[0], *z = zb, *mk = mb[2], *mkm1 = mb[1], *mkm2 = mb[0];
  k_minres_init<T><<<nbn, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, Nd, rhs, Nd.dv, q, z, mk, mkm1, qk,
                                       e->partial);
  k_minres_reduce<<<1, 256, 0, s>>>(e->partial, nbn, e->sc, e->ci, MR_INIT, o->global_tolerance, 0.5, 0,
                                    o->minres_max_iterations);
  k_minres_scale<T><<<nbn, 256, 0, s>>>(nn, e->gm, C, q, z, e->sc, e->ci);
  ZR_LAUNCH_CHECK(e);
  int next_poll = 1;
  for (int k = 0; k < o->minres_max_iterations; ++k) {
    k_minres_begin<<<1, 64, 0, s>>>(e->sc, e->ci);
    // m_{k-2} <- m_{k-1} <- m_k <- z; q_{k-1} <- q_k <- q_{k+1}
    T* t = mkm2;
    mkm2 = mkm1;
    mkm1 = mk;
    mk = z;
    z = t;
    t = qkm1;
    qkm1 = qk;
    qk = q;
    q = t;
    admm_apply<T>(e, op, mk);
    k_minres_apply<T><<<nbn, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, Nd, mk, q, e->ci, e->partial);
    k_minres_reduce<<<1, 256, 0, s>>>(e->partial, nbn, e->sc, e->ci, MR_ALPHA, 0.0, 0.0, k, o->minres_max_iterations);
    k_minres_lanczos<T><<<nbn, 256, 0, s>>>(nn, e->gm, Nd, q, qk, qkm1, z, e->sc, e->ci, e->partial);
    k_minres_reduce<<<1, 256, 0, s>>>(e->partial, nbn, e->sc, e->ci, MR_GIVENS, 0.0, 0.0, k,
                                      o->minres_max_iterations);
    k_minres_update<T><<<nbn, 256, 0, s>>>(nn, e->gm, Nd, q, z, mk, mkm1, mkm2, Nd.dv, e->sc, e->ci);
    ZR_LAUNCH_CHECK(e);
    if (k + 1 == next_poll || k + 1 == o->minres_max_iterations) {
      next_poll *= 2;
      ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      ZR_HIP(e, hipStreamSynchronize(s));
      if (e->hci[AI_DONE]) break;
    }
  }
  ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ZR_HIP(e, hipStreamSynchronize(s));
  return ZR_OK;
}

int do_admm(zr_engine* e, double dt, const zr_admm_opts* o, zr_admm_stats* st)
{
  hipStream_t s = e->stream;
  const int64_t n = e->n;
  const size_t N = (size_t)n;
  int rc = ensure_scalars(e);
  if (rc) return rc;
  const bool visco = e->visco;
  const bool first = e->admm_steps == 0 || e->admm_n != n || e->admm_visco != visco;
  if (e->admm_n != n || e->admm_visco != visco) {
    const size_t rows = visco ? 41 : 22;
    dfree(e->admm_p);
    ZR_HIP(e, hipMalloc(&e->admm_p, rows * N * sizeof(double)));
    ZR_HIP(e, hipMemsetAsync(e->admm_p, 0, rows * N * sizeof(double), s));
    e->admm_n = n;
    e->admm_visco = visco;
    e->admm_steps = 0;
  }
  if (e->admm_cap < e->grid_cap) {
    dfree(e->admm_node);
    dfree(e->admm_coll);
    ZR_HIP(e, hipMalloc(&e->admm_node, kAdmmNodeRows * e->grid_cap * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->admm_coll, e->grid_cap * sizeof(int32_t)));
    e->admm_cap = e->grid_cap;
  }
  size_t nb = (e->grid_cap + 255) / 256 + 1;
  if (nb > e->partial_cap) {
    dfree(e->partial);
    ZR_HIP(e, hipMalloc(&e->partial, nb * sizeof(double)));
    e->partial_cap = nb;
  }
  const int64_t C = (int64_t)e->admm_cap;
  double* base = e->admm_node;
  AdmmNodes Nd{ base, base + 3 * C, base + 6 * C, e->admm_coll, base + 15 * C, C };
  double* rhs = base + 16 * C;
  double* qb[3] = { base + 19 * C, base + 22 * C, base + 25 * C };
  double* zb = base + 28 * C;
  double* mb[3] = { base + 31 * C, base + 34 * C, base + 37 * C };
  double* pa = e->admm_p;
  AdmmParticles Ap{ pa, pa + 9 * N, pa + 18 * N, pa + 19 * N, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                    e->visc_d, e->visc_v };
  if (visco) {
    Ap.F2 = pa + 22 * N;
    Ap.y2 = pa + 31 * N;
    Ap.om2 = pa + 40 * N;
    Ap.fneq = e->fneq;
    Ap.mu2 = e->mu_neq;
    Ap.lam2 = e->lam_neq;
  }
  const uint32_t nn = e->S * 32, nbn = blocks_for(nn, 256);
  const double dx = e->cfg.dx, inv_dx = 1.0 / dx, rho = o->rho_scale;
  ColliderSet cs = collider_set(e);
  ParticlesView P = pview(e);
  const int32_t* mat = e->mat;
  // initStep
  if (!first)
    k_admm_lastdv_p2g<<<e->A, 256, 0, s>>>(P, Ap, e->order, e->group_start, e->group_page, e->stage, inv_dx);
  k_admm_init_nodes<<<nbn, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, e->gv, Nd, first ? 0 : 1, dt);
  k_admm_collide<<<nbn, 256, 0, s>>>(nn, e->slot_page, e->gm, Nd, dx, cs);
  if (visco)
    k_admm_init_particles<true><<<e->A, 256, 0, s>>>(P, Ap, Nd, mat, e->pparam, e->pstate, e->order, e->group_start,
                                                     e->group_page, e->group_nslot, e->gm, e->stage, first ? 1 : 0,
                                                     dt, rho, inv_dx);
  else
    k_admm_init_particles<false><<<e->A, 256, 0, s>>>(P, Ap, Nd, mat, e->pparam, e->pstate, e->order, e->group_start,
                                                      e->group_page, e->group_nslot, e->gm, e->stage, first ? 1 : 0,
                                                      dt, rho, inv_dx);
  k_admm_diag<<<nbn, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, Nd.diag);
  ZR_LAUNCH_CHECK(e);
  // the FP32-bulk global step: per-particle operator data in FP32, sorted,
  // and FP32 Krylov vectors
  const bool fp32 = o->precision == ZR_FP32_REFINED;
  AdmmOp op{ fp32, visco, P, Ap, Nd, AdmmCache32{}, dt, rho, inv_dx };
  float *qb32[3] = { nullptr, nullptr, nullptr }, *zb32 = nullptr, *mb32[3] = { nullptr, nullptr, nullptr };
  if (fp32) {
    const size_t rows = visco ? 23 : 13;
    if (e->admm_c32_cap < N || e->admm_c32_rows < rows) {
      dfree(e->admm_c32);
      dfree(e->admm_cell32);
      e->admm_c32_cap = 0;
      ZR_HIP(e, hipMalloc(&e->admm_c32, rows * N * sizeof(float)));
      ZR_HIP(e, hipMalloc(&e->admm_cell32, N));
      e->admm_c32_cap = N;
      e->admm_c32_rows = rows;
    }
    if (e->admm_node32_cap < (size_t)C) {
      dfree(e->admm_node32);
      e->admm_node32_cap = 0;
      ZR_HIP(e, hipMalloc(&e->admm_node32, 21 * (size_t)C * sizeof(float)));
      e->admm_node32_cap = (size_t)C;
    }
    float* c = e->admm_c32;
    op.C = AdmmCache32{ c, c + 3 * N, c + 12 * N, visco ? c + 13 * N : nullptr, visco ? c + 22 * N : nullptr,
                        e->admm_cell32, n };
    float* b32 = e->admm_node32;
    for (int k = 0; k < 3; ++k) {
      qb32[k] = b32 + 3 * k * C;
      mb32[k] = b32 + 3 * (4 + k) * C;
    }
    zb32 = b32 + 9 * C;
    if (visco)
      k_admm_cache32<true><<<blocks_for((uint64_t)n, 256), 256, 0, s>>>(P, Ap, e->order, (uint32_t)n, op.C, inv_dx,
                                                                       rho, dt);
    else
      k_admm_cache32<false><<<blocks_for((uint64_t)n, 256), 256, 0, s>>>(P, Ap, e->order, (uint32_t)n, op.C, inv_dx,
                                                                        rho, dt);
    ZR_LAUNCH_CHECK(e);
  }
  // the local step's per-particle grad(vn + dv) (9 rows of n)
  if (e->be_fb_cap < N) {
    dfree(e->be_fb);
    e->be_fb_cap = 0;
    ZR_HIP(e, hipMalloc(&e->be_fb, 9 * N * sizeof(double)));
    e->be_fb_cap = N;
  }
  const uint32_t npb = blocks_for((uint64_t)n, 256);
  int total_minres = 0;
  for (int tim = 0; tim < o->admm_iterations; ++tim) {
    k_admm_collide<<<nbn, 256, 0, s>>>(nn, e->slot_page, e->gm, Nd, dx, cs);
    k_admm_fgrad<<<e->A, 256, 0, s>>>(P, Nd, e->order, e->group_start, e->group_page, e->group_nslot, e->gm, inv_dx,
                                      e->be_fb);
    if (visco) {
      k_admm_flocal<true><<<npb, 256, 0, s>>>(P, Ap, mat, e->pparam, e->pstate, (uint32_t)n, e->be_fb, dt, rho,
                                              o->local_tolerance);
      k_admm_scatter<0, true><<<e->A, 256, 0, s>>>(P, Ap, Nd, nullptr, e->order, e->group_start, e->group_page,
                                                   e->group_nslot, e->gm, e->stage, dt, rho, inv_dx, e->ci);
    } else {
      k_admm_flocal<false><<<npb, 256, 0, s>>>(P, Ap, mat, e->pparam, e->pstate, (uint32_t)n, e->be_fb, dt, rho,
                                               o->local_tolerance);
      k_admm_scatter<0, false><<<e->A, 256, 0, s>>>(P, Ap, Nd, nullptr, e->order, e->group_start, e->group_page,
                                                    e->group_nslot, e->gm, e->stage, dt, rho, inv_dx, e->ci);
    }
    k_admm_rhs<<<nbn, 256, 0, s>>>(nn, e->slot_src, e->stage, e->gm, rhs, C, dt, e->cfg.gravity[0],
                                   e->cfg.gravity[1], e->cfg.gravity[2]);
    // MINRES on the global system, warm-started from dv (Minres.h:69-149),
    // relative tolerance min(0.5, sqrt(max(1, 1e-6))) = 0.5 (:785-787)
    int rc2 = fp32 ? admm_minres<float>(e, op, rhs, qb32, zb32, mb32, o, nn, nbn)
                   : admm_minres<double>(e, op, rhs, qb, zb, mb, o, nn, nbn);
    if (rc2) return rc2;
    total_minres += e->hci[AI_ITERS];
    if (visco)
      k_admm_ustep<true><<<e->A, 256, 0, s>>>(P, Ap, Nd, e->order, e->group_start, e->group_page, e->group_nslot,
                                              e->gm, dt, rho, inv_dx);
    else
      k_admm_ustep<false><<<e->A, 256, 0, s>>>(P, Ap, Nd, e->order, e->group_start, e->group_page, e->group_nslot,
                                               e->gm, dt, rho, inv_dx);
    ZR_LAUNCH_CHECK(e);
  }
  k_admm_finish_nodes<<<nbn, 256, 0, s>>>(nn, e->gm, Nd, e->gvs);
  k_admm_lastdv<<<e->A, 256, 0, s>>>(P, Ap, Nd, e->order, e->group_start, e->group_page, e->group_nslot, e->gm, dt,
                                     inv_dx);
  ZR_LAUNCH_CHECK(e);
  e->admm_steps++;
  e->updated = true;
  if (st) st->minres_iterations = total_minres;
  return ZR_OK;
}

// ---------------------------------------------------------------- implicit
int ensure_implicit(zr_engine* e, int32_t prec)
{
  if (e->imp_cap < e->grid_cap) {
    double** arrs[] = { &e->ivn, &e->idv, &e->ir, &e->ix, &e->irc, &e->iq, &e->ip, &e->iap };
    for (auto a : arrs) {
      dfree(*a);
      ZR_HIP(e, hipMalloc(a, 3 * e->grid_cap * sizeof(double)));
    }
    e->imp_cap = e->grid_cap;
  }
  if (prec == ZR_FP32_REFINED && e->imp32_cap < e->grid_cap) {
    float** arrs[] = { &e->fx, &e->frc, &e->fq, &e->fp, &e->fap };
    for (auto a : arrs) {
      dfree(*a);
      ZR_HIP(e, hipMalloc(a, 3 * e->grid_cap * sizeof(float)));
    }
    e->imp32_cap = e->grid_cap;
  }
  size_t nb = (e->grid_cap + 255) / 256 + 1;
  if (nb > e->partial_cap) {
    dfree(e->partial);
    ZR_HIP(e, hipMalloc(&e->partial, nb * sizeof(double)));
    e->partial_cap = nb;
  }
  if (!e->partial2) ZR_HIP(e, hipMalloc(&e->partial2, kCgBlocks * sizeof(double)));
  int rc = ensure_scalars(e);
  if (rc) return rc;
  const size_t n = (size_t)(e->n > 0 ? e->n : 1);
  if (e->ccell_n != e->n) {
    dfree(e->ccell);
    ZR_HIP(e, hipMalloc(&e->ccell, n * sizeof(uint32_t)));
    e->ccell_n = e->n;
  }
  if (prec == ZR_FP64 && e->c64_n != e->n) {
    dfree(e->c64);
    ZR_HIP(e, hipMalloc(&e->c64, 26 * n * sizeof(double)));
    e->c64_n = e->n;
  }
  if (prec == ZR_FP32_REFINED && e->c32_n != e->n) {
    dfree(e->c32);
    ZR_HIP(e, hipMalloc(&e->c32, 26 * n * sizeof(float)));
    e->c32_n = e->n;
  }
  return ZR_OK;
}

template <class T>
HvpCache<T> hvp_cache(zr_engine* e);
template <>
HvpCache<double> hvp_cache<double>(zr_engine* e)
{
  const int64_t S = e->n;
  double* b = e->c64;
  return HvpCache<double>{ b, b + 9 * S, b + 18 * S, b + 21 * S, b + 23 * S, e->ccell, S };
}
template <>
HvpCache<float> hvp_cache<float>(zr_engine* e)
{
  const int64_t S = e->n;
  float* b = e->c32;
  return HvpCache<float>{ b, b + 9 * S, b + 18 * S, b + 21 * S, b + 23 * S, e->ccell, S };
}

struct CgVecs64 {
  using T = double;
  static double*& extra(zr_engine* e) { return e->mr64; }
  static size_t& extra_cap(zr_engine* e) { return e->mr64_cap; }
  static double* x(zr_engine* e) { return e->ix; }
  static double* r(zr_engine* e) { return e->irc; }
  static double* q(zr_engine* e) { return e->iq; }
  static double* p(zr_engine* e) { return e->ip; }
  static double* ap(zr_engine* e) { return e->iap; }
};
struct CgVecs32 {
  using T = float;
  static float*& extra(zr_engine* e) { return e->mr32; }
  static size_t& extra_cap(zr_engine* e) { return e->mr32_cap; }
  static float* x(zr_engine* e) { return e->fx; }
  static float* r(zr_engine* e) { return e->frc; }
  static float* q(zr_engine* e) { return e->fq; }
  static float* p(zr_engine* e) { return e->fp; }
  static float* ap(zr_engine* e) { return e->fap; }
};

NodeVec3 nv(zr_engine* e, double* p) { return NodeVec3{ p, (int64_t)e->grid_cap }; }
template <class T>
NodeVec<T> nvT(zr_engine* e, T* p)
{
  return NodeVec<T>{ p, (int64_t)e->grid_cap };
}

int read_scalars(zr_engine* e)
{
  ZR_HIP(e, hipMemcpyAsync(e->hsc, e->sc, SC_N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

// Newton state at the current dv + residual + its squared mass norm (device);
// with jac, also the operator diagonal's particle sums for the Jacobi
// preconditioner (staged in e->jstage)
template <class T>
int be_state_residual(zr_engine* e, bool jac = false)
{
  const double dx = e->cfg.dx, dt = e->be_dt;
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256);
  ColliderSet cs = collider_set(e);
  uint8_t* co = e->cell_order_set ? nullptr : e->cell_order;
  prof_mark(e, ZR_PROF_STATE);
  if (jac)
    k_be_state<T, true><<<e->A, 256, 0, e->stream>>>(pview(e), hvp_cache<T>(e), e->order, e->group_start,
                                                     e->group_page, e->page_slot, e->ivn, e->idv,
                                                     (int64_t)e->grid_cap, dx, 1.0 / dx, dt, e->stage, co, e->jstage);
  else {
    // the same state in three passes (k_be_grad, k_be_svd, k_be_scatter):
    // the SVD at a higher occupancy than the fused kernel's
    const size_t N = (size_t)e->n;
    if (e->be_fb_cap < N) {
      dfree(e->be_fb);
      e->be_fb_cap = 0;
      ZR_HIP(e, hipMalloc(&e->be_fb, 9 * N * sizeof(double)));
      e->be_fb_cap = N;
    }
    const HvpCache<T> hc = hvp_cache<T>(e);
    k_be_grad<T><<<e->A, 256, 0, e->stream>>>(pview(e), hc, e->order, e->group_start, e->group_page, e->page_slot,
                                              e->ivn, e->idv, (int64_t)e->grid_cap, 1.0 / dx, dt, e->be_fb, co);
    k_be_svd<T><<<blocks_for(N, 256), 256, 0, e->stream>>>(pview(e), hc, e->order, (uint32_t)N, e->be_fb);
    k_be_scatter<<<e->A, 256, 0, e->stream>>>(pview(e), e->order, e->group_start, e->group_page, hc.stride,
                                              1.0 / dx, e->be_fb, e->stage);
  }
  ZR_LAUNCH_CHECK(e);
  e->cell_order_set = true;  // the cells do not move within a step
  prof_mark(e, ZR_PROF_STATE);
  k_be_residual<<<nb, 256, 0, e->stream>>>(nn, e->slot_page, e->slot_src, e->stage, e->gm, nv(e, e->idv),
                                           nv(e, e->ir), dx, dt, e->cfg.gravity[0], e->cfg.gravity[1],
                                           e->cfg.gravity[2], cs, e->partial);
  k_reduce_cg<<<1, 256, 0, e->stream>>>(e->partial, nb, e->sc, e->ci, RED_STORE, SC_NORM2, 0.0, 0.0);
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

// out = A p, or A (q + beta p) with p updated to that direction when q is
// given (the CG step); projected when requested; the CG flags gate it
// cg_iter >= 1: the CG iteration whose direction is qvec + beta pvec, beta
// from the z^T r partials of the previous update (e->partial2); 0 otherwise
template <class T>
int be_apply(zr_engine* e, T* qvec, T* pvec, T* out, int project, int32_t* ci, int cg_iter)
{
  const double dx = e->cfg.dx, dt = e->be_dt;
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256), ncg = nb < kCgBlocks ? nb : kCgBlocks;
  ColliderSet cs = collider_set(e);
  T* stage = reinterpret_cast<T*>(e->stage);
  prof_mark(e, ZR_PROF_HVP);
  const uint32_t resident = e->hvp_resident[sizeof(T) == sizeof(double) ? 1 : 0];
  const uint32_t nblk = e->A < resident ? e->A : resident;
  // the CG step (qvec given) forms A p over the listed nodes with mass; the
  // plain product writes every slot node (zeros without mass)
  const uint32_t* act = qvec ? e->act : nullptr;
  const uint32_t nap = qvec ? e->ncg_act : ncg;
  k_hvp<T><<<nblk, HvpChunk<T>::threads, 0, e->stream>>>(hvp_cache<T>(e), e->group_start, e->group_nslot, qvec, pvec, e->sc,
                                        (int64_t)e->grid_cap, (T)(1.0 / dx), stage, ci,
                                        e->hvp_bounds + (sizeof(T) == sizeof(double) ? 9 : 0),
                                        e->hvp_order + (sizeof(T) == sizeof(double) ? (size_t)e->A : 0), e->hvp_ctr,
                                        e->partial2, ncg, cg_iter, e->sc, e->cell_order);
  ZR_LAUNCH_CHECK(e);
  prof_mark(e, ZR_PROF_HVP);
  k_cg_ap<T><<<nap, 256, 0, e->stream>>>(nn, e->slot_page, e->slot_src, stage, e->gm, nvT(e, qvec), nvT(e, pvec),
                                         nvT(e, out), dx, dt, cs, project, e->partial, e->sc, ci,
                                         e->hvp_ctr, act, e->nact);
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

// buffers of the Jacobi preconditioner for the current grid and groups
int ensure_jacobi(zr_engine* e)
{
  if (e->jac_cap < e->grid_cap) {
    dfree(e->pinv);
    e->jac_cap = 0;
    ZR_HIP(e, hipMalloc(&e->pinv, 3 * e->grid_cap * sizeof(double)));
    e->jac_cap = e->grid_cap;
  }
  if (e->jstage_cap < (size_t)e->A || !e->jstage) {
    dfree(e->jstage);
    e->jstage_cap = 0;
    const size_t cap = (size_t)e->A + e->A / 4 + 64;
    ZR_HIP(e, hipMalloc(&e->jstage, cap * kImpStage * sizeof(double)));
    e->jstage_cap = cap;
  }
  return ZR_OK;
}

// the inverse diagonal from the sums that the last k_be_state<T, true> staged
template <class T>
int jacobi_inverse(zr_engine* e)
{
  const uint32_t nn = e->S * 32;
  k_jacobi<T><<<blocks_for(nn, 256), 256, 0, e->stream>>>(nn, e->slot_src, e->jstage, e->gm, e->collided, e->be_dt,
                                                          nvT(e, reinterpret_cast<T*>(e->pinv)));
  ZR_LAUNCH_CHECK(e);
  return ZR_OK;
}

// ConjugateGradient::solve (ConjugateGradient.h:99-147) on A x = r with the
// mass-lumped or the Jacobi preconditioner in the vectors' precision, FP64
// inner products;
// termination is decided on the device and the host polls the flag every
// check_interval iterations
template <class V>
int be_cg(zr_engine* e, double rel, const zr_implicit_opts* o)
{
  using T = typename V::T;
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256);
  hipStream_t s = e->stream;
  const NodeVec<T> pinv = nvT(e, e->jacobi ? reinterpret_cast<T*>(e->pinv) : (T*)nullptr);
  k_cg_init<T><<<nb, 256, 0, s>>>(nn, e->gm, nv(e, e->ir), nvT(e, V::x(e)), nvT(e, V::r(e)), nvT(e, V::q(e)),
                                  nvT(e, V::p(e)), pinv, e->partial);
  k_reduce_cg<<<1, 256, 0, s>>>(e->partial, nb, e->sc, e->ci, RED_CG_INIT, 0, rel, o->cg_tolerance);
  ZR_LAUNCH_CHECK(e);
  const int check = o->cg_check_interval > 0 ? o->cg_check_interval : 8;
  const uint32_t ncg = nb < kCgBlocks ? nb : kCgBlocks;
  // per iteration: the operator apply (which first closes the previous
  // update: beta, count, termination), A p with p^T A p partials, and the
  // update (alpha from those partials) with z^T r partials
  int k = 0;
  for (; k < o->cg_max_iterations; ++k) {
    int rc = be_apply<T>(e, V::q(e), V::p(e), V::ap(e), 1, e->ci, k);
    if (rc) return rc;
    k_cg_update<T><<<ncg, 256, 0, s>>>(nn, e->gm, nvT(e, V::x(e)), nvT(e, V::r(e)), nvT(e, V::q(e)),
                                       nvT(e, V::p(e)), nvT(e, V::ap(e)), pinv, e->sc, e->partial, e->ncg_act,
                                       e->partial2, e->ci, k);
    ZR_LAUNCH_CHECK(e);
    if ((k + 1) % check == 0 && k + 1 < o->cg_max_iterations) {
      ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      ZR_HIP(e, hipStreamSynchronize(s));
      if (e->hci[CI_DONE]) {
        ++k;
        break;
      }
    }
  }
  // closes the last update when the loop ran out of iterations (a no-op
  // once the solve has terminated)
  if (k > 0) {
    k_reduce_cg<<<1, 256, 0, s>>>(e->partial2, ncg, e->sc, e->ci, RED_FINISH, k - 1, 0.0, 0.0);
    ZR_LAUNCH_CHECK(e);
  }
  return ZR_OK;
}

// Minres::solve (Minres.h:69-178) on A x = r for the Newton step, the
// reference's default linear solver (BackwardEuler.h:55-62): vectors in the
// operator's precision, FP64 inner products and scalars on the device (the
// Givens QR in one thread), the host polling the termination flag every
// check_interval iterations.  The seven Krylov vectors (q_{k+1}, q_k, q_{k-1},
// z, m_k, m_{k-1}, m_{k-2}) rotate by pointer; `iters` returns the count.
template <class V>
int be_minres(zr_engine* e, double rel, const zr_implicit_opts* o, int& iters)
{
  using T = typename V::T;
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256), ncg = nb < kCgBlocks ? nb : kCgBlocks;
  hipStream_t s = e->stream;
  if (V::extra_cap(e) < e->grid_cap) {
    dfree(V::extra(e));
    V::extra_cap(e) = 0;
    ZR_HIP(e, hipMalloc(&V::extra(e), 9 * e->grid_cap * sizeof(T)));
    V::extra_cap(e) = e->grid_cap;
  }
  T* ex = V::extra(e);
  const size_t vs = 3 * e->grid_cap;
  T *qkp1 = V::r(e), *qk = V::q(e), *qkm1 = V::p(e), *z = V::ap(e);
  T *mk = ex, *mkm1 = ex + vs, *mkm2 = ex + 2 * vs;
  const NodeVec<T> pinv = nvT(e, e->jacobi ? reinterpret_cast<T*>(e->pinv) : (T*)nullptr);
  const int max_it = o->cg_max_iterations;
  k_imr_init<T><<<nb, 256, 0, s>>>(nn, e->gm, nv(e, e->ir), nvT(e, V::x(e)), nvT(e, qkp1), nvT(e, z), nvT(e, qk),
                                   nvT(e, mk), nvT(e, mkm1), pinv, e->partial);
  k_minres_reduce<<<1, 256, 0, s>>>(e->partial, nb, e->sc, e->ci, MR_INIT, o->cg_tolerance, rel, 0, max_it);
  k_imr_scale<T><<<nb, 256, 0, s>>>(nn, e->gm, nvT(e, qkp1), nvT(e, z), e->sc, e->ci);
  ZR_LAUNCH_CHECK(e);
  const int check = o->cg_check_interval > 0 ? o->cg_check_interval : 8;
  for (int k = 0; k < max_it; ++k) {
    k_minres_begin<<<1, 64, 0, s>>>(e->sc, e->ci);
    T* t = mkm2;  // m_{k-2} <- m_{k-1} <- m_k <- z; z takes the freed vector
    mkm2 = mkm1;
    mkm1 = mk;
    mk = z;
    z = t;
    t = qkm1;  // q_{k-1} <- q_k <- q_{k+1}; q_{k+1} takes the freed vector
    qkm1 = qk;
    qk = qkp1;
    qkp1 = t;
    // q_{k+1} = P A m_k with the partials of alpha = m_k . q_{k+1}
    int rc = be_apply<T>(e, nullptr, mk, qkp1, 1, e->ci, 0);
    if (rc) return rc;
    k_minres_reduce<<<1, 256, 0, s>>>(e->partial, ncg, e->sc, e->ci, MR_ALPHA, 0.0, 0.0, k, max_it);
    k_imr_lanczos<T><<<nb, 256, 0, s>>>(nn, e->gm, nvT(e, qkp1), nvT(e, qk), nvT(e, qkm1), nvT(e, z), pinv, e->sc,
                                        e->ci, e->partial);
    k_minres_reduce<<<1, 256, 0, s>>>(e->partial, nb, e->sc, e->ci, MR_GIVENS, 0.0, 0.0, k, max_it);
    k_imr_update<T><<<nb, 256, 0, s>>>(nn, e->gm, nvT(e, qkp1), nvT(e, z), nvT(e, mk), nvT(e, mkm1),
                                       nvT(e, mkm2), nvT(e, V::x(e)), e->sc, e->ci);
    ZR_LAUNCH_CHECK(e);
    if ((k + 1) % check == 0 && k + 1 < max_it) {
      ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      ZR_HIP(e, hipStreamSynchronize(s));
      if (e->hci[AI_DONE]) break;
    }
  }
  ZR_HIP(e, hipMemcpyAsync(e->hci, e->ci, CI_N * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  ZR_HIP(e, hipStreamSynchronize(s));
  iters = e->hci[AI_ITERS];
  return ZR_OK;
}

template <class V>
int be_newton(zr_engine* e, const zr_implicit_opts* o, zr_implicit_stats& local)
{
  using T = typename V::T;
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256);
  // NewtonsMethod::solve (NewtonsMethod.h:42-66)
  for (int it = 0; it < o->newton_max_iterations; ++it) {
    int rc = be_state_residual<T>(e, e->jacobi);
    if (rc) return rc;
    rc = read_scalars(e);
    if (rc) return rc;
    const double norm = sqrt(e->hsc[SC_NORM2]);
    if (local.num_residuals < 8) local.residual_norm[local.num_residuals++] = norm;
    if (norm < o->newton_tolerance) {
      local.converged = 1;
      break;
    }
    const double rel = o->cg_relative_tolerance > 0.0 ? o->cg_relative_tolerance
                                                      : fmin(0.5, sqrt(fmax(norm, o->newton_tolerance)));
    if (e->jacobi) {
      rc = jacobi_inverse<T>(e);
      if (rc) return rc;
    }
    if (o->solver == ZR_SOLVER_MINRES) {
      int its = 0;
      rc = be_minres<V>(e, rel, o, its);
      local.cg_iterations += its;
    } else {
      rc = be_cg<V>(e, rel, o);
    }
    if (rc) return rc;
    k_axpy3<T><<<nb, 256, 0, e->stream>>>(nn, nv(e, e->idv), nvT(e, V::x(e)));
    ZR_LAUNCH_CHECK(e);
    local.newton_iterations++;
  }
  return ZR_OK;
}

int do_be(zr_engine* e, double dt, const zr_implicit_opts* o, zr_implicit_stats* st)
{
  if (!e->rasterised || e->mls_rasterised)
    return fail(e, ZR_ERR_STATE, "backwardEulerStep needs the APIC particlesToGrid of this step");
  int rc = ensure_implicit(e, o->precision);
  if (rc) return rc;
  if (e->A > 0) {  // XCD ranges of the operator apply for this step's groups
    const uint32_t n32 = e->A < e->hvp_resident[0] ? e->A : e->hvp_resident[0];
    const uint32_t n64 = e->A < e->hvp_resident[1] ? e->A : e->hvp_resident[1];
    if (e->hvp_order_cap < e->A) {
      dfree(e->hvp_order);
      dfree(e->cell_order);
      e->hvp_order_cap = 0;
      const size_t cap = (size_t)e->A + e->A / 4 + 64;
      ZR_HIP(e, hipMalloc(&e->hvp_order, 2 * cap * sizeof(uint32_t)));
      ZR_HIP(e, hipMalloc(&e->cell_order, cap * kCells));
      e->hvp_order_cap = cap;
    }
    k_xcd_split<<<1, 1024, 0, e->stream>>>(e->group_start, e->A, n32, n64, e->hvp_bounds, e->hvp_order);
    ZR_LAUNCH_CHECK(e);
  }
  zr_implicit_stats local;
  memset(&local, 0, sizeof(local));
  e->be_dt = dt;
  e->be_prec = o->precision;
  e->jacobi = o->preconditioner == ZR_PRECOND_JACOBI;
  if (e->jacobi) {
    rc = ensure_jacobi(e);
    if (rc) return rc;
  }
  if (e->collided_cap < e->grid_cap) {  // nodes a collider touches, for the Jacobi preconditioner
    dfree(e->collided);
    e->collided_cap = 0;
    ZR_HIP(e, hipMalloc(&e->collided, e->grid_cap));
    e->collided_cap = e->grid_cap;
  }
  const uint32_t nn = e->S * 32, nb = blocks_for(nn, 256);
  hipStream_t s = e->stream;
  ColliderSet cs = collider_set(e);
  k_be_init<<<nb, 256, 0, s>>>(nn, e->slot_page, e->gm, e->gv, nv(e, e->ivn), nv(e, e->idv), e->cfg.dx, dt,
                               e->cfg.gravity[0], e->cfg.gravity[1], e->cfg.gravity[2], cs,
                               e->collided);
  ZR_HIP(e, hipMemsetAsync(e->ci, 0, CI_N * sizeof(int32_t), s));
  ZR_LAUNCH_CHECK(e);
  // the nodes with mass, for the CG node kernels (the masses hold for the step)
  if (e->act_cap < e->grid_cap) {
    dfree(e->act);
    dfree(e->act_scan);
    e->act_cap = 0;
    ZR_HIP(e, hipMalloc(&e->act, e->grid_cap * sizeof(uint32_t)));
    ZR_HIP(e, hipMalloc(&e->act_scan, e->grid_cap * sizeof(uint32_t)));
    e->act_cap = e->grid_cap;
  }
  rc = run_scan(e, LdNonzeroMass{ e->gm }, nn, e->act_scan, e->dcount + CNT_TOTAL);
  if (rc) return rc;
  if (nn) k_active_list<<<nb, 256, 0, s>>>(nn, e->gm, e->act_scan, e->act);
  ZR_LAUNCH_CHECK(e);
  rc = sync_counts(e);
  if (rc) return rc;
  e->nact = e->hcount[CNT_TOTAL];
  {
    const uint32_t b = blocks_for(e->nact, 256);
    e->ncg_act = b == 0 ? 1u : (b < kCgBlocks ? b : kCgBlocks);
  }
  e->be_ready = true;
  rc = o->precision == ZR_FP32_REFINED ? be_newton<CgVecs32>(e, o, local) : be_newton<CgVecs64>(e, o, local);
  if (rc) return rc;
  k_be_finish<<<nb, 256, 0, s>>>(nn, e->gm, e->gv, nv(e, e->idv), e->gvs);
  ZR_LAUNCH_CHECK(e);
  rc = read_scalars(e);
  if (rc) return rc;
  if (o->solver != ZR_SOLVER_MINRES) local.cg_iterations = e->hci[CI_ITERS];
  e->updated = true;
  if (st) *st = local;
  return ZR_OK;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
// evalMaxParticleSpeed without allocation (zr_eval_max_particle_speed_noalloc):
// the block maxima live in a file-scope device array; each device holds its
// own copy, whose address is taken once per device (hipGetSymbolAddress, with
// that device current) and passed to the kernels.  Every block writes its
// partials before the final pass reads them, so nothing needs clearing;
// calls that overlap on one device share the array and race.
namespace {
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) double scratch[kSpeedFields * kSpeedBlocks];
std::mutex g_scratch_mu;
double* g_scratch[kNoAllocDevices] = {};
// message of the last failed call without an engine on this thread (a
// string literal: nothing is allocated)
thread_local const char* g_noalloc_err = "";

int noalloc_fail(int code, const char* msg)
{
  g_noalloc_err = msg;
  return code;
}
}  // namespace

extern "C" {

int zr_create(const zr_config* cfg, zr_engine** out)
{
  if (!cfg || !out) return ZR_ERR_ARG;
  *out = nullptr;
  if (!(cfg->dx > 0.0) || cfg->grid_log2 < 6 || cfg->grid_log2 > 12 ||
      (cfg->precision != ZR_FP64 && cfg->precision != ZR_FP32_REFINED) || cfg->transfer != ZR_TRANSFER_MLS)
    return ZR_ERR_ARG;
  zr_engine* e = new zr_engine();
  e->cfg = *cfg;
  e->L = (uint32_t)cfg->grid_log2;
  e->lim = (1u << e->L) - 8u;
  e->ndense = 1u << (3 * e->L - 5);
  e->nwords = (e->ndense + 31u) / 32u;
  e->nsum = (e->nwords + 31u) / 32u;
  hipError_t st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (st != hipSuccess) {
    delete e;
    return ZR_ERR_HIP;
  }
  {
    // the persistent operator apply launches one wave of resident blocks
    // (its blocks never wait on each other, so fewer resident is still correct)
    int dev = 0, ncu = 0, b32 = 0, b64 = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b32, k_hvp<float>, HvpChunk<float>::threads, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b64, k_hvp<double>, HvpChunk<double>::threads, 0) !=
            hipSuccess ||
        ncu <= 0 ||
        b32 <= 0 || b64 <= 0) {
      (void)hipStreamDestroy(e->stream);
      delete e;
      return ZR_ERR_HIP;
    }
    e->hvp_resident[0] = (uint32_t)(b32 * ncu);
    e->hvp_resident[1] = (uint32_t)(b64 * ncu);
  }
  bool ok = hipMalloc(&e->page_count, e->ndense * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->page_group, e->ndense * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&e->page_touch, e->ndense * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->page_slot, e->ndense * sizeof(int32_t)) == hipSuccess &&
            hipMalloc(&e->page_bits, e->nwords * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->page_sum, e->nsum * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->page_sumcnt, e->nsum * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->sum_base, (e->nsum + 1) * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->dcount, CNT_N * sizeof(uint32_t)) == hipSuccess &&
            hipMalloc(&e->ticket, sizeof(uint32_t)) == hipSuccess &&
            hipHostMalloc(&e->hcount, CNT_N * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMemset(e->dcount, 0, CNT_N * sizeof(uint32_t)) == hipSuccess;
  ok = ok && hipMemset(e->page_count, 0, e->ndense * sizeof(uint32_t)) == hipSuccess &&
       hipMemset(e->page_group, 0xff, e->ndense * sizeof(int32_t)) == hipSuccess &&
       hipMemset(e->page_touch, 0xff, e->ndense * sizeof(uint32_t)) == hipSuccess &&
       hipMemset(e->page_slot, 0xff, e->ndense * sizeof(int32_t)) == hipSuccess &&
       hipMemset(e->page_bits, 0, e->nwords * sizeof(uint32_t)) == hipSuccess &&
       hipMemset(e->page_sum, 0, e->nsum * sizeof(uint32_t)) == hipSuccess &&
       hipMemset(e->page_sumcnt, 0, e->nsum * sizeof(uint32_t)) == hipSuccess &&
       hipMemset(e->ticket, 0, sizeof(uint32_t)) == hipSuccess &&
       hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    zr_destroy(e);
    return ZR_ERR_HIP;
  }
  *out = e;
  return ZR_OK;
}

void zr_destroy(zr_engine* e)
{
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  free_particles(e);
  dfree(e->page_count); dfree(e->page_group); dfree(e->page_touch); dfree(e->page_slot); dfree(e->page_bits);
  dfree(e->page_sum); dfree(e->page_sumcnt); dfree(e->sum_base); dfree(e->stage); dfree(e->gm); dfree(e->gv); dfree(e->gf); dfree(e->gvs);
  dfree(e->scan_tmp); dfree(e->speed); dfree(e->dcount); dfree(e->ticket); dfree(e->slot_src);
  dfree(e->ivn); dfree(e->idv); dfree(e->ir); dfree(e->ix); dfree(e->irc); dfree(e->iq); dfree(e->ip); dfree(e->iap);
  dfree(e->partial); dfree(e->partial2); dfree(e->sc); dfree(e->ci); dfree(e->hvp_ctr); dfree(e->hvp_bounds); dfree(e->hvp_order); dfree(e->cell_order); dfree(e->jstage); dfree(e->pinv); dfree(e->collided); dfree(e->c64); dfree(e->c32); dfree(e->ccell);
  dfree(e->pf_node); dfree(e->pf_partial); dfree(e->adj_grid); dfree(e->admm_node); dfree(e->admm_coll); dfree(e->act); dfree(e->act_scan); dfree(e->be_fb); dfree(e->admm_c32); dfree(e->admm_cell32); dfree(e->admm_node32);
  dfree(e->fx); dfree(e->frc); dfree(e->fq); dfree(e->fp); dfree(e->fap);
  dfree(e->mr64); dfree(e->mr32);
  if (e->hsc) (void)hipHostFree(e->hsc);
  if (e->hci) (void)hipHostFree(e->hci);
  for (auto& v : e->prof_ev)
    for (auto h : v) (void)hipEventDestroy(h);
  if (e->hcount) (void)hipHostFree(e->hcount);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* zr_last_error(const zr_engine* e) { return e ? e->err.c_str() : g_noalloc_err; }

int zr_set_colliders(zr_engine* e, const zr_collider* c, int32_t n)
{
  if (!e || n < 0 || n > kMaxColliders || (n > 0 && !c)) return fail(e, ZR_ERR_ARG, "bad collider set");
  e->colliders.assign(c, c + n);
  for (auto& k : e->colliders) {
    if (k.shape == ZR_HALF_SPACE) {  // HalfSpace normalises its normal (AnalyticLevelSet.cpp:262-266)
      double l = sqrt(k.normal[0] * k.normal[0] + k.normal[1] * k.normal[1] + k.normal[2] * k.normal[2]);
      if (!(l > 0.0)) return fail(e, ZR_ERR_ARG, "half-space normal is zero");
      for (int d = 0; d < 3; ++d) k.normal[d] /= l;
    } else if (k.shape != ZR_SPHERE) {
      return fail(e, ZR_ERR_ARG, "unknown collider shape");
    }
    if (k.type < ZR_STICKY || k.type > ZR_SEPARATE) return fail(e, ZR_ERR_ARG, "unknown collider type");
    if (!(k.scale == 0.0 || k.scale > 0.0)) return fail(e, ZR_ERR_ARG, "collider scale must be positive");
  }
  return ZR_OK;
}

int zr_set_time(zr_engine* e, double t)
{
  if (!e) return ZR_ERR_ARG;
  e->time = t;
  return ZR_OK;
}
double zr_get_time(const zr_engine* e) { return e ? e->time : 0.0; }

int zr_eval_max_particle_speed(zr_engine* e, double* out)
{
  if (!e || !out) return ZR_ERR_ARG;
  if (!e->speed) {
    ZR_HIP(e, hipMalloc(&e->speed, (kSpeedFields * kSpeedBlocks + kSpeedFields) * sizeof(double)));
  }
  const uint32_t n = (uint32_t)e->n;
  uint32_t nb = blocks_for((uint64_t)n, 256);
  if (nb > kSpeedBlocks) nb = kSpeedBlocks;
  if (nb == 0) nb = 1;
  double* res = e->speed + kSpeedFields * kSpeedBlocks;
  k_max_speed<<<nb, 256, 0, e->stream>>>(e->x, e->v, 1, e->n, n, e->speed);
  k_max_speed_final<<<1, 256, 0, e->stream>>>(e->speed, nb, res);
  ZR_LAUNCH_CHECK(e);
  ZR_HIP(e, hipMemcpyAsync(out, res, kSpeedFields * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

int zr_calculate_dt(zr_engine* e, double max_dt, double* dt)
{
  if (!e || !dt || !(max_dt > 0.0)) return ZR_ERR_ARG;
  double r[kSpeedFields];
  const int rc = zr_eval_max_particle_speed(e, r);
  if (rc) return rc;
  double speed = r[0];
  // AnalyticCollisionObject::evalMaxSpeed (CollisionObject.cpp:158-194) over
  // the particle box widened by (degree + 2) dx (MpmSimulationBase.cpp:774-781)
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    hi[d] = r[1 + d] + 4.0 * e->cfg.dx;
    lo[d] = r[4 + d] - 4.0 * e->cfg.dx;
  }
  for (const zr_collider& c : e->colliders) {
#pragma clang fp contract(off)
    const double* u = c.velocity;
    double cs = 0.0;
    if (c.dscale_dt != 0.0 || c.omega[0] != 0.0 || c.omega[1] != 0.0 || c.omega[2] != 0.0) {
      if (c.shape != ZR_SPHERE) return fail(e, ZR_ERR_ARG, "a spinning or scaling half-space has no bounds");
      double mn[3], mx[3];
      ColliderFrame fr;
      collider_frame(c, e->time, fr);
      const double *R = fr.R, *b = fr.b;
      const double s = fr.s, inv_s = fr.inv_s;
      for (int d = 0; d < 3; ++d) {
        mn[d] = c.origin[d] - c.radius;
        mx[d] = c.origin[d] + c.radius;
      }
      std::vector<std::array<double, 3>> corners;
      for (int i = 0; i < 8; ++i) {  // particle-box corners inside the object's box (either test per axis)
        double x[3], X[3];
        for (int d = 0; d < 3; ++d) x[d] = (i & (1 << d)) ? lo[d] : hi[d];
        for (int d = 0; d < 3; ++d)
          X[d] = (R[d] * (x[0] - b[0]) + R[3 + d] * (x[1] - b[1]) + R[6 + d] * (x[2] - b[2])) * inv_s;
        const bool a1 = (mn[0] < X[0]) || (mn[1] < X[1]) || (mn[2] < X[2]);
        const bool a2 = (X[0] < mx[0]) || (X[1] < mx[1]) || (X[2] < mx[2]);
        if (a1 && a2) corners.push_back({ x[0], x[1], x[2] });
      }
      for (int i = 0; i < 8; ++i) {  // object-box corners inside the particle box
        double X[3], x[3];
        for (int d = 0; d < 3; ++d) X[d] = (i & (1 << d)) ? mn[d] : mx[d];
        for (int d = 0; d < 3; ++d) x[d] = s * (R[3 * d] * X[0] + R[3 * d + 1] * X[1] + R[3 * d + 2] * X[2]) + b[d];
        const bool a1 = (lo[0] < x[0]) || (lo[1] < x[1]) || (lo[2] < x[2]);
        const bool a2 = (x[0] < hi[0]) || (x[1] < hi[1]) || (x[2] < hi[2]);
        if (a1 && a2) corners.push_back({ x[0], x[1], x[2] });
      }
      const double g = c.dscale_dt * inv_s;
      for (const auto& x : corners) {
        const double xb[3] = { x[0] - b[0], x[1] - b[1], x[2] - b[2] };
        const double v[3] = { (c.omega[1] * xb[2] - c.omega[2] * xb[1]) + g * xb[0] + u[0],
                              (c.omega[2] * xb[0] - c.omega[0] * xb[2]) + g * xb[1] + u[1],
                              (c.omega[0] * xb[1] - c.omega[1] * xb[0]) + g * xb[2] + u[2] };
        const double sp = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        if (sp > cs) cs = sp;
      }
    } else {
      cs = std::sqrt((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
    }
    if (cs > speed) speed = cs;
  }
  *dt = speed != 0.0 ? e->cfg.cfl * e->cfg.dx / speed : max_dt;
  return ZR_OK;
}

// evalMaxParticleSpeed on caller-owned device arrays, allocating nothing
int zr_eval_max_particle_speed_noalloc(const double* x, const double* v, int64_t n, double* out, void* stream)
{
  if (!out || n < 0 || (n > 0 && (!x || !v))) return noalloc_fail(ZR_ERR_ARG, "null pointer or negative count");
  if (n > (int64_t)UINT32_MAX) return noalloc_fail(ZR_ERR_ARG, "more than 2^32 - 1 particles");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return noalloc_fail(ZR_ERR_HIP, "hipGetDevice failed");
  if (dev < 0 || dev >= kNoAllocDevices) return noalloc_fail(ZR_ERR_ARG, "device index beyond 63");
  double* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        return noalloc_fail(ZR_ERR_HIP, "hipGetSymbolAddress of the scratch array failed");
      g_scratch[dev] = static_cast<double*>(p);
    }
    ptr = g_scratch[dev];
  }
  uint32_t nb = blocks_for((uint64_t)n, 256);
  if (nb > kSpeedBlocks) nb = kSpeedBlocks;
  if (nb == 0) nb = 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  k_max_speed<<<nb, 256, 0, s>>>(x, v, 3, 1, (uint32_t)n, ptr);
  k_max_speed_final<<<1, 256, 0, s>>>(ptr, nb, out);
  if (hipGetLastError() != hipSuccess) return noalloc_fail(ZR_ERR_HIP, "kernel launch failed");
  return ZR_OK;
}

int zr_upload_particles(zr_engine* e, const zr_particles* p)
{
  if (!e || !p || p->n < 0) return fail(e, ZR_ERR_ARG, "bad particle view");
  // per-particle arrays are addressed with 32-bit byte offsets (8-byte values)
  if (p->n > kMaxParticles) return fail(e, ZR_ERR_ARG, "more than 2^29 - 1 particles");
  if (p->n >= (1ll << 27)) return fail(e, ZR_ERR_ARG, "particle count must be below 2^27");
  if (p->n > 0 && (!p->X || !p->V || !p->C || !p->F || !p->mass || !p->vol || !p->mu || !p->lambda))
    return fail(e, ZR_ERR_ARG, "all particle attributes are required");
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  // reset tables owned by the previous particle set
  if (e->A || e->S) {
    uint32_t rst = e->A > e->S ? e->A : e->S;
    k_reset_tables<<<blocks_for(rst, 256), 256, 0, e->stream>>>(e->group_page, e->A, e->slot_page, e->S,
                                                                e->page_count, e->page_group, e->page_bits,
                                                                e->page_sum, e->page_touch, e->page_slot,
                                                                nullptr);
    ZR_LAUNCH_CHECK(e);
    ZR_HIP(e, hipStreamSynchronize(e->stream));
    e->A = e->S = 0;
  }
  free_particles(e);
  e->sorted = e->rasterised = e->updated = false;
  const int64_t n = p->n;
  e->n = n;
  if (n == 0) return ZR_OK;
  size_t N = (size_t)n;
  ZR_HIP(e, hipMalloc(&e->x, 3 * N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->v, 3 * N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->C, 9 * N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->F, 9 * N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->m, N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->vol, N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->mu, N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->lam, N * sizeof(double)));
  ZR_HIP(e, hipMalloc(&e->p_page, N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->p_re, N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->order, N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->keys, N * sizeof(uint64_t)));
  ZR_HIP(e, hipMalloc(&e->group_page, N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->group_count, N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->group_start, (N + 1) * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->touch_page, 8 * N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->touch_slot, 8 * N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->slot_page, 8 * N * sizeof(uint32_t)));
  ZR_HIP(e, hipMalloc(&e->group_nslot, 8 * N * sizeof(int32_t)));
  double* tmp = nullptr;
  ZR_HIP(e, hipMalloc(&tmp, 9 * N * sizeof(double)));
  struct Item {
    const double* src;
    double* dst;
    int nc;
  } items[] = { { p->X, e->x, 3 }, { p->V, e->v, 3 }, { p->C, e->C, 9 }, { p->F, e->F, 9 },
                { p->mass, e->m, 1 }, { p->vol, e->vol, 1 }, { p->mu, e->mu, 1 }, { p->lambda, e->lam, 1 } };
  for (auto& it : items) {
    hipError_t st = hipMemcpyAsync(tmp, it.src, it.nc * N * sizeof(double), hipMemcpyHostToDevice, e->stream);
    if (st == hipSuccess) {
      k_aos_to_soa<<<blocks_for((uint64_t)N * it.nc, 256), 256, 0, e->stream>>>(tmp, it.dst, (uint32_t)N, it.nc, n);
      st = hipGetLastError();
    }
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) {
      (void)hipFree(tmp);
      return fail(e, ZR_ERR_HIP, std::string("upload: ") + hipGetErrorString(st));
    }
  }
  ZR_HIP(e, hipFree(tmp));
  return ZR_OK;
}

int zr_download_particles(zr_engine* e, zr_particles* p)
{
  if (!e || !p) return fail(e, ZR_ERR_ARG, "bad particle view");
  if (p->n != e->n) return fail(e, ZR_ERR_ARG, "particle count mismatch");
  if (e->n == 0) return ZR_OK;
  size_t N = (size_t)e->n;
  double* tmp = nullptr;
  ZR_HIP(e, hipMalloc(&tmp, 9 * N * sizeof(double)));
  struct Item {
    double* dst;
    const double* src;
    int nc;
  } items[] = { { p->X, e->x, 3 }, { p->V, e->v, 3 }, { p->C, e->C, 9 }, { p->F, e->F, 9 },
                { p->mass, e->m, 1 }, { p->vol, e->vol, 1 }, { p->mu, e->mu, 1 }, { p->lambda, e->lam, 1 } };
  for (auto& it : items) {
    if (!it.dst) continue;
    k_soa_to_aos<<<blocks_for((uint64_t)N * it.nc, 256), 256, 0, e->stream>>>(it.src, tmp, (uint32_t)N, it.nc, e->n);
    hipError_t st = hipGetLastError();
    if (st == hipSuccess) st = hipMemcpyAsync(it.dst, tmp, it.nc * N * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) {
      (void)hipFree(tmp);
      return fail(e, ZR_ERR_HIP, std::string("download: ") + hipGetErrorString(st));
    }
  }
  ZR_HIP(e, hipFree(tmp));
  return ZR_OK;
}

int zr_set_materials(zr_engine* e, const zr_materials* mt)
{
  if (!e || !mt) return fail(e, ZR_ERR_ARG, "bad materials view");
  if (mt->n != e->n) return fail(e, ZR_ERR_ARG, "material count differs from the particle count");
  const int64_t n = e->n;
  if (n == 0) return ZR_OK;
  const size_t N = (size_t)n;
  std::vector<int32_t> code(N, 0);
  bool models = false, plastic = false;
  for (size_t i = 0; i < N; ++i) {
    const int32_t md = mt->model ? mt->model[i] : ZR_MODEL_FIXED_COROTATED;
    const int32_t pl = mt->plasticity ? mt->plasticity[i] : ZR_PLASTIC_NONE;
    if (md < ZR_MODEL_FIXED_COROTATED || md > ZR_MODEL_NEO_HOOKEAN_BORDEN || pl < ZR_PLASTIC_NONE ||
        pl > ZR_PLASTIC_CAM_CLAY)
      return fail(e, ZR_ERR_ARG, "unknown model or plasticity code");
    if (pl == ZR_PLASTIC_CAM_CLAY && md != ZR_MODEL_NEO_HOOKEAN_BORDEN)
      return fail(e, ZR_ERR_ARG, "Cam-Clay plasticity needs the NeoHookeanBorden model");
    models = models || md != ZR_MODEL_FIXED_COROTATED;
    plastic = plastic || pl != ZR_PLASTIC_NONE;
    code[i] = md | (pl << 4);
  }
  if (plastic && (!mt->plastic_param || !mt->plastic_state))
    return fail(e, ZR_ERR_ARG, "plasticity needs plastic_param and plastic_state");
  if (!e->mat) {
    ZR_HIP(e, hipMalloc(&e->mat, N * sizeof(int32_t)));
    ZR_HIP(e, hipMalloc(&e->pparam, 4 * N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->pstate, N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->gdeg, N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->broken, N * sizeof(int32_t)));
  }
  std::vector<double> pp(4 * N, 0.0), ps(N, 0.0), g(N, 1.0);
  for (size_t i = 0; i < N; ++i) {
    for (int k = 0; k < 4; ++k)
      if (mt->plastic_param) pp[k * N + i] = mt->plastic_param[4 * i + k];  // particle-major -> [4][n]
    if (mt->plastic_state) ps[i] = mt->plastic_state[i];
    if (mt->degradation) g[i] = mt->degradation[i];
  }
  ZR_HIP(e, hipMemcpyAsync(e->mat, code.data(), N * sizeof(int32_t), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipMemcpyAsync(e->pparam, pp.data(), 4 * N * sizeof(double), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipMemcpyAsync(e->pstate, ps.data(), N * sizeof(double), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipMemcpyAsync(e->gdeg, g.data(), N * sizeof(double), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipMemsetAsync(e->broken, 0, N * sizeof(int32_t), e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  e->has_models = models;
  e->has_plastic = plastic;
  e->all_hencky = true;
  for (size_t i = 0; i < N; ++i) e->all_hencky = e->all_hencky && (code[i] & 15) == ZR_MODEL_STVK_HENCKY;
  e->uniform_model = N ? (code[0] & 15) : ZR_MODEL_FIXED_COROTATED;
  for (size_t i = 0; i < N; ++i)
    if ((code[i] & 15) != e->uniform_model) e->uniform_model = kAnyModel;
  return ZR_OK;
}

int zr_set_viscoelastic(zr_engine* e, const double* F_neq, const double* mu, const double* lambda,
                        double viscosity_d, double viscosity_v)
{
  if (!e) return ZR_ERR_ARG;
  if (!mu && !lambda) {  // switch the branch off
    e->visco = false;
    return ZR_OK;
  }
  if (!mu || !lambda) return fail(e, ZR_ERR_ARG, "the viscoelastic branch needs mu and lambda");
  if (!(viscosity_d > 0.0) || !(viscosity_v > 0.0)) return fail(e, ZR_ERR_ARG, "viscosities must be positive");
  const size_t N = (size_t)e->n;
  if (N == 0) return fail(e, ZR_ERR_STATE, "upload the particles before the viscoelastic branch");
  if (!e->fneq) {
    ZR_HIP(e, hipMalloc(&e->fneq, 9 * N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->vtau_neq, 9 * N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->mu_neq, N * sizeof(double)));
    ZR_HIP(e, hipMalloc(&e->lam_neq, N * sizeof(double)));
  }
  std::vector<double> f(9 * N);
  for (size_t i = 0; i < N; ++i)
    for (int k = 0; k < 9; ++k) f[k * N + i] = F_neq ? F_neq[9 * i + k] : (k % 4 == 0 ? 1.0 : 0.0);
  ZR_HIP(e, hipMemcpyAsync(e->fneq, f.data(), 9 * N * sizeof(double), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipMemcpyAsync(e->mu_neq, mu, N * sizeof(double), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipMemcpyAsync(e->lam_neq, lambda, N * sizeof(double), hipMemcpyHostToDevice, e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  e->visc_d = viscosity_d;
  e->visc_v = viscosity_v;
  e->visco = true;
  return ZR_OK;
}

int zr_get_viscoelastic(zr_engine* e, double* F_neq)
{
  if (!e || !F_neq) return ZR_ERR_ARG;
  if (!e->visco) return fail(e, ZR_ERR_STATE, "no viscoelastic branch");
  const size_t N = (size_t)e->n;
  std::vector<double> f(9 * N);
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  ZR_HIP(e, hipMemcpy(f.data(), e->fneq, 9 * N * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < N; ++i)
    for (int k = 0; k < 9; ++k) F_neq[9 * i + k] = f[k * N + i];
  return ZR_OK;
}

int zr_get_materials(zr_engine* e, zr_materials* mt)
{
  if (!e || !mt) return fail(e, ZR_ERR_ARG, "bad materials view");
  if (mt->n != e->n) return fail(e, ZR_ERR_ARG, "material count differs from the particle count");
  const size_t N = (size_t)e->n;
  if (N == 0) return ZR_OK;
  if (!e->mat) {  // never set: defaults
    for (size_t i = 0; i < N; ++i) {
      if (mt->plastic_state) mt->plastic_state[i] = 0.0;
      if (mt->degradation) mt->degradation[i] = 1.0;
      if (mt->broken) mt->broken[i] = 0;
    }
    return ZR_OK;
  }
  if (mt->plastic_state)
    ZR_HIP(e, hipMemcpyAsync(mt->plastic_state, e->pstate, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  if (mt->degradation)
    ZR_HIP(e, hipMemcpyAsync(mt->degradation, e->gdeg, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  if (mt->broken)
    ZR_HIP(e, hipMemcpyAsync(mt->broken, e->broken, N * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

int zr_set_phase_field(zr_engine* e, const zr_phase_field* pf, const zr_phase_field_opts* o)
{
  if (!e || !pf || !o) return fail(e, ZR_ERR_ARG, "bad phase-field view");
  if (pf->n != e->n) return fail(e, ZR_ERR_ARG, "phase-field count differs from the particle count");
  if (o->naive != ZR_DAMAGE_PHASE_FIELD && o->naive != ZR_DAMAGE_RANKINE && o->naive != ZR_DAMAGE_PHASE_FIELD_RANKINE)
    return fail(e, ZR_ERR_ARG, "unknown damage rule");
  if (o->naive != ZR_DAMAGE_RANKINE && (o->cg_max_iterations < 0 || !(o->cg_tolerance >= 0.0)))
    return fail(e, ZR_ERR_ARG, "bad phase-field CG options");
  if (o->naive != ZR_DAMAGE_RANKINE && (!pf->l0 || !pf->one_over_sigma_c))
    return fail(e, ZR_ERR_ARG, "the phase field needs l0 and one_over_sigma_c");
  const size_t N = (size_t)e->n;
  if (N == 0) return ZR_OK;
  // every particle becomes NeoHookeanBorden (keeping its plasticity code)
  std::vector<int32_t> code(N, ZR_MODEL_NEO_HOOKEAN_BORDEN);
  if (e->mat) {
    ZR_HIP(e, hipMemcpy(code.data(), e->mat, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (auto& c : code) c = (c & ~15) | ZR_MODEL_NEO_HOOKEAN_BORDEN;
  } else {
    zr_materials m{};
    m.n = e->n;
    int rc = zr_set_materials(e, &m);
    if (rc) return rc;
  }
  ZR_HIP(e, hipMemcpy(e->mat, code.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
  e->has_models = true;
  e->uniform_model = ZR_MODEL_NEO_HOOKEAN_BORDEN;
  e->all_hencky = false;
  double** dst[] = { &e->pf_c, &e->pf_H, &e->pf_l0, &e->pf_osc, &e->pf_Fp, &e->pf_Hmax, &e->pf_k, &e->pf_vol };
  const double* src[] = { pf->c, pf->H, pf->l0, pf->one_over_sigma_c, pf->pf_Fp, pf->H_max, pf->residual_phase,
                          nullptr };
  const double dflt[] = { 1.0, 0.0, 0.0, 0.0, 1.0, HUGE_VAL, 0.001, 0.0 };
  std::vector<double> tmp(N);
  for (int a = 0; a < 8; ++a) {
    if (!*dst[a]) ZR_HIP(e, hipMalloc(dst[a], N * sizeof(double)));
    for (size_t i = 0; i < N; ++i) tmp[i] = src[a] ? src[a][i] : dflt[a];
    ZR_HIP(e, hipMemcpy(*dst[a], tmp.data(), N * sizeof(double), hipMemcpyHostToDevice));
  }
  if (!e->pf_allow) ZR_HIP(e, hipMalloc(&e->pf_allow, N * sizeof(int32_t)));
  std::vector<int32_t> al(N, 1);
  if (pf->allow_damage)
    for (size_t i = 0; i < N; ++i) al[i] = pf->allow_damage[i] ? 1 : 0;
  ZR_HIP(e, hipMemcpy(e->pf_allow, al.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
  e->pf_opts = *o;
  e->pf_on = true;
  return ZR_OK;
}

int zr_get_phase_field(zr_engine* e, zr_phase_field* pf)
{
  if (!e || !pf) return fail(e, ZR_ERR_ARG, "bad phase-field view");
  if (pf->n != e->n) return fail(e, ZR_ERR_ARG, "phase-field count differs from the particle count");
  if (!e->pf_on) return fail(e, ZR_ERR_STATE, "no phase field was set");
  const size_t N = (size_t)e->n;
  if (N == 0) return ZR_OK;
  if (pf->c) ZR_HIP(e, hipMemcpyAsync(pf->c, e->pf_c, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  if (pf->H) ZR_HIP(e, hipMemcpyAsync(pf->H, e->pf_H, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  if (pf->pf_Fp) ZR_HIP(e, hipMemcpyAsync(pf->pf_Fp, e->pf_Fp, N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

int zr_phase_field_solve(zr_engine* e, double dt, zr_phase_field_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  if (!(dt > 0.0)) return fail(e, ZR_ERR_ARG, "dt must be positive");
  if (!e->pf_on) return fail(e, ZR_ERR_STATE, "no phase field was set");
  int rc = do_pf(e, dt);
  if (rc) return rc;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  if (st) *st = e->pf_stats;
  return ZR_OK;
}

int zr_phase_field_last_stats(zr_engine* e, zr_phase_field_stats* st)
{
  if (!e || !st) return ZR_ERR_ARG;
  *st = e->pf_stats;
  return ZR_OK;
}

int zr_adjoint_set(zr_engine* e, const zr_particles_adjoint* a)
{
  if (!e || !a) return fail(e, ZR_ERR_ARG, "bad adjoint view");
  if (a->n != e->n) return fail(e, ZR_ERR_ARG, "adjoint count differs from the particle count");
  if (!adjoint_supported(e)) return ZR_ERR_ARG;
  if (e->n == 0) return ZR_OK;
  int rc = ensure_adjoint(e);
  if (rc) return rc;
  const size_t N = (size_t)e->n;
  AdjointView A = adj_view(e);
  double* tmp = nullptr;
  ZR_HIP(e, hipMalloc(&tmp, 9 * N * sizeof(double)));
  struct Item {
    const double* src;
    double* dst;
    int nc;
  } items[] = { { a->X, A.x, 3 }, { a->V, A.v, 3 }, { a->C, A.C, 9 }, { a->F, A.F, 9 }, { a->mu, A.mu, 1 },
                { a->lambda, A.lam, 1 } };
  for (auto& it : items) {
    hipError_t st;
    if (it.src) {
      st = hipMemcpyAsync(tmp, it.src, it.nc * N * sizeof(double), hipMemcpyHostToDevice, e->stream);
      if (st == hipSuccess) {
        k_aos_to_soa<<<blocks_for((uint64_t)N * it.nc, 256), 256, 0, e->stream>>>(tmp, it.dst, (uint32_t)N, it.nc,
                                                                                  e->n);
        st = hipGetLastError();
      }
    } else {
      st = hipMemsetAsync(it.dst, 0, it.nc * N * sizeof(double), e->stream);
    }
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) {
      (void)hipFree(tmp);
      return fail(e, ZR_ERR_HIP, std::string("adjoint upload: ") + hipGetErrorString(st));
    }
  }
  ZR_HIP(e, hipFree(tmp));
  return ZR_OK;
}

int zr_adjoint_get(zr_engine* e, zr_particles_adjoint* a)
{
  if (!e || !a) return fail(e, ZR_ERR_ARG, "bad adjoint view");
  if (a->n != e->n) return fail(e, ZR_ERR_ARG, "adjoint count differs from the particle count");
  if (e->n == 0) return ZR_OK;
  if (!e->adj) return fail(e, ZR_ERR_STATE, "no adjoint was set");
  const size_t N = (size_t)e->n;
  AdjointView A = adj_view(e);
  double* tmp = nullptr;
  ZR_HIP(e, hipMalloc(&tmp, 9 * N * sizeof(double)));
  struct Item {
    double* dst;
    const double* src;
    int nc;
  } items[] = { { a->X, A.x, 3 }, { a->V, A.v, 3 }, { a->C, A.C, 9 }, { a->F, A.F, 9 }, { a->mu, A.mu, 1 },
                { a->lambda, A.lam, 1 } };
  for (auto& it : items) {
    if (!it.dst) continue;
    k_soa_to_aos<<<blocks_for((uint64_t)N * it.nc, 256), 256, 0, e->stream>>>(it.src, tmp, (uint32_t)N, it.nc, e->n);
    hipError_t st = hipGetLastError();
    if (st == hipSuccess) st = hipMemcpyAsync(it.dst, tmp, it.nc * N * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) {
      (void)hipFree(tmp);
      return fail(e, ZR_ERR_HIP, std::string("adjoint download: ") + hipGetErrorString(st));
    }
  }
  ZR_HIP(e, hipFree(tmp));
  return ZR_OK;
}

int zr_adjoint_step(zr_engine* e, double dt)
{
  if (!e) return ZR_ERR_ARG;
  if (!(dt > 0.0)) return fail(e, ZR_ERR_ARG, "dt must be positive");
  if (!adjoint_supported(e)) return ZR_ERR_ARG;
  if (!e->adj || e->adj_n != e->n) return fail(e, ZR_ERR_STATE, "zr_adjoint_set was not called");
  ZR_HIP(e, hipMemsetAsync(e->dcount + CNT_ADJ_BAD, 0, sizeof(uint32_t), e->stream));
  return do_adjoint_step(e, dt, true);
}

namespace {
// the checkpoint arrays X, V, C, F and a 24n-double checkpoint, copied in
// one pass (to: state -> checkpoint, else back)
__global__ __launch_bounds__(256) void k_ckpt_copy(double* x, double* v, double* C, double* F, double* c, uint64_t N,
                                                   int to)
{
  const uint64_t total = 24 * N, stride = (uint64_t)gridDim.x * 256u;
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < total; i += stride) {
    double* a;
    uint64_t j;
    if (i < 3 * N) {
      a = x;
      j = i;
    } else if (i < 6 * N) {
      a = v;
      j = i - 3 * N;
    } else if (i < 15 * N) {
      a = C;
      j = i - 6 * N;
    } else {
      a = F;
      j = i - 15 * N;
    }
    if (to)
      c[i] = a[j];
    else
      a[j] = c[i];
  }
}

// checkpoint save / restore, enqueued on the engine's stream
int ckpt_save(zr_engine* e, int32_t slot)
{
  if (slot < 0 || slot > 4096) return fail(e, ZR_ERR_ARG, "bad checkpoint slot");
  const size_t N = (size_t)e->n;
  if ((size_t)slot >= e->ckpt.size()) {
    e->ckpt.resize((size_t)slot + 1, nullptr);
    e->ckpt_time.resize((size_t)slot + 1, 0.0);
  }
  if (!e->ckpt[slot]) ZR_HIP(e, hipMalloc(&e->ckpt[slot], 24 * (N ? N : 1) * sizeof(double)));
  if (N) {
    const uint64_t b = blocks_for(24 * (uint64_t)N, 256);
    k_ckpt_copy<<<b < 2048 ? (uint32_t)b : 2048u, 256, 0, e->stream>>>(e->x, e->v, e->C, e->F, e->ckpt[slot], N, 1);
    ZR_LAUNCH_CHECK(e);
  }
  e->ckpt_time[slot] = e->time;
  return ZR_OK;
}

int ckpt_restore(zr_engine* e, int32_t slot)
{
  if (slot < 0 || (size_t)slot >= e->ckpt.size() || !e->ckpt[slot])
    return fail(e, ZR_ERR_ARG, "no such checkpoint");
  const size_t N = (size_t)e->n;
  if (N) {
    const uint64_t b = blocks_for(24 * (uint64_t)N, 256);
    k_ckpt_copy<<<b < 2048 ? (uint32_t)b : 2048u, 256, 0, e->stream>>>(e->x, e->v, e->C, e->F, e->ckpt[slot], N, 0);
    ZR_LAUNCH_CHECK(e);
  }
  e->time = e->ckpt_time[slot];
  e->sorted = e->rasterised = e->updated = e->be_ready = false;
  return ZR_OK;
}
}  // namespace

int zr_checkpoint_save(zr_engine* e, int32_t slot)
{
  if (!e) return ZR_ERR_ARG;
  int rc = ckpt_save(e, slot);
  if (rc) return rc;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

int zr_checkpoint_restore(zr_engine* e, int32_t slot)
{
  if (!e) return ZR_ERR_ARG;
  int rc = ckpt_restore(e, slot);
  if (rc) return rc;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

// the checkpointed rollout's loops (the same calls as the Python driver's,
// without a host round trip per call)
namespace {
// the steps of a rollout defer their CFL flags to the next step's sort
struct DeferCfl {
  zr_engine* e;
  explicit DeferCfl(zr_engine* en) : e(en) { e->defer_cfl = true; }
  ~DeferCfl() { e->defer_cfl = false; }
};

// the flags of the rollout's last step
int rollout_cfl(zr_engine* e)
{
  int rc = sync_counts(e);
  if (rc) return rc;
  return (e->hcount[CNT_CFL_RUN] & 1u) ? ZR_CFL_RESTART : ZR_OK;
}
}  // namespace

int zr_rollout_forward(zr_engine* e, int32_t steps, double dt, int32_t every, int32_t slot0)
{
  if (!e) return ZR_ERR_ARG;
  if (steps < 0 || every < 1 || slot0 < 0 || !(dt > 0.0)) return fail(e, ZR_ERR_ARG, "bad rollout arguments");
  ZR_HIP(e, hipMemsetAsync(e->dcount + CNT_CFL_RUN, 0, sizeof(uint32_t), e->stream));
  DeferCfl defer(e);
  for (int32_t k = 0; k < steps; ++k) {
    if (k % every == 0) {
      int rc = ckpt_save(e, slot0 + k / every);
      if (rc) return rc;
    }
    // an error, or ZR_CFL_RESTART from the sort after the step that asked
    // for it: the rollout stops there
    int rc = zr_advance_one_time_step(e, dt, nullptr);
    if (rc) return rc;
  }
  return rollout_cfl(e);
}

int zr_rollout_backward(zr_engine* e, int32_t steps, double dt, int32_t every, int32_t slot0, int32_t scratch0)
{
  if (!e) return ZR_ERR_ARG;
  if (steps < 0 || every < 1 || slot0 < 0 || scratch0 < 0 || !(dt > 0.0))
    return fail(e, ZR_ERR_ARG, "bad rollout arguments");
  if (!adjoint_supported(e)) return ZR_ERR_ARG;
  if (!e->adj || e->adj_n != e->n) return fail(e, ZR_ERR_STATE, "zr_adjoint_set was not called");
  static_assert(CNT_CFL_RUN == CNT_ADJ_BAD + 1, "cleared together");
  ZR_HIP(e, hipMemsetAsync(e->dcount + CNT_ADJ_BAD, 0, 2 * sizeof(uint32_t), e->stream));
  DeferCfl defer(e);
  const int32_t nseg = (steps + every - 1) / every;
  for (int32_t seg = nseg - 1; seg >= 0; --seg) {
    const int32_t k0 = seg * every, k1 = k0 + every < steps ? k0 + every : steps;
    int rc = ckpt_restore(e, slot0 + seg);
    if (rc) return rc;
    // the segment's step inputs, and the sort tables and grid of each
    // recomputed step, which its adjoint step reads instead of sorting and
    // rasterising the same input again
    if (e->step_rec.size() < (size_t)(k1 - k0)) e->step_rec.resize((size_t)(k1 - k0));
    for (int32_t k = k0; k < k1; ++k) {
      if (k > k0) {  // the first step's input is the segment's checkpoint
        rc = ckpt_save(e, scratch0 + (k - k0));
        if (rc) return rc;
      }
      if (k < k1 - 1) {
        rc = advance_recorded(e, dt, e->step_rec[(size_t)(k - k0)]);
        if (rc) return rc;
      }
    }
    for (int32_t k = k1 - 1; k >= k0; --k) {  // the adjoint of each step, last first
      rc = ckpt_restore(e, k > k0 ? scratch0 + (k - k0) : slot0 + seg);
      if (rc) return rc;
      if (k == k1 - 1)
        rc = do_adjoint_step(e, dt, false);
      else
        rc = adjoint_kernels(e, dt, record_tables(e, e->step_rec[(size_t)(k - k0)]));
      if (rc) return rc;
    }
  }
  int rc = adjoint_flags(e);
  if (rc) return rc;
  return (e->hcount[CNT_CFL_RUN] & 1u) ? ZR_CFL_RESTART : ZR_OK;
}

int zr_advance_one_time_step_admm(zr_engine* e, double dt, const zr_admm_opts* o, zr_admm_stats* ast,
                                  zr_step_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  if (!(dt > 0.0) || !o || o->admm_iterations < 0 || o->minres_max_iterations < 0 || !(o->rho_scale > 0.0) ||
      (o->precision != ZR_FP64 && o->precision != ZR_FP32_REFINED))
    return fail(e, ZR_ERR_ARG, "bad dt or ADMM options");
  if (e->n > 0 && (!e->mat || !e->all_hencky || e->pf_on))
    return fail(e, ZR_ERR_ARG, "the ADMM step needs StvkWithHencky particles (zr_set_materials)");
  if (e->visco && e->has_plastic)
    return fail(e, ZR_ERR_ARG, "the ADMM step's viscoelastic branch needs particles without plasticity");
  if (ast) ast->minres_iterations = 0;
  int rc = do_sort(e, st);
  if (rc) return rc;
  rc = do_p2g(e, false);
  if (rc) return rc;
  rc = do_grid(e, GRID_COMBINE, 0.0);
  if (rc) return rc;
  rc = do_admm(e, dt, o, ast);
  if (rc) return rc;
  // gridToParticles: with the viscoelastic branch the APIC G2P also writes
  // grad v, from which that helper's evolveStrain moves and relaxes Fneq
  e->g2p_gradv = e->visco ? e->vtau_neq : nullptr;
  rc = do_g2p(e, dt, st, false);
  e->g2p_gradv = nullptr;
  if (rc < 0) return rc;
  if (e->visco) {
    k_visco_strain<<<blocks_for((uint64_t)e->n, 256), 256, 0, e->stream>>>((uint32_t)e->n, e->vtau_neq, e->fneq,
                                                                           e->mu_neq, e->lam_neq, e->n, dt, e->visc_d,
                                                                           e->visc_v);
    ZR_LAUNCH_CHECK(e);
  }
  e->time += dt;
  return rc;
}

int zr_get_admm_state(zr_engine* e, double* F, double* y, double* omega)
{
  if (!e) return ZR_ERR_ARG;
  if (!e->admm_p || e->admm_n != e->n) return fail(e, ZR_ERR_STATE, "no ADMM step has run");
  const size_t N = (size_t)e->n;
  double* tmp = nullptr;
  ZR_HIP(e, hipMalloc(&tmp, 9 * N * sizeof(double)));
  struct Item {
    double* dst;
    const double* src;
    int nc;
  } items[] = { { F, e->admm_p, 9 }, { y, e->admm_p + 9 * N, 9 }, { omega, e->admm_p + 18 * N, 1 } };
  for (auto& it : items) {
    if (!it.dst) continue;
    k_soa_to_aos<<<blocks_for((uint64_t)N * it.nc, 256), 256, 0, e->stream>>>(it.src, tmp, (uint32_t)N, it.nc, e->n);
    hipError_t stt = hipGetLastError();
    if (stt == hipSuccess)
      stt = hipMemcpyAsync(it.dst, tmp, it.nc * N * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (stt == hipSuccess) stt = hipStreamSynchronize(e->stream);
    if (stt != hipSuccess) {
      (void)hipFree(tmp);
      return fail(e, ZR_ERR_HIP, std::string("admm download: ") + hipGetErrorString(stt));
    }
  }
  ZR_HIP(e, hipFree(tmp));
  return ZR_OK;
}

int zr_sort_particles_and_pollute_grid(zr_engine* e, zr_step_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  return do_sort(e, st);
}

int zr_particles_to_grid(zr_engine* e)
{
  if (!e) return ZR_ERR_ARG;
  int rc = do_p2g(e, true);
  if (rc) return rc;
  rc = do_grid(e, GRID_COMBINE, 0.0);
  if (rc) return rc;
  e->updated = false;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

int zr_grid_velocity_explicit_update(zr_engine* e, double dt)
{
  if (!e) return ZR_ERR_ARG;
  if (!e->rasterised) return fail(e, ZR_ERR_STATE, "grid update before particlesToGrid");
  int rc = do_grid(e, GRID_UPDATE, dt);
  if (rc) return rc;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

int zr_grid_to_particles(zr_engine* e, double dt, zr_step_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  return do_g2p(e, dt, st, true);
}

int zr_advance_one_time_step(zr_engine* e, double dt, zr_step_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  if (!(dt > 0.0)) return fail(e, ZR_ERR_ARG, "dt must be positive");
  if (e->pf_on && e->n > 0) {  // reinitialize(): FractureSimulation::deleteParticles
    k_pf_delete<<<blocks_for((uint64_t)e->n, 256), 256, 0, e->stream>>>(pview(e), e->pf_c, e->pf_allow,
                                                                        e->pf_opts.delete_threshold);
    ZR_LAUNCH_CHECK(e);
  }
  int rc = do_sort(e, st);
  if (rc) return rc;
  rc = do_p2g(e, true, dt);
  if (rc) return rc;
  if (e->pf_on) {  // FractureSimulation::particlesToGrid: the damage solve after P2G
    rc = do_grid(e, GRID_COMBINE, dt);
    if (rc) return rc;
    rc = do_pf(e, dt);
    if (rc) return rc;
    rc = do_grid(e, GRID_UPDATE, dt);
  } else {
    rc = do_grid(e, GRID_COMBINE_UPDATE, dt);
  }
  if (rc) return rc;
  rc = do_g2p(e, dt, st, true);
  if (rc < 0) return rc;
  e->time += dt;
  return rc;
}

static bool implicit_materials_ok(zr_engine* e);

int zr_particles_to_grid_apic(zr_engine* e)
{
  if (!e) return ZR_ERR_ARG;
  if (!implicit_materials_ok(e)) return ZR_ERR_ARG;
  int rc = do_p2g(e, false);
  if (rc) return rc;
  rc = do_grid(e, GRID_COMBINE, 0.0);
  if (rc) return rc;
  e->updated = false;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  return ZR_OK;
}

static bool valid_opts(const zr_implicit_opts* o)
{
  return o && o->newton_max_iterations >= 0 && o->cg_max_iterations >= 0 &&
         (o->precision == ZR_FP64 || o->precision == ZR_FP32_REFINED) &&
         (o->preconditioner == ZR_PRECOND_MASS || o->preconditioner == ZR_PRECOND_JACOBI) &&
         (o->solver == ZR_SOLVER_CG || o->solver == ZR_SOLVER_MINRES);
}

static bool implicit_materials_ok(zr_engine* e)
{
  if (e->has_models || e->has_plastic || e->visco) {
    fail(e, ZR_ERR_ARG,
         "the implicit step supports fixed-corotated particles without plasticity or a viscoelastic branch only");
    return false;
  }
  return true;
}

int zr_backward_euler_step(zr_engine* e, double dt, const zr_implicit_opts* o, zr_implicit_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  if (!implicit_materials_ok(e)) return ZR_ERR_ARG;
  if (!(dt > 0.0) || !valid_opts(o)) return fail(e, ZR_ERR_ARG, "bad dt or implicit options");
  return do_be(e, dt, o, st);
}

int zr_grid_to_particles_apic(zr_engine* e, double dt, zr_step_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  if (!implicit_materials_ok(e)) return ZR_ERR_ARG;
  return do_g2p(e, dt, st, false);
}

int zr_advance_one_time_step_implicit(zr_engine* e, double dt, const zr_implicit_opts* o, zr_implicit_stats* ist,
                                      zr_step_stats* st)
{
  if (!e) return ZR_ERR_ARG;
  if (!(dt > 0.0) || !valid_opts(o)) return fail(e, ZR_ERR_ARG, "bad dt or implicit options");
  if (!implicit_materials_ok(e)) return ZR_ERR_ARG;
  int rc = do_sort(e, st);
  if (rc) return rc;
  rc = do_p2g(e, false);
  if (rc) return rc;
  rc = do_grid(e, GRID_COMBINE, 0.0);
  if (rc) return rc;
  rc = do_be(e, dt, o, ist);
  if (rc) return rc;
  rc = do_g2p(e, dt, st, false);
  if (rc < 0) return rc;
  e->time += dt;
  return rc;
}

// host node-major (3 per node) <-> device [3][grid_cap]
static int upload_nodevec(zr_engine* e, const double* host, double* dev)
{
  const size_t nn = (size_t)e->S * 32;
  std::vector<double> h(3 * nn);
  for (size_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d) h[d * nn + k] = host[3 * k + d];
  for (int d = 0; d < 3; ++d)
    ZR_HIP(e, hipMemcpy(dev + d * e->grid_cap, h.data() + d * nn, nn * sizeof(double), hipMemcpyHostToDevice));
  return ZR_OK;
}
static int download_nodevec(zr_engine* e, const double* dev, double* host)
{
  const size_t nn = (size_t)e->S * 32;
  std::vector<double> h(3 * nn);
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  for (int d = 0; d < 3; ++d)
    ZR_HIP(e, hipMemcpy(h.data() + d * nn, dev + d * e->grid_cap, nn * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d) host[3 * k + d] = h[d * nn + k];
  return ZR_OK;
}

int zr_implicit_residual(zr_engine* e, const double* dv, double* r)
{
  if (!e || !dv || !r) return ZR_ERR_ARG;
  if (!e->be_ready) return fail(e, ZR_ERR_STATE, "no backward-Euler state");
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  int rc = upload_nodevec(e, dv, e->idv);
  if (rc) return rc;
  rc = e->be_prec == ZR_FP32_REFINED ? be_state_residual<float>(e) : be_state_residual<double>(e);
  if (rc) return rc;
  return download_nodevec(e, e->ir, r);
}

int zr_implicit_multiply(zr_engine* e, const double* x, double* b)
{
  if (!e || !x || !b) return ZR_ERR_ARG;
  if (!e->be_ready) return fail(e, ZR_ERR_STATE, "no backward-Euler state");
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  int rc = upload_nodevec(e, x, e->ip);
  if (rc) return rc;
  if (e->be_prec == ZR_FP32_REFINED) {
    // through the FP32 operator: cast in, apply, cast out
    const size_t nn = (size_t)e->S * 32;
    std::vector<double> h(3 * nn);
    std::vector<float> hf(3 * nn);
    for (size_t k = 0; k < 3 * nn; ++k) hf[k] = (float)x[k];
    std::vector<float> soa(3 * nn);
    for (size_t k = 0; k < nn; ++k)
      for (int d = 0; d < 3; ++d) soa[d * nn + k] = hf[3 * k + d];
    for (int d = 0; d < 3; ++d)
      ZR_HIP(e, hipMemcpy(e->fp + d * e->grid_cap, soa.data() + d * nn, nn * sizeof(float), hipMemcpyHostToDevice));
    rc = be_apply<float>(e, nullptr, e->fp, e->fap, 0, nullptr, 0);
    if (rc) return rc;
    ZR_HIP(e, hipStreamSynchronize(e->stream));
    for (int d = 0; d < 3; ++d)
      ZR_HIP(e, hipMemcpy(soa.data() + d * nn, e->fap + d * e->grid_cap, nn * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < nn; ++k)
      for (int d = 0; d < 3; ++d) b[3 * k + d] = (double)soa[d * nn + k];
    return ZR_OK;
  }
  rc = be_apply<double>(e, nullptr, e->ip, e->iap, 0, nullptr, 0);
  if (rc) return rc;
  return download_nodevec(e, e->iap, b);
}

int64_t zr_debug_size(zr_engine* e, int32_t what)
{
  if (!e) return -1;
  switch (what) {
    case ZR_DUMP_SORTER: return e->n * 8;
    case ZR_DUMP_ORDER: return e->n * 4;
    case ZR_DUMP_BASE_OFFSET: return e->n * 8;
    case ZR_DUMP_GROUP_FIRST:
    case ZR_DUMP_GROUP_LAST:
    case ZR_DUMP_BLOCK_OFFSET: return (int64_t)e->A * 8;
    case ZR_DUMP_PAGES: return (int64_t)e->S * 8;
    case ZR_DUMP_GRID_M: return (int64_t)e->S * 32 * 8;
    case ZR_DUMP_GRID_V:
    case ZR_DUMP_GRID_F:
    case ZR_DUMP_GRID_VSTAR: return (int64_t)e->S * 32 * 3 * 8;
    case ZR_DUMP_NODE_IDX: return (int64_t)e->S * 32 * 8;
    case ZR_DUMP_DV: return e->be_ready ? (int64_t)e->S * 32 * 3 * 8 : 0;
    case ZR_DUMP_HVP_XCD: return e->be_ready ? 20 * 4 : 0;
    case ZR_DUMP_JACOBI: return e->be_ready ? (int64_t)e->S * 32 * 3 * 8 : 0;
    default: return -1;
  }
}

int zr_debug_dump(zr_engine* e, int32_t what, void* host, size_t bytes)
{
  if (!e) return ZR_ERR_ARG;
  int64_t need = zr_debug_size(e, what);
  if (need < 0 || (size_t)need != bytes || (need > 0 && !host)) return fail(e, ZR_ERR_ARG, "bad dump request");
  if (need == 0) return ZR_OK;
  hipStream_t s = e->stream;
  const uint32_t n = (uint32_t)e->n, A = e->A, nn = e->S * 32;
  ZR_HIP(e, hipStreamSynchronize(s));
  if (what == ZR_DUMP_SORTER || what == ZR_DUMP_BASE_OFFSET) {
    uint64_t* tmp = nullptr;
    ZR_HIP(e, hipMalloc(&tmp, n * sizeof(uint64_t)));
    if (what == ZR_DUMP_SORTER)
      k_dump_sorter<<<blocks_for(n, 256), 256, 0, s>>>(n, e->order, e->p_page, e->p_re, tmp);
    else
      k_dump_base_offset<<<blocks_for(n, 256), 256, 0, s>>>(n, e->p_page, e->p_re, tmp);
    hipError_t st = hipMemcpyAsync(host, tmp, bytes, hipMemcpyDeviceToHost, s);
    if (st == hipSuccess) st = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    if (st != hipSuccess) return fail(e, ZR_ERR_HIP, hipGetErrorString(st));
    return ZR_OK;
  }
  if (what == ZR_DUMP_ORDER) {
    ZR_HIP(e, hipMemcpy(host, e->order, bytes, hipMemcpyDeviceToHost));
    return ZR_OK;
  }
  if (what == ZR_DUMP_GROUP_FIRST || what == ZR_DUMP_GROUP_LAST || what == ZR_DUMP_BLOCK_OFFSET ||
      what == ZR_DUMP_PAGES) {
    const uint32_t* src = what == ZR_DUMP_PAGES ? e->slot_page : (what == ZR_DUMP_BLOCK_OFFSET ? e->group_page : e->group_start);
    size_t cnt = what == ZR_DUMP_PAGES ? e->S : (what == ZR_DUMP_BLOCK_OFFSET ? A : A + 1);
    std::vector<uint32_t> h(cnt);
    ZR_HIP(e, hipMemcpy(h.data(), src, cnt * sizeof(uint32_t), hipMemcpyDeviceToHost));
    int64_t* o64 = (int64_t*)host;
    uint64_t* u64 = (uint64_t*)host;
    if (what == ZR_DUMP_GROUP_FIRST)
      for (uint32_t g = 0; g < A; ++g) o64[g] = h[g];
    else if (what == ZR_DUMP_GROUP_LAST)
      for (uint32_t g = 0; g < A; ++g) o64[g] = (int64_t)h[g + 1] - 1;
    else if (what == ZR_DUMP_BLOCK_OFFSET)
      for (uint32_t g = 0; g < A; ++g) u64[g] = h[g];
    else
      for (uint32_t k = 0; k < e->S; ++k) u64[k] = (uint64_t)h[k] << 12;
    return ZR_OK;
  }
  if (what == ZR_DUMP_GRID_M) {
    ZR_HIP(e, hipMemcpy(host, e->gm, bytes, hipMemcpyDeviceToHost));
    return ZR_OK;
  }
  if (what == ZR_DUMP_GRID_V || what == ZR_DUMP_GRID_F || what == ZR_DUMP_GRID_VSTAR || what == ZR_DUMP_DV) {
    const double* src = what == ZR_DUMP_GRID_V ? e->gv
                        : what == ZR_DUMP_GRID_F ? e->gf
                        : what == ZR_DUMP_DV     ? e->idv
                                                 : e->gvs;
    std::vector<double> h(3 * (size_t)nn);
    for (int d = 0; d < 3; ++d)
      ZR_HIP(e, hipMemcpy(h.data() + (size_t)d * nn, src + (size_t)d * e->grid_cap, nn * sizeof(double),
                          hipMemcpyDeviceToHost));
    double* o = (double*)host;
    for (uint32_t k = 0; k < nn; ++k)
      for (int d = 0; d < 3; ++d) o[3 * (size_t)k + d] = h[(size_t)d * nn + k];
    return ZR_OK;
  }
  if (what == ZR_DUMP_JACOBI) {
    int rc = ensure_jacobi(e);
    if (rc) return rc;
    const bool f32 = e->be_prec == ZR_FP32_REFINED;
    rc = f32 ? be_state_residual<float>(e, true) : be_state_residual<double>(e, true);
    if (rc) return rc;
    rc = f32 ? jacobi_inverse<float>(e) : jacobi_inverse<double>(e);
    if (rc) return rc;
    ZR_HIP(e, hipStreamSynchronize(s));
    std::vector<double> h(3 * (size_t)nn);
    if (f32) {
      std::vector<float> hf(nn);
      for (int d = 0; d < 3; ++d) {
        ZR_HIP(e, hipMemcpy(hf.data(), reinterpret_cast<const float*>(e->pinv) + (size_t)d * e->grid_cap,
                            nn * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nn; ++k) h[(size_t)d * nn + k] = hf[k];
      }
    } else {
      for (int d = 0; d < 3; ++d)
        ZR_HIP(e, hipMemcpy(h.data() + (size_t)d * nn, e->pinv + (size_t)d * e->grid_cap, nn * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    double* out = static_cast<double*>(host);
    for (size_t k = 0; k < nn; ++k)
      for (int d = 0; d < 3; ++d) out[3 * k + d] = h[(size_t)d * nn + k];
    return ZR_OK;
  }
  if (what == ZR_DUMP_HVP_XCD) {
    uint32_t* h = static_cast<uint32_t*>(host);
    ZR_HIP(e, hipMemcpy(h, e->hvp_bounds, 18 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    h[18] = A < e->hvp_resident[0] ? A : e->hvp_resident[0];
    h[19] = A < e->hvp_resident[1] ? A : e->hvp_resident[1];
    return ZR_OK;
  }
  if (what == ZR_DUMP_NODE_IDX) {
    uint32_t* sc = nullptr;
    int64_t* idx = nullptr;
    ZR_HIP(e, hipMalloc(&sc, (size_t)nn * sizeof(uint32_t)));
    ZR_HIP(e, hipMalloc(&idx, (size_t)nn * sizeof(int64_t)));
    int rc = run_scan(e, LdNonzeroMass{ e->gm }, nn, sc, e->dcount + CNT_TOTAL);
    if (rc == ZR_OK) {
      k_node_idx<<<blocks_for(nn, 256), 256, 0, s>>>(nn, e->gm, sc, idx);
      hipError_t st = hipGetLastError();
      if (st == hipSuccess) st = hipMemcpyAsync(host, idx, bytes, hipMemcpyDeviceToHost, s);
      if (st == hipSuccess) st = hipStreamSynchronize(s);
      if (st != hipSuccess) rc = fail(e, ZR_ERR_HIP, hipGetErrorString(st));
    }
    (void)hipFree(sc);
    (void)hipFree(idx);
    return rc;
  }
  return fail(e, ZR_ERR_ARG, "unknown dump");
}

int zr_profile_enable(zr_engine* e, int32_t on)
{
  if (!e) return ZR_ERR_ARG;
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  e->prof = on != 0;
  e->prof_period = on > 0 ? on : 1;
  for (int k = 0; k < ZR_PROF_COUNT; ++k) {
    e->prof_used[k] = 0;
    e->prof_seen[k] = 0;
    e->prof_phase[k] = false;
    e->prof_take[k] = false;
  }
  return ZR_OK;
}

int zr_profile_read(zr_engine* e, int32_t which, double* total_ms, int64_t* launches)
{
  if (!e || which < 0 || which >= ZR_PROF_COUNT) return fail(e, ZR_ERR_ARG, "bad profile slot");
  ZR_HIP(e, hipStreamSynchronize(e->stream));
  double t = 0.0;
  size_t pairs = e->prof_used[which] / 2;
  for (size_t k = 0; k < pairs; ++k) {
    float ms = 0.f;
    ZR_HIP(e, hipEventElapsedTime(&ms, e->prof_ev[which][2 * k], e->prof_ev[which][2 * k + 1]));
    t += ms;
  }
  if (total_ms) *total_ms = t;
  if (launches) *launches = (int64_t)pairs;
  return ZR_OK;
}

}  // extern "C"
