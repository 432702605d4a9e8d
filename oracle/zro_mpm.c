/*
 * zro_mpm.c -- oracle (TEST INFRASTRUCTURE ONLY): SPGrid indexing, particle
 * sort + grid pollution, quadratic B-splines and the explicit MLS-MPM step,
 * FP64, restated from the reference:
 *   Lib/SPGrid/Core/SPGrid_Mask.h:22-305, SPGrid_Utilities.{h:80-358,cpp:43-62}
 *   Lib/SPGrid/Core/SPGrid_Page_Map.h:53-97      (first-touch block list)
 *   Lib/MPM/MpmGrid.h:14-34,55-78,123-271        (GridState, weights, iteration)
 *   Lib/Ziran/Math/Splines/BSplines.h:10-81, Lib/Ziran/Math/MathTools.h:15-25
 *   Lib/MPM/MpmSimulationBase.cpp:256-312,437-561,681-709,818-865,906-1103
 *   Lib/MPM/Force/FBasedMpmForceHelper.cpp:44-64,95-110
 *   Lib/Ziran/Math/Geometry/CollisionObject.cpp:48-105,340-405
 *   Lib/Ziran/Math/Geometry/AnalyticLevelSet.cpp:111-119,259-287,435-450
 * P2G runs the reference's 8-colour page schedule with OpenMP
 * (MpmSimulationBase.h:252-265, MpmSimulationBase.cpp:499-502).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "zr_oracle.h"

/* ---------------- SPGrid masks for GridState<double,3> ----------------
 * sizeof(GridState<double,3>) = 128 B -> data_bits = 7, block_bits = 5,
 * block (x,y,z) bits = (1,2,2), lob = 1 (SPGrid_Mask.h:27-47,76-123):
 *   page_xmask = (0x4924..24 << 1) & ~0xfff, element_xmask = 1 << 11
 *   page_ymask = (0x2492..92 << 1) & ~0xfff, element_ymask = 3 << 9
 *   page_zmask = (0x9249..49 << 1) & ~0xfff, element_zmask = 3 << 7       */
#define ZRO_XMASK 0x9249249249249800ULL
#define ZRO_YMASK 0x4924924924924600ULL
#define ZRO_ZMASK 0x2492492492492180ULL
#define ZRO_WMASK (ZRO_XMASK | ZRO_YMASK | ZRO_ZMASK)

/* Bit_Spread_u64 / pdep (SPGrid_Utilities.h:80-358) */
static uint64_t spread(uint64_t data, uint64_t mask)
{
  uint64_t r = 0;
  for (uint64_t bit = 1; bit; bit <<= 1)
    if (mask & bit) { r |= (data & 1ULL) ? bit : 0; data >>= 1; }
  return r;
}
/* Bit_Pack (SPGrid_Utilities.cpp:43-62) */
static int64_t pack(uint64_t data, uint64_t mask)
{
  uint64_t r = 0;
  int count = 0;
  for (uint64_t bit = 1; bit; bit <<= 1)
    if (mask & bit) r |= (data & bit) >> count;
    else count++;
  return (int64_t)r;
}

/* Linear_Offset (SPGrid_Mask.h:150-174); coordinates are sign-extended */
uint64_t zro_linear_offset(int i, int j, int k)
{
  return spread((uint64_t)(int64_t)i, ZRO_XMASK) | spread((uint64_t)(int64_t)j, ZRO_YMASK) |
         spread((uint64_t)(int64_t)k, ZRO_ZMASK);
}
void zro_linear_to_coord(uint64_t off, int* ijk)
{
  ijk[0] = (int)pack(off, ZRO_XMASK);
  ijk[1] = (int)pack(off, ZRO_YMASK);
  ijk[2] = (int)pack(off, ZRO_ZMASK);
}
/* Packed_Add (SPGrid_Mask.h:237-245) */
uint64_t zro_packed_add(uint64_t i, uint64_t j)
{
  uint64_t x = ((i | ~ZRO_XMASK) + (j & ZRO_XMASK)) & ZRO_XMASK;
  uint64_t y = ((i | ~ZRO_YMASK) + (j & ZRO_YMASK)) & ZRO_YMASK;
  uint64_t z = ((i | ~ZRO_ZMASK) + (j & ZRO_ZMASK)) & ZRO_ZMASK;
  uint64_t w = ((i | ZRO_WMASK) + (j & ~ZRO_WMASK)) & ~ZRO_WMASK;
  return x | y | z | w;
}

/* MathTools.h:21-25 */
static int int_floor(double x)
{
  int i = (int)x;
  return i - (i > x);
}

/* BSplines.h:16-20,55-81 (quadratic) */
void zro_bspline_weights(double x, int* base, double* w, double* dw)
{
  int b = int_floor(x - 0.5);
  double d0 = x - b;
  double z = 1.5 - d0;
  w[0] = 0.5 * z * z;
  double d1 = d0 - 1.0;
  w[1] = 0.75 - d1 * d1;
  double d2 = 1.0 - d1;
  double zz = 1.5 - d2;
  w[2] = 0.5 * zz * zz;
  if (dw) { dw[0] = -z; dw[1] = -2.0 * d1; dw[2] = zz; }
  *base = b;
}

/* ---------------- simulation state ---------------- */
typedef struct {
  double v[3];
  double m;
  double new_v[3];
  int64_t idx;
} zro_node;

#define ZRO_MAX_COLLIDERS 16

struct zro_sim {
  double dx, one_over_dx, D_inverse, gravity[3], cfl, time;
  int nthreads;
  int64_t n;
  double *X, *V, *C, *F, *mass, *vol, *mu, *lam;
  /* materials: elasticity model and plasticity per particle (NULL: all
     fixed-corotated, no plasticity), plasticity parameters (4 per particle),
     plastic state (VM yield stress / DP logJp), NeoHookeanBorden degradation g */
  int32_t *model, *plastic, *broken;
  double *pparam, *pstate, *gdeg;
  /* phase field (PhaseField.h:13-21) per particle and per node slot */
  int pf_on;
  zro_pf_opts pf_o;
  double *pf_c, *pf_H, *pf_l0, *pf_osc, *pf_Fp, *pf_Hmax, *pf_k, *pf_vol;
  int32_t* pf_allow;
  double *pf_phase, *pf_mult, *pf_rhs, *pf_diag, *pf_x;
  int64_t pf_cap;
  double pf_dt;
  /* ADMM (AdmmSimulationFoam.h:48-70): per particle F, y, omega, last_dv;
     per node slot vn, dv, collision projection, preconditioner diagonal */
  double *admm_F, *admm_y, *admm_omega, *admm_last_dv;
  int64_t admm_n;
  int admm_visco;
  int admm_steps;
  double *admm_vn, *admm_dv, *admm_proj, *admm_diag;
  char* admm_coll;
  int64_t admm_cap;
  double *vtau, *gradV;
  uint64_t* sorter;
  int32_t* order;
  uint64_t* base_offset;
  int64_t ngroups;
  int64_t *gfirst, *glast;
  uint64_t* block_offset;
  int64_t npages, pages_cap;
  uint64_t* pages;             /* page base offsets, first-touch order */
  uint64_t* hkeys;             /* hash: page code + 1 (0 = empty) */
  int64_t* hvals;
  int64_t hcap;
  zro_node* nodes;             /* npages * 32 */
  int64_t num_nodes;
  char* explicit_collided;     /* per node slot */
  /* implicit state: Fn backup, node vectors (3 per node slot), projections */
  double* Fn;
  double *vn, *dv, *proj;      /* proj: 9 per node, P = I - K K^T (identity when free) */
  char* collided;
  double be_dt;
  int ncol;
  zro_collider col[ZRO_MAX_COLLIDERS];
  /* FElasticNonequilibratedBasedMpmForceHelper: Fe_Nonequilibrated and its
     StvkWithHencky parameters per particle, viscosities */
  double *fneq, *mu_neq, *lam_neq;
  double visc_d, visc_v;
};

zro_sim* zro_create(double dx, const double* gravity, double cfl, int nthreads)
{
  zro_sim* s = (zro_sim*)calloc(1, sizeof(zro_sim));
  s->dx = dx;
  s->one_over_dx = 1.0 / dx;
  s->D_inverse = 4.0 / (dx * dx); /* MpmSimulationBase.cpp:114-116 */
  for (int d = 0; d < 3; ++d) s->gravity[d] = gravity ? gravity[d] : 0.0;
  s->cfl = cfl;
  s->nthreads = nthreads > 0 ? nthreads : 1;
  return s;
}

static void free_particles(zro_sim* s)
{
  free(s->X); free(s->V); free(s->C); free(s->F); free(s->mass); free(s->vol);
  free(s->mu); free(s->lam); free(s->vtau); free(s->gradV); free(s->sorter); free(s->Fn);
  s->Fn = NULL;
  free(s->model); free(s->plastic); free(s->broken); free(s->pparam); free(s->pstate); free(s->gdeg);
  s->model = s->plastic = s->broken = NULL;
  s->pparam = s->pstate = s->gdeg = NULL;
  free(s->fneq); free(s->mu_neq); free(s->lam_neq);
  s->fneq = s->mu_neq = s->lam_neq = NULL;
  free(s->pf_c); free(s->pf_H); free(s->pf_l0); free(s->pf_osc); free(s->pf_Fp); free(s->pf_Hmax);
  free(s->pf_k); free(s->pf_vol); free(s->pf_allow);
  s->pf_c = s->pf_H = s->pf_l0 = s->pf_osc = s->pf_Fp = s->pf_Hmax = s->pf_k = s->pf_vol = NULL;
  s->pf_allow = NULL;
  s->pf_on = 0;
  free(s->order); free(s->base_offset); free(s->gfirst); free(s->glast);
  free(s->block_offset);
}

void zro_destroy(zro_sim* s)
{
  if (!s) return;
  free_particles(s);
  free(s->pages); free(s->hkeys); free(s->hvals); free(s->nodes);
  free(s->explicit_collided);
  free(s->vn); free(s->dv); free(s->proj); free(s->collided);
  free(s->pf_phase); free(s->pf_mult); free(s->pf_rhs); free(s->pf_diag); free(s->pf_x);
  free(s->admm_F); free(s->admm_y); free(s->admm_omega); free(s->admm_last_dv);
  free(s->admm_vn); free(s->admm_dv); free(s->admm_proj); free(s->admm_diag); free(s->admm_coll);
  free(s);
}

static double* dup(const double* a, int64_t n)
{
  double* r = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  if (a) memcpy(r, a, sizeof(double) * (size_t)n);
  else memset(r, 0, sizeof(double) * (size_t)(n > 0 ? n : 1));
  return r;
}

void zro_set_particles(zro_sim* s, int64_t n, const double* X, const double* V,
                       const double* C, const double* F, const double* mass,
                       const double* vol, const double* mu, const double* lam)
{
  free_particles(s);
  s->n = n;
  s->X = dup(X, 3 * n); s->V = dup(V, 3 * n); s->C = dup(C, 9 * n); s->F = dup(F, 9 * n);
  s->mass = dup(mass, n); s->vol = dup(vol, n); s->mu = dup(mu, n); s->lam = dup(lam, n);
  s->vtau = dup(NULL, 9 * n); s->gradV = dup(NULL, 9 * n);
  size_t nn = (size_t)(n > 0 ? n : 1);
  s->sorter = (uint64_t*)malloc(sizeof(uint64_t) * nn);
  s->order = (int32_t*)malloc(sizeof(int32_t) * nn);
  s->base_offset = (uint64_t*)malloc(sizeof(uint64_t) * nn);
  s->gfirst = (int64_t*)malloc(sizeof(int64_t) * nn);
  s->glast = (int64_t*)malloc(sizeof(int64_t) * nn);
  s->block_offset = (uint64_t*)malloc(sizeof(uint64_t) * nn);
  s->ngroups = 0;
}

void zro_get_particles(const zro_sim* s, double* X, double* V, double* C, double* F)
{
  if (X) memcpy(X, s->X, sizeof(double) * 3 * (size_t)s->n);
  if (V) memcpy(V, s->V, sizeof(double) * 3 * (size_t)s->n);
  if (C) memcpy(C, s->C, sizeof(double) * 9 * (size_t)s->n);
  if (F) memcpy(F, s->F, sizeof(double) * 9 * (size_t)s->n);
}

int zro_add_collider(zro_sim* s, const zro_collider* c)
{
  if (s->ncol >= ZRO_MAX_COLLIDERS) return -1;
  zro_collider k = *c;
  if (k.shape == 0) { /* HalfSpace normalises its normal (AnalyticLevelSet.cpp:262-266) */
    double l = sqrt(k.normal[0] * k.normal[0] + k.normal[1] * k.normal[1] + k.normal[2] * k.normal[2]);
    for (int d = 0; d < 3; ++d) k.normal[d] /= l;
  }
  s->col[s->ncol++] = k;
  return 0;
}

void zro_set_time(zro_sim* s, double t) { s->time = t; }

/* ---------------- page hash ---------------- */
static int64_t hash_find(const zro_sim* s, uint64_t code)
{
  uint64_t key = code + 1, h = (code * 0x9E3779B97F4A7C15ULL) & (uint64_t)(s->hcap - 1);
  while (s->hkeys[h]) {
    if (s->hkeys[h] == key) return s->hvals[h];
    h = (h + 1) & (uint64_t)(s->hcap - 1);
  }
  return -1;
}
/* SPGrid_Page_Map::Set_Page (SPGrid_Page_Map.h:61-70): first touch appends */
static void set_page(zro_sim* s, uint64_t offset)
{
  uint64_t code = offset >> 12;
  uint64_t key = code + 1, h = (code * 0x9E3779B97F4A7C15ULL) & (uint64_t)(s->hcap - 1);
  while (s->hkeys[h]) {
    if (s->hkeys[h] == key) return;
    h = (h + 1) & (uint64_t)(s->hcap - 1);
  }
  s->hkeys[h] = key;
  s->hvals[h] = s->npages;
  s->pages[s->npages++] = (offset >> 12) << 12;
}

static int cmp_u64(const void* a, const void* b)
{
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y);
}

/* MpmSimulationBase.cpp:1032-1103 */
void zro_sort_and_pollute(zro_sim* s)
{
  const int index_bits = 32 - 5; /* 32 - block_bits */
  int64_t n = s->n;
  double one_over_dx = s->one_over_dx;
#pragma omp parallel for num_threads(s->nthreads)
  for (int64_t i = 0; i < n; ++i) {
    int b[3];
    for (int d = 0; d < 3; ++d) b[d] = int_floor(s->X[3 * i + d] * one_over_dx - 0.5);
    uint64_t off = zro_linear_offset(b[0], b[1], b[2]);
    s->sorter[i] = ((off >> 7) << index_bits) + (uint64_t)i;
  }
  qsort(s->sorter, (size_t)n, sizeof(uint64_t), cmp_u64);

  s->ngroups = 0;
  int64_t last = 0;
  for (int64_t i = 0; i < n; ++i)
    if (i == n - 1 || (s->sorter[i] >> 32) != (s->sorter[i + 1] >> 32)) {
      s->gfirst[s->ngroups] = last;
      s->glast[s->ngroups] = i;
      s->block_offset[s->ngroups] = s->sorter[i] >> 32;
      s->ngroups++;
      last = i + 1;
    }

  /* page map: worst case 8 pages per group */
  int64_t need = 8 * (s->ngroups > 0 ? s->ngroups : 1);
  if (need > s->pages_cap) {
    free(s->pages); free(s->hkeys); free(s->hvals);
    s->pages_cap = need;
    s->pages = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)need);
    s->hcap = 1;
    while (s->hcap < 2 * need) s->hcap <<= 1;
    s->hkeys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)s->hcap);
    s->hvals = (int64_t*)malloc(sizeof(int64_t) * (size_t)s->hcap);
  }
  memset(s->hkeys, 0, sizeof(uint64_t) * (size_t)s->hcap);
  s->npages = 0;
  for (int64_t i = 0; i < n; ++i) {
    s->order[i] = (int32_t)(s->sorter[i] & ((1ULL << index_bits) - 1));
    uint64_t offset = (s->sorter[i] >> index_bits) << 7;
    s->base_offset[s->order[i]] = offset;
    if (i == n - 1 || (s->sorter[i] >> 32) != (s->sorter[i + 1] >> 32)) {
      set_page(s, offset);
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int c = 0; c < 2; ++c)
            set_page(s, zro_packed_add(offset, zro_linear_offset(2 * a, 4 * b, 4 * c)));
    }
  }
  free(s->nodes);
  free(s->explicit_collided);
  s->nodes = (zro_node*)calloc((size_t)(s->npages * 32 + 1), sizeof(zro_node));
  s->explicit_collided = (char*)calloc((size_t)(s->npages * 32 + 1), 1);
  for (int64_t k = 0; k < s->npages * 32; ++k) s->nodes[k].idx = -1;
  s->num_nodes = 0;
}

int64_t zro_num_groups(const zro_sim* s) { return s->ngroups; }
int64_t zro_num_pages(const zro_sim* s) { return s->npages; }
void zro_get_sort(const zro_sim* s, uint64_t* sorter, int32_t* order, uint64_t* base_offset)
{
  if (sorter) memcpy(sorter, s->sorter, sizeof(uint64_t) * (size_t)s->n);
  if (order) memcpy(order, s->order, sizeof(int32_t) * (size_t)s->n);
  if (base_offset) memcpy(base_offset, s->base_offset, sizeof(uint64_t) * (size_t)s->n);
}
void zro_get_groups(const zro_sim* s, int64_t* first, int64_t* last, uint64_t* block_offset)
{
  if (first) memcpy(first, s->gfirst, sizeof(int64_t) * (size_t)s->ngroups);
  if (last) memcpy(last, s->glast, sizeof(int64_t) * (size_t)s->ngroups);
  if (block_offset) memcpy(block_offset, s->block_offset, sizeof(uint64_t) * (size_t)s->ngroups);
}
void zro_get_pages(const zro_sim* s, uint64_t* pages)
{
  memcpy(pages, s->pages, sizeof(uint64_t) * (size_t)s->npages);
}

/* node storage lookup for a node coordinate (grid_array(offset)) */
static zro_node* node_at(zro_sim* s, int i, int j, int k, int64_t* slot_out)
{
  uint64_t off = zro_linear_offset(i, j, k);
  int64_t p = hash_find(s, off >> 12);
  if (p < 0) return NULL;
  int64_t slot = p * 32 + (int64_t)((off >> 7) & 31);
  if (slot_out) *slot_out = slot;
  return &s->nodes[slot];
}

/* ---------------- materials ---------------- */
void zro_set_materials(zro_sim* s, const int32_t* model, const int32_t* plastic, const double* pparam,
                       const double* pstate, const double* g)
{
  int64_t n = s->n;
  size_t nn = (size_t)(n > 0 ? n : 1);
  free(s->model); free(s->plastic); free(s->broken); free(s->pparam); free(s->pstate); free(s->gdeg);
  s->model = (int32_t*)calloc(nn, sizeof(int32_t));
  s->plastic = (int32_t*)calloc(nn, sizeof(int32_t));
  s->broken = (int32_t*)calloc(nn, sizeof(int32_t));
  if (model) memcpy(s->model, model, sizeof(int32_t) * (size_t)n);
  if (plastic) memcpy(s->plastic, plastic, sizeof(int32_t) * (size_t)n);
  s->pparam = dup(pparam, 4 * n);
  s->pstate = dup(pstate, n);
  s->gdeg = dup(NULL, n);
  for (int64_t i = 0; i < n; ++i) s->gdeg[i] = g ? g[i] : 1.0;
}

void zro_set_viscoelastic(zro_sim* s, const double* Fneq, const double* mu, const double* lam, double visc_d,
                          double visc_v)
{
  free(s->fneq); free(s->mu_neq); free(s->lam_neq);
  s->fneq = s->mu_neq = s->lam_neq = NULL;
  if (!mu || !lam) return;
  const int64_t n = s->n;
  s->fneq = dup(Fneq, 9 * n);
  if (!Fneq)
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 9; ++k) s->fneq[9 * i + k] = (k % 4 == 0) ? 1.0 : 0.0;
  s->mu_neq = dup(mu, n);
  s->lam_neq = dup(lam, n);
  s->visc_d = visc_d;
  s->visc_v = visc_v;
}

void zro_get_viscoelastic(const zro_sim* s, double* Fneq)
{
  if (s->fneq && Fneq) memcpy(Fneq, s->fneq, sizeof(double) * 9 * (size_t)s->n);
}

void zro_get_materials(const zro_sim* s, double* pstate, int32_t* broken, double* g)
{
  if (!s->model) return;
  if (pstate) memcpy(pstate, s->pstate, sizeof(double) * (size_t)s->n);
  if (broken) memcpy(broken, s->broken, sizeof(int32_t) * (size_t)s->n);
  if (g) memcpy(g, s->gdeg, sizeof(double) * (size_t)s->n);
}

/* constitutive_model.kirchhoff for the particle's model */
static void particle_kirchhoff(const zro_sim* s, int64_t p, double* tau)
{
  int m = s->model ? s->model[p] : ZRO_MODEL_FIXED_COROTATED;
  if (m == ZRO_MODEL_STVK_HENCKY)
    zro_stvk_hencky_kirchhoff(s->mu[p], s->lam[p], s->F + 9 * p, tau);
  else if (m == ZRO_MODEL_NEO_HOOKEAN_BORDEN)
    zro_nhb_kirchhoff(s->mu[p], s->lam[p], s->gdeg[p], s->F + 9 * p, tau);
  else
    zro_fcr_kirchhoff(s->mu[p], s->lam[p], s->F + 9 * p, tau);
}

/* PlasticityApplier::applyPlasticity for one particle (PlasticityApplier.h:39-56) */
static void apply_plasticity(zro_sim* s, int64_t i)
{
  if (!s->plastic) return;
  const double* q = s->pparam + 4 * i;
  if (s->plastic[i] == ZRO_PLASTIC_VON_MISES) {
    int b = 0;
    zro_von_mises_project(s->mu[i], s->lam[i], q[0], q[1], &s->pstate[i], s->F + 9 * i, &b);
    if (b) s->broken[i] = 1;
  } else if (s->plastic[i] == ZRO_PLASTIC_DRUCKER_PRAGER) {
    zro_drucker_prager_project(s->mu[i], s->lam[i], q[0], q[1], q[2], q[3] != 0.0, &s->pstate[i], s->F + 9 * i);
  } else if (s->plastic[i] == ZRO_PLASTIC_CAM_CLAY) {
    zro_cam_clay_project(s->mu[i], s->lam[i], q[0], q[1], q[2], &s->pstate[i], s->F + 9 * i);
  }
}

/* ---------------- stress: FBasedMpmForceHelper::updateState ---------------- */
void zro_update_stress(zro_sim* s)
{
#pragma omp parallel for num_threads(s->nthreads)
  for (int64_t p = 0; p < s->n; ++p) {
    double tau[9];
    particle_kirchhoff(s, p, tau);
    for (int c = 0; c < 9; ++c) s->vtau[9 * p + c] = s->vol[p] * tau[c];
    if (s->fneq) { /* FElasticNonequilibratedBasedMpmForceHelper::updateState (:36-60) */
      zro_stvk_hencky_kirchhoff(s->mu_neq[p], s->lam_neq[p], s->fneq + 9 * p, tau);
      for (int c = 0; c < 9; ++c) s->vtau[9 * p + c] += s->vol[p] * tau[c];
    }
  }
}

/* ---------------- P2G (MLS, APIC) ---------------- */
static void p2g_particle(zro_sim* s, int i)
{
  const double* Xp = s->X + 3 * i;
  double mass = s->mass[i];
  double mom[3], Cm[3][3], st[3][3];
  for (int d = 0; d < 3; ++d) mom[d] = mass * s->V[3 * i + d];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) {
      Cm[r][c] = mass * s->C[9 * i + c * 3 + r];
      st[r][c] = s->vtau[9 * i + c * 3 + r];
    }
  int base[3];
  double w[3][3];
  for (int d = 0; d < 3; ++d) zro_bspline_weights(Xp[d] * s->one_over_dx, &base[d], w[d], NULL);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double wab = w[0][a] * w[1][b];
      for (int c = 0; c < 3; ++c) {
        double wt = wab * w[2][c];
        int nc[3] = { base[0] + a, base[1] + b, base[2] + c };
        zro_node* g = node_at(s, nc[0], nc[1], nc[2], NULL);
        double dxp[3];
        for (int d = 0; d < 3; ++d) dxp[d] = (double)nc[d] * s->dx - Xp[d];
        g->m += mass * wt;
        for (int r = 0; r < 3; ++r) {
          double vd = Cm[r][0] * dxp[0] + Cm[r][1] * dxp[1] + Cm[r][2] * dxp[2] + mom[r];
          g->v[r] += vd * wt;
          double fd = st[r][0] * dxp[0] + st[r][1] * dxp[1] + st[r][2] * dxp[2];
          g->new_v[r] -= fd * (wt * s->D_inverse);
        }
      }
    }
}

/* getNumNodes (MpmGrid.h:123-136) + normalisation (MpmSimulationBase.cpp:474-485) */
static void number_and_normalise(zro_sim* s)
{
  int64_t total = 0;
  for (int64_t k = 0; k < s->npages * 32; ++k)
    if (s->nodes[k].m != 0.0) s->nodes[k].idx = total++;
  s->num_nodes = total;
  for (int64_t k = 0; k < s->npages * 32; ++k) {
    zro_node* g = &s->nodes[k];
    if (g->idx < 0) continue;
    if (g->m != 0.0) { for (int d = 0; d < 3; ++d) g->v[d] /= g->m; }
    else { for (int d = 0; d < 3; ++d) g->v[d] = 0.0; }
  }
}

int64_t zro_p2g_mls(zro_sim* s)
{
  for (int color = 0; color < 8; ++color) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(s->nthreads)
    for (int64_t g = 0; g < s->ngroups; ++g) {
      if ((int)(s->block_offset[g] & 7) != color) continue;
      for (int64_t idx = s->gfirst[g]; idx <= s->glast[g]; ++idx) p2g_particle(s, s->order[idx]);
    }
  }
  number_and_normalise(s);
  return s->num_nodes;
}

/* ---------------- adaptive time step ---------------- */
/* evalMaxParticleSpeed (MpmSimulationBase.cpp:1143-1170): a map-reduce of
   (|v|, x, -x) under the component-wise max, from (0, lowest, lowest) with
   lowest = numeric_limits<float>::lowest() */
void zro_eval_max_particle_speed(const zro_sim* s, double out[7])
{
  const double lowest = -(double)FLT_MAX;
  out[0] = 0.0;
  for (int k = 1; k < 7; ++k) out[k] = lowest;
  for (int64_t i = 0; i < s->n; ++i) {
    const double* v = s->V + 3 * i;
    const double* x = s->X + 3 * i;
    const double sq = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const double sp = sqrt(sq);
    if (sp > out[0]) out[0] = sp;
    for (int d = 0; d < 3; ++d) {
      if (x[d] > out[1 + d]) out[1 + d] = x[d];
      if (-x[d] > out[4 + d]) out[4 + d] = -x[d];
    }
  }
  /* the min corner is the negated max of -x */
  for (int d = 0; d < 3; ++d) out[4 + d] = -out[4 + d];
}

/* the object's frame at a time t: x = R(t) s(t) X + b(t) */
typedef struct {
  double b[3], R[9], s, inv_s;
} zro_frame;
static void collider_frame(const zro_collider* c, double t, zro_frame* f);
static void collider_velocity(const zro_collider* c, const zro_frame* f, const double* xb, double* vo);

/* calculateDt (MpmSimulationBase.cpp:766-790): the corners widened by
   (degree + 2) dx only bound the colliders' speed queries, which for a
   constant translation is |dbdt| anywhere (CollisionObject.cpp:158-194) */
double zro_calculate_dt(const zro_sim* s, double max_dt)
{
  double r[7];
  zro_eval_max_particle_speed(s, r);
  double speed = r[0];
  /* AnalyticCollisionObject::evalMaxSpeed (CollisionObject.cpp:158-194) over
     the particle box widened by (degree + 2) dx (MpmSimulationBase.cpp:774-781):
     with spin or scaling, the object's velocity at the corners of either box
     that the other box's "any axis" test keeps; a sphere's box is its centre
     +- radius, a half-space has none (-1) */
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) { hi[d] = r[1 + d] + 4.0 * s->dx; lo[d] = r[4 + d] - 4.0 * s->dx; }
  for (int k = 0; k < s->ncol; ++k) {
    const zro_collider* c = &s->col[k];
    const double* u = c->velocity;
    double cs = 0.0;
    if (c->dscale_dt != 0.0 || c->omega[0] != 0.0 || c->omega[1] != 0.0 || c->omega[2] != 0.0) {
      if (c->shape != 1) return -1.0;
      double mn[3], mx[3], pts[16][3];
      int np = 0;
      zro_frame fr;
      collider_frame(c, s->time, &fr);
      const double *R = fr.R, *b = fr.b;
      double sc = fr.s, inv_s = fr.inv_s;
      for (int d = 0; d < 3; ++d) { mn[d] = c->origin[d] - c->radius; mx[d] = c->origin[d] + c->radius; }
      for (int i = 0; i < 8; ++i) {
        double x[3], X[3];
        for (int d = 0; d < 3; ++d) x[d] = (i & (1 << d)) ? lo[d] : hi[d];
        for (int d = 0; d < 3; ++d)
          X[d] = (R[d] * (x[0] - b[0]) + R[3 + d] * (x[1] - b[1]) + R[6 + d] * (x[2] - b[2])) * inv_s;
        int a1 = mn[0] < X[0] || mn[1] < X[1] || mn[2] < X[2];
        int a2 = X[0] < mx[0] || X[1] < mx[1] || X[2] < mx[2];
        if (a1 && a2) { memcpy(pts[np], x, sizeof(x)); ++np; }
      }
      for (int i = 0; i < 8; ++i) {
        double X[3], x[3];
        for (int d = 0; d < 3; ++d) X[d] = (i & (1 << d)) ? mn[d] : mx[d];
        for (int d = 0; d < 3; ++d) x[d] = sc * (R[3 * d] * X[0] + R[3 * d + 1] * X[1] + R[3 * d + 2] * X[2]) + b[d];
        int a1 = lo[0] < x[0] || lo[1] < x[1] || lo[2] < x[2];
        int a2 = x[0] < hi[0] || x[1] < hi[1] || x[2] < hi[2];
        if (a1 && a2) { memcpy(pts[np], x, sizeof(x)); ++np; }
      }
      for (int i = 0; i < np; ++i) {
        double xb[3], v[3];
        for (int d = 0; d < 3; ++d) xb[d] = pts[i][d] - b[d];
        collider_velocity(c, &fr, xb, v);
        double sp = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        if (sp > cs) cs = sp;
      }
    } else {
      cs = sqrt((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
    }
    if (cs > speed) speed = cs;
  }
  double dt = max_dt;
  if (speed != 0.0) dt = s->cfl * s->dx / speed;
  return dt;
}

/* TimeStepping::nextDt (TimeStepping.h:45-61) */
double zro_next_dt(double dt, double min_dt, double max_dt, double max_dt_original, double frame_dt,
                   double time_since_last_frame)
{
  if (max_dt_original < max_dt) max_dt = max_dt_original;
  dt = (dt < min_dt) ? min_dt : (dt > max_dt) ? max_dt : dt;
  if (time_since_last_frame + dt >= frame_dt)
    dt = frame_dt - time_since_last_frame;
  else if (time_since_last_frame + 2 * dt > frame_dt)
    dt = (frame_dt - time_since_last_frame) / 2;
  return dt;
}

/* ---------------- colliders ---------------- */
/* R (row-major) from the quaternion <w,x,y,z>, Eigen's toRotationMatrix
   (Rotation.h:53-57): R = I + 2w[q]x + 2[q]x^2 written entrywise */
static void quat_to_matrix(const double q[4], double R[9])
{
  double w = q[0], x = q[1], y = q[2], z = q[3];
  double tx = 2 * x, ty = 2 * y, tz = 2 * z;
  R[0] = 1 - (ty * y + tz * z); R[1] = ty * x - tz * w;       R[2] = tz * x + ty * w;
  R[3] = ty * x + tz * w;       R[4] = 1 - (tx * x + tz * z); R[5] = tz * y - tx * w;
  R[6] = tz * x - ty * w;       R[7] = tz * y + tx * w;       R[8] = 1 - (tx * x + ty * y);
}

/* The state updateState(t) sets for an object moving at constant rates from
   its time-0 state (CollisionObject.h:81-172; the reference's scene callbacks
   write the same form, e.g. AdmmInit3D.h:298-313: rotation
   (cos(theta t/2), 0, sin(theta t/2), 0) with omega (0, theta, 0)):
   b(t) = translation + velocity t, s(t) = s0 + ds/dt t, and
   R(t) = exp(t [omega]x) R0, as the quaternion product
   (cos(|omega| t/2), sin(|omega| t/2) omega/|omega|) q0.  A zero quaternion
   means the identity, scale 0 means 1. */
static void collider_frame(const zro_collider* c, double t, zro_frame* f)
{
  double q0[4] = { c->rotation[0], c->rotation[1], c->rotation[2], c->rotation[3] };
  if (q0[0] == 0.0 && q0[1] == 0.0 && q0[2] == 0.0 && q0[3] == 0.0) q0[0] = 1.0;
  const double* w = c->omega;
  double wn = sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  double q[4] = { q0[0], q0[1], q0[2], q0[3] };
  if (wn != 0.0 && t != 0.0) {
    double h = 0.5 * wn * t, sh = sin(h) / wn;
    double a[4] = { cos(h), sh * w[0], sh * w[1], sh * w[2] };
    q[0] = a[0] * q0[0] - a[1] * q0[1] - a[2] * q0[2] - a[3] * q0[3];
    q[1] = a[0] * q0[1] + a[1] * q0[0] + a[2] * q0[3] - a[3] * q0[2];
    q[2] = a[0] * q0[2] - a[1] * q0[3] + a[2] * q0[0] + a[3] * q0[1];
    q[3] = a[0] * q0[3] + a[1] * q0[2] - a[2] * q0[1] + a[3] * q0[0];
  }
  quat_to_matrix(q, f->R);
  for (int d = 0; d < 3; ++d) f->b[d] = c->translation[d] + c->velocity[d] * t;
  f->s = (c->scale == 0.0 ? 1.0 : c->scale) + c->dscale_dt * t;
  f->inv_s = 1.0 / f->s;
}

void zro_collider_frame(const zro_collider* c, double t, double* b, double* R, double* s)
{
  zro_frame f;
  collider_frame(c, t, &f);
  memcpy(b, f.b, sizeof(f.b));
  memcpy(R, f.R, sizeof(f.R));
  *s = f.s;
}

/* level-set queryInside in material space (AnalyticLevelSet.cpp:111-119, 435-450) */
static int ls_query_inside(const zro_collider* c, const double* X, double* phi, double* n)
{
  if (c->shape == 0) {
    double sd = 0.0;
    for (int d = 0; d < 3; ++d) sd += c->normal[d] * (X[d] - c->origin[d]);
    *phi = sd;
    if (sd <= 0.0) { for (int d = 0; d < 3; ++d) n[d] = c->normal[d]; return 1; }
    return 0;
  } else {
    double tc[3], d2 = 0.0;
    for (int d = 0; d < 3; ++d) { tc[d] = X[d] - c->origin[d]; d2 += tc[d] * tc[d]; }
    double r2 = c->radius * c->radius;
    if (d2 < r2) {
      double dist = sqrt(d2);
      *phi = dist - c->radius;
      if (dist < 1e-7) { n[0] = 1.0; n[1] = 0.0; n[2] = 0.0; }
      else for (int d = 0; d < 3; ++d) n[d] = tc[d] / dist;
      return 1;
    }
    return 0;
  }
}

/* AnalyticCollisionObject::queryInside in world space (CollisionObject.cpp:
   279-296): xb = x - b, X = R^T xb / s; inside: n = R N, phi scaled by s */
static int collider_query(const zro_collider* c, const zro_frame* f, const double* x, double* phi, double* n,
                          double* xb)
{
  double X[3], N[3];
  const double* R = f->R;
  for (int d = 0; d < 3; ++d) xb[d] = x[d] - f->b[d];
  for (int d = 0; d < 3; ++d) X[d] = (R[d] * xb[0] + R[3 + d] * xb[1] + R[6 + d] * xb[2]) * f->inv_s;
  if (!ls_query_inside(c, X, phi, N)) return 0;
  for (int d = 0; d < 3; ++d) n[d] = R[3 * d] * N[0] + R[3 * d + 1] * N[1] + R[3 * d + 2] * N[2];
  *phi *= f->s;
  return 1;
}

/* the object's velocity at x: omega x (x - b) + (ds/dt / s)(x - b) + db/dt */
static void collider_velocity(const zro_collider* c, const zro_frame* f, const double* xb, double* vo)
{
  double g = c->dscale_dt * f->inv_s;
  const double* w = c->omega;
  vo[0] = (w[1] * xb[2] - w[2] * xb[1]) + g * xb[0] + c->velocity[0];
  vo[1] = (w[2] * xb[0] - w[0] * xb[2]) + g * xb[1] + c->velocity[1];
  vo[2] = (w[0] * xb[1] - w[1] * xb[0]) + g * xb[2] + c->velocity[2];
}

/* AnalyticCollisionObject::detectAndResolveCollision (CollisionObject.cpp:340-405) */
static int detect_and_resolve(const zro_collider* c, const zro_frame* f, const double* x, double* v)
{
  double xb[3], phi, N[3], vo[3];
  if (!collider_query(c, f, x, &phi, N, xb)) return 0;
  collider_velocity(c, f, xb, vo);
  for (int d = 0; d < 3; ++d) v[d] -= vo[d];
  if (c->type == 1) {
    v[0] = v[1] = v[2] = 0.0;
  } else if (c->type == 2 || c->type == 3) {
    double dot = v[0] * N[0] + v[1] * N[1] + v[2] * N[2];
    if (c->type == 2 || dot < 0) {
      for (int d = 0; d < 3; ++d) v[d] -= N[d] * dot;
      if (c->friction != 0.0 && dot < 0) {
        double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (-dot * c->friction < vn) {
          for (int d = 0; d < 3; ++d) v[d] += v[d] / vn * dot * c->friction;
        } else {
          v[0] = v[1] = v[2] = 0.0;
        }
      }
    }
  }
  for (int d = 0; d < 3; ++d) v[d] += vo[d];
  return 1;
}

/* multiObjectCollision (CollisionObject.cpp:66-105); the normal basis is only
   needed by the implicit path */
static int multi_object_collision(const zro_sim* s, double t, const double* x, double* v)
{
  int any = 0, nslip = 0;
  double basis[3][3];
  for (int k = 0; k < s->ncol; ++k) {
    double n[3] = { 0, 0, 0 };
    zro_frame fr;
    collider_frame(&s->col[k], t, &fr);
    int hit = detect_and_resolve(&s->col[k], &fr, x, v);
    any = hit || any;
    if (hit) {
      if (s->col[k].type == 1) break;
      double xb[3], phi;
      collider_query(&s->col[k], &fr, x, &phi, n, xb);
      for (int c2 = 0; c2 < nslip; ++c2) {
        double dot = basis[c2][0] * n[0] + basis[c2][1] * n[1] + basis[c2][2] * n[2];
        for (int d = 0; d < 3; ++d) n[d] -= dot * basis[c2][d];
      }
      double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (l != 0.0) {
        for (int d = 0; d < 3; ++d) basis[nslip][d] = n[d] / l;
        if (++nslip == 3) break;
      }
    }
  }
  return any;
}

/* trialCollision (CollisionObject.cpp:48-64) */
static void trial_collision(const zro_sim* s, double t, const double* x, double dt, double* v)
{
  for (int k = 0; k < s->ncol; ++k) {
    double xt[3], xb[3], phi, n[3];
    zro_frame fr;
    collider_frame(&s->col[k], t, &fr);
    for (int d = 0; d < 3; ++d) xt[d] = x[d] + dt * v[d];
    if (collider_query(&s->col[k], &fr, xt, &phi, n, xb))
      for (int d = 0; d < 3; ++d) v[d] -= (phi / dt) * n[d];
  }
}

/* gridVelocityExplicitUpdate (MpmSimulationBase.cpp:818-865), collider
   updateState(t+dt) and trialCollision (MpmSimulationBase.cpp:267-271,298-312) */
void zro_grid_update(zro_sim* s, double dt)
{
  double dvg[3];
  for (int d = 0; d < 3; ++d) dvg[d] = s->gravity[d] * dt;
#pragma omp parallel for num_threads(s->nthreads)
  for (int64_t p = 0; p < s->npages; ++p) {
    int base[3];
    zro_linear_to_coord(s->pages[p], base);
    for (int e = 0; e < 32; ++e) {
      int64_t slot = p * 32 + e;
      zro_node* g = &s->nodes[slot];
      if (g->idx < 0) continue;
      double xi[3] = { (double)(base[0] + ((e >> 4) & 1)) * s->dx, (double)(base[1] + ((e >> 2) & 3)) * s->dx,
                       (double)(base[2] + (e & 3)) * s->dx };
      double dt_over_mi = dt / g->m, vi[3];
      for (int d = 0; d < 3; ++d) vi[d] = g->v[d] + dvg[d] + g->new_v[d] * dt_over_mi;
      s->explicit_collided[slot] = (char)multi_object_collision(s, s->time, xi, vi);
      for (int d = 0; d < 3; ++d) g->new_v[d] = vi[d];
    }
  }
  /* trial collision with colliders at t + dt */
#pragma omp parallel for num_threads(s->nthreads)
  for (int64_t p = 0; p < s->npages; ++p) {
    int base[3];
    zro_linear_to_coord(s->pages[p], base);
    for (int e = 0; e < 32; ++e) {
      int64_t slot = p * 32 + e;
      zro_node* g = &s->nodes[slot];
      if (g->idx < 0 || s->explicit_collided[slot]) continue;
      double xi[3] = { (double)(base[0] + ((e >> 4) & 1)) * s->dx, (double)(base[1] + ((e >> 2) & 3)) * s->dx,
                       (double)(base[2] + (e & 3)) * s->dx };
      trial_collision(s, s->time + dt, xi, dt, g->new_v);
    }
  }
}

/* gridToParticlesHelper<APIC,deg2,MLS> (MpmSimulationBase.cpp:906-1016) +
   evolveStrain (FBasedMpmForceHelper.cpp:95-110).  Returns bit0 = a particle
   moved more than one cell, bit1 = more than half a cell scaled by cfl. */
int zro_g2p_mls(zro_sim* s, double dt)
{
  int flags = 0;
  double dx2 = s->dx * s->dx;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : flags) num_threads(s->nthreads)
  for (int64_t g = 0; g < s->ngroups; ++g) {
    for (int64_t idx = s->gfirst[g]; idx <= s->glast[g]; ++idx) {
      int i = s->order[idx];
      double* Xp = s->X + 3 * i;
      int base[3];
      double w[3][3];
      for (int d = 0; d < 3; ++d) zro_bspline_weights(Xp[d] * s->one_over_dx, &base[d], w[d], NULL);
      double picV[3] = { 0, 0, 0 }, B[3][3] = { { 0 } };
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double wab = w[0][a] * w[1][b];
          for (int c = 0; c < 3; ++c) {
            double wt = wab * w[2][c];
            int nc[3] = { base[0] + a, base[1] + b, base[2] + c };
            zro_node* gn = node_at(s, nc[0], nc[1], nc[2], NULL);
            double dxp[3];
            for (int d = 0; d < 3; ++d) dxp[d] = (double)nc[d] * s->dx - Xp[d];
            for (int r = 0; r < 3; ++r) {
              picV[r] += wt * gn->new_v[r];
              for (int cc = 0; cc < 3; ++cc) B[r][cc] += wt * gn->new_v[r] * dxp[cc];
            }
          }
        }
      for (int d = 0; d < 3; ++d) s->V[3 * i + d] = picV[d];
      double Cn[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          Cn[r][c] = B[r][c] * s->D_inverse;
          s->C[9 * i + c * 3 + r] = Cn[r][c];
          s->gradV[9 * i + c * 3 + r] = Cn[r][c];
        }
      double inc2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        double inc = dt * picV[d];
        Xp[d] += inc;
        inc2 += inc * inc;
      }
      if (inc2 > dx2) flags |= 1;
      if (inc2 > dx2 * 0.25 * (s->cfl * s->cfl)) flags |= 2;
      /* F = (I + dt gradV) F */
      double Fo[3][3], Fn[3][3];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) Fo[r][c] = s->F[9 * i + c * 3 + r];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = Fo[r][c];
          for (int k = 0; k < 3; ++k) acc += dt * Cn[r][k] * Fo[k][c];
          Fn[r][c] = acc;
        }
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) s->F[9 * i + c * 3 + r] = Fn[r][c];
      if (s->fneq) { /* the viscoelastic helper's evolveStrain with gradV = C */
        double g9[9];
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r) g9[c * 3 + r] = Cn[r][c];
        zro_visco_relax(s->mu_neq[i], s->lam_neq[i], s->visc_d, s->visc_v, dt, g9, s->fneq + 9 * i);
      }
      apply_plasticity(s, i); /* MpmSimulationBase.cpp:1015 */
    }
  }
  return flags;
}

/* advanceOneTimeStep, symplectic branch (MpmSimulationBase.cpp:256-281) */
int zro_step_explicit(zro_sim* s, double dt)
{
  zro_delete_particles(s);
  zro_sort_and_pollute(s);
  zro_update_stress(s);
  zro_p2g_mls(s);
  if (s->pf_on) zro_phase_field_solve(s, dt, NULL); /* FractureSimulation::particlesToGrid (:149-160) */
  zro_grid_update(s, dt);
  int flags = zro_g2p_mls(s, dt);
  s->time += dt;
  return flags;
}

void zro_get_grid(const zro_sim* s, double* m, double* v, double* new_v, int64_t* idx)
{
  for (int64_t k = 0; k < s->npages * 32; ++k) {
    const zro_node* g = &s->nodes[k];
    if (m) m[k] = g->m;
    for (int d = 0; d < 3; ++d) {
      if (v) v[3 * k + d] = g->v[d];
      if (new_v) new_v[3 * k + d] = g->new_v[d];
    }
    if (idx) idx[k] = g->idx;
  }
}

/* ====================================================================== */
/* implicit backward Euler (APIC transfers)                               */
/* ====================================================================== */

/* weights and weight gradients in iterateKernel order (MpmGrid.h:231-268) */
typedef struct {
  int base[3];
  double w[3][3], dw[3][3];
} zro_spline;

static void spline_of(const zro_sim* s, const double* Xp, zro_spline* sp)
{
  for (int d = 0; d < 3; ++d) zro_bspline_weights(Xp[d] * s->one_over_dx, &sp->base[d], sp->w[d], sp->dw[d]);
}

/* w_abc and grad w_abc exactly as iterateKernel forms them */
static double spline_node(const zro_sim* s, const zro_spline* sp, int a, int b, int c, double* gw)
{
  double wi = sp->w[0][a], wj = sp->w[1][b], wk = sp->w[2][c];
  double dwidxi = s->one_over_dx * sp->dw[0][a];
  double wij = wi * wj;
  double dwijdxi = dwidxi * wj;
  double dwijdxj = wi * s->one_over_dx * sp->dw[1][b];
  gw[0] = dwijdxi * wk;
  gw[1] = dwijdxj * wk;
  gw[2] = wij * s->one_over_dx * sp->dw[2][c];
  return wij * wk;
}

static int64_t node_slot(zro_sim* s, int i, int j, int k)
{
  int64_t slot = -1;
  node_at(s, i, j, k, &slot);
  return slot;
}

static void p2g_apic_particle(zro_sim* s, int i)
{
  const double* Xp = s->X + 3 * i;
  double mass = s->mass[i], mom[3], Cm[3][3];
  for (int d = 0; d < 3; ++d) mom[d] = mass * s->V[3 * i + d];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) Cm[r][c] = mass * s->C[9 * i + c * 3 + r];
  zro_spline sp;
  spline_of(s, Xp, &sp);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double gw[3];
        double wt = spline_node(s, &sp, a, b, c, gw);
        int nc[3] = { sp.base[0] + a, sp.base[1] + b, sp.base[2] + c };
        zro_node* g = node_at(s, nc[0], nc[1], nc[2], NULL);
        double dxp[3];
        for (int d = 0; d < 3; ++d) dxp[d] = (double)nc[d] * s->dx - Xp[d];
        g->m += mass * wt;
        for (int r = 0; r < 3; ++r)
          g->v[r] += (Cm[r][0] * dxp[0] + Cm[r][1] * dxp[1] + Cm[r][2] * dxp[2] + mom[r]) * wt;
      }
}

int64_t zro_p2g_apic(zro_sim* s)
{
  for (int color = 0; color < 8; ++color) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(s->nthreads)
    for (int64_t g = 0; g < s->ngroups; ++g) {
      if ((int)(s->block_offset[g] & 7) != color) continue;
      for (int64_t idx = s->gfirst[g]; idx <= s->glast[g]; ++idx) p2g_apic_particle(s, s->order[idx]);
    }
  }
  number_and_normalise(s);
  return s->num_nodes;
}

/* multiObjectCollision with its normal basis (CollisionObject.cpp:66-105);
   returns the projection P = I - K K^T of CollisionNode (CollisionObject.h:17-40) */
static int collision_projection(const zro_sim* s, double t, const double* x, double* v, double P[3][3])
{
  int any = 0, nslip = 0, sticky = 0;
  double K[3][3] = { { 0 } };
  for (int k = 0; k < s->ncol; ++k) {
    zro_frame fr;
    collider_frame(&s->col[k], t, &fr);
    int hit = detect_and_resolve(&s->col[k], &fr, x, v);
    any = hit || any;
    if (!hit) continue;
    if (s->col[k].type == 1) { sticky = 1; break; }
    double xb[3], phi, n[3];
    collider_query(&s->col[k], &fr, x, &phi, n, xb);
    for (int c2 = 0; c2 < nslip; ++c2) {
      double dot = K[c2][0] * n[0] + K[c2][1] * n[1] + K[c2][2] * n[2];
      for (int d = 0; d < 3; ++d) n[d] -= dot * K[c2][d];
    }
    double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l != 0.0) {
      for (int d = 0; d < 3; ++d) K[nslip][d] = n[d] / l;
      if (++nslip == 3) break;
    }
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double kk = 0.0;
      for (int q = 0; q < nslip; ++q) kk += K[q][r] * K[q][c];
      P[r][c] = sticky ? 0.0 : ((r == c ? 1.0 : 0.0) - kk);
    }
  return any;
}

static void project_vec(const zro_sim* s, double* v)
{
  int64_t nn = s->npages * 32;
  for (int64_t k = 0; k < nn; ++k) {
    if (!s->collided[k]) continue;
    const double* P = s->proj + 9 * k;
    double a[3] = { v[3 * k], v[3 * k + 1], v[3 * k + 2] };
    for (int r = 0; r < 3; ++r) v[3 * k + r] = P[3 * r] * a[0] + P[3 * r + 1] * a[1] + P[3 * r + 2] * a[2];
  }
}

static int node_active(const zro_sim* s, int64_t k) { return s->nodes[k].idx >= 0; }

static double dot_active(const zro_sim* s, const double* a, const double* b)
{
  double acc = 0.0;
  int64_t nn = s->npages * 32;
  for (int64_t k = 0; k < nn; ++k)
    if (node_active(s, k)) acc += a[3 * k] * b[3 * k] + a[3 * k + 1] * b[3 * k + 1] + a[3 * k + 2] * b[3 * k + 2];
  return acc;
}

/* computeNorm (BackwardEuler.h:127-140) */
static double mass_norm(const zro_sim* s, const double* r)
{
  double acc = 0.0;
  int64_t nn = s->npages * 32;
  for (int64_t k = 0; k < nn; ++k)
    if (node_active(s, k))
      acc += (r[3 * k] * r[3 * k] + r[3 * k + 1] * r[3 * k + 1] + r[3 * k + 2] * r[3 * k + 2]) / s->nodes[k].m;
  return sqrt(acc);
}

/* grad f at particle i for a node field f: sum f_i grad w_ip^T (MpmForceBase.cpp:212-244) */
static void particle_gradient(zro_sim* s, int i, const double* f, double G[3][3])
{
  zro_spline sp;
  spline_of(s, s->X + 3 * i, &sp);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) G[r][c] = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double gw[3];
        spline_node(s, &sp, a, b, c, gw);
        int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
        for (int r = 0; r < 3; ++r)
          for (int q = 0; q < 3; ++q) G[r][q] += f[3 * k + r] * gw[q];
      }
}

/* out_i -= scale * S grad w_ip  (rasterizeForceToTVStack<false>, MpmForceBase.cpp:100-152) */
static void particle_scatter(zro_sim* s, int i, double S[3][3], double scale, double* out)
{
  zro_spline sp;
  spline_of(s, s->X + 3 * i, &sp);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double gw[3];
        spline_node(s, &sp, a, b, c, gw);
        int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
        for (int r = 0; r < 3; ++r) out[3 * k + r] -= scale * (S[r][0] * gw[0] + S[r][1] * gw[1] + S[r][2] * gw[2]);
      }
}

static void mat_from_cm(const double* a, double m[3][3])
{
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) m[r][c] = a[c * 3 + r];
}
static void mat_to_cm(double m[3][3], double* a)
{
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) a[c * 3 + r] = m[r][c];
}

/* updateState: F = (I + dt grad(vn + dv)) Fn, vPFnT = V0 P(F) Fn^T
   (MpmForceBase.cpp:305-344, FBasedMpmForceHelper.cpp:68-91) */
static void implicit_update_state(zro_sim* s, const double* dv)
{
  int64_t nn = s->npages * 32;
  double* vel = (double*)malloc(sizeof(double) * 3 * (size_t)nn);
  for (int64_t k = 0; k < 3 * nn; ++k) vel[k] = s->vn[k] + dv[k];
  double dt = s->be_dt;
#pragma omp parallel for num_threads(s->nthreads)
  for (int64_t i = 0; i < s->n; ++i) {
    double G[3][3], Fn[3][3], F[3][3], P[3][3], vp[3][3];
    particle_gradient(s, (int)i, vel, G);
    mat_from_cm(s->Fn + 9 * i, Fn);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = Fn[r][c];
        for (int k = 0; k < 3; ++k) acc += dt * G[r][k] * Fn[k][c];
        F[r][c] = acc;
      }
    mat_to_cm(F, s->F + 9 * i);
    double Pcm[9];
    zro_fcr_first_piola(s->mu[i], s->lam[i], s->F + 9 * i, Pcm);
    mat_from_cm(Pcm, P);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += P[r][k] * Fn[c][k];
        vp[r][c] = s->vol[i] * acc;
      }
    mat_to_cm(vp, s->vtau + 9 * i);
  }
  free(vel);
}

static void scatter_colored(zro_sim* s, const double* stress_cm, double scale, double* out)
{
  for (int color = 0; color < 8; ++color) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(s->nthreads)
    for (int64_t g = 0; g < s->ngroups; ++g) {
      if ((int)(s->block_offset[g] & 7) != color) continue;
      for (int64_t idx = s->gfirst[g]; idx <= s->glast[g]; ++idx) {
        int i = s->order[idx];
        double S[3][3];
        mat_from_cm(stress_cm + 9 * (int64_t)i, S);
        particle_scatter(s, i, S, scale, out);
      }
    }
  }
}

/* computeResidual (BackwardEuler.h:101-124) with MassLumpedInertia
   (Inertia.cpp:33-43) at the state left by implicit_update_state */
static void implicit_residual_at_state(zro_sim* s, const double* dv, double* r)
{
  int64_t nn = s->npages * 32;
  double dt = s->be_dt;
  for (int64_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d) r[3 * k + d] = node_active(s, k) ? dt * s->gravity[d] * s->nodes[k].m : 0.0;
  scatter_colored(s, s->vtau, dt, r);
  for (int64_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d) {
      if (node_active(s, k)) r[3 * k + d] -= s->nodes[k].m * dv[3 * k + d];
      else r[3 * k + d] = 0.0;
    }
  project_vec(s, r);
}

void zro_implicit_residual(zro_sim* s, const double* dv, double* r)
{
  implicit_update_state(s, dv);
  implicit_residual_at_state(s, dv, r);
}

/* multiply (BackwardEuler.h:322-351): M x + dt^2 sum V0 dP(grad x Fn) Fn^T grad w */
void zro_implicit_multiply(zro_sim* s, const double* x, double* b)
{
  int64_t nn = s->npages * 32;
  double dt = s->be_dt;
  double* ds = (double*)malloc(sizeof(double) * 9 * (size_t)(s->n > 0 ? s->n : 1));
#pragma omp parallel for num_threads(s->nthreads)
  for (int64_t i = 0; i < s->n; ++i) {
    double G[3][3], Fn[3][3], dF[3][3], dP[3][3], out[3][3], dFcm[9], dPcm[9];
    particle_gradient(s, (int)i, x, G);
    mat_from_cm(s->Fn + 9 * i, Fn);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += G[r][k] * Fn[k][c];
        dF[r][c] = acc;
      }
    mat_to_cm(dF, dFcm);
    zro_fcr_first_piola_differential(s->mu[i], s->lam[i], s->F + 9 * i, dFcm, dPcm);
    mat_from_cm(dPcm, dP);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) acc += dP[r][k] * Fn[c][k];
        out[r][c] = acc * s->vol[i];
      }
    mat_to_cm(out, ds + 9 * i);
  }
  for (int64_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d) b[3 * k + d] = node_active(s, k) ? s->nodes[k].m * x[3 * k + d] : 0.0;
  scatter_colored(s, ds, -dt * dt, b);
  free(ds);
}

/* GivensRotation::compute / rowRotation on a 2-vector (Givens.h:56-141) */
typedef struct { double c, s; } zro_g2;
static zro_g2 g2_make(void) { zro_g2 g = { 1.0, 0.0 }; return g; }
static void g2_compute(zro_g2* g, double a, double b)
{
  double d = a * a + b * b;
  g->c = 1.0;
  g->s = 0.0;
  double sqrtd = sqrt(d);
  if (sqrtd) {
    double t = 1.0 / sqrtd;
    g->c = a * t;
    g->s = -b * t;
  }
}
static void g2_rot(const zro_g2* g, double* v)
{
  double t1 = v[0], t2 = v[1];
  v[0] = g->c * t1 - g->s * t2;
  v[1] = g->s * t1 + g->c * t2;
}

/* Minres::solve (Minres.h:69-178) on the backward-Euler operator for the
   Newton step: x starts at zero (computeStep, BackwardEuler.h:202), the
   residual is projected (BackwardEuler::project) and preconditioned with
   in / m or the Jacobi inverse diagonal; LinearSolver::dotProduct
   (LinearSolver.h:34-37).  Returns the iteration count as Minres::solve does. */
static int minres_solve(zro_sim* s, double* x, const double* b, double rel_tol, const zro_implicit_opts* o,
                        const double* inv_diag)
{
  int64_t nn = s->npages * 32, n3 = 3 * nn;
  size_t bytes = sizeof(double) * (size_t)n3;
  double *mk = calloc(n3, 8), *mkm1 = calloc(n3, 8), *mkm2 = calloc(n3, 8), *z = calloc(n3, 8);
  double *qkm1 = calloc(n3, 8), *qk = calloc(n3, 8), *qkp1 = calloc(n3, 8), *tmp;
  zro_g2 Gk = g2_make(), Gkm1 = g2_make(), Gkm2 = g2_make();
  double gamma = 0, delta = 0, epsilon = 0, beta_kp1 = 0, alpha_k = 0, beta_k = 0, tk = 0, lt[2];
#define PRECOND(dst, src)                                                               \
  for (int64_t k = 0; k < nn; ++k)                                                      \
    for (int d = 0; d < 3; ++d)                                                         \
      dst[3 * k + d] = !node_active(s, k) ? 0.0                                         \
                     : inv_diag ? src[3 * k + d] * inv_diag[3 * k + d]                  \
                                : src[3 * k + d] / s->nodes[k].m;
  for (int64_t k = 0; k < n3; ++k) x[k] = 0.0;
  for (int64_t k = 0; k < n3; ++k) qkp1[k] = b[k];
  project_vec(s, qkp1);
  PRECOND(z, qkp1);
  double res = sqrt(dot_active(s, z, qkp1));
  beta_kp1 = res;
  double local_tol = rel_tol * res < o->cg_tolerance ? rel_tol * res : o->cg_tolerance;
  int it = 0;
  if (res < local_tol) goto done;
  if (res > 0)
    for (int64_t k = 0; k < n3; ++k) { qkp1[k] /= beta_kp1; z[k] /= beta_kp1; }
  lt[0] = res;
  lt[1] = 0.0;
  for (it = 0; it < o->cg_max_iterations; ++it) {
    if (res < local_tol) break;
    tmp = mkm2; mkm2 = mkm1; mkm1 = mk; mk = tmp;
    memcpy(mk, z, bytes);
    beta_k = beta_kp1;
    tmp = qkm1; qkm1 = qk; qk = qkp1; qkp1 = tmp;
    zro_implicit_multiply(s, mk, qkp1);
    project_vec(s, qkp1);
    alpha_k = dot_active(s, mk, qkp1);
    for (int64_t k = 0; k < n3; ++k) qkp1[k] -= alpha_k * qk[k];
    for (int64_t k = 0; k < n3; ++k) qkp1[k] -= beta_k * qkm1[k];
    PRECOND(z, qkp1);
    double bb = dot_active(s, z, qkp1);
    beta_kp1 = sqrt(bb > 0.0 ? bb : 0.0);
    if (beta_kp1 > 0)
      for (int64_t k = 0; k < n3; ++k) { qkp1[k] /= beta_kp1; z[k] /= beta_kp1; }
    /* applyAllPreviousGivensRotationsAndDetermineNewGivens (Minres.h:151-178) */
    Gkm2 = Gkm1;
    Gkm1 = Gk;
    double ek[2] = { 0.0, beta_k };
    g2_rot(&Gkm2, ek);
    epsilon = ek[0];
    double dk[2] = { ek[1], alpha_k };
    g2_rot(&Gkm1, dk);
    delta = dk[0];
    double tv[2] = { dk[1], beta_kp1 };
    g2_compute(&Gk, tv[0], tv[1]);
    g2_rot(&Gk, tv);
    gamma = tv[0];
    g2_rot(&Gk, lt);
    tk = lt[0];
    double r = lt[1];
    lt[0] = r;
    lt[1] = 0.0;
    res = r < 0 ? -r : r;
    for (int64_t k = 0; k < n3; ++k) mk[k] = (mk[k] - delta * mkm1[k] - epsilon * mkm2[k]) / gamma;
    for (int64_t k = 0; k < n3; ++k) x[k] += tk * mk[k];
  }
done:
#undef PRECOND
  free(mk); free(mkm1); free(mkm2); free(z); free(qkm1); free(qk); free(qkp1);
  return it;
}

/* ConjugateGradient::solve (ConjugateGradient.h:99-147), mass-lumped
   preconditioner out = in / m (MpmSimulationBase.cpp:617-623) */
static int cg_solve(zro_sim* s, double* x, const double* b, double rel_tol, const zro_implicit_opts* o,
                    const double* inv_diag)
{
  int64_t nn = s->npages * 32, n3 = 3 * nn;
  double* r = (double*)calloc((size_t)n3, sizeof(double));
  double* q = (double*)calloc((size_t)n3, sizeof(double));
  double* p = (double*)calloc((size_t)n3, sizeof(double));
  double* t = (double*)calloc((size_t)n3, sizeof(double));
  for (int64_t k = 0; k < n3; ++k) x[k] = 0.0;
  for (int64_t k = 0; k < n3; ++k) r[k] = b[k];
  project_vec(s, r);
#define PRECOND(dst, src)                                                               \
  for (int64_t k = 0; k < nn; ++k)                                                      \
    for (int d = 0; d < 3; ++d)                                                         \
      dst[3 * k + d] = !node_active(s, k) ? 0.0                                         \
                     : inv_diag ? src[3 * k + d] * inv_diag[3 * k + d]                  \
                                : src[3 * k + d] / s->nodes[k].m;
  PRECOND(q, r);
  for (int64_t k = 0; k < n3; ++k) p[k] = q[k];
  double zTr = dot_active(s, r, q);
  double rnorm = sqrt(zTr);
  double local_tol = rel_tol * rnorm < o->cg_tolerance ? rel_tol * rnorm : o->cg_tolerance;
  int cnt;
  for (cnt = 0; cnt < o->cg_max_iterations; ++cnt) {
    if (rnorm < local_tol) break;
    zro_implicit_multiply(s, p, t);
    project_vec(s, t);
    double alpha = zTr / dot_active(s, t, p);
    for (int64_t k = 0; k < n3; ++k) {
      x[k] += p[k] * alpha;
      r[k] -= t[k] * alpha;
    }
    PRECOND(q, r);
    double last = zTr;
    zTr = dot_active(s, q, r);
    double beta = zTr / last;
    for (int64_t k = 0; k < n3; ++k) p[k] = q[k] + beta * p[k];
    rnorm = sqrt(zTr);
  }
#undef PRECOND
  free(r); free(q); free(p); free(t);
  return cnt;
}

int zro_backward_euler(zro_sim* s, double dt, const zro_implicit_opts* o, zro_implicit_stats* st)
{
  int64_t nn = s->npages * 32;
  s->be_dt = dt;
  free(s->vn); free(s->dv); free(s->proj); free(s->collided);
  s->vn = (double*)calloc((size_t)(3 * nn + 1), sizeof(double));
  s->dv = (double*)calloc((size_t)(3 * nn + 1), sizeof(double));
  s->proj = (double*)calloc((size_t)(9 * nn + 1), sizeof(double));
  s->collided = (char*)calloc((size_t)(nn + 1), 1);
  if (!s->Fn) s->Fn = (double*)malloc(sizeof(double) * 9 * (size_t)(s->n > 0 ? s->n : 1));
  /* buildInitialDvAndVnForNewton (MpmSimulationBase.cpp:1105-1136) */
  for (int64_t p = 0; p < s->npages; ++p) {
    int base[3];
    zro_linear_to_coord(s->pages[p], base);
    for (int e = 0; e < 32; ++e) {
      int64_t k = p * 32 + e;
      if (!node_active(s, k)) continue;
      double xi[3] = { (double)(base[0] + ((e >> 4) & 1)) * s->dx, (double)(base[1] + ((e >> 2) & 3)) * s->dx,
                       (double)(base[2] + (e & 3)) * s->dx };
      double vi[3] = { s->nodes[k].v[0], s->nodes[k].v[1], s->nodes[k].v[2] };
      double P[3][3];
      int hit = collision_projection(s, s->time, xi, vi, P);
      s->collided[k] = (char)hit;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s->proj[9 * k + 3 * r + c] = P[r][c];
      for (int d = 0; d < 3; ++d) {
        s->dv[3 * k + d] = hit ? vi[d] - s->nodes[k].v[d] : s->gravity[d] * dt;
        s->vn[3 * k + d] = s->nodes[k].v[d];
      }
    }
  }
  memcpy(s->Fn, s->F, sizeof(double) * 9 * (size_t)s->n); /* backupStrain */
  zro_implicit_stats local;
  memset(&local, 0, sizeof(local));
  double* r = (double*)calloc((size_t)(3 * nn + 1), sizeof(double));
  double* step = (double*)calloc((size_t)(3 * nn + 1), sizeof(double));
  /* NewtonsMethod::solve (NewtonsMethod.h:42-66) */
  for (int it = 0; it < o->newton_max_iterations; ++it) {
    implicit_update_state(s, s->dv);
    implicit_residual_at_state(s, s->dv, r);
    double norm = mass_norm(s, r);
    if (local.num_residuals < 8) local.residual_norm[local.num_residuals++] = norm;
    if (norm < o->newton_tolerance) { local.converged = 1; break; }
    double rel = o->cg_relative_tolerance > 0 ? o->cg_relative_tolerance
               : fmin(0.5, sqrt(fmax(norm, o->newton_tolerance)));
    double* inv_diag = NULL;
    if (o->preconditioner == 1) {
      inv_diag = (double*)calloc((size_t)(3 * nn + 1), sizeof(double));
      zro_implicit_diagonal(s, inv_diag);
      for (int64_t k = 0; k < nn; ++k) {
        /* JacobiPreconditioner: 1 / max(16 eps, |a_ii|); isotropic (largest
           entry) on slip-collided nodes so z stays in the projected space */
        double mx = 0.0;
        for (int d = 0; d < 3; ++d) {
          double a = fabs(inv_diag[3 * k + d]);
          inv_diag[3 * k + d] = a;
          if (a > mx) mx = a;
        }
        for (int d = 0; d < 3; ++d) {
          double a = s->collided[k] ? mx : inv_diag[3 * k + d];
          inv_diag[3 * k + d] = 1.0 / fmax(16.0 * DBL_EPSILON, a);
        }
      }
    }
    local.cg_iterations += o->solver == 1 ? minres_solve(s, step, r, rel, o, inv_diag)
                                          : cg_solve(s, step, r, rel, o, inv_diag);
    free(inv_diag);
    for (int64_t k = 0; k < 3 * nn; ++k) s->dv[k] += step[k];
    local.newton_iterations++;
  }
  free(r); free(step);
  memcpy(s->F, s->Fn, sizeof(double) * 9 * (size_t)s->n); /* restoreStrain */
  /* constructNewVelocityFromNewtonResult (MpmSimulationBase.cpp:867-877) */
  for (int64_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d)
      s->nodes[k].new_v[d] = node_active(s, k) ? s->nodes[k].v[d] + s->dv[3 * k + d] : 0.0;
  if (st) *st = local;
  return local.converged;
}

/* gridToParticlesHelper<APIC_blend_RPIC, deg 2, !MLS> (MpmSimulationBase.cpp:941-956)
   + evolveStrain: gradV = sum v* grad w^T, C = B D^-1 */
int zro_g2p_apic(zro_sim* s, double dt)
{
  int flags = 0;
  double dx2 = s->dx * s->dx;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : flags) num_threads(s->nthreads)
  for (int64_t g = 0; g < s->ngroups; ++g) {
    for (int64_t idx = s->gfirst[g]; idx <= s->glast[g]; ++idx) {
      int i = s->order[idx];
      double* Xp = s->X + 3 * i;
      zro_spline sp;
      spline_of(s, Xp, &sp);
      double picV[3] = { 0, 0, 0 }, B[3][3] = { { 0 } }, G[3][3] = { { 0 } };
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            double gw[3];
            double wt = spline_node(s, &sp, a, b, c, gw);
            int nc[3] = { sp.base[0] + a, sp.base[1] + b, sp.base[2] + c };
            zro_node* gn = node_at(s, nc[0], nc[1], nc[2], NULL);
            double dxp[3];
            for (int d = 0; d < 3; ++d) dxp[d] = (double)nc[d] * s->dx - Xp[d];
            for (int r = 0; r < 3; ++r) {
              picV[r] += wt * gn->new_v[r];
              for (int cc = 0; cc < 3; ++cc) {
                B[r][cc] += wt * gn->new_v[r] * dxp[cc];
                G[r][cc] += gn->new_v[r] * gw[cc];
              }
            }
          }
      for (int d = 0; d < 3; ++d) s->V[3 * i + d] = picV[d];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          s->C[9 * i + c * 3 + r] = B[r][c] * s->D_inverse;
          s->gradV[9 * i + c * 3 + r] = G[r][c];
        }
      double inc2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        double inc = dt * picV[d];
        Xp[d] += inc;
        inc2 += inc * inc;
      }
      if (inc2 > dx2) flags |= 1;
      if (inc2 > dx2 * 0.25 * (s->cfl * s->cfl)) flags |= 2;
      double Fo[3][3], Fnew[3][3];
      mat_from_cm(s->F + 9 * (int64_t)i, Fo);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = Fo[r][c];
          for (int k = 0; k < 3; ++k) acc += dt * G[r][k] * Fo[k][c];
          Fnew[r][c] = acc;
        }
      mat_to_cm(Fnew, s->F + 9 * (int64_t)i);
      apply_plasticity(s, i); /* MpmSimulationBase.cpp:1015 */
    }
  }
  return flags;
}

int zro_step_implicit(zro_sim* s, double dt, const zro_implicit_opts* o, zro_implicit_stats* st)
{
  zro_sort_and_pollute(s);
  zro_p2g_apic(s);
  zro_backward_euler(s, dt, o, st);
  int flags = zro_g2p_apic(s, dt);
  s->time += dt;
  return flags;
}

/* total energy of the backward-Euler objective at dv (BackwardEuler.h:183-197,
   Inertia.cpp:13-31): sum_i m |dv_i|^2 / 2 - dt sum_i m g.dv_i + sum_p V0 psi(F_p(dv)).
   Its negative gradient is the residual (before projection). */
double zro_implicit_energy(zro_sim* s, const double* dv)
{
  implicit_update_state(s, dv);
  double e = 0.0;
  int64_t nn = s->npages * 32;
  for (int64_t k = 0; k < nn; ++k) {
    if (!node_active(s, k)) continue;
    double m = s->nodes[k].m;
    for (int d = 0; d < 3; ++d) e += 0.5 * m * dv[3 * k + d] * dv[3 * k + d] - s->be_dt * m * s->gravity[d] * dv[3 * k + d];
  }
  double pe = 0.0;
#pragma omp parallel for reduction(+ : pe) num_threads(s->nthreads)
  for (int64_t i = 0; i < s->n; ++i) pe += s->vol[i] * zro_fcr_psi(s->mu[i], s->lam[i], s->F + 9 * i);
  return e + pe;
}

int64_t zro_num_node_slots(const zro_sim* s) { return s->npages * 32; }
void zro_get_implicit(const zro_sim* s, double* dv, double* vn, char* collided)
{
  int64_t nn = s->npages * 32;
  if (dv && s->dv) memcpy(dv, s->dv, sizeof(double) * 3 * (size_t)nn);
  if (vn && s->vn) memcpy(vn, s->vn, sizeof(double) * 3 * (size_t)nn);
  if (collided && s->collided) memcpy(collided, s->collided, (size_t)nn);
}

/* diagonal of multiply(): a_(i,d) = m_i + dt^2 sum_p (V0 dP(e_d (Fn^T grad w_ip)^T) Fn^T grad w_ip)_d,
   the entries JacobiPreconditioner reads from the assembled matrix
   (KrylovSolvers.h:32-57, BackwardEuler.h:322-351) */
void zro_implicit_diagonal(zro_sim* s, double* diag)
{
  int64_t nn = s->npages * 32;
  double dt = s->be_dt;
  for (int64_t k = 0; k < nn; ++k)
    for (int d = 0; d < 3; ++d) diag[3 * k + d] = node_active(s, k) ? s->nodes[k].m : 0.0;
  for (int color = 0; color < 8; ++color) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(s->nthreads)
    for (int64_t g = 0; g < s->ngroups; ++g) {
      if ((int)(s->block_offset[g] & 7) != color) continue;
      for (int64_t idx = s->gfirst[g]; idx <= s->glast[g]; ++idx) {
        int i = s->order[idx];
        double Fn[3][3];
        mat_from_cm(s->Fn + 9 * (int64_t)i, Fn);
        zro_spline sp;
        spline_of(s, s->X + 3 * i, &sp);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int c = 0; c < 3; ++c) {
              double gw[3];
              spline_node(s, &sp, a, b, c, gw);
              int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
              double h[3];
              for (int q = 0; q < 3; ++q) h[q] = Fn[0][q] * gw[0] + Fn[1][q] * gw[1] + Fn[2][q] * gw[2];
              for (int d = 0; d < 3; ++d) {
                double dF[3][3] = { { 0 } }, dFcm[9], dPcm[9], dP[3][3];
                for (int q = 0; q < 3; ++q) dF[d][q] = h[q];
                mat_to_cm(dF, dFcm);
                zro_fcr_first_piola_differential(s->mu[i], s->lam[i], s->F + 9 * (int64_t)i, dFcm, dPcm);
                mat_from_cm(dPcm, dP);
                double v = dP[d][0] * h[0] + dP[d][1] * h[1] + dP[d][2] * h[2];
                diag[3 * k + d] += dt * dt * s->vol[i] * v;
              }
            }
      }
    }
  }
}

/* ====================================================================== */
/* CD-MPM phase-field damage (Projects/fracture/FractureSimulation.h)      */
/* ====================================================================== */

/* largest eigenvalue of a symmetric 3x3 matrix: the trigonometric closed
   form of the characteristic cubic (O. K. Smith, CACM 4(4):168, 1961) */
double zro_sym3_max_eigenvalue(const double* A)
{
  double a00 = A[0], a11 = A[4], a22 = A[8];
  double a01 = 0.5 * (A[3] + A[1]), a02 = 0.5 * (A[6] + A[2]), a12 = 0.5 * (A[7] + A[5]);
  double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  double q = (a00 + a11 + a22) / 3.0;
  double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  if (p2 <= 0.0) return q;
  double p = sqrt(p2 / 6.0);
  double ip = 1.0 / p;
  double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip, c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
  double r = 0.5 * det;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  double phi = acos(r) / 3.0;
  return q + 2.0 * p * cos(phi);
}

static double* pf_dup(const double* a, int64_t n, double dflt)
{
  double* r = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) r[i] = a ? a[i] : dflt;
  return r;
}

void zro_set_phase_field(zro_sim* s, const double* c, const double* H, const double* l0,
                         const double* one_over_sigma_c, const double* pf_Fp, const double* H_max,
                         const double* residual_phase, const int32_t* allow_damage, const zro_pf_opts* o)
{
  int64_t n = s->n;
  if (!s->model) zro_set_materials(s, NULL, NULL, NULL, NULL, NULL);
  for (int64_t i = 0; i < n; ++i) s->model[i] = ZRO_MODEL_NEO_HOOKEAN_BORDEN;
  free(s->pf_c); free(s->pf_H); free(s->pf_l0); free(s->pf_osc); free(s->pf_Fp); free(s->pf_Hmax);
  free(s->pf_k); free(s->pf_vol); free(s->pf_allow);
  s->pf_c = pf_dup(c, n, 1.0);
  s->pf_H = pf_dup(H, n, 0.0);
  s->pf_l0 = pf_dup(l0, n, 0.0);
  s->pf_osc = pf_dup(one_over_sigma_c, n, 0.0);
  s->pf_Fp = pf_dup(pf_Fp, n, 1.0);
  s->pf_Hmax = pf_dup(H_max, n, HUGE_VAL);
  s->pf_k = pf_dup(residual_phase, n, 0.001);
  s->pf_vol = pf_dup(NULL, n, 0.0);
  s->pf_allow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; ++i) s->pf_allow[i] = allow_damage ? allow_damage[i] : 1;
  s->pf_o = *o;
  s->pf_on = 1;
}

void zro_get_phase_field(const zro_sim* s, double* c, double* H, double* pf_Fp)
{
  if (!s->pf_on) return;
  if (c) memcpy(c, s->pf_c, sizeof(double) * (size_t)s->n);
  if (H) memcpy(H, s->pf_H, sizeof(double) * (size_t)s->n);
  if (pf_Fp) memcpy(pf_Fp, s->pf_Fp, sizeof(double) * (size_t)s->n);
}

static double det_cm(const double* F)
{
  return F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) + F[6] * (F[1] * F[5] - F[4] * F[2]);
}

/* loops over the particles of the groups in sort order, as the reference's
   particle_group iteration (the colour phases only order its TBB tasks) */
#define PF_FOR_PARTICLES(s, i)                                           \
  for (int64_t g_ = 0; g_ < (s)->ngroups; ++g_)                         \
    for (int64_t k_ = (s)->gfirst[g_]; k_ <= (s)->glast[g_]; ++k_)       \
      for (int i = (s)->order[k_], once_ = 1; once_; once_ = 0)

/* the lumped operator (FractureSimulation.h:338-403) on node-slot vectors */
static void pf_multiply(zro_sim* s, const double* x, double* b)
{
  const double M = s->pf_o.parabolic_M, dt = s->pf_dt;
  int64_t ns = s->npages * 32;
  memset(b, 0, sizeof(double) * (size_t)ns);
  PF_FOR_PARTICLES(s, i)
  {
    zro_spline sp;
    spline_of(s, s->X + 3 * i, &sp);
    double gc[3] = { 0, 0, 0 };
    for (int a = 0; a < 3; ++a)
      for (int bb = 0; bb < 3; ++bb)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          spline_node(s, &sp, a, bb, c, gw);
          int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + bb, sp.base[2] + c);
          if (s->nodes[k].idx < 0) continue;
          for (int d = 0; d < 3; ++d) gc[d] += x[k] * gw[d];
        }
    double vol = s->pf_vol[i], l0 = s->pf_l0[i];
    for (int d = 0; d < 3; ++d) gc[d] *= vol * 4.0 * l0 * l0;
    double lump = M > 0 ? vol * (s->pf_Fp[i] * M + 1.0 / dt) : vol * s->pf_Fp[i];
    double gscale = M > 0 ? M : 1.0;
    for (int a = 0; a < 3; ++a)
      for (int bb = 0; bb < 3; ++bb)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, bb, c, gw);
          int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + bb, sp.base[2] + c);
          if (s->nodes[k].idx < 0) continue;
          b[k] += lump * w * x[k] + (gc[0] * gw[0] + gc[1] * gw[1] + gc[2] * gw[2]) * gscale;
        }
  }
}

static double pf_dot(const zro_sim* s, const double* a, const double* b)
{
  double acc = 0.0;
  for (int64_t k = 0; k < s->npages * 32; ++k)
    if (s->nodes[k].idx >= 0) acc += a[k] * b[k];
  return acc;
}

/* Projects/fracture/ConjugateGradient.h:50-97 with x = 0 on entry */
static int pf_cg(zro_sim* s, double* x, const double* rhs, const double* diag, double* res)
{
  int64_t ns = s->npages * 32;
  double* r = (double*)calloc((size_t)ns, sizeof(double));
  double* mr = (double*)calloc((size_t)ns, sizeof(double));
  double* sd = (double*)calloc((size_t)ns, sizeof(double));
  double* q = (double*)calloc((size_t)ns, sizeof(double));
  memset(x, 0, sizeof(double) * (size_t)ns);
  memcpy(r, rhs, sizeof(double) * (size_t)ns); /* r = b - A*0 */
  double rho_old = DBL_MAX;
  double norm = sqrt(pf_dot(s, r, r));
  double rel = norm * s->pf_o.cg_relative;
  int it;
  for (it = 0;; ++it) {
    for (int64_t k = 0; k < ns; ++k) mr[k] = s->nodes[k].idx >= 0 ? r[k] / diag[k] : 0.0;
    double rho = pf_dot(s, mr, r);
    if (it == 0)
      memcpy(sd, mr, sizeof(double) * (size_t)ns);
    else
      for (int64_t k = 0; k < ns; ++k) sd[k] = rho / rho_old * sd[k] + mr[k];
    pf_multiply(s, sd, q);
    double sq = pf_dot(s, sd, q);
    double alpha = sq != 0.0 ? rho / sq : DBL_MAX;
    for (int64_t k = 0; k < ns; ++k) {
      x[k] = alpha * sd[k] + x[k];
      r[k] = -alpha * q[k] + r[k];
    }
    rho_old = rho;
    norm = sqrt(pf_dot(s, r, r));
    if (norm < s->pf_o.cg_tolerance || norm < rel || it == s->pf_o.cg_max_iterations) break;
  }
  free(r); free(mr); free(sd); free(q);
  if (res) *res = norm;
  return it;
}

static void pf_update_model(zro_sim* s, int i)
{
  double c = s->pf_c[i], k = s->pf_k[i];
  s->gdeg[i] = c * c * (1.0 - k) + k;
}

/* solvePhaseFieldSystemNaive (:509-554): c cut to sigma_F / sigma_max, the
   largest principal Cauchy stress of the undegraded model */
static void pf_rankine_cut(zro_sim* s)
{
  const zro_pf_opts* o = &s->pf_o;
  for (int64_t i = 0; i < s->n; ++i) {
    const double* F = s->F + 9 * i;
    double J = det_cm(F), tau[9], cauchy[9];
    zro_nhb_kirchhoff(s->mu[i], s->lam[i], 1.0, F, tau);
    for (int c = 0; c < 9; ++c) cauchy[c] = tau[c] / J;
    double smax = zro_sym3_max_eigenvalue(cauchy);
    double newc = s->pf_c[i];
    if (smax > o->sigma_F) newc = fmin(newc, o->sigma_F / smax);
    s->pf_c[i] = newc;
    pf_update_model(s, (int)i);
  }
}

/* FractureSimulation::deleteParticles (:162-207), which reinitialize() runs
   at the start of every step (MpmSimulationBase.cpp:262) */
void zro_delete_particles(zro_sim* s)
{
  if (!s->pf_on) return;
  for (int64_t i = 0; i < s->n; ++i) {
    if (!(s->pf_c[i] < s->pf_o.delete_threshold)) continue;
    for (int d = 0; d < 3; ++d) {
      s->X[3 * i + d] = 4.0;
      s->V[3 * i + d] = 0.0;
    }
    for (int k = 0; k < 9; ++k) s->F[9 * i + k] = (k % 4 == 0) ? 1.0 : 0.0;
    s->pf_allow[i] = 0;
  }
}

void zro_phase_field_solve(zro_sim* s, double dt, zro_pf_stats* st)
{
  if (!s->pf_on) return;
  const zro_pf_opts* o = &s->pf_o;
  if (st) memset(st, 0, sizeof(*st));
  if (o->naive == 1) {
    pf_rankine_cut(s);
    return;
  }
  s->pf_dt = dt;
  int64_t ns = s->npages * 32;
  if (s->pf_cap < ns) {
    free(s->pf_phase); free(s->pf_mult); free(s->pf_rhs); free(s->pf_diag); free(s->pf_x);
    s->pf_phase = (double*)malloc(sizeof(double) * (size_t)ns);
    s->pf_mult = (double*)malloc(sizeof(double) * (size_t)ns);
    s->pf_rhs = (double*)malloc(sizeof(double) * (size_t)ns);
    s->pf_diag = (double*)malloc(sizeof(double) * (size_t)ns);
    s->pf_x = (double*)malloc(sizeof(double) * (size_t)ns);
    s->pf_cap = ns;
  }
  memset(s->pf_phase, 0, sizeof(double) * (size_t)ns);
  memset(s->pf_mult, 0, sizeof(double) * (size_t)ns);
  memset(s->pf_rhs, 0, sizeof(double) * (size_t)ns);
  memset(s->pf_diag, 0, sizeof(double) * (size_t)ns);
  const double M = o->parabolic_M;
  /* phaseP2G (:209-257) */
  if (M > 0) {
    PF_FOR_PARTICLES(s, i)
    {
      zro_spline sp;
      spline_of(s, s->X + 3 * i, &sp);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            double gw[3];
            double w = spline_node(s, &sp, a, b, c, gw);
            int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
            s->pf_mult[k] += w;
            s->pf_phase[k] += s->pf_c[i] * w;
          }
    }
    for (int64_t k = 0; k < ns; ++k)
      if (s->nodes[k].idx >= 0) s->pf_phase[k] /= s->pf_mult[k];
  }
  for (int64_t i = 0; i < s->n; ++i) s->pf_vol[i] = s->vol[i] * det_cm(s->F + 9 * i);
  /* rhs (:303-325) and the lumped preconditioner's diagonal (:405-437) */
  PF_FOR_PARTICLES(s, i)
  {
    zro_spline sp;
    spline_of(s, s->X + 3 * i, &sp);
    double vol = s->pf_vol[i];
    double dg = M > 0 ? vol * (M * s->pf_Fp[i] + 1.0 / dt) : vol * s->pf_Fp[i];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, b, c, gw);
          int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
          if (s->nodes[k].idx < 0) continue;
          s->pf_rhs[k] += (M > 0 ? (M + s->pf_phase[k] / dt) * vol * w : vol * w);
          s->pf_diag[k] += dg * w;
        }
  }
  double res = 0.0;
  int its = pf_cg(s, s->pf_x, s->pf_rhs, s->pf_diag, &res);
  if (st) {
    st->cg_iterations = its;
    st->residual_norm = res;
  }
  for (int64_t k = 0; k < ns; ++k)
    if (s->nodes[k].idx >= 0) s->pf_phase[k] = M > 0 ? s->pf_x[k] - s->pf_phase[k] : s->pf_x[k];
  /* phaseG2P (:457-507) */
  PF_FOR_PARTICLES(s, i)
  {
    if (!s->pf_allow[i]) continue;
    zro_spline sp;
    spline_of(s, s->X + 3 * i, &sp);
    double pf = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, b, c, gw);
          int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
          if (s->nodes[k].idx >= 0) pf += w * s->pf_phase[k];
        }
    double c0 = s->pf_c[i];
    double nc = M > 0 ? c0 + pf : pf;
    double cm = c0 < nc ? c0 : nc;
    s->pf_c[i] = cm > 0.0 ? cm : 0.0;
  }
  for (int64_t i = 0; i < s->n; ++i) {
    double psi = zro_nhb_psi_pos(s->mu[i], s->lam[i], s->F + 9 * i);
    if (psi > s->pf_H[i]) { /* PhaseField::Update_Phase_Field_Fp (PhaseField.h:70-76) */
      s->pf_H[i] = psi < s->pf_Hmax[i] ? psi : s->pf_Hmax[i];
      s->pf_Fp[i] = 4.0 * s->pf_l0[i] * (1.0 - s->pf_k[i]) * s->pf_H[i] * s->pf_osc[i] + 1.0;
    }
    pf_update_model(s, (int)i);
  }
  if (o->naive == 2) pf_rankine_cut(s);
}

void zro_phase_field_system(const zro_sim* s, double* rhs, double* diag, double* x)
{
  size_t ns = (size_t)(s->npages * 32);
  if (!s->pf_rhs) return;
  if (rhs) memcpy(rhs, s->pf_rhs, sizeof(double) * ns);
  if (diag) memcpy(diag, s->pf_diag, sizeof(double) * ns);
  if (x) memcpy(x, s->pf_x, sizeof(double) * ns);
}

void zro_phase_field_multiply(zro_sim* s, const double* x, double* b) { pf_multiply(s, x, b); }

/* ====================================================================== */
/* Adjoint of the explicit MLS step (no counterpart in the reference): the  */
/* reverse of stress -> P2G -> grid update -> G2P + F update for fixed-     */
/* corotated particles, with the grid topology of the step held fixed.      */
/* ====================================================================== */

typedef double M3[3][3];

/* Jacobian d v_after / d v_before of the collider handling of one node
   (multi_object_collision in the frame at t, then trial_collision in the
   frame at t + dt when nothing collided) for objects moving as x = R s X + b.
   The object's velocity at the node does not depend on v, so each collider
   acts on v - v_obj through its world normal n = R N: sticky gives 0, a
   slipping or approaching contact the projection P = I - n n^T.  Coulomb
   friction on the sliding branch (|t| > -mu (v - v_obj).n, t the tangential
   part) scales t by k = 1 + mu dot / |t|: d t' / d rel = k P + t (mu n / |t|
   - mu dot t / |t|^3)^T; the sticking branch gives 0.  The trial collision
   v -= (phi / dt) n at x + dt v has I - n n^T - phi grad n, with grad n =
   P / (phi + s r) for a sphere of radius s r.  Returns 0 for a node at a
   sphere's centre in the trial collision (no normal). */
static int collide_jacobian(const zro_sim* s, double dt, const double* x, const double* vpre, double J[3][3])
{
  double v[3] = { vpre[0], vpre[1], vpre[2] };
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[i][j] = i == j ? 1.0 : 0.0;
  int any = 0;
  for (int k = 0; k < s->ncol; ++k) {
    const zro_collider* c = &s->col[k];
    double xb[3], phi, N[3], vo[3];
    zro_frame fr;
    collider_frame(c, s->time, &fr);
    if (!collider_query(c, &fr, x, &phi, N, xb)) continue;
    any = 1;
    if (c->type == 1) {
      memset(J, 0, sizeof(double) * 9);
      return 1;
    }
    collider_velocity(c, &fr, xb, vo);
    double rel[3], dot = 0.0;
    for (int d = 0; d < 3; ++d) rel[d] = v[d] - vo[d];
    for (int d = 0; d < 3; ++d) dot += rel[d] * N[d];
    if (c->type == 2 || dot < 0) {
      double P[3][3], Jr[3][3], T[3][3], t[3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) P[i][j] = (i == j ? 1.0 : 0.0) - N[i] * N[j];
      for (int d = 0; d < 3; ++d) t[d] = rel[d] - N[d] * dot;
      memcpy(Jr, P, sizeof(P));
      if (c->friction != 0.0 && dot < 0) {
        const double mu = c->friction;
        const double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (-dot * mu < tn) {
          const double kf = 1.0 + mu * dot / tn;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Jr[i][j] = kf * P[i][j] + t[i] * (mu * N[j] / tn - mu * dot * t[j] / (tn * tn * tn));
          for (int d = 0; d < 3; ++d) t[d] *= kf;
        } else {
          memset(Jr, 0, sizeof(Jr));
          t[0] = t[1] = t[2] = 0.0;
        }
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[i][j] = Jr[i][0] * J[0][j] + Jr[i][1] * J[1][j] + Jr[i][2] * J[2][j];
      memcpy(J, T, sizeof(T));
      for (int d = 0; d < 3; ++d) rel[d] = t[d];
    }
    for (int d = 0; d < 3; ++d) v[d] = rel[d] + vo[d];
  }
  if (any) return 1;
  for (int k = 0; k < s->ncol; ++k) { /* trial collision at t + dt */
    const zro_collider* c = &s->col[k];
    double xt[3], xb[3], phi, n[3];
    zro_frame fr;
    collider_frame(c, s->time + dt, &fr);
    for (int d = 0; d < 3; ++d) xt[d] = x[d] + dt * v[d];
    if (!collider_query(c, &fr, xt, &phi, n, xb)) continue;
    double A[3][3], T[3][3];
    if (c->shape == 0) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = (i == j ? 1.0 : 0.0) - n[i] * n[j];
    } else {
      double dist = phi + fr.s * c->radius;
      if (dist < 1e-7 * fr.s) return 0;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double pr = (i == j ? 1.0 : 0.0) - n[i] * n[j];
          A[i][j] = pr - phi / dist * pr;
        }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[i][j] = A[i][0] * J[0][j] + A[i][1] * J[1][j] + A[i][2] * J[2][j];
    memcpy(J, T, sizeof(T));
    for (int d = 0; d < 3; ++d) v[d] -= (phi / dt) * n[d];
  }
  return 1;
}

/* One reverse step.  The sim holds the step's INPUT particle state; the
   forward sort, stress, P2G and grid update are recomputed here.  On entry
   aX, aV, aC, aF hold the adjoint of the step's output state (particle-major,
   matrices column-major); on return the adjoint of its input.  amu, alam
   accumulate.  Returns 0, or -1 when a collider branch has no derivative. */
int zro_adjoint_step(zro_sim* s, double dt, double* aX, double* aV, double* aC, double* aF, double* amu,
                     double* alam)
{
  if (!s->n) return 0;
  zro_sort_and_pollute(s);
  zro_update_stress(s);
  zro_p2g_mls(s);
  /* forward grid quantities before the update: v = p/m, f */
  int64_t ns = s->npages * 32;
  double* gv = (double*)calloc((size_t)(3 * ns), sizeof(double));
  double* gf = (double*)calloc((size_t)(3 * ns), sizeof(double));
  for (int64_t k = 0; k < ns; ++k) {
    if (s->nodes[k].idx < 0) continue;
    for (int d = 0; d < 3; ++d) { gv[3 * k + d] = s->nodes[k].v[d]; gf[3 * k + d] = s->nodes[k].new_v[d]; }
  }
  zro_grid_update(s, dt);
  const double Dinv = s->D_inverse;
  double* avs = (double*)calloc((size_t)(3 * ns), sizeof(double));  /* adjoint of v* */
  double* ag = (double*)calloc((size_t)(7 * ns), sizeof(double));   /* adjoint of m, p, f */
  double* xin = (double*)calloc((size_t)(3 * s->n), sizeof(double)); /* adjoint of x (input) */
  double* Fin = (double*)calloc((size_t)(9 * s->n), sizeof(double));
  /* ---- G2P + F update reverse ---- */
  for (int64_t i = 0; i < s->n; ++i) {
    const double* Xp = s->X + 3 * i;
    zro_spline sp;
    spline_of(s, Xp, &sp);
    /* forward recompute of v', C' */
    double pv[3] = { 0, 0, 0 }, B[3][3] = { { 0 } };
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, b, c, gw);
          int nc[3] = { sp.base[0] + a, sp.base[1] + b, sp.base[2] + c };
          zro_node* g = node_at(s, nc[0], nc[1], nc[2], NULL);
          double d3[3];
          for (int d = 0; d < 3; ++d) d3[d] = (double)nc[d] * s->dx - Xp[d];
          for (int r = 0; r < 3; ++r) {
            pv[r] += w * g->new_v[r];
            for (int cc = 0; cc < 3; ++cc) B[r][cc] += w * g->new_v[r] * d3[cc];
          }
        }
    M3 Cn, Fo, Fb, Cb, Bb;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Cn[r][c] = B[r][c] * Dinv;
    mat_from_cm(s->F + 9 * i, Fo);
    mat_from_cm(aF + 9 * i, Fb);
    mat_from_cm(aC + 9 * i, Cb);
    /* F' = F + dt C' F */
    M3 Fin_m;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = Fb[r][c];
        for (int k = 0; k < 3; ++k) acc += dt * Cn[k][r] * Fb[k][c];
        Fin_m[r][c] = acc;
        double cb = 0.0;
        for (int k = 0; k < 3; ++k) cb += Fb[r][k] * Fo[c][k];
        Cb[r][c] += dt * cb;
      }
    mat_to_cm(Fin_m, Fin + 9 * i);
    double vb[3];
    for (int d = 0; d < 3; ++d) {
      xin[3 * i + d] = aX[3 * i + d];
      vb[d] = aV[3 * i + d] + dt * aX[3 * i + d];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Bb[r][c] = Dinv * Cb[r][c];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, b, c, gw);
          int nc[3] = { sp.base[0] + a, sp.base[1] + b, sp.base[2] + c };
          int64_t slot;
          zro_node* g = node_at(s, nc[0], nc[1], nc[2], &slot);
          double d3[3];
          for (int d = 0; d < 3; ++d) d3[d] = (double)nc[d] * s->dx - Xp[d];
          const double* vs = g->new_v;
          double bd[3];
          for (int r = 0; r < 3; ++r) bd[r] = Bb[r][0] * d3[0] + Bb[r][1] * d3[1] + Bb[r][2] * d3[2];
          double sc = 0.0;
          for (int r = 0; r < 3; ++r) {
            avs[3 * slot + r] += w * (vb[r] + bd[r]);
            sc += vs[r] * (vb[r] + bd[r]);
          }
          for (int d = 0; d < 3; ++d) {
            double btv = Bb[0][d] * vs[0] + Bb[1][d] * vs[1] + Bb[2][d] * vs[2];
            xin[3 * i + d] += gw[d] * sc - w * btv;
          }
        }
  }
  /* ---- grid update reverse ---- */
  int ok = 1;
  for (int64_t p = 0; p < s->npages; ++p) {
    int base[3];
    zro_linear_to_coord(s->pages[p], base);
    for (int e = 0; e < 32; ++e) {
      int64_t k = p * 32 + e;
      zro_node* g = &s->nodes[k];
      if (g->idx < 0) continue;
      double xi[3] = { (double)(base[0] + ((e >> 4) & 1)) * s->dx, (double)(base[1] + ((e >> 2) & 3)) * s->dx,
                       (double)(base[2] + (e & 3)) * s->dx };
      double m = g->m, vpre[3], J[3][3], vb[3];
      for (int d = 0; d < 3; ++d) vpre[d] = gv[3 * k + d] + s->gravity[d] * dt + gf[3 * k + d] * (dt / m);
      if (!collide_jacobian(s, dt, xi, vpre, J)) ok = 0;
      for (int d = 0; d < 3; ++d) vb[d] = J[0][d] * avs[3 * k] + J[1][d] * avs[3 * k + 1] + J[2][d] * avs[3 * k + 2];
      double mb = 0.0;
      for (int d = 0; d < 3; ++d) {
        ag[7 * k + 1 + d] = vb[d] / m;
        ag[7 * k + 4 + d] = dt * vb[d] / m;
        mb -= (gv[3 * k + d] + gf[3 * k + d] * (dt / m)) * vb[d] / m;
      }
      ag[7 * k] = mb;
    }
  }
  /* ---- P2G + stress reverse ---- */
  for (int64_t i = 0; i < s->n; ++i) {
    const double* Xp = s->X + 3 * i;
    zro_spline sp;
    spline_of(s, Xp, &sp);
    double mass = s->mass[i];
    M3 Cm, vt, Cb, vtb;
    mat_from_cm(s->C + 9 * i, Cm);
    mat_from_cm(s->vtau + 9 * i, vt);
    memset(Cb, 0, sizeof(Cb));
    memset(vtb, 0, sizeof(vtb));
    double vbar[3] = { 0, 0, 0 }, xb[3] = { 0, 0, 0 };
    const double* vp = s->V + 3 * i;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, b, c, gw);
          int nc[3] = { sp.base[0] + a, sp.base[1] + b, sp.base[2] + c };
          int64_t slot;
          node_at(s, nc[0], nc[1], nc[2], &slot);
          if (s->nodes[slot].idx < 0) continue;
          const double* A = ag + 7 * slot;
          double d3[3];
          for (int d = 0; d < 3; ++d) d3[d] = (double)nc[d] * s->dx - Xp[d];
          double mom[3], td[3];
          for (int r = 0; r < 3; ++r) {
            mom[r] = mass * vp[r] + mass * (Cm[r][0] * d3[0] + Cm[r][1] * d3[1] + Cm[r][2] * d3[2]);
            td[r] = vt[r][0] * d3[0] + vt[r][1] * d3[1] + vt[r][2] * d3[2];
          }
          double sp_ = A[1] * mom[0] + A[2] * mom[1] + A[3] * mom[2];
          double sf = A[4] * td[0] + A[5] * td[1] + A[6] * td[2];
          for (int r = 0; r < 3; ++r) {
            vbar[r] += w * mass * A[1 + r];
            for (int cc = 0; cc < 3; ++cc) {
              Cb[r][cc] += w * mass * A[1 + r] * d3[cc];
              vtb[r][cc] += -w * Dinv * A[4 + r] * d3[cc];
            }
          }
          for (int d = 0; d < 3; ++d) {
            double ctp = Cm[0][d] * A[1] + Cm[1][d] * A[2] + Cm[2][d] * A[3];
            double vtf = vt[0][d] * A[4] + vt[1][d] * A[5] + vt[2][d] * A[6];
            xb[d] += gw[d] * (A[0] * mass + sp_ - Dinv * sf) - w * mass * ctp + w * Dinv * vtf;
          }
        }
    for (int d = 0; d < 3; ++d) {
      aX[3 * i + d] = xin[3 * i + d] + xb[d];
      aV[3 * i + d] = vbar[d];
    }
    mat_to_cm(Cb, aC + 9 * i);
    /* stress reverse: vtau = V0 tau(F) */
    double tbar[9], mub = 0.0, lab = 0.0;
    M3 tbm;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) tbm[r][c] = s->vol[i] * vtb[r][c];
    mat_to_cm(tbm, tbar);
    zro_fcr_kirchhoff_adjoint(s->mu[i], s->lam[i], s->F + 9 * i, tbar, Fin + 9 * i, &mub, &lab);
    memcpy(aF + 9 * i, Fin + 9 * i, sizeof(double) * 9);
    amu[i] += mub;
    alam[i] += lab;
  }
  free(gv); free(gf); free(avs); free(ag); free(xin); free(Fin);
  return ok ? 0 : -1;
}

/* ====================================================================== */
/* ADMM step (Projects/admm/AdmmSimulationFoam.h, AdmmFormula.h)           */
/* ====================================================================== */

/* eigenvalues of a symmetric 3x3 matrix, ascending (the trigonometric closed
   form of the characteristic cubic, O. K. Smith 1961) */
void zro_sym3_eigenvalues(const double* A, double* lam)
{
  double a00 = A[0], a11 = A[4], a22 = A[8];
  double a01 = 0.5 * (A[3] + A[1]), a02 = 0.5 * (A[6] + A[2]), a12 = 0.5 * (A[7] + A[5]);
  double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  double q = (a00 + a11 + a22) / 3.0;
  double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  if (p2 <= 0.0) {
    lam[0] = lam[1] = lam[2] = q;
    return;
  }
  double p = sqrt(p2 / 6.0), ip = 1.0 / p;
  double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip, c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  double det = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02);
  double r = 0.5 * det;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  double phi = acos(r) / 3.0;
  double e0 = q + 2.0 * p * cos(phi);
  double e2 = q + 2.0 * p * cos(phi + 2.0943951023931957);
  lam[0] = e2;
  lam[1] = 3.0 * q - e0 - e2;
  lam[2] = e0;
}

/* |makePD(A)|_F^2 = sum of the squared nonnegative parts of the eigenvalues
   (AdmmFormula.h:12-35): the eigenvectors are orthonormal */
static double pd_norm2_3(const double A[3][3])
{
  double a[9], l[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[j * 3 + i] = A[i][j];
  zro_sym3_eigenvalues(a, l);
  double t = 0.0;
  for (int i = 0; i < 3; ++i) t += l[i] > 0.0 ? l[i] * l[i] : 0.0;
  return t;
}
static double pd_norm2_2(double a, double b, double d)  /* [[a b][b d]] */
{
  double m = 0.5 * (a + d), r = sqrt(0.25 * (a - d) * (a - d) + b * b);
  double l0 = m - r, l1 = m + r, t = 0.0;
  if (l0 > 0.0) t += l0 * l0;
  if (l1 > 0.0) t += l1 * l1;
  return t;
}

/* StvkWithHencky::firstPiolaDiagonal / firstPiolaDerivativeDiagonal (StvkWithHencky.cpp:150-204) */
static void hencky_dE(double mu, double lam, const double* s, double* dE)
{
  double l[3] = { log(fabs(s[0])), log(fabs(s[1])), log(fabs(s[2])) };
  double sl = l[0] + l[1] + l[2];
  for (int i = 0; i < 3; ++i) dE[i] = (2.0 * mu * l[i] + lam * sl) / s[i];
}
static void hencky_ddE(double mu, double lam, const double* s, double d[3][3])
{
  double l[3] = { log(fabs(s[0])), log(fabs(s[1])), log(fabs(s[2])) };
  double g = 2.0 * mu + lam;
  d[0][0] = (g * (1.0 - l[0]) - lam * (l[1] + l[2])) / (s[0] * s[0]);
  d[1][1] = (g * (1.0 - l[1]) - lam * (l[0] + l[2])) / (s[1] * s[1]);
  d[2][2] = (g * (1.0 - l[2]) - lam * (l[0] + l[1])) / (s[2] * s[2]);
  d[0][1] = d[1][0] = lam / (s[0] * s[1]);
  d[0][2] = d[2][0] = lam / (s[0] * s[2]);
  d[1][2] = d[2][1] = lam / (s[1] * s[2]);
}
/* StvkWithHencky::Bij, 3-D branch (StvkWithHencky.cpp:34-76) with
   MathTools::diff_interlock_log_over_diff (MathTools.h:197-262) */
static void hencky_Bij(double mu, double lam, const double* s, int i, int j, double clamp, double B[2][2])
{
  double dE[3];
  hencky_dE(mu, lam, s, dE);
  double sum = s[i] + s[j];
  double mag = fabs(sum), sign = sum > 0 ? 1.0 : -1.0;
  double cl = (mag > clamp ? mag : clamp) * sign;
  double BP = 0.5 * (dE[i] + dE[j]) / cl;
  double l[3] = { log(fabs(s[0])), log(fabs(s[1])), log(fabs(s[2])) };
  double sl = l[0] + l[1] + l[2];
  double eps = 1e-8;
  double p = s[i] / s[j] - 1.0;
  double l1p = fabs(p) < eps ? 1.0 : log1p(p) / p;
  double dlod = l1p / s[j];
  double interlock = l[j] - s[j] * dlod;
  double BM = -0.5 * (lam * sl + 2.0 * mu * interlock) / (s[i] * s[j]);
  B[0][0] = BP + BM;
  B[0][1] = -BP + BM;
  B[1][0] = -BP + BM;
  B[1][1] = BP + BM;
}

/* projectSigma of the particle's plasticity (PlasticityApplier.cpp:321-346, 528-547;
   DummyPlasticity returns sigma) */
static void project_sigma(double mu, double lam, int pl, const double* q, double ty, const double* sig, double* out)
{
  for (int d = 0; d < 3; ++d) out[d] = sig[d];
  if (pl == ZRO_PLASTIC_VON_MISES) {
    double e[3], tr = 0.0, n2 = 0.0, h[3];
    for (int d = 0; d < 3; ++d) { e[d] = log(sig[d] > 1e-4 ? sig[d] : 1e-4); tr += e[d]; }
    for (int d = 0; d < 3; ++d) { h[d] = e[d] - tr / 3.0; n2 += h[d] * h[d]; }
    double nrm = sqrt(n2), dg = nrm - ty / (2.0 * mu);
    if (dg <= 0.0) return;
    for (int d = 0; d < 3; ++d) out[d] = exp(e[d] - (dg / nrm) * h[d]);
  } else if (pl == ZRO_PLASTIC_DRUCKER_PRAGER) {
    double alpha = q[0], coh = q[2];
    double e[3], tr = 0.0, n2 = 0.0, h[3];
    for (int d = 0; d < 3; ++d) {
      double a = fabs(sig[d]);
      e[d] = log(a > 1e-4 ? a : 1e-4) - coh;
      tr += e[d];
    }
    for (int d = 0; d < 3; ++d) { h[d] = e[d] - tr / 3.0; n2 += h[d] * h[d]; }
    if (tr >= 0.0) { for (int d = 0; d < 3; ++d) out[d] = exp(coh); return; }
    double nrm = sqrt(n2);
    double dg = nrm + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * alpha;
    for (int d = 0; d < 3; ++d) out[d] = exp(dg <= 0.0 ? e[d] + coh : e[d] - (dg / nrm) * h[d] + coh);
  }
}

static void admm_project_sigma(const zro_sim* s, int64_t i, const double* sig, double* out)
{
  project_sigma(s->mu[i], s->lam[i], s->plastic ? s->plastic[i] : ZRO_PLASTIC_NONE,
                s->pparam ? s->pparam + 4 * i : NULL, s->pstate ? s->pstate[i] : 0.0, sig, out);
}

/* projectSigmaAndDerivative (PlasticityApplier.cpp:389-423, 580-606; Dummy: identity) */
static void project_sigma_d(double mu, double lam, int pl, const double* q, double ty, const double* sig, double* zp,
                            double D[3][3])
{
  for (int a = 0; a < 3; ++a) {
    zp[a] = sig[a];
    for (int b = 0; b < 3; ++b) D[a][b] = a == b ? 1.0 : 0.0;
  }
  if (pl == ZRO_PLASTIC_VON_MISES) {
    double e[3], tr = 0.0, n2 = 0.0, h[3];
    for (int d = 0; d < 3; ++d) { e[d] = log(sig[d] > 1e-4 ? sig[d] : 1e-4); tr += e[d]; }
    for (int d = 0; d < 3; ++d) { h[d] = e[d] - tr / 3.0; n2 += h[d] * h[d]; }
    double nrm = sqrt(n2), dg = nrm - ty / (2.0 * mu);
    if (dg <= 0.0) return;
    double w[3], sv[3], sn = 0.0, sh[3], H[3];
    for (int d = 0; d < 3; ++d) { w[d] = 1.0 / sig[d]; sv[d] = e[d] - tr / 3.0; sn += sv[d] * sv[d]; }
    sn = sqrt(sn);
    for (int d = 0; d < 3; ++d) { sh[d] = sv[d] / sn; H[d] = e[d] - (dg / nrm) * h[d]; zp[d] = exp(H[d]); }
    double a1 = 1.0 - ty / (2.0 * mu * sn), a2 = ty / (2.0 * mu * sn);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double Iw = r == c ? w[c] : 0.0;
        double inner = Iw - a1 * (Iw - w[c] / 3.0) - a2 * sh[r] * sh[c] * w[c];
        D[r][c] = zp[r] * inner;
      }
  } else if (pl == ZRO_PLASTIC_DRUCKER_PRAGER) {
    double alpha = q[0], coh = q[2];
    double e[3], tr = 0.0, n2 = 0.0, h[3];
    for (int d = 0; d < 3; ++d) {
      double a = fabs(sig[d]);
      e[d] = log(a > 1e-4 ? a : 1e-4) - coh;
      tr += e[d];
    }
    for (int d = 0; d < 3; ++d) { h[d] = e[d] - tr / 3.0; n2 += h[d] * h[d]; }
    double nrm = sqrt(n2);
    double dg = nrm + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * alpha;
    if (dg <= 0.0) {
      for (int d = 0; d < 3; ++d) zp[d] = exp(e[d] + coh);
      return;
    }
    if (tr > 0.0 || nrm == 0.0) {
      for (int d = 0; d < 3; ++d) {
        zp[d] = exp(coh);
        for (int c = 0; c < 3; ++c) D[d][c] = 0.0;
      }
      return;
    }
    double w[3], sv[3], sn = 0.0, sh[3];
    double k = tr;
    for (int d = 0; d < 3; ++d) { w[d] = 1.0 / sig[d]; sv[d] = e[d] - k / 3.0; sn += sv[d] * sv[d]; }
    sn = sqrt(sn);
    for (int d = 0; d < 3; ++d) sh[d] = sv[d] / sn;
    double pp = alpha * k * (3.0 * lam + 2.0 * mu) / (2.0 * mu * sn);
    for (int d = 0; d < 3; ++d) zp[d] = exp(e[d] - (dg / nrm) * h[d] + coh);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double t1 = ((1.0 + 2.0 * pp) / 3.0 - pp / k * e[r]) * w[c];
        double t2 = pp * ((r == c ? 1.0 : 0.0) - sh[r] * sh[c]) * w[c];
        D[r][c] = zp[r] * (t1 - t2);
      }
  }
}

static void admm_project_sigma_d(const zro_sim* s, int64_t i, const double* sig, double* zp, double D[3][3])
{
  project_sigma_d(s->mu[i], s->lam[i], s->plastic ? s->plastic[i] : ZRO_PLASTIC_NONE,
                  s->pparam ? s->pparam + 4 * i : NULL, s->pstate ? s->pstate[i] : 0.0, sig, zp, D);
}

void zro_project_sigma(double mu, double lam, int plastic, const double* pparam, double pstate, const double* sig,
                       double* out)
{
  project_sigma(mu, lam, plastic, pparam, pstate, sig, out);
}

void zro_project_sigma_and_derivative(double mu, double lam, int plastic, const double* pparam, double pstate,
                                      const double* sig, double* zp, double* D)
{
  double M[3][3];
  project_sigma_d(mu, lam, plastic, pparam, pstate, sig, zp, M);
  memcpy(D, M, sizeof(M));
}

/* initOmega (AdmmSimulationFoam.h:222-274) for one particle with F = U S V^T */
static double admm_omega_of(const zro_sim* s, int64_t i, const double* sig)
{
  double mu = s->mu[i], lam = s->lam[i];
  double A[3][3], ps[3], B[2][2];
  hencky_ddE(mu, lam, sig, A);
  double total = pd_norm2_3(A);
  admm_project_sigma(s, i, sig, ps);
  const int pr[3][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 } };
  for (int k = 0; k < 3; ++k) {
    hencky_Bij(mu, lam, ps, pr[k][0], pr[k][1], 1e-6, B);
    total += pd_norm2_2(B[0][0], B[0][1], B[1][1]);
  }
  return sqrt(s->vol[i] * sqrt(total));
}

/* initOmega's second half (AdmmSimulationFoam.h:254-270) for the viscoelastic
   branch's F2 = U S V^T: the StvkWithHencky model of that branch, no
   projection */
static double admm_visco_omega_of(const zro_sim* s, int64_t i, const double* sig)
{
  double mu = s->mu_neq[i], lam = s->lam_neq[i];
  double A[3][3], B[2][2];
  hencky_ddE(mu, lam, sig, A);
  double total = pd_norm2_3(A);
  const int pr[3][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 } };
  for (int k = 0; k < 3; ++k) {
    hencky_Bij(mu, lam, sig, pr[k][0], pr[k][1], 1e-6, B);
    total += pd_norm2_2(B[0][0], B[0][1], B[1][1]);
  }
  return sqrt(s->vol[i] * sqrt(total));
}

/* The viscoelastic branch's return map inside FStepSolve (AdmmSimulationFoam.h:
   583-603): the relaxed singular values zp = exp(eps) with
   eps = (log|z| - c tr(log|z|) 1) / (1 + dt alpha), c = dt beta / (1 + dt (alpha
   + 3 beta)), where tr of the 3-vector is Eigen's trace() of a 3x1 matrix, its
   first entry; and the derivative the reference forms,
   D = diag(zp) (diag(1/z) - c 1 (1/z)^T) / (1 + dt alpha) (the full sum of the
   log's derivatives, as written there). */
void zro_admm_visco_map(double mu, double lam, double visc_d, double visc_v, double dt, const double* z, double* zp,
                        double D[3][3])
{
  const double alpha = 2.0 * mu / visc_d;
  const double beta = 2.0 * (2.0 * mu + lam * 3.0) / (9.0 * visc_v) - 2.0 * mu / (visc_d * 3.0);
  const double c = dt * beta / (1.0 + dt * (alpha + 3.0 * beta));
  const double h = 1.0 / (1.0 + dt * alpha);
  double e[3];
  for (int d = 0; d < 3; ++d) e[d] = log(fabs(z[d]));
  const double tr = e[0];
  for (int d = 0; d < 3; ++d) zp[d] = exp(h * (e[d] - c * tr));
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) D[a][b] = zp[a] * h * ((a == b ? 1.0 / z[b] : 0.0) - c / z[b]);
}

/* multiObjectCollisionWithFakeSeparate (CollisionObject.cpp:108-160): resolves v,
   returns 1 and the projection P = I - K K^T when any collider acts.

   A separating object releases the node when its relative normal velocity
   (v - v_obj).n is positive.  Inside the ADMM loop that test is repeated on
   nodes the previous iteration projected: their relative normal velocity is
   then exactly zero (the projection removed it and the global step, whose
   preconditioner is a scalar per node, AdmmSimulationFoam.h:476-503, only
   moves dv within the projected space), so in floating point the reference
   decides it by rounding noise, and a 1e-15 perturbation of the input changes
   the step's result by percent.  The test here is against a rounding bound,
   (v - v_obj).n > kSeparateRoundoff sum_d (|v_d| + |v_obj,d|), which gives the
   exact-arithmetic outcome: a node with zero relative normal velocity stays
   in contact. */
#define ZRO_SEPARATE_ROUNDOFF 1e-12
int zro_fake_separate(const zro_collider* cols, int ncol, double t, const double* x, double* v, double* Pout)
{
  int any = 0, nslip = 0;
  double K[3][3], P[3][3];
  for (int k = 0; k < ncol; ++k) {
    const zro_collider* c = &cols[k];
    double xb[3], phi, n[3], vo[3], old[3] = { v[0], v[1], v[2] };
    zro_frame fr;
    collider_frame(c, t, &fr);
    if (!collider_query(c, &fr, x, &phi, n, xb)) continue;
    collider_velocity(c, &fr, xb, vo);
    int col = detect_and_resolve(c, &fr, x, v);
    if (col && c->type == 3) {
      double dot = 0.0, scale = 0.0;
      for (int d = 0; d < 3; ++d) {
        dot += (old[d] - vo[d]) * n[d];
        scale += fabs(old[d]) + fabs(vo[d]);
      }
      if (dot > ZRO_SEPARATE_ROUNDOFF * scale) {
        for (int d = 0; d < 3; ++d) v[d] = old[d];
        col = 0;
      }
    }
    any = any || col;
    if (!col) continue;
    if (c->type == 1) {
      memset(Pout, 0, sizeof(double) * 9);
      return 1;
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
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      P[i][j] = i == j ? 1.0 : 0.0;
      for (int q = 0; q < nslip; ++q) P[i][j] -= K[q][i] * K[q][j];
    }
  memcpy(Pout, P, sizeof(P));
  return any;
}

static int admm_collision(const zro_sim* s, double t, const double* x, double* v, double P[3][3])
{
  return zro_fake_separate(s->col, s->ncol, t, x, v, &P[0][0]);
}

/* updateCollisionNodes (AdmmSimulationFoam.h:276-294) */
static void admm_update_collisions(zro_sim* s)
{
  for (int64_t p = 0; p < s->npages; ++p) {
    int base[3];
    zro_linear_to_coord(s->pages[p], base);
    for (int e = 0; e < 32; ++e) {
      int64_t k = p * 32 + e;
      s->admm_coll[k] = 0;
      if (s->nodes[k].idx < 0) continue;
      double xi[3] = { (double)(base[0] + ((e >> 4) & 1)) * s->dx, (double)(base[1] + ((e >> 2) & 3)) * s->dx,
                       (double)(base[2] + (e & 3)) * s->dx };
      double vi[3], P[3][3];
      for (int d = 0; d < 3; ++d) vi[d] = s->admm_dv[3 * k + d] + s->admm_vn[3 * k + d];
      if (admm_collision(s, s->time, xi, vi, P)) {
        s->admm_coll[k] = 1;
        memcpy(s->admm_proj + 9 * k, P, sizeof(P));
        for (int d = 0; d < 3; ++d) s->admm_dv[3 * k + d] = vi[d] - s->admm_vn[3 * k + d];
      }
    }
  }
}

static void admm_project(const zro_sim* s, double* r)
{
  int64_t ns = s->npages * 32;
  for (int64_t k = 0; k < ns; ++k) {
    if (!s->admm_coll[k]) continue;
    const double* P = s->admm_proj + 9 * k;
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = P[3 * i] * r[3 * k] + P[3 * i + 1] * r[3 * k + 1] + P[3 * i + 2] * r[3 * k + 2];
    for (int i = 0; i < 3; ++i) r[3 * k + i] = t[i];
  }
}

/* sum over the particle's active stencil nodes of vel_i grad w_i^T (3x3) */
static void admm_grad(zro_sim* s, int64_t i, const double* vel, double G[3][3])
{
  zro_spline sp;
  spline_of(s, s->X + 3 * i, &sp);
  memset(G, 0, sizeof(double) * 9);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double gw[3];
        spline_node(s, &sp, a, b, c, gw);
        int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
        if (s->nodes[k].idx < 0) continue;
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) G[r][cc] += vel[3 * k + r] * gw[cc];
      }
}

/* out_i += S Fn^T grad w_i over the particle's active nodes */
static void admm_scatter(zro_sim* s, int64_t i, const double S[3][3], const double Fn[3][3], double* out)
{
  zro_spline sp;
  spline_of(s, s->X + 3 * i, &sp);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double gw[3], fg[3];
        spline_node(s, &sp, a, b, c, gw);
        int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
        if (s->nodes[k].idx < 0) continue;
        for (int d = 0; d < 3; ++d) fg[d] = Fn[0][d] * gw[0] + Fn[1][d] * gw[1] + Fn[2][d] * gw[2];
        for (int r = 0; r < 3; ++r) out[3 * k + r] += S[r][0] * fg[0] + S[r][1] * fg[1] + S[r][2] * fg[2];
      }
}

/* the global matrix (initMultiplierAndPreconditioner, :407-505) applied
   matrix-free: b_i = m_i x_i + sum_p c_p (sum_j x_j grad w_j^T Fn) Fn^T grad w_i */
static void admm_multiply(zro_sim* s, double dt, double rho, const double* x, double* b)
{
  int64_t ns = s->npages * 32;
  memset(b, 0, sizeof(double) * (size_t)(3 * ns));
  for (int64_t k = 0; k < ns; ++k)
    if (s->nodes[k].idx >= 0)
      for (int d = 0; d < 3; ++d) b[3 * k + d] = s->nodes[k].m * x[3 * k + d];
  PF_FOR_PARTICLES(s, i)
  {
    double G[3][3], Fn[3][3], Y[3][3];
    mat_from_cm(s->F + 9 * (int64_t)i, Fn);
    admm_grad(s, i, x, G);
    double cp = rho * s->admm_omega[i] * s->admm_omega[i] * dt * dt;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Y[r][c] = cp * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
    admm_scatter(s, i, Y, Fn, b);
    if (s->fneq) { /* the viscoelastic branch's term with FFn (:449-471) */
      double Fv[3][3];
      mat_from_cm(s->fneq + 9 * (int64_t)i, Fv);
      double om2 = s->admm_omega[s->n + i], c2 = rho * om2 * om2 * dt * dt;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Y[r][c] = c2 * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
      admm_scatter(s, i, Y, Fv, b);
    }
  }
}

static double admm_dot(const zro_sim* s, const double* a, const double* b)
{
  double acc = 0.0;
  for (int64_t k = 0; k < s->npages * 32; ++k)
    if (s->nodes[k].idx >= 0) acc += a[3 * k] * b[3 * k] + a[3 * k + 1] * b[3 * k + 1] + a[3 * k + 2] * b[3 * k + 2];
  return acc;
}

/* Minres::solve (Minres.h:69-178) with the preconditioner in / diag */
static int admm_minres(zro_sim* s, double dt, double rho, double* x, const double* b, double tol, double rel,
                       int max_it)
{
  int64_t n3 = 3 * s->npages * 32;
  size_t bytes = sizeof(double) * (size_t)n3;
  double *mk = calloc(n3, 8), *mkm1 = calloc(n3, 8), *mkm2 = calloc(n3, 8), *z = calloc(n3, 8);
  double *qkm1 = calloc(n3, 8), *qk = calloc(n3, 8), *qkp1 = calloc(n3, 8), *tmp;
  zro_g2 Gk = g2_make(), Gkm1 = g2_make(), Gkm2 = g2_make();
  double gamma = 0, delta = 0, epsilon = 0, beta_kp1 = 0, alpha_k = 0, beta_k = 0, tk = 0, lt[2];
  admm_multiply(s, dt, rho, x, qkp1);
  for (int64_t k = 0; k < n3; ++k) qkp1[k] = b[k] - qkp1[k];
  admm_project(s, qkp1);
  for (int64_t k = 0; k < n3; ++k) z[k] = s->nodes[k / 3].idx >= 0 ? qkp1[k] / s->admm_diag[k / 3] : 0.0;
  double res = sqrt(admm_dot(s, z, qkp1));
  beta_kp1 = res;
  double local_tol = rel * res < tol ? rel * res : tol;
  int it = 0;
  if (res < local_tol) goto done;
  if (res > 0)
    for (int64_t k = 0; k < n3; ++k) { qkp1[k] /= beta_kp1; z[k] /= beta_kp1; }
  lt[0] = res;
  lt[1] = 0.0;
  for (it = 0; it < max_it; ++it) {
    if (res < local_tol) break;
    tmp = mkm2; mkm2 = mkm1; mkm1 = mk; mk = tmp;
    memcpy(mk, z, bytes);
    beta_k = beta_kp1;
    tmp = qkm1; qkm1 = qk; qk = qkp1; qkp1 = tmp;
    admm_multiply(s, dt, rho, mk, qkp1);
    admm_project(s, qkp1);
    alpha_k = admm_dot(s, mk, qkp1);
    for (int64_t k = 0; k < n3; ++k) qkp1[k] -= alpha_k * qk[k];
    for (int64_t k = 0; k < n3; ++k) qkp1[k] -= beta_k * qkm1[k];
    for (int64_t k = 0; k < n3; ++k) z[k] = s->nodes[k / 3].idx >= 0 ? qkp1[k] / s->admm_diag[k / 3] : 0.0;
    double bb = admm_dot(s, z, qkp1);
    beta_kp1 = sqrt(bb > 0.0 ? bb : 0.0);
    if (beta_kp1 > 0)
      for (int64_t k = 0; k < n3; ++k) { qkp1[k] /= beta_kp1; z[k] /= beta_kp1; }
    /* applyAllPreviousGivensRotationsAndDetermineNewGivens */
    Gkm2 = Gkm1;
    Gkm1 = Gk;
    double ek[2] = { 0.0, beta_k };
    g2_rot(&Gkm2, ek);
    epsilon = ek[0];
    double dk[2] = { ek[1], alpha_k };
    g2_rot(&Gkm1, dk);
    delta = dk[0];
    double tv[2] = { dk[1], beta_kp1 };
    g2_compute(&Gk, tv[0], tv[1]);
    g2_rot(&Gk, tv);
    gamma = tv[0];
    g2_rot(&Gk, lt);
    tk = lt[0];
    double r = lt[1];
    lt[0] = r;
    lt[1] = 0.0;
    res = r < 0 ? -r : r;
    for (int64_t k = 0; k < n3; ++k) mk[k] = (mk[k] - delta * mkm1[k] - epsilon * mkm2[k]) / gamma;
    for (int64_t k = 0; k < n3; ++k) x[k] += tk * mk[k];
  }
done:
  free(mk); free(mkm1); free(mkm2); free(z); free(qkm1); free(qk); free(qkp1);
  return it;
}

/* 3x3 inverse by cofactors */
static void inv3(const double P[3][3], double Pi[3][3])
{
  double det = P[0][0] * (P[1][1] * P[2][2] - P[2][1] * P[1][2]) - P[0][1] * (P[1][0] * P[2][2] - P[2][0] * P[1][2]) +
               P[0][2] * (P[1][0] * P[2][1] - P[2][0] * P[1][1]);
  Pi[0][0] = (P[1][1] * P[2][2] - P[1][2] * P[2][1]) / det;
  Pi[0][1] = (P[0][2] * P[2][1] - P[0][1] * P[2][2]) / det;
  Pi[0][2] = (P[0][1] * P[1][2] - P[0][2] * P[1][1]) / det;
  Pi[1][0] = (P[1][2] * P[2][0] - P[1][0] * P[2][2]) / det;
  Pi[1][1] = (P[0][0] * P[2][2] - P[0][2] * P[2][0]) / det;
  Pi[1][2] = (P[0][2] * P[1][0] - P[0][0] * P[1][2]) / det;
  Pi[2][0] = (P[1][0] * P[2][1] - P[1][1] * P[2][0]) / det;
  Pi[2][1] = (P[0][1] * P[2][0] - P[0][0] * P[2][1]) / det;
  Pi[2][2] = (P[0][0] * P[1][1] - P[0][1] * P[1][0]) / det;
}

/* FStepSolve's viscoelastic branch (AdmmSimulationFoam.h:566-611): F2 = U
   diag(z) V^T with z the Newton solution (at most 20 iterations) of
   vol dE(zp(z)) + rho omega2^2 (z - sigma) = 0, zp the relaxation map of
   zro_admm_visco_map with the branch's StvkWithHencky model, sigma the
   singular values of FFn - y2 / (rho omega2) + dt G FFn; G is the particle's
   grad(vn + dv). */
static void admm_visco_fstep(zro_sim* s, int64_t i, const double G[3][3], double dt, double rho, double local_tol)
{
  const int64_t j = s->n + i;
  double Fv[3][3], FD[3][3], Fa[3][3];
  mat_from_cm(s->fneq + 9 * i, Fv);
  mat_from_cm(s->admm_F + 9 * j, Fa);
  double om = s->admm_omega[j];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      FD[r][c] = Fv[r][c] - s->admm_y[9 * j + c * 3 + r] / (om * rho) +
                 dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
  double fdc[9], U[9], sg[3], V[9], Um[3][3], Vm[3][3], zi[3];
  mat_to_cm(FD, fdc);
  zro_svd3(fdc, U, sg, V);
  mat_from_cm(U, Um);
  mat_from_cm(V, Vm);
  for (int d = 0; d < 3; ++d) {
    double acc = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) acc += Um[a][d] * Fa[a][b] * Vm[b][d];
    zi[d] = acc;
  }
  const double mu = s->mu_neq[i], lam = s->lam_neq[i], coeff = rho * om * om, vol = s->vol[i];
  for (int it = 0; it < 20; ++it) {
    double zp[3], D[3][3], dE[3], ddE[3][3], g[3], P[3][3], Pi[3][3], step[3];
    zro_admm_visco_map(mu, lam, s->visc_d, s->visc_v, dt, zi, zp, D);
    hencky_dE(mu, lam, zp, dE);
    hencky_ddE(mu, lam, zp, ddE);
    for (int a = 0; a < 3; ++a) {
      g[a] = vol * dE[a] + coeff * (zi[a] - sg[a]);
      for (int b = 0; b < 3; ++b)
        P[a][b] = vol * (ddE[a][0] * D[0][b] + ddE[a][1] * D[1][b] + ddE[a][2] * D[2][b]) + (a == b ? coeff : 0.0);
    }
    inv3(P, Pi);
    double sn = 0.0;
    for (int a = 0; a < 3; ++a) {
      step[a] = -(Pi[a][0] * g[0] + Pi[a][1] * g[1] + Pi[a][2] * g[2]);
      sn += step[a] * step[a];
    }
    for (int a = 0; a < 3; ++a) zi[a] += step[a];
    if (sqrt(sn) < local_tol) break;
  }
  double Fo[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Fo[r][c] = Um[r][0] * zi[0] * Vm[c][0] + Um[r][1] * zi[1] * Vm[c][1] + Um[r][2] * zi[2] * Vm[c][2];
  mat_to_cm(Fo, s->admm_F + 9 * j);
}

static double* admm_grow(double* p, int64_t n) { free(p); return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

int zro_step_admm(zro_sim* s, double dt, const zro_admm_opts* o, zro_admm_stats* st)
{
  const double rho = o->rho_scale;
  int64_t n = s->n;
  if (st) memset(st, 0, sizeof(*st));
  zro_sort_and_pollute(s);
  zro_p2g_apic(s);
  int64_t ns = s->npages * 32;
  if (s->admm_cap < ns) {
    s->admm_vn = admm_grow(s->admm_vn, 3 * ns);
    s->admm_dv = admm_grow(s->admm_dv, 3 * ns);
    s->admm_proj = admm_grow(s->admm_proj, 9 * ns);
    s->admm_diag = admm_grow(s->admm_diag, ns);
    free(s->admm_coll);
    s->admm_coll = (char*)calloc((size_t)ns, 1);
    s->admm_cap = ns;
  }
  /* the viscoelastic branch (enable_visco) keeps a second F, y and omega per
     particle at i + n, as the reference's arrays of 2 * particles.count; its
     elastic branch is StvkWithHenckyIsotropic without plasticity
     (use_elasticity_plasticity = false) */
  const int visco = s->fneq != NULL;
  if (visco && s->plastic)
    for (int64_t i = 0; i < n; ++i)
      if (s->plastic[i] != ZRO_PLASTIC_NONE) return -2;
  const int64_t nb = visco ? 2 * n : n;
  int first = s->admm_steps == 0 || s->admm_n != n || s->admm_visco != visco;
  if (s->admm_n != n || s->admm_visco != visco) {
    s->admm_F = admm_grow(s->admm_F, 9 * nb);
    s->admm_y = admm_grow(s->admm_y, 9 * nb);
    s->admm_omega = admm_grow(s->admm_omega, nb);
    s->admm_last_dv = admm_grow(s->admm_last_dv, 3 * n);
    s->admm_n = n;
    s->admm_visco = visco;
  }
  /* initStep (:296-401) */
  memset(s->admm_vn, 0, sizeof(double) * 3 * ns);
  memset(s->admm_dv, 0, sizeof(double) * 3 * ns);
  for (int64_t k = 0; k < ns; ++k)
    if (s->nodes[k].idx >= 0)
      for (int d = 0; d < 3; ++d) s->admm_vn[3 * k + d] = s->nodes[k].v[d];
  if (!first) {
    double* wsum = (double*)calloc((size_t)ns, sizeof(double));
    PF_FOR_PARTICLES(s, i)
    {
      zro_spline sp;
      spline_of(s, s->X + 3 * i, &sp);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int c = 0; c < 3; ++c) {
            double gw[3];
            double w = spline_node(s, &sp, a, b, c, gw);
            int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
            if (s->nodes[k].idx < 0) continue;
            for (int d = 0; d < 3; ++d) s->admm_dv[3 * k + d] += s->admm_last_dv[3 * i + d] * w;
            wsum[k] += w;
          }
    }
    for (int64_t k = 0; k < ns; ++k)
      if (s->nodes[k].idx >= 0)
        for (int d = 0; d < 3; ++d) s->admm_dv[3 * k + d] = s->admm_dv[3 * k + d] * dt / wsum[k];
    free(wsum);
  }
  admm_update_collisions(s);
  double* nv = (double*)calloc((size_t)(3 * ns), sizeof(double));
  for (int64_t k = 0; k < 3 * ns; ++k) nv[k] = s->admm_vn[k] + s->admm_dv[k];
  for (int64_t i = 0; i < n; ++i) {
    double G[3][3], Fn[3][3], F[3][3];
    mat_from_cm(s->F + 9 * i, Fn);
    admm_grad(s, i, nv, G);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) F[r][c] = Fn[r][c] + dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
    mat_to_cm(F, s->admm_F + 9 * i);
    double U[9], sg[3], V[9];
    zro_svd3(s->admm_F + 9 * i, U, sg, V);
    double om = admm_omega_of(s, i, sg);
    double old = first ? 0.0 : s->admm_omega[i];
    for (int c = 0; c < 9; ++c) s->admm_y[9 * i + c] = first ? 0.0 : s->admm_y[9 * i + c] * old / om;
    s->admm_omega[i] = om;
    if (visco) { /* F2 = (I + dt grad(vn + dv)) FFn, its omega and y (:372-400) */
      const int64_t j = n + i;
      double Fv[3][3], F2[3][3];
      mat_from_cm(s->fneq + 9 * i, Fv);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) F2[r][c] = Fv[r][c] + dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
      mat_to_cm(F2, s->admm_F + 9 * j);
      zro_svd3(s->admm_F + 9 * j, U, sg, V);
      double om2 = admm_visco_omega_of(s, i, sg);
      double old2 = first ? 0.0 : s->admm_omega[j];
      for (int c = 0; c < 9; ++c) s->admm_y[9 * j + c] = first ? 0.0 : s->admm_y[9 * j + c] * old2 / om2;
      s->admm_omega[j] = om2;
    }
  }
  /* preconditioner diagonal */
  for (int64_t k = 0; k < ns; ++k) s->admm_diag[k] = s->nodes[k].idx >= 0 ? s->nodes[k].m : 0.0;
  PF_FOR_PARTICLES(s, i)
  {
    double Fn[3][3];
    mat_from_cm(s->F + 9 * (int64_t)i, Fn);
    double cp = rho * s->admm_omega[i] * s->admm_omega[i] * dt * dt;
    zro_spline sp;
    spline_of(s, s->X + 3 * i, &sp);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3], fg[3];
          spline_node(s, &sp, a, b, c, gw);
          int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
          if (s->nodes[k].idx < 0) continue;
          for (int d = 0; d < 3; ++d) fg[d] = Fn[0][d] * gw[0] + Fn[1][d] * gw[1] + Fn[2][d] * gw[2];
          s->admm_diag[k] += cp * (fg[0] * fg[0] + fg[1] * fg[1] + fg[2] * fg[2]);
          if (visco) { /* :488-497 */
            double Fv[3][3], om2 = s->admm_omega[n + i];
            mat_from_cm(s->fneq + 9 * (int64_t)i, Fv);
            for (int d = 0; d < 3; ++d) fg[d] = Fv[0][d] * gw[0] + Fv[1][d] * gw[1] + Fv[2][d] * gw[2];
            s->admm_diag[k] += rho * om2 * om2 * dt * dt * (fg[0] * fg[0] + fg[1] * fg[1] + fg[2] * fg[2]);
          }
        }
  }
  double* rhs = (double*)calloc((size_t)(3 * ns), sizeof(double));
  for (int tim = 0; tim < o->admm_iterations; ++tim) {
    admm_update_collisions(s);
    /* FStep (:507-636) */
    for (int64_t k = 0; k < 3 * ns; ++k) nv[k] = s->admm_vn[k] + s->admm_dv[k];
    for (int64_t i = 0; i < n; ++i) {
      double G[3][3], Fn[3][3], FD[3][3], Fa[3][3];
      mat_from_cm(s->F + 9 * i, Fn);
      mat_from_cm(s->admm_F + 9 * i, Fa);
      admm_grad(s, i, nv, G);
      double om = s->admm_omega[i];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          FD[r][c] = Fn[r][c] - s->admm_y[9 * i + c * 3 + r] / (om * rho) +
                     dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
      double fdc[9], U[9], sg[3], V[9];
      mat_to_cm(FD, fdc);
      zro_svd3(fdc, U, sg, V);
      double Um[3][3], Vm[3][3], zi[3];
      mat_from_cm(U, Um);
      mat_from_cm(V, Vm);
      for (int d = 0; d < 3; ++d) {
        double acc = 0.0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) acc += Um[a][d] * Fa[a][b] * Vm[b][d];
        zi[d] = acc;
      }
      double coeff = rho * om * om, vol = s->vol[i];
      for (int j = 0; j < 5; ++j) {
        double zp[3], D[3][3], dE[3], ddE[3][3], g[3], P[3][3], Pi[3][3], step[3];
        admm_project_sigma_d(s, i, zi, zp, D);
        hencky_dE(s->mu[i], s->lam[i], zp, dE);
        hencky_ddE(s->mu[i], s->lam[i], zp, ddE);
        for (int a = 0; a < 3; ++a) {
          g[a] = vol * dE[a] + coeff * (zi[a] - sg[a]);
          for (int b = 0; b < 3; ++b)
            P[a][b] = vol * (ddE[a][0] * D[0][b] + ddE[a][1] * D[1][b] + ddE[a][2] * D[2][b]) + (a == b ? coeff : 0.0);
        }
        /* 3x3 inverse by cofactors */
        double det = P[0][0] * (P[1][1] * P[2][2] - P[2][1] * P[1][2]) - P[0][1] * (P[1][0] * P[2][2] - P[2][0] * P[1][2]) +
                     P[0][2] * (P[1][0] * P[2][1] - P[2][0] * P[1][1]);
        Pi[0][0] = (P[1][1] * P[2][2] - P[1][2] * P[2][1]) / det;
        Pi[0][1] = (P[0][2] * P[2][1] - P[0][1] * P[2][2]) / det;
        Pi[0][2] = (P[0][1] * P[1][2] - P[0][2] * P[1][1]) / det;
        Pi[1][0] = (P[1][2] * P[2][0] - P[1][0] * P[2][2]) / det;
        Pi[1][1] = (P[0][0] * P[2][2] - P[0][2] * P[2][0]) / det;
        Pi[1][2] = (P[0][2] * P[1][0] - P[0][0] * P[1][2]) / det;
        Pi[2][0] = (P[1][0] * P[2][1] - P[1][1] * P[2][0]) / det;
        Pi[2][1] = (P[0][1] * P[2][0] - P[0][0] * P[2][1]) / det;
        Pi[2][2] = (P[0][0] * P[1][1] - P[0][1] * P[1][0]) / det;
        double sn = 0.0;
        for (int a = 0; a < 3; ++a) {
          step[a] = -(Pi[a][0] * g[0] + Pi[a][1] * g[1] + Pi[a][2] * g[2]);
          sn += step[a] * step[a];
        }
        for (int a = 0; a < 3; ++a) zi[a] += step[a];
        if (sqrt(sn) < o->local_tolerance) break;
      }
      double Fo[3][3];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Fo[r][c] = Um[r][0] * zi[0] * Vm[c][0] + Um[r][1] * zi[1] * Vm[c][1] + Um[r][2] * zi[2] * Vm[c][2];
      mat_to_cm(Fo, s->admm_F + 9 * i);
      if (visco) admm_visco_fstep(s, i, G, dt, rho, o->local_tolerance);
    }
    /* dvStep (:163-220) */
    memset(rhs, 0, sizeof(double) * 3 * ns);
    for (int64_t k = 0; k < ns; ++k)
      if (s->nodes[k].idx >= 0)
        for (int d = 0; d < 3; ++d) rhs[3 * k + d] = dt * s->nodes[k].m * s->gravity[d];
    PF_FOR_PARTICLES(s, i)
    {
      double G[3][3], Fn[3][3], bF[3][3];
      mat_from_cm(s->F + 9 * (int64_t)i, Fn);
      admm_grad(s, i, s->admm_vn, G);
      double om = s->admm_omega[i];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double fbu = s->admm_F[9 * (int64_t)i + c * 3 + r] - Fn[r][c] + s->admm_y[9 * (int64_t)i + c * 3 + r] / (om * rho) -
                       dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
          bF[r][c] = dt * rho * om * om * fbu;
        }
      admm_scatter(s, i, bF, Fn, rhs);
      if (visco) { /* :183-201 */
        const int64_t j = n + i;
        double Fv[3][3], om2 = s->admm_omega[j];
        mat_from_cm(s->fneq + 9 * (int64_t)i, Fv);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double fbu = s->admm_F[9 * j + c * 3 + r] - Fv[r][c] + s->admm_y[9 * j + c * 3 + r] / (om2 * rho) -
                         dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
            bF[r][c] = dt * rho * om2 * om2 * fbu;
          }
        admm_scatter(s, i, bF, Fv, rhs);
      }
    }
    double rel = 0.5;  /* min(0.5, sqrt(max(residual = 1, 1e-6))) (:785-787) */
    int its = admm_minres(s, dt, rho, s->admm_dv, rhs, o->global_tolerance, rel, o->minres_max_iterations);
    if (st) st->minres_iterations += its;
    /* uStep (:638-671) */
    for (int64_t k = 0; k < 3 * ns; ++k) nv[k] = s->admm_vn[k] + s->admm_dv[k];
    for (int64_t i = 0; i < n; ++i) {
      double G[3][3], Fn[3][3];
      mat_from_cm(s->F + 9 * i, Fn);
      admm_grad(s, i, nv, G);
      double om = s->admm_omega[i];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double t = s->admm_F[9 * i + c * 3 + r] - Fn[r][c] - dt * (G[r][0] * Fn[0][c] + G[r][1] * Fn[1][c] + G[r][2] * Fn[2][c]);
          s->admm_y[9 * i + c * 3 + r] += rho * om * t;
        }
      if (visco) { /* :661-668 */
        const int64_t j = n + i;
        double Fv[3][3], om2 = s->admm_omega[j];
        mat_from_cm(s->fneq + 9 * i, Fv);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double t = s->admm_F[9 * j + c * 3 + r] - Fv[r][c] - dt * (G[r][0] * Fv[0][c] + G[r][1] * Fv[1][c] + G[r][2] * Fv[2][c]);
            s->admm_y[9 * j + c * 3 + r] += rho * om2 * t;
          }
      }
    }
  }
  /* constructNewVelocityFromNewtonResult and last_dv (:789-800) */
  for (int64_t k = 0; k < ns; ++k)
    for (int d = 0; d < 3; ++d)
      s->nodes[k].new_v[d] = s->nodes[k].idx >= 0 ? s->admm_vn[3 * k + d] + s->admm_dv[3 * k + d] : 0.0;
  for (int64_t i = 0; i < n; ++i) {
    zro_spline sp;
    spline_of(s, s->X + 3 * i, &sp);
    double acc[3] = { 0, 0, 0 }, ws = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          double gw[3];
          double w = spline_node(s, &sp, a, b, c, gw);
          int64_t k = node_slot(s, sp.base[0] + a, sp.base[1] + b, sp.base[2] + c);
          if (s->nodes[k].idx < 0) continue;
          for (int d = 0; d < 3; ++d) acc[d] += s->admm_dv[3 * k + d] * w;
          ws += w;
        }
    for (int d = 0; d < 3; ++d) s->admm_last_dv[3 * i + d] = acc[d] / (ws * dt);
  }
  free(rhs);
  free(nv);
  s->admm_steps++;
  int flags = zro_g2p_apic(s, dt);
  if (visco) /* gridToParticles' evolveStrain of the viscoelastic helper, with gradV */
    for (int64_t i = 0; i < n; ++i)
      zro_visco_relax(s->mu_neq[i], s->lam_neq[i], s->visc_d, s->visc_v, dt, s->gradV + 9 * i, s->fneq + 9 * i);
  s->time += dt;
  return flags;
}

void zro_get_admm(const zro_sim* s, double* F, double* y, double* omega)
{
  if (!s->admm_F) return;
  if (F) memcpy(F, s->admm_F, sizeof(double) * 9 * (size_t)s->n);
  if (y) memcpy(y, s->admm_y, sizeof(double) * 9 * (size_t)s->n);
  if (omega) memcpy(omega, s->admm_omega, sizeof(double) * (size_t)s->n);
}
