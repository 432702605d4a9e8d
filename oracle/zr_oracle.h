/*
 * zr_oracle.h -- FP64 CPU oracle for the MPM time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (penn-graphics-research/ziran2019) used as the checker for the
 * HIP engine.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * checks each function against something other than its own formula) lives in
 * tests/test_oracle_*.py; see the per-function notes there.
 *
 * Conventions (shared with the C-ABI in include/zr_mpm.h):
 *   - vectors are 3 doubles per particle, particle-major (Eigen StdVector
 *     layout of Vector<double,3>);
 *   - 3x3 matrices are 9 doubles per particle in COLUMN-MAJOR order (Eigen's
 *     default storage of Matrix<double,3,3>): m[c*3+r] = M(r,c).
 */
#ifndef ZR_ORACLE_H
#define ZR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SPGrid (Lib/SPGrid/Core/SPGrid_Mask.h, GridState<double,3>) ---- */
uint64_t zro_linear_offset(int i, int j, int k);
void zro_linear_to_coord(uint64_t off, int* ijk);
uint64_t zro_packed_add(uint64_t a, uint64_t b);

/* ---- quadratic B-spline (Lib/Ziran/Math/Splines/BSplines.h:55-81) ---- */
void zro_bspline_weights(double x, int* base, double* w3, double* dw3);

/* ---- 3x3 SVD (Lib/Ziran/Math/Linear/ImplicitQRSVD.h:355-533) ---- */
int zro_svd3(const double* F, double* U, double* sigma, double* V);
void zro_polar3(const double* F, double* R, double* S);

/* ---- fixed-corotated model (textbook; no counterpart in the reference) --- */
double zro_fcr_psi(double mu, double lam, const double* F);
void zro_fcr_first_piola(double mu, double lam, const double* F, double* P);
void zro_fcr_kirchhoff(double mu, double lam, const double* F, double* tau);
void zro_fcr_first_piola_differential(double mu, double lam, const double* F,
                                      const double* dF, double* dP);
void zro_lame(double E, double nu, double* mu, double* lam);
/* reverse of tau(F; mu, lam): Fbar += (dtau/dF)^T taubar; returns mubar, lambar */
void zro_fcr_kirchhoff_adjoint(double mu, double lam, const double* F, const double* taubar, double* Fbar,
                               double* mubar, double* lambar);

/* ---- StvkWithHencky (StvkWithHencky.cpp:16-143) ---- */
double zro_stvk_hencky_psi(double mu, double lam, const double* F);
void zro_stvk_hencky_first_piola(double mu, double lam, const double* F, double* P);
void zro_stvk_hencky_kirchhoff(double mu, double lam, const double* F, double* tau);

/* ---- NeoHookeanBorden with degradation g (NeoHookeanBorden.cpp:16-86) ---- */
double zro_nhb_psi(double mu, double lam, double g, const double* F);
double zro_nhb_psi_pos(double mu, double lam, const double* F);
void zro_nhb_kirchhoff(double mu, double lam, double g, const double* F, double* tau);

/* ---- return mapping (PlasticityApplier.cpp:249-316, 443-501) ---- */
int zro_von_mises_project(double mu, double lam, double xi, double fail_stress, double* yield_stress, double* F,
                          int* broken);
double zro_drucker_prager_alpha(double friction_angle_deg);
int zro_drucker_prager_project(double mu, double lam, double alpha, double beta, double cohesion,
                               int volume_correction, double* logJp, double* F);
/* NonAssociativeCamClay (PlasticityApplier.cpp:10-19, 36-135) on a
   NeoHookeanBorden model; M from the friction angle */
double zro_cam_clay_M(double friction_angle_deg, int dim);
int zro_cam_clay_project(double mu, double lam, double M, double beta, double xi, double* logJp, double* F);

/* material codes (same values as ZR_MODEL_* / ZR_PLASTIC_* of the engine) */
#define ZRO_MODEL_FIXED_COROTATED 0
#define ZRO_MODEL_STVK_HENCKY 1
#define ZRO_MODEL_NEO_HOOKEAN_BORDEN 2
#define ZRO_PLASTIC_NONE 0
#define ZRO_PLASTIC_VON_MISES 1
#define ZRO_PLASTIC_DRUCKER_PRAGER 2
#define ZRO_PLASTIC_CAM_CLAY 3

/* ---- simulation state ---- */
typedef struct zro_sim zro_sim;

/* collider description, identical field meaning to zr_collider */
typedef struct {
  int32_t shape;        /* 0 half-space, 1 sphere */
  int32_t type;         /* 1 STICKY, 2 SLIP, 3 SEPARATE (CollisionObject.h:47-52) */
  double friction;
  double origin[3];     /* half-space point / sphere centre at t=0 */
  double normal[3];     /* half-space outward normal (normalised here) */
  double radius;
  double velocity[3];   /* constant translation velocity (dbdt) */
  double translation[3];/* b at t = 0: b(t) = translation + velocity t */
  double rotation[4];   /* R as a quaternion <w,x,y,z>; all 0: identity */
  double omega[3];      /* angular velocity */
  double scale;         /* s; 0: 1 */
  double dscale_dt;     /* ds/dt */
} zro_collider;

zro_sim* zro_create(double dx, const double* gravity, double cfl, int nthreads);
void zro_destroy(zro_sim* s);
void zro_set_particles(zro_sim* s, int64_t n, const double* X, const double* V,
                       const double* C, const double* F, const double* mass,
                       const double* vol, const double* mu, const double* lam);
/* per-particle materials; pparam 4 per particle: von Mises {xi, fail_stress,
   0, 0}, Drucker-Prager {alpha, beta, cohesion, volume_correction};
   pstate: von Mises yield stress, Drucker-Prager logJp; g: degradation of
   NeoHookeanBorden (NULL: 1) */
void zro_set_materials(zro_sim* s, const int32_t* model, const int32_t* plastic, const double* pparam,
                       const double* pstate, const double* g);
void zro_get_materials(const zro_sim* s, double* pstate, int32_t* broken, double* g);
void zro_get_particles(const zro_sim* s, double* X, double* V, double* C, double* F);
/* FElasticNonequilibratedBasedMpmForceHelper (:27-87): Fe_Nonequilibrated
   (9 per particle, column-major; NULL: identity) with a StvkWithHencky model
   (mu, lam per particle) and viscosities; mu == NULL switches it off */
void zro_set_viscoelastic(zro_sim* s, const double* Fneq, const double* mu, const double* lam, double visc_d,
                          double visc_v);
void zro_get_viscoelastic(const zro_sim* s, double* Fneq);
void zro_visco_relax(double mu, double lam, double visc_d, double visc_v, double dt, const double* gradV_cm,
                     double* Fneq_cm);
int zro_add_collider(zro_sim* s, const zro_collider* c);
/* updateState(t) of an object moving at constant rates from its time-0
   state: b(t) = translation + velocity t, R(t) = exp(t [omega]x) R0 (row-major),
   s(t) = s0 + ds/dt t */
void zro_collider_frame(const zro_collider* c, double t, double* b, double* R, double* s);
/* multiObjectCollisionWithFakeSeparate (CollisionObject.cpp:108-160) on one
   node at x with velocity v (resolved in place); P = I - K K^T (9 doubles,
   row-major); returns 1 when any object acts */
int zro_fake_separate(const zro_collider* cols, int ncol, double t, const double* x, double* v, double* P);
void zro_set_time(zro_sim* s, double t);

/* sortParticlesAndPolluteGrid (MpmSimulationBase.cpp:1032-1103) */
void zro_sort_and_pollute(zro_sim* s);
int64_t zro_num_groups(const zro_sim* s);
int64_t zro_num_pages(const zro_sim* s);
void zro_get_sort(const zro_sim* s, uint64_t* sorter, int32_t* order, uint64_t* base_offset);
void zro_get_groups(const zro_sim* s, int64_t* first, int64_t* last, uint64_t* block_offset);
void zro_get_pages(const zro_sim* s, uint64_t* page_offsets);

/* forcesUpdateParticleState + particlesToGrid (MLS) + getNumNodes +
   normalisation: MpmSimulationBase.cpp:437-561,681-709 */
void zro_update_stress(zro_sim* s);
int64_t zro_p2g_mls(zro_sim* s);
/* gridVelocityExplicitUpdate + collider update + trialCollision
   (MpmSimulationBase.cpp:298-312,818-865) */
void zro_grid_update(zro_sim* s, double dt);
/* gridToParticles MLS/APIC + evolveStrain (MpmSimulationBase.cpp:906-1016) */
int zro_g2p_mls(zro_sim* s, double dt);
/* the whole symplectic step (MpmSimulationBase.cpp:256-281) */
int zro_step_explicit(zro_sim* s, double dt);

/* ---- adaptive time step ---- */
/* evalMaxParticleSpeed (MpmSimulationBase.cpp:1143-1170): out[0] = max |v|
   (the APIC speed, out[0] of the reference's pair), out[1..3] = max corner,
   out[4..6] = min corner of the positions; the reduction starts from
   (0, float lowest, -float lowest) as the reference's does */
void zro_eval_max_particle_speed(const zro_sim* s, double out[7]);
/* calculateDt (MpmSimulationBase.cpp:766-790) with quadratic splines and
   colliders translating at constant velocity (evalMaxSpeed = |dbdt|,
   CollisionObject.cpp:158-194): cfl dx / max speed, or max_dt when it is 0 */
double zro_calculate_dt(const zro_sim* s, double max_dt);
/* TimeStepping::nextDt (TimeStepping.h:45-61) on its state */
double zro_next_dt(double dt, double min_dt, double max_dt, double max_dt_original, double frame_dt,
                   double time_since_last_frame);

/* ---- implicit backward Euler (APIC transfers) ---- */
typedef struct {
  int32_t newton_max_iterations;   /* NewtonsMethod max_iterations (MpmSimulationBase.cpp:51: 3) */
  double newton_tolerance;         /* NewtonsMethod tolerance (1e-5) */
  int32_t cg_max_iterations;       /* ConjugateGradient(20) (BackwardEuler.h:90-91) */
  double cg_tolerance;             /* LinearSolver::tolerance, default 16 eps (LinearSolver.h:26) */
  double cg_relative_tolerance;    /* <= 0: Newton-adaptive min(0.5, sqrt(max(|r|, tol))) */
  int32_t preconditioner;          /* 0: mass-lumped in/m (MpmSimulationBase.cpp:617-623),
                                      1: Jacobi on the operator diagonal (KrylovSolvers.h:32-57) */
  int32_t solver;                  /* 0: ConjugateGradient (switchLinearSolver(CG)),
                                      1: Minres, the reference's default (BackwardEuler.h:55-62) */
} zro_implicit_opts;

typedef struct {
  int32_t newton_iterations;       /* Newton iterations that computed a step */
  int32_t cg_iterations;           /* total CG iterations */
  double residual_norm[8];         /* mass-weighted residual norm per Newton iteration */
  int32_t num_residuals;
  int32_t converged;
} zro_implicit_stats;

/* particlesToGridHelper<APIC> (MpmSimulationBase.cpp:563-608) + normalisation */
int64_t zro_p2g_apic(zro_sim* s);
/* backwardEulerStep (MpmSimulationBase.cpp:610-657) */
int zro_backward_euler(zro_sim* s, double dt, const zro_implicit_opts* o, zro_implicit_stats* st);
/* gridToParticlesHelper<APIC, deg2, !MLS> + evolveStrain */
int zro_g2p_apic(zro_sim* s, double dt);
/* advanceOneTimeStep, implicit branch (MpmSimulationBase.cpp:282-295) */
int zro_step_implicit(zro_sim* s, double dt, const zro_implicit_opts* o, zro_implicit_stats* st);

/* ---- CD-MPM phase-field damage (Projects/fracture/FractureSimulation.h) ---- */
typedef struct {
  double parabolic_M;        /* FractureSimulation::parabolic_M (<= 0: elliptic form) */
  int32_t cg_max_iterations; /* ConjugateGradient(100) (FractureSimulation.h:70-73) */
  int32_t naive;             /* 0: phase field, 1: solvePhaseFieldSystemNaive (Rankine),
                                2: the phase field, then the Rankine cut of its result */
  double cg_tolerance;       /* 1e-6 (FractureSimulation.h:439) */
  double cg_relative;        /* 1e-12 of the initial residual (fracture/ConjugateGradient.h:64) */
  double sigma_F;            /* Rankine threshold of the naive form */
  double delete_threshold;   /* deleteParticles' delete_particle_threshold (:162-207) */
} zro_pf_opts;

typedef struct {
  int32_t cg_iterations;     /* the value the fracture CG returns */
  int32_t reserved;
  double residual_norm;      /* final unpreconditioned residual norm */
} zro_pf_stats;

/* per-particle PhaseField attributes (PhaseField.h:13-21); NULL arrays keep
   the addFBasedMpmForceWithPhaseField defaults (MpmParticleHandleBase.cpp:105-114):
   c 1, H 0, pf_Fp 1, H_max +inf, residual_phase 0.001, allow_damage 1.
   Particles with a phase field must use the NeoHookeanBorden model. */
void zro_set_phase_field(zro_sim* s, const double* c, const double* H, const double* l0,
                         const double* one_over_sigma_c, const double* pf_Fp, const double* H_max,
                         const double* residual_phase, const int32_t* allow_damage, const zro_pf_opts* o);
void zro_get_phase_field(const zro_sim* s, double* c, double* H, double* pf_Fp);
/* solvePhaseFieldSystem / solvePhaseFieldSystemNaive: after zro_p2g_mls of the
   same step; zro_step_explicit calls it once a phase field is set */
void zro_phase_field_solve(zro_sim* s, double dt, zro_pf_stats* st);
/* FractureSimulation::deleteParticles (:162-207); zro_step_explicit runs it first */
void zro_delete_particles(zro_sim* s);
/* pieces for tests: per node slot (npages*32) the system of the last solve */
void zro_phase_field_system(const zro_sim* s, double* rhs, double* diag, double* x);
void zro_phase_field_multiply(zro_sim* s, const double* x, double* b);
/* ---- adjoint of the explicit MLS step (no reference counterpart) ---- */
/* The sim holds the step's input state; a* hold the adjoint of the step's
   output (particle-major) on entry and of its input on return; amu / alam
   accumulate.  Returns -1 if a collider branch has no derivative (friction). */
int zro_adjoint_step(zro_sim* s, double dt, double* aX, double* aV, double* aC, double* aF, double* amu,
                     double* alam);

/* ---- ADMM step (Projects/admm/AdmmSimulationFoam.h:163-801) ----
   StvkWithHencky particles with DummyPlasticity (ZRO_PLASTIC_NONE), von
   Mises or Drucker-Prager; no viscoelastic branch.  Runs the whole
   AdmmSimulation::advanceOneTimeStep: sort, APIC P2G, admmStep, APIC G2P
   with the F update and the return mapping. */
typedef struct {
  int32_t admm_iterations;        /* admm_max_iterations (reference default 100) */
  int32_t minres_max_iterations;  /* global_max_iteration 10000 */
  double rho_scale;               /* 0.5 */
  double global_tolerance;        /* 1e-6 */
  double local_tolerance;         /* 1e-8 */
} zro_admm_opts;

typedef struct {
  int32_t minres_iterations;      /* total over the ADMM iterations */
  int32_t reserved;
} zro_admm_stats;

int zro_step_admm(zro_sim* s, double dt, const zro_admm_opts* o, zro_admm_stats* st);
/* The viscoelastic branch's relaxation map in the ADMM local step
   (AdmmSimulationFoam.h:583-603): relaxed singular values zp of z and the
   derivative D the reference forms (row-major). */
void zro_admm_visco_map(double mu, double lam, double visc_d, double visc_v, double dt, const double* z, double* zp,
                        double D[3][3]);
/* The ADMM local step's plastic projection of singular values, projectSigma
   (PlasticityApplier.cpp:319-346 Drucker-Prager, :526-547 von Mises; other
   plasticity: identity), and projectSigmaAndDerivative (:389-423, :580-606):
   pparam as zr_materials.plastic_param (4 doubles), pstate the von Mises
   yield stress; D is d out / d sig, row-major. */
void zro_project_sigma(double mu, double lam, int plastic, const double* pparam, double pstate, const double* sig,
                       double* out);
void zro_project_sigma_and_derivative(double mu, double lam, int plastic, const double* pparam, double pstate,
                                      const double* sig, double* zp, double* D);
/* per particle ADMM state after the last step: F (9n), y (9n), omega (n) */
void zro_get_admm(const zro_sim* s, double* F, double* y, double* omega);
/* all three eigenvalues of a symmetric 3x3 matrix, ascending */
void zro_sym3_eigenvalues(const double* A, double* lam);

/* largest eigenvalue of a symmetric 3x3 matrix (column-major) */
double zro_sym3_max_eigenvalue(const double* A);
/* matrix-free operator of the last backward-Euler solve at the final Newton
   iterate, on node vectors in page-list x element order (3 per node):
   b = M x + dt^2 sum_p V0 dP(grad x_p Fn) Fn^T grad w_ip   (BackwardEuler.h:322-351) */
void zro_implicit_multiply(zro_sim* s, const double* x, double* b);
/* residual at dv: dt m g + dt f(dv) - m dv, projected (BackwardEuler.h:101-124) */
void zro_implicit_residual(zro_sim* s, const double* dv, double* r);
/* exact diagonal of the operator at the current Newton state (3 per node) */
void zro_implicit_diagonal(zro_sim* s, double* diag);
/* objective energy whose negative gradient is the unprojected residual */
double zro_implicit_energy(zro_sim* s, const double* dv);
int64_t zro_num_node_slots(const zro_sim* s);
/* dv (Newton result), vn and collision flags of the last solve, per node slot */
void zro_get_implicit(const zro_sim* s, double* dv, double* vn, char* collided);

/* grid fields, per node in page-list order x 32 elements:
   m[npages*32], v[npages*32*3] (normalised), new_v (force, then v*),
   idx[npages*32] (node numbering, -1 for m==0) */
void zro_get_grid(const zro_sim* s, double* m, double* v, double* new_v, int64_t* idx);

#ifdef __cplusplus
}
#endif
#endif
