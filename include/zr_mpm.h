/*
 * zr_mpm.h -- C-ABI of the MI355X MPM time-step engine.
 *
 * The engine replaces the CPU/TBB MPM step of penn-graphics-research/ziran2019
 * behind its own simulation interface.  Each entry point names the reference
 * virtual or function it stands in for (file:line in the reference tree).
 * Plain pointers and sizes only; no exceptions cross this boundary; every call
 * is stream-ordered on the engine's own HIP stream and returns when its results
 * are visible to the host.  One host thread per engine.
 *
 * Host data layout (the layout of the reference's per-attribute arrays,
 * DataArray<T>::array, Lib/Ziran/CS/DataStructure/DataArray.h:43-53):
 *   vectors  : 3 doubles per particle, particle-major (Vector<double,3>)
 *   matrices : 9 doubles per particle, COLUMN-MAJOR (Matrix<double,3,3>)
 * The engine keeps its own structure-of-arrays copy in HBM.
 */
#ifndef ZR_MPM_H
#define ZR_MPM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: 0 = OK, > 0 control signals, < 0 errors */
enum {
  ZR_OK = 0,
  /* a particle moved more than one cell in the step: the reference throws
     RestartException (MpmSimulationBase.cpp:983-1002, SimulationBase.h:21-29) */
  ZR_CFL_RESTART = 1,
  ZR_ERR_ARG = -1,
  ZR_ERR_HIP = -2,
  /* a particle's base node left [0, 2^grid_log2 - 8] on some axis */
  ZR_ERR_DOMAIN = -3,
  ZR_ERR_STATE = -4
};

/* ZR_FP32_REFINED: the implicit operator and the Krylov vectors in FP32
   (FP32 accumulation), inside the FP64 Newton residual that refines each
   step; the SVD, the stress, inner products and norms stay FP64. */
enum { ZR_FP64 = 0, ZR_FP32_REFINED = 1 };
enum { ZR_TRANSFER_MLS = 0, ZR_TRANSFER_APIC = 1 };
/* CG preconditioner of the implicit step: the lumped mass, z = r / m
   (BackwardEuler::precondition, MpmSimulationBase.cpp:617-623), or
   JacobiPreconditioner (KrylovSolvers.h:32-57) on the operator's diagonal
   m + dt^2 sum_p (V0 dP(e_d h^T) h)_d, h = Fn^T grad w, recomputed every
   Newton iteration; a node a collider touches gets its largest entry on
   every axis. */
enum { ZR_PRECOND_MASS = 0, ZR_PRECOND_JACOBI = 1 };
/* linear solver of the Newton step: ConjugateGradient (switchLinearSolver(CG),
   ConjugateGradient.h:99-147) or Minres (Minres.h:69-178), the reference's
   default for the backward-Euler objective (BackwardEuler.h:55-62: Minres(20),
   tolerance 1), which also solves with the indefinite Hessians of the
   elasticity models */
enum { ZR_SOLVER_CG = 0, ZR_SOLVER_MINRES = 1 };

/* elasticity models (Lib/Ziran/Physics/ConstitutiveModel/): fixed-corotated
   (no reference counterpart; the default), StvkWithHencky, NeoHookeanBorden.
   StvkWithHenckyIsotropic has StvkWithHencky's energy and stress
   (StvkWithHenckyIsotropic.h: psi = mu tr(log S)^2 + lambda/2 (tr log S)^2)
   and uses ZR_MODEL_STVK_HENCKY. */
enum { ZR_MODEL_FIXED_COROTATED = 0, ZR_MODEL_STVK_HENCKY = 1, ZR_MODEL_NEO_HOOKEAN_BORDEN = 2 };
/* plasticity (Lib/Ziran/Physics/PlasticityApplier.h): VonMisesStvkHencky,
   DruckerPragerStvkHencky (both assume the StvkWithHencky model) and
   NonAssociativeCamClay (on the NeoHookeanBorden model, whose kappa it uses) */
enum { ZR_PLASTIC_NONE = 0, ZR_PLASTIC_VON_MISES = 1, ZR_PLASTIC_DRUCKER_PRAGER = 2, ZR_PLASTIC_CAM_CLAY = 3 };

/* collision object types, AnalyticCollisionObject::COLLISION_OBJECT_TYPE
   (Lib/Ziran/Math/Geometry/CollisionObject.h:47-52) */
enum { ZR_STICKY = 1, ZR_SLIP = 2, ZR_SEPARATE = 3 };
enum { ZR_HALF_SPACE = 0, ZR_SPHERE = 1 };

typedef struct zr_config {
  double dx;            /* MpmSimulationBase::dx */
  double gravity[3];    /* MpmSimulationBase::gravity */
  double cfl;           /* MpmSimulationBase::cfl (used by the CFL flags) */
  int32_t grid_log2;    /* node coordinates live in [0, 2^grid_log2); 6..12 (12: the
                           reference's 4096^3 SPGrid extent; the dense page tables
                           take 4 x 2^(3 grid_log2 - 3) bytes, 32 GiB at 12) */
  int32_t precision;    /* ZR_FP64; ZR_FP32_REFINED: the explicit P2G's per-particle
                           terms (weights, offsets, products; a NeoHookeanBorden
                           stress) in FP32, the node sums, grid, G2P, damage and
                           particle state FP64 */
  int32_t transfer;     /* ZR_TRANSFER_MLS (mls_mpm = true) */
  int32_t reserved;
} zr_config;

/* An analytic collision object (AnalyticCollisionObject + HalfSpace / Sphere
   level sets, CollisionObject.cpp:48-105,340-405; AnalyticLevelSet.cpp:
   259-287,435-450) moving as x = R(t) s(t) X + b(t) (CollisionObject.h:81-172)
   at constant rates from its state at time 0, which is the state the
   reference's updateState(t) callbacks set for such motion
   (e.g. AdmmInit3D.h:298-313):
     b(t) = translation + velocity t,
     R(t) = exp(t [omega]x) R0   (R0 from the quaternion `rotation`),
     s(t) = scale + dscale_dt t.
   The level set is queried at X = R^T (x - b) / s, its normal is R N, and the
   object's velocity at x is omega x (x - b) + (ds/dt / s)(x - b) + velocity
   (CollisionObject.cpp:340-405).  A step from t to t + dt uses the state at
   t for the grid update and the collision projections and the state at
   t + dt for the trial collision (updateState(t + dt) before trialCollision,
   MpmSimulationBase.cpp:267-278).  Other motions are set again (with
   zr_set_colliders) between steps.  Zero-initialised extra fields give the
   static frame (R = I, s = 1). */
typedef struct zr_collider {
  int32_t shape;        /* ZR_HALF_SPACE or ZR_SPHERE */
  int32_t type;         /* ZR_STICKY, ZR_SLIP or ZR_SEPARATE */
  double friction;
  double origin[3];     /* half-space point / sphere centre (material space) */
  double normal[3];     /* half-space outward normal (material space) */
  double radius;        /* sphere radius */
  double velocity[3];   /* dbdt */
  double translation[3];/* b at time 0 */
  double rotation[4];   /* R as a quaternion <w,x,y,z> (setRotation); all 0: identity */
  double omega[3];      /* angular velocity (setAngularVelocity) */
  double scale;         /* s (setScaling); 0: 1 */
  double dscale_dt;     /* ds/dt */
} zr_collider;

/* Host views of the particle attributes (Particles m/X/V plus the F-based
   force helper attributes, Particles.h:8-45, FBasedMpmForceHelper.h:17-60). */
typedef struct zr_particles {
  int64_t n;
  double* X;       /* 3n  */
  double* V;       /* 3n  */
  double* C;       /* 9n, APIC/MLS affine matrix (C_name<TM>()) */
  double* F;       /* 9n, deformation gradient */
  double* mass;    /* n   */
  double* vol;     /* n, element measure V0 */
  double* mu;      /* n, per-particle constitutive model (fixed-corotated) */
  double* lambda;  /* n   */
} zr_particles;

/* Per-particle material choice (the constitutive-model and plasticity
   attributes the reference registers per particle, FBasedMpmForceHelper.h:17-60,
   PlasticityApplier.h:19-56).  NULL model / plasticity mean all
   ZR_MODEL_FIXED_COROTATED / ZR_PLASTIC_NONE.  plastic_param holds 4 doubles
   per particle: von Mises {xi, fail_stress, 0, 0} (PlasticityApplier.cpp:443-462),
   Drucker-Prager {alpha, beta, cohesion, volume_correction != 0} with alpha
   from the friction angle as in PlasticityApplier.cpp:249-260, Cam-Clay
   {M, beta, xi, 0} with M from the friction angle as in
   PlasticityApplier.cpp:10-19 (Cam-Clay needs ZR_MODEL_NEO_HOOKEAN_BORDEN).
   plastic_state is updated by every step: von Mises yield stress (hardened),
   Drucker-Prager and Cam-Clay logJp.  degradation is NeoHookeanBorden's g (NULL: 1). */
typedef struct zr_materials {
  int64_t n;
  int32_t* model;          /* n */
  int32_t* plasticity;     /* n */
  double* plastic_param;   /* 4n */
  double* plastic_state;   /* n */
  double* degradation;     /* n */
  int32_t* broken;         /* n, download only: von Mises fail_stress reached */
} zr_materials;

/* CD-MPM phase field (Projects/fracture/PhaseField.h:13-21), per particle.
   NULL arrays take the defaults of addFBasedMpmForceWithPhaseField
   (MpmParticleHandleBase.cpp:105-114): c 1, H 0, pf_Fp 1, H_max +inf,
   residual_phase 0.001, allow_damage 1 (l0 and one_over_sigma_c have none). */
typedef struct zr_phase_field {
  int64_t n;
  double* c;                 /* damage, 1 intact .. 0 broken (updated) */
  double* H;                 /* history of psi+ (updated) */
  double* l0;
  double* one_over_sigma_c;
  double* pf_Fp;             /* (updated) */
  double* H_max;
  double* residual_phase;
  int32_t* allow_damage;
} zr_phase_field;

/* FractureSimulation options (FractureSimulation.h:50-73,439;
   fracture/ConjugateGradient.h:64): parabolic_M, CG(100), tolerance 1e-6,
   relative 1e-12 of the initial residual.  naive chooses the damage rule:
   ZR_DAMAGE_PHASE_FIELD (solvePhaseFieldSystem), ZR_DAMAGE_RANKINE
   (useNaiveDamage: solvePhaseFieldSystemNaive with threshold sigma_F), or
   ZR_DAMAGE_PHASE_FIELD_RANKINE: the phase-field solve and then the Rankine
   cut c <- min(c, sigma_F / sigma_max) on its result (the reference runs the
   two as alternatives, FractureSimulation.h:149-160; this rule applies both,
   keeps c monotone and takes the smaller of the two damages' c).
   delete_threshold is deleteParticles' delete_particle_threshold
   (FractureSimulation.h:162-207, run at the start of every explicit step):
   particles with c below it are parked at (4, 4, 4) at rest with F = I and
   allow_damage off; 0 (the reference's default) deletes nothing, and the
   grid must reach (4, 4, 4) (grid_log2 with 2^grid_log2 - 8 > 4 / dx).
   Only the lumped form (lumping = true) exists. */
enum { ZR_DAMAGE_PHASE_FIELD = 0, ZR_DAMAGE_RANKINE = 1, ZR_DAMAGE_PHASE_FIELD_RANKINE = 2 };
typedef struct zr_phase_field_opts {
  double parabolic_M;
  int32_t cg_max_iterations;
  int32_t naive;
  double cg_tolerance;
  double cg_relative;
  double sigma_F;
  double delete_threshold;
} zr_phase_field_opts;

typedef struct zr_phase_field_stats {
  int32_t cg_iterations;     /* the fracture CG's return value */
  int32_t reserved;
  double residual_norm;
} zr_phase_field_stats;

/* Adjoints of the particle state (same layout as zr_particles) and of the
   per-particle Lame parameters (accumulated over reverse steps). */
typedef struct zr_particles_adjoint {
  int64_t n;
  double* X;       /* 3n */
  double* V;       /* 3n */
  double* C;       /* 9n */
  double* F;       /* 9n */
  double* mu;      /* n  */
  double* lambda;  /* n  */
} zr_particles_adjoint;

/* AdmmSimulation options (Projects/admm/AdmmSimulationFoam.h:55-68):
   admm_max_iterations (reference default 100), Minres(10000), rho_scale 0.5,
   global tolerance 1e-6, local Newton tolerance 1e-8.  precision: ZR_FP64
   (the reference's arithmetic) or ZR_FP32_REFINED (the global step's
   operator data and Krylov vectors in FP32; the local step, projections,
   inner products and the velocity update dv stay FP64). */
typedef struct zr_admm_opts {
  int32_t admm_iterations;
  int32_t minres_max_iterations;
  double rho_scale;
  double global_tolerance;
  double local_tolerance;
  int32_t precision;   /* ZR_FP64 */
  int32_t reserved;
} zr_admm_opts;

typedef struct zr_admm_stats {
  int32_t minres_iterations;   /* total over the ADMM iterations of the step */
  int32_t reserved;
} zr_admm_stats;

typedef struct zr_step_stats {
  int64_t num_groups;   /* particle pages (particle_group.size()) */
  int64_t num_pages;    /* touched SPGrid pages (page_map->Get_Blocks().second) */
  int32_t cfl_flags;    /* bit0: > 1 cell, bit1: > cfl/2 cell */
  int32_t reserved;
} zr_step_stats;

/* Options of the implicit backward-Euler step.  Reference defaults:
   NewtonsMethod(tol 1e-5, 3 iterations) (MpmSimulationBase.cpp:51),
   Minres(20) with tolerance 1 (BackwardEuler.h:55-62), or after
   switchLinearSolver(CG) ConjugateGradient(20) (BackwardEuler.h:90-91) with
   LinearSolver::tolerance 16 eps (LinearSolver.h:26); the Newton-adaptive
   relative tolerance min(0.5, sqrt(max(|r|, tol))) (NewtonsMethod.h:61).
   The cg_* fields set whichever solver runs. */
typedef struct zr_implicit_opts {
  int32_t newton_max_iterations;
  int32_t cg_max_iterations;
  double newton_tolerance;
  double cg_tolerance;           /* absolute cap of the CG local tolerance */
  double cg_relative_tolerance;  /* <= 0: Newton-adaptive (reference behaviour) */
  int32_t precision;             /* ZR_FP64 or ZR_FP32_REFINED */
  int32_t cg_check_interval;     /* CG iterations between host polls of the
                                    device-side termination flag (0: 8) */
  int32_t preconditioner;        /* ZR_PRECOND_MASS or ZR_PRECOND_JACOBI */
  int32_t solver;                /* ZR_SOLVER_CG or ZR_SOLVER_MINRES */
} zr_implicit_opts;

typedef struct zr_implicit_stats {
  int32_t newton_iterations;     /* Newton iterations that computed a step */
  int32_t cg_iterations;         /* total CG iterations of the step */
  double residual_norm[8];       /* mass-weighted residual norm per Newton iteration */
  int32_t num_residuals;
  int32_t converged;             /* residual fell below newton_tolerance */
} zr_implicit_stats;

typedef struct zr_engine zr_engine;

/* MpmSimulationBase constructor + setDx (MpmSimulationBase.cpp:33-72) */
int zr_create(const zr_config* cfg, zr_engine** out);
void zr_destroy(zr_engine* e);
/* message of the last failed call on this engine; with NULL, of the last
   failed call without an engine (zr_eval_max_particle_speed_noalloc) on the
   calling thread */
const char* zr_last_error(const zr_engine* e);

/* addCollisionObject (MpmSimulationBase.cpp:84-88); replaces the whole set */
int zr_set_colliders(zr_engine* e, const zr_collider* c, int32_t n);
/* SimulationBase::step.time, used for moving colliders */
int zr_set_time(zr_engine* e, double t);
double zr_get_time(const zr_engine* e);

/* Upload / download of the particle attributes: Scene/DataManager state
   (DataManager.h:67-346).  Download fills X, V, C, F (NULL pointers skip).
   At most 2^29 - 1 particles per engine (ZR_ERR_ARG beyond). */
int zr_upload_particles(zr_engine* e, const zr_particles* p);
int zr_download_particles(zr_engine* e, zr_particles* p);
/* Per-particle models and plastic state, after zr_upload_particles (which
   resets every particle to fixed-corotated without plasticity).  Download
   fills plastic_state, degradation and broken (NULL pointers skip). */
int zr_set_materials(zr_engine* e, const zr_materials* m);
int zr_get_materials(zr_engine* e, zr_materials* m);
/* FElasticNonequilibratedBasedMpmForceHelper
   (Lib/MPM/Force/FElasticNonequilibratedBasedMpmForceHelper.h:27-87): a
   second elastic strain per particle (F_neq, 9 doubles column-major; NULL:
   identity) with a StvkWithHencky model of its own (mu, lambda per
   particle), whose V0 tau adds to the particle's stress and which follows
   the flow and relaxes with viscosities viscosity_d (deviatoric) and
   viscosity_v (volumetric), setParameters (:31-35).  Explicit MLS step only
   (the implicit, ADMM and adjoint steps reject it).  mu = lambda = NULL
   switches it off; an upload of particles clears it. */
int zr_set_viscoelastic(zr_engine* e, const double* F_neq, const double* mu, const double* lambda,
                        double viscosity_d, double viscosity_v);
/* the current F_neq (9 doubles per particle, column-major) */
int zr_get_viscoelastic(zr_engine* e, double* F_neq);

/* FractureSimulation: attaches a phase field to every particle and switches
   them to NeoHookeanBorden (addFBasedMpmForceWithPhaseField,
   MpmParticleHandleBase.cpp:95-118).  From then on zr_advance_one_time_step
   runs the damage solve after P2G (FractureSimulation::particlesToGrid,
   FractureSimulation.h:149-160). */
int zr_set_phase_field(zr_engine* e, const zr_phase_field* pf, const zr_phase_field_opts* o);
/* downloads c, H and pf_Fp (NULL pointers skip) */
int zr_get_phase_field(zr_engine* e, zr_phase_field* pf);
/* solvePhaseFieldSystem / solvePhaseFieldSystemNaive (FractureSimulation.h:259-554)
   on the grid of the last zr_particles_to_grid */
int zr_phase_field_solve(zr_engine* e, double dt, zr_phase_field_stats* st);
/* statistics of the last damage solve (also the one inside a step) */
int zr_phase_field_last_stats(zr_engine* e, zr_phase_field_stats* st);

/* MpmSimulationBase::sortParticlesAndPolluteGrid (MpmSimulationBase.cpp:1032-1103) */
int zr_sort_particles_and_pollute_grid(zr_engine* e, zr_step_stats* st);
/* forcesUpdateParticleState + particlesToGrid, MLS with force, then getNumNodes
   and velocity normalisation (MpmSimulationBase.cpp:437-561,681-709) */
int zr_particles_to_grid(zr_engine* e);
/* gridVelocityExplicitUpdate + collider updateState(t+dt) + trialCollision
   (MpmSimulationBase.cpp:267-271,298-312,818-865) */
int zr_grid_velocity_explicit_update(zr_engine* e, double dt);
/* gridToParticles (MLS/APIC) + evolveStrain (MpmSimulationBase.cpp:879-1016);
   returns ZR_CFL_RESTART when a particle moved more than one cell */
int zr_grid_to_particles(zr_engine* e, double dt, zr_step_stats* st);
/* MpmSimulationBase::advanceOneTimeStep, symplectic branch
   (MpmSimulationBase.cpp:256-281); advances the engine time by dt */
int zr_advance_one_time_step(zr_engine* e, double dt, zr_step_stats* st);

/* MpmSimulationBase::evalMaxParticleSpeed (MpmSimulationBase.cpp:1143-1170)
   on the engine's particles: out[0] = max |v| (Eigen norm), out[1..3] = max
   corner, out[4..6] = min corner of the positions.  With no particles:
   0, -FLT_MAX (x3), FLT_MAX (x3), as the reference's reduction start. */
int zr_eval_max_particle_speed(zr_engine* e, double out[7]);
/* MpmSimulationBase::calculateDt (MpmSimulationBase.cpp:766-790): cfl dx /
   the larger of the particles' max speed and the colliders' speed (constant
   translations: |velocity|, AnalyticCollisionObject::evalMaxSpeed,
   CollisionObject.cpp:158-194), or max_dt (the reference's step.max_dt)
   when both are 0.  The caller applies TimeStepping::nextDt. */
int zr_calculate_dt(zr_engine* e, double max_dt, double* dt);

/* zr_eval_max_particle_speed on caller-owned DEVICE arrays, allocating
   nothing: no counterpart in the reference; for callers that cannot allocate
   device memory (code under HIP graph capture, serving loops, harnesses that
   forbid hipMalloc).  x, v: n particles in the reference layout (3 doubles
   per particle, particle-major); out: 7 doubles on the device, with the
   meaning of zr_eval_max_particle_speed's out; stream: a hipStream_t (NULL:
   the null stream).  Enqueues two kernels and returns; out is ready when the
   stream reaches them.  The block maxima live in a fixed device array of
   512 x 7 doubles (28 KiB) per device, held until the process ends and
   shared by every call on that device: calls that overlap on one device
   race.  n <= 2^32 - 1 (ZR_ERR_ARG beyond); errors are described by
   zr_last_error(NULL). */
int zr_eval_max_particle_speed_noalloc(const double* x, const double* v, int64_t n, double* out, void* stream);

/* particlesToGrid of the implicit path: APIC mass and momentum only
   (particlesToGridHelper<APIC>, MpmSimulationBase.cpp:563-608) + getNumNodes +
   normalisation (MpmSimulationBase.cpp:474-485) */
int zr_particles_to_grid_apic(zr_engine* e);
/* MpmSimulationBase::backwardEulerStep (MpmSimulationBase.cpp:610-657):
   buildInitialDvAndVnForNewton, Newton with matrix-free preconditioned CG on
   BackwardEulerLagrangianForceObjective (BackwardEuler.h:101-351),
   constructNewVelocityFromNewtonResult */
int zr_backward_euler_step(zr_engine* e, double dt, const zr_implicit_opts* o, zr_implicit_stats* st);
/* gridToParticles with APIC (grad v = sum v grad w^T) + evolveStrain
   (MpmSimulationBase.cpp:906-1016 with !mls_mpm) */
int zr_grid_to_particles_apic(zr_engine* e, double dt, zr_step_stats* st);
/* advanceOneTimeStep, implicit branch (MpmSimulationBase.cpp:282-295) */
int zr_advance_one_time_step_implicit(zr_engine* e, double dt, const zr_implicit_opts* o, zr_implicit_stats* ist,
                                      zr_step_stats* st);
/* At the state of the last backward-Euler solve, on host node vectors in
   page-list x element order (3 doubles per node, 32 nodes per page):
   residual at dv (BackwardEuler.h:101-124; also moves the Newton state to dv)
   and the unprojected operator b = A x (BackwardEuler.h:322-351). */
int zr_implicit_residual(zr_engine* e, const double* dv, double* r);
int zr_implicit_multiply(zr_engine* e, const double* x, double* b);

/* AdmmSimulation::advanceOneTimeStep (AdmmSimulationFoam.h:803-818): APIC
   P2G, admmStep (:766-801), APIC G2P with the F update and the return map.
   Every particle must be StvkWithHencky (zr_set_materials) with no, von Mises
   or Drucker-Prager plasticity (no plasticity = DummyPlasticity). */
int zr_advance_one_time_step_admm(zr_engine* e, double dt, const zr_admm_opts* o, zr_admm_stats* ast,
                                  zr_step_stats* st);
/* the ADMM variables after the last step, host layout (F, y: 9n; omega: n) */
int zr_get_admm_state(zr_engine* e, double* F, double* y, double* omega);

/* Differentiable form of the explicit MLS step.  No counterpart in the
   reference; for gradient-based callers (parameter estimation, control) that
   need d loss / d (state, mu, lambda) of a rollout.  Fixed-corotated
   particles without plasticity or damage; every collider kind, in moving
   frames, with Coulomb friction (differentiated on its sliding branch; the
   sticking branch has a zero derivative).
   zr_adjoint_set: the adjoint of the current state (NULL arrays: zero; mu and
   lambda set the accumulators).  zr_adjoint_step: with the engine holding
   the INPUT state of a step of size dt (and its time), replaces the adjoint
   of that step's output by the adjoint of its input, accumulating mu and
   lambda; the particle state is left unchanged.  Checkpoints keep X, V, C, F
   and the time on the device for recompute. */
int zr_adjoint_set(zr_engine* e, const zr_particles_adjoint* a);
int zr_adjoint_get(zr_engine* e, zr_particles_adjoint* a);
int zr_adjoint_step(zr_engine* e, double dt);
int zr_checkpoint_save(zr_engine* e, int32_t slot);
int zr_checkpoint_restore(zr_engine* e, int32_t slot);
/* The checkpointed rollout's two loops in one call each.  No counterpart in
   the reference; for rollout drivers (ziran2019_amd/autodiff.py) that would
   otherwise pay a host round trip per step and per checkpoint.
   zr_rollout_forward: steps explicit steps of size dt from the current
   state, saving checkpoint slot0 + k / every before step k when k is a
   multiple of every; stops at the first error or ZR_CFL_RESTART.  A step's
   CFL flags are read with the next step's sort (one host wait per step in
   place of two), so ZR_CFL_RESTART leaves the state after the step that
   asked for it, as a call per step would, and a checkpoint due before the
   next step may already hold that state.
   zr_rollout_backward: after zr_adjoint_set on the rollout's output, per
   segment from the last: restore its checkpoint, save the inputs of its
   steps after the first in slots scratch0 + 1 .. scratch0 + every - 1
   (recomputing them; the first step's input is the checkpoint), then
   run zr_adjoint_step for each, last first; the adjoint then holds that of
   the rollout's input.  The recomputed steps keep their sort tables and grid
   on the device (4 B per particle, 40 B per particle group and 36 B per grid
   page, 80 B per grid node; up to every - 1 records) for their adjoint steps, which do not sort again.  The
   adjoint steps' collider flags are read once, at the end. */
int zr_rollout_forward(zr_engine* e, int32_t steps, double dt, int32_t every, int32_t slot0);
int zr_rollout_backward(zr_engine* e, int32_t steps, double dt, int32_t every, int32_t slot0, int32_t scratch0);

/* Introspection for parity tests (reference state after the named call):
   sizes in bytes via zr_debug_size, contents via zr_debug_dump. */
enum {
  ZR_DUMP_SORTER = 0,       /* u64[n]  particle_sorter after the sort */
  ZR_DUMP_ORDER = 1,        /* i32[n]  particle_order */
  ZR_DUMP_BASE_OFFSET = 2,  /* u64[n]  particle_base_offset (by particle) */
  ZR_DUMP_GROUP_FIRST = 3,  /* i64[G]  particle_group[g].first */
  ZR_DUMP_GROUP_LAST = 4,   /* i64[G]  particle_group[g].second */
  ZR_DUMP_BLOCK_OFFSET = 5, /* u64[G]  block_offset (page code) */
  ZR_DUMP_PAGES = 6,        /* u64[P]  page_map block offsets, first-touch order */
  ZR_DUMP_GRID_M = 7,       /* f64[32P] node mass, page-list order x in-page order */
  ZR_DUMP_GRID_V = 8,       /* f64[3*32P] normalised node velocity */
  ZR_DUMP_GRID_F = 9,       /* f64[3*32P] rasterised force (new_v after P2G) */
  ZR_DUMP_GRID_VSTAR = 10,  /* f64[3*32P] new_v after the grid update */
  ZR_DUMP_NODE_IDX = 11,    /* i64[32P] GridState::idx (getNumNodes numbering) */
  ZR_DUMP_DV = 12,          /* f64[3*32P] Newton result dv of the last backward-Euler step */
  ZR_DUMP_HVP_XCD = 13,     /* u32[20] operator-apply group ranges per XCD (FP32, then FP64;
                               9 bounds each) of the last backward-Euler step, then the block
                               counts of the two launches; no reference counterpart: a
                               scheduling detail of this implementation */
  ZR_DUMP_JACOBI = 14       /* f64[3*32P] JacobiPreconditioner's inverse diagonal (ZR_PRECOND_JACOBI)
                               at the Newton state of the last backward-Euler solve, computed by
                               the call in that solve's precision (after zr_implicit_residual:
                               at its dv) */
};
int64_t zr_debug_size(zr_engine* e, int32_t what);
int zr_debug_dump(zr_engine* e, int32_t what, void* host, size_t bytes);

/* Kernel timing with HIP events recorded on the engine's stream around each
   launch.  Instrumentation for measurement; it stands in for the reference's
   ZIRAN_TIMER scoped timers (Lib/Ziran/CS/Util/Timer.cpp:34-95), which time
   host code.  on = 0 turns timing off; on = k > 0 times every k-th launch of
   each kernel class (k = 1: all), so that sampled timing barely perturbs the
   step.  Enabling resets the accumulated times; zr_profile_read reports the
   timed launches only. */
enum {
  ZR_PROF_P2G = 0,
  ZR_PROF_GRID = 1,
  ZR_PROF_G2P = 2,
  ZR_PROF_SORT = 3,
  ZR_PROF_STATE = 4, /* Newton state: G2P gather + SVD + stress scatter */
  ZR_PROF_HVP = 5,   /* operator apply, particle part */
  ZR_PROF_COUNT = 6
};
int zr_profile_enable(zr_engine* e, int32_t on);
int zr_profile_read(zr_engine* e, int32_t which, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
