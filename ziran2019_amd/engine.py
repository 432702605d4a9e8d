"""Python binding over the C-ABI of ``libzr_mpm.so`` (include/zr_mpm.h).

``MpmEngine`` mirrors the step API of the reference's
``MpmSimulationBase<double,3>`` (Lib/MPM/MpmSimulationBase.h:143-232):
``sortParticlesAndPolluteGrid``, ``particlesToGrid``,
``gridVelocityExplicitUpdate``, ``gridToParticles`` and
``advanceOneTimeStep``.  A CFL violation surfaces as ``RestartException``
like the reference's (Lib/Ziran/Sim/SimulationBase.h:21-29).

The native library is required: importing works without it, but creating an
engine raises if ``ziran2019_amd/lib/libzr_mpm.so`` is missing or no GPU is
visible.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libzr_mpm.so")

ZR_OK = 0
ZR_CFL_RESTART = 1
ZR_ERR_ARG = -1
ZR_ERR_HIP = -2
ZR_ERR_DOMAIN = -3
ZR_ERR_STATE = -4

ZR_STICKY, ZR_SLIP, ZR_SEPARATE = 1, 2, 3
ZR_HALF_SPACE, ZR_SPHERE = 0, 1
ZR_FP64, ZR_FP32_REFINED = 0, 1

DUMP = dict(sorter=0, order=1, base_offset=2, group_first=3, group_last=4, block_offset=5, pages=6,
            grid_m=7, grid_v=8, grid_f=9, grid_vstar=10, node_idx=11, dv=12, hvp_xcd=13, jacobi=14)
_DUMP_DTYPE = dict(sorter=np.uint64, order=np.int32, base_offset=np.uint64, group_first=np.int64,
                   group_last=np.int64, block_offset=np.uint64, pages=np.uint64, grid_m=np.float64,
                   grid_v=np.float64, grid_f=np.float64, grid_vstar=np.float64, node_idx=np.int64, dv=np.float64,
                   hvp_xcd=np.uint32, jacobi=np.float64)

# every symbol include/zr_mpm.h declares
EXPORTS = (
    "zr_create", "zr_destroy", "zr_last_error", "zr_set_colliders", "zr_set_time", "zr_get_time",
    "zr_upload_particles", "zr_download_particles", "zr_sort_particles_and_pollute_grid",
    "zr_particles_to_grid", "zr_grid_velocity_explicit_update", "zr_grid_to_particles",
    "zr_advance_one_time_step", "zr_debug_size", "zr_debug_dump", "zr_profile_enable", "zr_profile_read",
    "zr_particles_to_grid_apic", "zr_backward_euler_step", "zr_grid_to_particles_apic",
    "zr_advance_one_time_step_implicit", "zr_implicit_residual", "zr_implicit_multiply",
    "zr_set_materials", "zr_get_materials", "zr_set_phase_field", "zr_get_phase_field", "zr_phase_field_solve",
    "zr_phase_field_last_stats", "zr_adjoint_set", "zr_adjoint_get", "zr_adjoint_step", "zr_checkpoint_save",
    "zr_checkpoint_restore", "zr_advance_one_time_step_admm", "zr_get_admm_state",
    "zr_eval_max_particle_speed", "zr_calculate_dt", "zr_set_viscoelastic", "zr_get_viscoelastic",
    "zr_eval_max_particle_speed_noalloc", "zr_rollout_forward", "zr_rollout_backward",
)
ZR_MODEL_FIXED_COROTATED, ZR_MODEL_STVK_HENCKY, ZR_MODEL_NEO_HOOKEAN_BORDEN = 0, 1, 2
ZR_PLASTIC_NONE, ZR_PLASTIC_VON_MISES, ZR_PLASTIC_DRUCKER_PRAGER, ZR_PLASTIC_CAM_CLAY = 0, 1, 2, 3
PRECOND_MASS, PRECOND_JACOBI = 0, 1
PROF = dict(p2g=0, grid=1, g2p=2, sort=3, state=4, hvp=5)


class RestartException(RuntimeError):
    """A particle travelled more than one grid cell in the step
    (MpmSimulationBase.cpp:983-1002); ``frame`` is the frame whose saved
    state a restart reads (SimulationBase.h:21-29), -1 when none is named."""

    def __init__(self, what, frame=-1):
        super().__init__(what)
        self.frame = frame


class TimeStepping:
    """The frame clock of SimulationBase::step (TimeStepping.h:8-103):
    substep sizes between min_dt and max_dt that land exactly on the frame
    times."""

    SUBSTEP, FRAME = 1, 0

    def __init__(self, frame_dt=1.0 / 24, min_dt=1e-6, max_dt=None):
        self.time = 0.0
        self.time_since_last_frame = 0.0
        self.frame_dt = frame_dt
        self.min_dt = min_dt
        self.max_dt = frame_dt if max_dt is None else max_dt
        self.max_dt_original = self.max_dt

    def nextDt(self, dt):
        """TimeStepping::nextDt (TimeStepping.h:45-61)."""
        if not dt > 0:
            raise ValueError("nextDt needs a positive dt")
        self.max_dt = min(self.max_dt, self.max_dt_original)
        dt = self.min_dt if dt < self.min_dt else (self.max_dt if dt > self.max_dt else dt)
        if self.time_since_last_frame + dt >= self.frame_dt:
            dt = self.frame_dt - self.time_since_last_frame
        elif self.time_since_last_frame + 2 * dt > self.frame_dt:
            dt = (self.frame_dt - self.time_since_last_frame) / 2
        return dt

    def advance(self, dt):
        """TimeStepping::advance (TimeStepping.h:63-80)."""
        self.time += dt
        self.time_since_last_frame += dt
        if self.time_since_last_frame >= self.frame_dt:
            self.time_since_last_frame = 0.0
            return self.FRAME
        return self.SUBSTEP

    def reset(self, frame):
        """TimeStepping::reset (TimeStepping.h:82-90)."""
        self.time = frame * self.frame_dt
        self.time_since_last_frame = 0.0


class ZrError(RuntimeError):
    pass


class ZrConfig(ctypes.Structure):
    _fields_ = [("dx", ctypes.c_double), ("gravity", ctypes.c_double * 3), ("cfl", ctypes.c_double),
                ("grid_log2", ctypes.c_int32), ("precision", ctypes.c_int32), ("transfer", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class ZrCollider(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("type", ctypes.c_int32), ("friction", ctypes.c_double),
                ("origin", ctypes.c_double * 3), ("normal", ctypes.c_double * 3), ("radius", ctypes.c_double),
                ("velocity", ctypes.c_double * 3), ("translation", ctypes.c_double * 3),
                ("rotation", ctypes.c_double * 4), ("omega", ctypes.c_double * 3), ("scale", ctypes.c_double),
                ("dscale_dt", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)


class ZrImplicitOpts(ctypes.Structure):
    _fields_ = [("newton_max_iterations", ctypes.c_int32), ("cg_max_iterations", ctypes.c_int32),
                ("newton_tolerance", ctypes.c_double), ("cg_tolerance", ctypes.c_double),
                ("cg_relative_tolerance", ctypes.c_double), ("precision", ctypes.c_int32),
                ("cg_check_interval", ctypes.c_int32), ("preconditioner", ctypes.c_int32),
                ("solver", ctypes.c_int32)]


class ZrImplicitStats(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int32), ("cg_iterations", ctypes.c_int32),
                ("residual_norm", ctypes.c_double * 8), ("num_residuals", ctypes.c_int32),
                ("converged", ctypes.c_int32)]


def implicit_opts(newton_max_iterations=3, newton_tolerance=1e-5, cg_max_iterations=20,
                  cg_tolerance=16 * 2.220446049250313e-16, cg_relative_tolerance=0.0, precision=0,
                  cg_check_interval=8, preconditioner=0, solver=0):
    """Reference defaults: NewtonsMethod(1e-5, 3) (MpmSimulationBase.cpp:51),
    ConjugateGradient(20) (BackwardEuler.h:90-91), LinearSolver::tolerance
    16 eps (LinearSolver.h:26), Newton-adaptive relative tolerance, the
    mass-lumped preconditioner (MpmSimulationBase.cpp:617-623);
    preconditioner=1 is JacobiPreconditioner (KrylovSolvers.h:32-57) on the
    operator's diagonal; solver=1 is Minres (Minres.h:69-178), the
    reference's default linear solver (BackwardEuler.h:55-62, with tolerance
    1: cg_tolerance=1.0), as ``minres_opts`` sets it up."""
    return ZrImplicitOpts(int(newton_max_iterations), int(cg_max_iterations), float(newton_tolerance),
                          float(cg_tolerance), float(cg_relative_tolerance), int(precision), int(cg_check_interval),
                          int(preconditioner), int(solver))


def minres_opts(**kw):
    """The reference's default implicit solve: Newton(1e-5, 3) with Minres(20)
    at tolerance 1 (BackwardEuler.h:55-62, MpmSimulationBase.cpp:51)."""
    d = dict(cg_max_iterations=20, cg_tolerance=1.0, solver=1)
    d.update(kw)
    return implicit_opts(**d)


class ZrParticles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("X", _dp), ("V", _dp), ("C", _dp), ("F", _dp), ("mass", _dp),
                ("vol", _dp), ("mu", _dp), ("lam", _dp)]


_ip = ctypes.POINTER(ctypes.c_int32)


class ZrMaterials(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("model", _ip), ("plasticity", _ip), ("plastic_param", _dp),
                ("plastic_state", _dp), ("degradation", _dp), ("broken", _ip)]


def cam_clay_M(friction_angle_deg, dim=3):
    """NonAssociativeCamClay's M from the friction angle with the reference's
    pi = 3.141592653 (PlasticityApplier.cpp:10-19)."""
    s = np.sin(friction_angle_deg / 180.0 * 3.141592653)
    return np.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s) * dim / np.sqrt(2.0 / (6.0 - dim))


def drucker_prager_alpha(friction_angle_deg):
    """alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)) with the reference's
    pi = 3.141592653 (PlasticityApplier.cpp:249-260)."""
    s = np.sin(friction_angle_deg / 180.0 * 3.141592653)
    return np.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


class ZrPhaseField(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("c", _dp), ("H", _dp), ("l0", _dp), ("one_over_sigma_c", _dp),
                ("pf_Fp", _dp), ("H_max", _dp), ("residual_phase", _dp), ("allow_damage", _ip)]


class ZrPhaseFieldOpts(ctypes.Structure):
    _fields_ = [("parabolic_M", ctypes.c_double), ("cg_max_iterations", ctypes.c_int32),
                ("naive", ctypes.c_int32), ("cg_tolerance", ctypes.c_double), ("cg_relative", ctypes.c_double),
                ("sigma_F", ctypes.c_double), ("delete_threshold", ctypes.c_double)]


class ZrPhaseFieldStats(ctypes.Structure):
    _fields_ = [("cg_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32), ("residual_norm", ctypes.c_double)]


# damage rules of zr_phase_field_opts.naive
ZR_DAMAGE_PHASE_FIELD, ZR_DAMAGE_RANKINE, ZR_DAMAGE_PHASE_FIELD_RANKINE = 0, 1, 2


def phase_field_opts(parabolic_M=0.0, cg_max_iterations=100, naive=False, cg_tolerance=1e-6, cg_relative=1e-12,
                     sigma_F=0.0, delete_threshold=0.0):
    """FractureSimulation defaults (FractureSimulation.h:50-73,439;
    fracture/ConjugateGradient.h:64).  naive: False / ZR_DAMAGE_PHASE_FIELD,
    True / ZR_DAMAGE_RANKINE (useNaiveDamage) or
    ZR_DAMAGE_PHASE_FIELD_RANKINE (the phase-field solve, then the Rankine
    cut of its c); delete_threshold: deleteParticles' threshold (0 deletes
    nothing)."""
    return ZrPhaseFieldOpts(float(parabolic_M), int(cg_max_iterations), int(naive), float(cg_tolerance),
                            float(cg_relative), float(sigma_F), float(delete_threshold))


def nhb_one_over_sigma_c(percentage, mu, lam):
    """PhaseField::Get_Sigma_C_Based_On_Max_Deformation (PhaseField.h:59-68):
    1 / psi(F = (1 + percentage) I) of NeoHookeanBorden with g = 1, i.e. the
    volumetric energy kappa/2 ((J^2 - 1)/2 - log J), kappa = 0.666666 mu + lambda."""
    J = (1.0 + percentage) ** 3
    kappa = 0.666666 * mu + lam
    return 1.0 / (kappa / 2.0 * ((J * J - 1.0) / 2.0 - np.log(J)))


class ZrParticlesAdjoint(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("X", _dp), ("V", _dp), ("C", _dp), ("F", _dp), ("mu", _dp), ("lam", _dp)]


class ZrAdmmOpts(ctypes.Structure):
    _fields_ = [("admm_iterations", ctypes.c_int32), ("minres_max_iterations", ctypes.c_int32),
                ("rho_scale", ctypes.c_double), ("global_tolerance", ctypes.c_double),
                ("local_tolerance", ctypes.c_double), ("precision", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ZrAdmmStats(ctypes.Structure):
    _fields_ = [("minres_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def admm_opts(admm_iterations=100, minres_max_iterations=10000, rho_scale=0.5, global_tolerance=1e-6,
              local_tolerance=1e-8, precision=ZR_FP64):
    """AdmmSimulation defaults (AdmmSimulationFoam.h:55-68); ``precision``
    ZR_FP32_REFINED runs the global step's operator and Krylov vectors in FP32."""
    return ZrAdmmOpts(int(admm_iterations), int(minres_max_iterations), float(rho_scale), float(global_tolerance),
                      float(local_tolerance), int(precision), 0)


class ZrStepStats(ctypes.Structure):
    _fields_ = [("num_groups", ctypes.c_int64), ("num_pages", ctypes.c_int64), ("cfl_flags", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


_lib = None


def load_library(path=LIB_PATH):
    """Load ``libzr_mpm.so``; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ZrError(f"{path} is missing: build it with `python -m ziran2019_amd.build`")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.zr_create.argtypes = [ctypes.POINTER(ZrConfig), ctypes.POINTER(vp)]
    L.zr_create.restype = ctypes.c_int
    L.zr_destroy.argtypes = [vp]
    L.zr_destroy.restype = None
    L.zr_last_error.argtypes = [vp]
    L.zr_last_error.restype = ctypes.c_char_p
    L.zr_set_colliders.argtypes = [vp, ctypes.POINTER(ZrCollider), ctypes.c_int32]
    L.zr_set_time.argtypes = [vp, ctypes.c_double]
    L.zr_get_time.argtypes = [vp]
    L.zr_get_time.restype = ctypes.c_double
    L.zr_upload_particles.argtypes = [vp, ctypes.POINTER(ZrParticles)]
    L.zr_download_particles.argtypes = [vp, ctypes.POINTER(ZrParticles)]
    L.zr_sort_particles_and_pollute_grid.argtypes = [vp, ctypes.POINTER(ZrStepStats)]
    L.zr_particles_to_grid.argtypes = [vp]
    L.zr_grid_velocity_explicit_update.argtypes = [vp, ctypes.c_double]
    L.zr_grid_to_particles.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrStepStats)]
    L.zr_advance_one_time_step.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrStepStats)]
    L.zr_eval_max_particle_speed.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.zr_eval_max_particle_speed_noalloc.argtypes = [vp, vp, ctypes.c_int64, vp, vp]
    L.zr_set_viscoelastic.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_double]
    L.zr_get_viscoelastic.argtypes = [vp, vp]
    L.zr_calculate_dt.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    L.zr_debug_size.argtypes = [vp, ctypes.c_int32]
    L.zr_debug_size.restype = ctypes.c_int64
    L.zr_debug_dump.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_size_t]
    L.zr_particles_to_grid_apic.argtypes = [vp]
    L.zr_backward_euler_step.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrImplicitOpts),
                                         ctypes.POINTER(ZrImplicitStats)]
    L.zr_grid_to_particles_apic.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrStepStats)]
    L.zr_advance_one_time_step_implicit.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrImplicitOpts),
                                                    ctypes.POINTER(ZrImplicitStats), ctypes.POINTER(ZrStepStats)]
    L.zr_implicit_residual.argtypes = [vp, vp, vp]
    L.zr_implicit_multiply.argtypes = [vp, vp, vp]
    L.zr_set_materials.argtypes = [vp, ctypes.POINTER(ZrMaterials)]
    L.zr_set_phase_field.argtypes = [vp, ctypes.POINTER(ZrPhaseField), ctypes.POINTER(ZrPhaseFieldOpts)]
    L.zr_get_phase_field.argtypes = [vp, ctypes.POINTER(ZrPhaseField)]
    L.zr_phase_field_solve.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrPhaseFieldStats)]
    L.zr_phase_field_last_stats.argtypes = [vp, ctypes.POINTER(ZrPhaseFieldStats)]
    L.zr_get_materials.argtypes = [vp, ctypes.POINTER(ZrMaterials)]
    L.zr_advance_one_time_step_admm.argtypes = [vp, ctypes.c_double, ctypes.POINTER(ZrAdmmOpts),
                                                ctypes.POINTER(ZrAdmmStats), ctypes.POINTER(ZrStepStats)]
    L.zr_get_admm_state.argtypes = [vp, vp, vp, vp]
    L.zr_adjoint_set.argtypes = [vp, ctypes.POINTER(ZrParticlesAdjoint)]
    L.zr_adjoint_get.argtypes = [vp, ctypes.POINTER(ZrParticlesAdjoint)]
    L.zr_adjoint_step.argtypes = [vp, ctypes.c_double]
    L.zr_checkpoint_save.argtypes = [vp, ctypes.c_int32]
    L.zr_checkpoint_restore.argtypes = [vp, ctypes.c_int32]
    L.zr_rollout_forward.argtypes = [vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32]
    L.zr_rollout_backward.argtypes = [vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32]
    L.zr_profile_enable.argtypes = [vp, ctypes.c_int32]
    L.zr_profile_read.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_int64)]
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def grid_log2_for(dx, extent=1.0):
    """Smallest power-of-two node range that holds [0, extent] plus the
    8-node margin the engine keeps."""
    need = int(np.ceil(extent / dx)) + 9
    L = 6
    while (1 << L) < need:
        L += 1
    return L


class MpmEngine:
    """One GPU MPM simulation (FP64 plain path, MLS transfers)."""

    def __init__(self, dx, gravity=(0.0, -9.8, 0.0), cfl=0.6, grid_log2=None, precision=ZR_FP64):
        self.L = load_library()
        cfg = ZrConfig()
        cfg.dx = float(dx)
        cfg.gravity = (ctypes.c_double * 3)(*gravity)
        cfg.cfl = float(cfl)
        cfg.grid_log2 = int(grid_log2 if grid_log2 is not None else grid_log2_for(dx))
        cfg.precision = int(precision)  # ZR_FP32_REFINED: the explicit P2G's per-particle terms in FP32
        cfg.transfer = 0
        h = ctypes.c_void_p()
        rc = self.L.zr_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != ZR_OK:
            raise ZrError(f"zr_create failed with status {rc}")
        self.h = h
        self.dx = dx
        self.n = 0
        self.stats = ZrStepStats()

    def close(self):
        if getattr(self, "h", None):
            self.L.zr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc == ZR_CFL_RESTART:
            raise RestartException(f"{what}: particle travelling more than a grid cell detected")
        if rc != ZR_OK:
            msg = self.L.zr_last_error(self.h).decode()
            raise ZrError(f"{what} failed ({rc}): {msg}")

    # ---------------------------------------------------------- scene setup
    def set_colliders(self, colliders):
        arr = (ZrCollider * max(1, len(colliders)))()
        for k, c in enumerate(colliders):
            arr[k].shape = int(c.get("shape", 0))
            arr[k].type = int(c.get("type", ZR_STICKY))
            arr[k].friction = float(c.get("friction", 0.0))
            arr[k].origin = (ctypes.c_double * 3)(*c.get("origin", (0, 0, 0)))
            arr[k].normal = (ctypes.c_double * 3)(*c.get("normal", (0, 1, 0)))
            arr[k].radius = float(c.get("radius", 0.0))
            arr[k].velocity = (ctypes.c_double * 3)(*c.get("velocity", (0, 0, 0)))
            arr[k].translation = (ctypes.c_double * 3)(*c.get("translation", (0, 0, 0)))
            arr[k].rotation = (ctypes.c_double * 4)(*c.get("rotation", (0, 0, 0, 0)))
            arr[k].omega = (ctypes.c_double * 3)(*c.get("omega", (0, 0, 0)))
            arr[k].scale = float(c.get("scale", 0.0))
            arr[k].dscale_dt = float(c.get("dscale_dt", 0.0))
        self._check(self.L.zr_set_colliders(self.h, arr, len(colliders)), "zr_set_colliders")

    def set_time(self, t):
        self._check(self.L.zr_set_time(self.h, float(t)), "zr_set_time")

    @property
    def time(self):
        return float(self.L.zr_get_time(self.h))

    def upload(self, scene):
        n = len(scene["mass"])
        keep = {}
        for key, nc in (("X", 3), ("V", 3), ("C", 9), ("F", 9), ("mass", 1), ("vol", 1), ("mu", 1), ("lam", 1)):
            a = np.ascontiguousarray(np.asarray(scene[key], dtype=np.float64).reshape(n * nc))
            keep[key] = a
        p = ZrParticles(n, *[_dptr(keep[k]) for k in ("X", "V", "C", "F", "mass", "vol", "mu", "lam")])
        self._check(self.L.zr_upload_particles(self.h, ctypes.byref(p)), "zr_upload_particles")
        self.n = n
        # the per-particle constants the steps do not change, for the restart state
        self.constants = {k: keep[k].copy() for k in ("mass", "vol", "mu", "lam")}

    def set_materials(self, model=None, plasticity=None, plastic_param=None, plastic_state=None, degradation=None):
        """Per-particle models (ZR_MODEL_*), plasticity (ZR_PLASTIC_*), its 4
        parameters per particle, its state and NeoHookeanBorden's g; scalars
        broadcast to every particle."""
        n = self.n

        def i32(a):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.int32)

        def f64(a, shape):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64),
                                                                               shape))
        m, p = i32(model), i32(plasticity)
        q = f64(None if plastic_param is None else np.asarray(plastic_param, dtype=np.float64).reshape(-1, 4),
                (n, 4))
        st, g = f64(plastic_state, (n,)), f64(degradation, (n,))
        mt = ZrMaterials(n, m.ctypes.data_as(_ip) if m is not None else None,
                         p.ctypes.data_as(_ip) if p is not None else None, _dptr(q), _dptr(st), _dptr(g), None)
        self._check(self.L.zr_set_materials(self.h, ctypes.byref(mt)), "set_materials")
        self._mat = dict(model=np.zeros(n, np.int32) if m is None else m.copy(),
                         plasticity=np.zeros(n, np.int32) if p is None else p.copy(),
                         plastic_param=np.zeros((n, 4)) if q is None else q.copy())

    def set_phase_field(self, opts, c=None, H=None, l0=None, one_over_sigma_c=None, pf_Fp=None, H_max=None,
                        residual_phase=None, allow_damage=None):
        n = self.n

        def f(a):
            return None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)))
        arrs = [f(c), f(H), f(l0), f(one_over_sigma_c), f(pf_Fp), f(H_max), f(residual_phase)]
        al = None if allow_damage is None else np.ascontiguousarray(np.broadcast_to(allow_damage, (n,)),
                                                                     dtype=np.int32)
        pf = ZrPhaseField(n, *[_dptr(a) for a in arrs], al.ctypes.data_as(_ip) if al is not None else None)
        self._check(self.L.zr_set_phase_field(self.h, ctypes.byref(pf), ctypes.byref(opts)), "set_phase_field")
        # the fields the steps do not change (defaults of addFBasedMpmForceWithPhaseField)
        dflt = dict(l0=0.0, one_over_sigma_c=0.0, H_max=np.inf, residual_phase=1e-3)
        self._pf = dict(opts=opts, **{k: (np.full(n, dflt[k]) if a is None else a.copy())
                                      for k, a in zip(("l0", "one_over_sigma_c", "H_max", "residual_phase"),
                                                      arrs[2:4] + arrs[5:7])},
                        allow_damage=np.ones(n, np.int32) if al is None else al.copy())

    def get_phase_field(self):
        n = self.n
        c, H, fp = np.zeros(n), np.zeros(n), np.zeros(n)
        pf = ZrPhaseField(n, _dptr(c), _dptr(H), None, None, _dptr(fp), None, None, None)
        self._check(self.L.zr_get_phase_field(self.h, ctypes.byref(pf)), "get_phase_field")
        return dict(c=c, H=H, pf_Fp=fp)

    def phase_field_solve(self, dt):
        st = ZrPhaseFieldStats()
        self._check(self.L.zr_phase_field_solve(self.h, float(dt), ctypes.byref(st)), "phase_field_solve")
        return st

    def phase_field_last_stats(self):
        st = ZrPhaseFieldStats()
        self._check(self.L.zr_phase_field_last_stats(self.h, ctypes.byref(st)), "phase_field_last_stats")
        return st

    # --------------------------------------------------------------- ADMM
    def advanceOneTimeStepAdmm(self, dt, opts=None):
        st = ZrAdmmStats()
        o = opts if opts is not None else admm_opts()
        self._check(self.L.zr_advance_one_time_step_admm(self.h, float(dt), ctypes.byref(o), ctypes.byref(st),
                                                         ctypes.byref(self.stats)), "advanceOneTimeStep (ADMM)")
        return st

    def get_admm_state(self):
        n = self.n
        F, y, om = np.zeros((n, 9)), np.zeros((n, 9)), np.zeros(n)
        self._check(self.L.zr_get_admm_state(self.h, F.ctypes.data_as(ctypes.c_void_p),
                                             y.ctypes.data_as(ctypes.c_void_p),
                                             om.ctypes.data_as(ctypes.c_void_p)), "get_admm_state")
        return dict(F=F, y=y, omega=om)

    # ------------------------------------------------- differentiable form
    def adjoint_set(self, X=None, V=None, C=None, F=None, mu=None, lam=None):
        n = self.n

        def f(a, c):
            return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n * c))
        arrs = [f(X, 3), f(V, 3), f(C, 9), f(F, 9), f(mu, 1), f(lam, 1)]
        a = ZrParticlesAdjoint(n, *[_dptr(x) for x in arrs])
        self._check(self.L.zr_adjoint_set(self.h, ctypes.byref(a)), "adjoint_set")

    def adjoint_get(self):
        n = self.n
        out = dict(X=np.zeros((n, 3)), V=np.zeros((n, 3)), C=np.zeros((n, 9)), F=np.zeros((n, 9)), mu=np.zeros(n),
                   lam=np.zeros(n))
        a = ZrParticlesAdjoint(n, *[_dptr(out[k]) for k in ("X", "V", "C", "F", "mu", "lam")])
        self._check(self.L.zr_adjoint_get(self.h, ctypes.byref(a)), "adjoint_get")
        return out

    def adjoint_step(self, dt):
        self._check(self.L.zr_adjoint_step(self.h, float(dt)), "adjoint_step")

    def checkpoint_save(self, slot):
        self._check(self.L.zr_checkpoint_save(self.h, int(slot)), "checkpoint_save")

    def checkpoint_restore(self, slot):
        self._check(self.L.zr_checkpoint_restore(self.h, int(slot)), "checkpoint_restore")

    def rollout_forward(self, steps, dt, every, slot0=0):
        """``steps`` explicit steps with a checkpoint (slot slot0 + k // every)
        before every every-th step, in one library call."""
        self._check(self.L.zr_rollout_forward(self.h, int(steps), float(dt), int(every), int(slot0)),
                    "rollout_forward")

    def rollout_backward(self, steps, dt, every, slot0=0, scratch0=None):
        """The adjoint of a rollout_forward (after adjoint_set on its output):
        the segments in reverse, each recomputed from its checkpoint into the
        scratch slots, in one library call."""
        nseg = (int(steps) + int(every) - 1) // int(every)
        scratch0 = slot0 + nseg if scratch0 is None else scratch0
        self._check(self.L.zr_rollout_backward(self.h, int(steps), float(dt), int(every), int(slot0),
                                               int(scratch0)), "rollout_backward")

    def get_materials(self):
        n = self.n
        st = np.zeros(n)
        g = np.zeros(n)
        b = np.zeros(n, dtype=np.int32)
        mt = ZrMaterials(n, None, None, None, _dptr(st), _dptr(g), b.ctypes.data_as(_ip))
        self._check(self.L.zr_get_materials(self.h, ctypes.byref(mt)), "get_materials")
        return dict(plastic_state=st, degradation=g, broken=b)

    def download(self):
        n = self.n
        X = np.zeros(3 * n); V = np.zeros(3 * n); C = np.zeros(9 * n); F = np.zeros(9 * n)
        p = ZrParticles(n, _dptr(X), _dptr(V), _dptr(C), _dptr(F), None, None, None, None)
        self._check(self.L.zr_download_particles(self.h, ctypes.byref(p)), "zr_download_particles")
        return dict(X=X.reshape(n, 3), V=V.reshape(n, 3), C=C.reshape(n, 9), F=F.reshape(n, 9))

    # ------------------------------------------------------- restart state
    def write_state(self, f):
        """MpmSimulationBase::writeState (MpmSimulationBase.cpp:731-745) of
        every particle attribute the engine holds, downloaded from the
        device, in DataManager::writeData's framing (ziran2019_amd/state.py);
        f is a binary stream or a path."""
        from . import state
        if isinstance(f, (str, os.PathLike)):
            with open(f, "wb") as fh:
                return self.write_state(fh)
        n, c, out = self.n, self.constants, self.download()
        arrays = {"m": c["mass"], "X": out["X"], "V": out["V"], "C": out["C"], "F": out["F"],
                  "element measure": c["vol"], "mu": c["mu"], "lambda": c["lam"]}
        if getattr(self, "_mat", None) is not None:
            g = self.get_materials()
            arrays.update({"material": self._mat["model"] | (self._mat["plasticity"] << 4),
                           "plastic parameters": self._mat["plastic_param"],
                           "plastic state": g["plastic_state"], "degradation": g["degradation"]})
        if getattr(self, "_visco_set", None) is not None:
            arrays["Fe_Nonequilibrated"] = self.get_viscoelastic()
        if getattr(self, "_pf", None) is not None:
            pf = dict(self.get_phase_field(), vol=c["vol"], **{k: v for k, v in self._pf.items() if k != "opts"})
            arrays["phase field"] = state.pack_phase_field(pf)
        state.write(f, n, arrays)

    def write_partio(self, f):
        """writePartio (Lib/Ziran/Math/Geometry/PartioIO.h:141-181), which
        MpmSimulationBase::writeState calls when write_partio is set: every
        particle's position in a .bgeo file (ziran2019_amd/partio.py)."""
        from . import partio
        partio.write_bgeo(f, self.download()["X"])

    def write_phase_field_partio(self, f):
        """The fracture project's frame output (Projects/fracture/
        FractureSimulation.h:111-144): position, "pf" = c and "sigmaC" =
        1 / one_over_sigma_c (the reciprocal of the float-rounded value),
        1 for particles without a phase field."""
        from . import partio
        X = self.download()["X"]
        pf, sc = np.ones(len(X)), np.ones(len(X))
        if getattr(self, "_pf", None) is not None:
            pf = self.get_phase_field()["c"]
            sc = np.asarray(self._pf["one_over_sigma_c"], dtype=np.float64) * np.ones(len(X))
        with np.errstate(divide="ignore"):
            sigma_c = 1.0 / sc.astype(np.float32).astype(np.float64)
        partio.write_bgeo(f, X, [("pf", partio.VECTOR, pf), ("sigmaC", partio.VECTOR, sigma_c)])

    def read_state(self, f):
        """MpmSimulationBase::readState: the particles of a state written by
        write_state (or by the C++ host) uploaded to the device, with the
        materials, the viscoelastic strain and the phase field when the file
        holds them.  As DataManager::readData needs its arrays created first,
        a file with the viscoelastic strain needs set_viscoelastic (for the
        viscosities) and the phase field keeps this engine's solver options."""
        from . import state
        if isinstance(f, (str, os.PathLike)):
            with open(f, "rb") as fh:
                return self.read_state(fh)
        n, a = state.read(f)
        scene = dict(X=a["X"], V=a["V"], C=a["C"], F=a["F"], mass=a["m"], vol=a["element measure"], mu=a["mu"],
                     lam=a["lambda"])
        pf_opts = self._pf["opts"] if getattr(self, "_pf", None) is not None else phase_field_opts()
        visco = getattr(self, "_visco_set", None)
        self.upload(scene)
        self._mat = self._pf = self._visco_set = None
        if "material" in a:
            self.set_materials(model=a["material"] & 15, plasticity=a["material"] >> 4,
                               plastic_param=a["plastic parameters"], plastic_state=a["plastic state"],
                               degradation=a["degradation"])
        if "Fe_Nonequilibrated" in a:
            if visco is None:
                raise ZrError("read_state: the file holds a viscoelastic strain; call set_viscoelastic first")
            self.set_viscoelastic(a["Fe_Nonequilibrated"], visco[0], visco[1], visco[2], visco[3])
        if "phase field" in a:
            pf = state.unpack_phase_field(a["phase field"])
            self.set_phase_field(pf_opts, c=pf["c"], H=pf["H"], l0=pf["l0"], one_over_sigma_c=pf["one_over_sigma_c"],
                                 pf_Fp=pf["pf_Fp"], H_max=pf["H_max"], residual_phase=pf["residual_phase"],
                                 allow_damage=pf["allow_damage"])
        return n

    # ---------------------------------------------------------- step API
    def sortParticlesAndPolluteGrid(self):
        self._check(self.L.zr_sort_particles_and_pollute_grid(self.h, ctypes.byref(self.stats)),
                    "sortParticlesAndPolluteGrid")

    def particlesToGrid(self):
        self._check(self.L.zr_particles_to_grid(self.h), "particlesToGrid")

    def gridVelocityExplicitUpdate(self, dt):
        self._check(self.L.zr_grid_velocity_explicit_update(self.h, float(dt)), "gridVelocityExplicitUpdate")

    def gridToParticles(self, dt):
        self._check(self.L.zr_grid_to_particles(self.h, float(dt), ctypes.byref(self.stats)), "gridToParticles")

    def advanceOneTimeStep(self, dt):
        self._check(self.L.zr_advance_one_time_step(self.h, float(dt), ctypes.byref(self.stats)),
                    "advanceOneTimeStep")

    # ---------------------------------------------------------- viscoelastic
    def set_viscoelastic(self, F_neq, mu, lam, viscosity_d, viscosity_v):
        """FElasticNonequilibratedBasedMpmForceHelper: F_neq (n x 9 column-major,
        None: identity), its StvkWithHencky mu / lambda (scalars or per
        particle) and the two viscosities; mu = lam = None switches it off."""
        n = self.n
        if mu is None and lam is None:
            self._check(self.L.zr_set_viscoelastic(self.h, None, None, None, 1.0, 1.0), "zr_set_viscoelastic")
            return
        self._visco = [None if F_neq is None else np.ascontiguousarray(np.asarray(F_neq, dtype=np.float64).reshape(9 * n)),
                       np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,))),
                       np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (n,)))]
        f, m, l = (None if a is None else a.ctypes.data_as(ctypes.c_void_p) for a in self._visco)
        self._check(self.L.zr_set_viscoelastic(self.h, f, m, l, float(viscosity_d), float(viscosity_v)),
                    "zr_set_viscoelastic")
        self._visco_set = (self._visco[1].copy(), self._visco[2].copy(), float(viscosity_d), float(viscosity_v))

    def get_viscoelastic(self):
        out = np.zeros(9 * self.n)
        self._check(self.L.zr_get_viscoelastic(self.h, out.ctypes.data_as(ctypes.c_void_p)), "zr_get_viscoelastic")
        return out.reshape(self.n, 9)

    # ---------------------------------------------------------- adaptive dt
    def evalMaxParticleSpeed(self):
        """(max |v|, max corner, min corner), MpmSimulationBase::evalMaxParticleSpeed."""
        out = (ctypes.c_double * 7)()
        self._check(self.L.zr_eval_max_particle_speed(self.h, out), "evalMaxParticleSpeed")
        return out[0], np.array(out[1:4]), np.array(out[4:7])

    def calculateDt(self, max_dt):
        """MpmSimulationBase::calculateDt: cfl dx / max speed, or max_dt."""
        dt = ctypes.c_double()
        self._check(self.L.zr_calculate_dt(self.h, float(max_dt), ctypes.byref(dt)), "calculateDt")
        return dt.value

    def advanceOneFrame(self, step, advance=None):
        """SimulationBase::advanceOneFrame (SimulationBase.h:288-321): substeps
        of step.nextDt(calculateDt()) until the frame time is reached; advance
        is the substep (advanceOneTimeStep by default).  A CFL violation raises
        RestartException with the substep's state on the device.  Returns the
        substep sizes."""
        advance = advance if advance is not None else self.advanceOneTimeStep
        dts = []
        while True:
            dt = step.nextDt(self.calculateDt(step.max_dt))
            advance(dt)
            dts.append(dt)
            if step.advance(dt) == TimeStepping.FRAME:
                return dts

    # ---------------------------------------------------------- implicit
    def particlesToGridApic(self):
        self._check(self.L.zr_particles_to_grid_apic(self.h), "particlesToGrid (APIC)")

    def backwardEulerStep(self, dt, opts=None):
        st = ZrImplicitStats()
        o = opts if opts is not None else implicit_opts()
        self._check(self.L.zr_backward_euler_step(self.h, float(dt), ctypes.byref(o), ctypes.byref(st)),
                    "backwardEulerStep")
        return st

    def gridToParticlesApic(self, dt):
        self._check(self.L.zr_grid_to_particles_apic(self.h, float(dt), ctypes.byref(self.stats)),
                    "gridToParticles (APIC)")

    def advanceOneTimeStepImplicit(self, dt, opts=None):
        st = ZrImplicitStats()
        o = opts if opts is not None else implicit_opts()
        self._check(self.L.zr_advance_one_time_step_implicit(self.h, float(dt), ctypes.byref(o), ctypes.byref(st),
                                                             ctypes.byref(self.stats)), "advanceOneTimeStep (implicit)")
        return st

    def implicit_residual(self, dv):
        dv = np.ascontiguousarray(dv, dtype=np.float64).reshape(-1)
        r = np.zeros_like(dv)
        self._check(self.L.zr_implicit_residual(self.h, dv.ctypes.data_as(ctypes.c_void_p),
                                                r.ctypes.data_as(ctypes.c_void_p)), "zr_implicit_residual")
        return r.reshape(-1, 3)

    def implicit_multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        b = np.zeros_like(x)
        self._check(self.L.zr_implicit_multiply(self.h, x.ctypes.data_as(ctypes.c_void_p),
                                                b.ctypes.data_as(ctypes.c_void_p)), "zr_implicit_multiply")
        return b.reshape(-1, 3)

    # ---------------------------------------------------------- timing
    def profile(self, on=True, every=1):
        """Time kernel launches with HIP events; every k-th launch of each
        kernel class when every = k."""
        self._check(self.L.zr_profile_enable(self.h, int(every) if on else 0), "zr_profile_enable")

    def kernel_time(self, which):
        """(total milliseconds, launches) of a kernel class since profiling was enabled."""
        t = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self.L.zr_profile_read(self.h, PROF[which], ctypes.byref(t), ctypes.byref(n)), "zr_profile_read")
        return t.value, n.value

    # ---------------------------------------------------------- introspection
    def dump(self, what):
        code = DUMP[what]
        size = int(self.L.zr_debug_size(self.h, code))
        if size < 0:
            raise ZrError(f"bad dump {what}")
        out = np.zeros(size // np.dtype(_DUMP_DTYPE[what]).itemsize, dtype=_DUMP_DTYPE[what])
        self._check(self.L.zr_debug_dump(self.h, code, out.ctypes.data_as(ctypes.c_void_p), size), "zr_debug_dump")
        if what in ("grid_v", "grid_f", "grid_vstar", "dv", "jacobi"):
            out = out.reshape(-1, 3)
        return out
