"""FP64 CPU oracle for the MPM time step -- TEST INFRASTRUCTURE ONLY.

ctypes loader for ``oracle/_build/libzr_oracle.so`` (built by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product path
(``ziran2019_amd``) never does.

The C sources restate penn-graphics-research/ziran2019 (file:line citations in
each function).  Arrays follow the C-ABI conventions: vectors are 3 doubles per
particle, matrices 9 doubles per particle in column-major (Eigen) order.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libzr_oracle.so")

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_p = ctypes.c_void_p


class ZroImplicitOpts(ctypes.Structure):
    _fields_ = [("newton_max_iterations", ctypes.c_int32), ("newton_tolerance", _d),
                ("cg_max_iterations", ctypes.c_int32), ("cg_tolerance", _d), ("cg_relative_tolerance", _d),
                ("preconditioner", ctypes.c_int32), ("solver", ctypes.c_int32)]


class ZroImplicitStats(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_int32), ("cg_iterations", ctypes.c_int32),
                ("residual_norm", _d * 8), ("num_residuals", ctypes.c_int32), ("converged", ctypes.c_int32)]


def implicit_opts(newton_max_iterations=3, newton_tolerance=1e-5, cg_max_iterations=20,
                  cg_tolerance=16 * 2.220446049250313e-16, cg_relative_tolerance=0.0, preconditioner=0, solver=0):
    """Defaults of the reference: NewtonsMethod(1e-5, 3) (MpmSimulationBase.cpp:51),
    ConjugateGradient(20) with LinearSolver::tolerance 16 eps (LinearSolver.h:26).
    solver=1 is Minres (the reference's default linear solver, BackwardEuler.h:55-62,
    with tolerance 1: pass cg_tolerance=1.0 for the reference's setting)."""
    return ZroImplicitOpts(int(newton_max_iterations), float(newton_tolerance), int(cg_max_iterations),
                           float(cg_tolerance), float(cg_relative_tolerance), int(preconditioner), int(solver))


class ZroPfOpts(ctypes.Structure):
    _fields_ = [("parabolic_M", _d), ("cg_max_iterations", ctypes.c_int32), ("naive", ctypes.c_int32),
                ("cg_tolerance", _d), ("cg_relative", _d), ("sigma_F", _d), ("delete_threshold", _d)]


class ZroPfStats(ctypes.Structure):
    _fields_ = [("cg_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32), ("residual_norm", _d)]


def pf_opts(parabolic_M=0.0, cg_max_iterations=100, naive=False, cg_tolerance=1e-6, cg_relative=1e-12,
            sigma_F=0.0, delete_threshold=0.0):
    """FractureSimulation defaults: ConjugateGradient(100), tolerance 1e-6,
    relative 1e-12 in double (FractureSimulation.h:70-73,439;
    fracture/ConjugateGradient.h:64).  naive: 0 phase field, 1 (True) the
    Rankine naive damage, 2 the phase field then the Rankine cut;
    delete_threshold: deleteParticles' threshold."""
    return ZroPfOpts(float(parabolic_M), int(cg_max_iterations), int(naive), float(cg_tolerance),
                     float(cg_relative), float(sigma_F), float(delete_threshold))


class ZroAdmmOpts(ctypes.Structure):
    _fields_ = [("admm_iterations", ctypes.c_int32), ("minres_max_iterations", ctypes.c_int32),
                ("rho_scale", _d), ("global_tolerance", _d), ("local_tolerance", _d)]


class ZroAdmmStats(ctypes.Structure):
    _fields_ = [("minres_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def admm_opts(admm_iterations=100, minres_max_iterations=10000, rho_scale=0.5, global_tolerance=1e-6,
              local_tolerance=1e-8):
    """AdmmSimulation defaults (AdmmSimulationFoam.h:55-68)."""
    return ZroAdmmOpts(int(admm_iterations), int(minres_max_iterations), float(rho_scale), float(global_tolerance),
                       float(local_tolerance))


class ZroCollider(ctypes.Structure):
    _fields_ = [
        ("shape", ctypes.c_int32),
        ("type", ctypes.c_int32),
        ("friction", _d),
        ("origin", _d * 3),
        ("normal", _d * 3),
        ("radius", _d),
        ("velocity", _d * 3),
        ("translation", _d * 3),
        ("rotation", _d * 4),
        ("omega", _d * 3),
        ("scale", _d),
        ("dscale_dt", _d),
    ]


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.zro_linear_offset.restype = _u64
        L.zro_linear_offset.argtypes = [_i, _i, _i]
        L.zro_linear_to_coord.argtypes = [_u64, _p]
        L.zro_packed_add.restype = _u64
        L.zro_packed_add.argtypes = [_u64, _u64]
        L.zro_bspline_weights.argtypes = [_d, _p, _p, _p]
        L.zro_svd3.restype = _i
        L.zro_svd3.argtypes = [_p, _p, _p, _p]
        L.zro_polar3.argtypes = [_p, _p, _p]
        L.zro_fcr_psi.restype = _d
        L.zro_fcr_psi.argtypes = [_d, _d, _p]
        for f in ("zro_fcr_first_piola", "zro_fcr_kirchhoff"):
            getattr(L, f).argtypes = [_d, _d, _p, _p]
        L.zro_fcr_first_piola_differential.argtypes = [_d, _d, _p, _p, _p]
        L.zro_collider_frame.argtypes = [ctypes.POINTER(ZroCollider), _d, _p, _p, _p]
        L.zro_project_sigma.argtypes = [_d, _d, _i, _p, _d, _p, _p]
        L.zro_project_sigma_and_derivative.argtypes = [_d, _d, _i, _p, _d, _p, _p, _p]
        L.zro_fake_separate.restype = _i
        L.zro_fake_separate.argtypes = [ctypes.POINTER(ZroCollider), _i, _d, _p, _p, _p]
        L.zro_lame.argtypes = [_d, _d, _p, _p]
        L.zro_create.restype = _p
        L.zro_create.argtypes = [_d, _p, _d, _i]
        L.zro_destroy.argtypes = [_p]
        L.zro_set_particles.argtypes = [_p, _i64] + [_p] * 8
        L.zro_get_particles.argtypes = [_p, _p, _p, _p, _p]
        L.zro_add_collider.restype = _i
        L.zro_add_collider.argtypes = [_p, ctypes.POINTER(ZroCollider)]
        L.zro_set_time.argtypes = [_p, _d]
        L.zro_sort_and_pollute.argtypes = [_p]
        L.zro_num_groups.restype = _i64
        L.zro_num_groups.argtypes = [_p]
        L.zro_num_pages.restype = _i64
        L.zro_num_pages.argtypes = [_p]
        L.zro_get_sort.argtypes = [_p, _p, _p, _p]
        L.zro_get_groups.argtypes = [_p, _p, _p, _p]
        L.zro_get_pages.argtypes = [_p, _p]
        L.zro_update_stress.argtypes = [_p]
        L.zro_p2g_mls.restype = _i64
        L.zro_p2g_mls.argtypes = [_p]
        L.zro_grid_update.argtypes = [_p, _d]
        L.zro_g2p_mls.restype = _i
        L.zro_g2p_mls.argtypes = [_p, _d]
        L.zro_step_explicit.restype = _i
        L.zro_step_explicit.argtypes = [_p, _d]
        L.zro_get_grid.argtypes = [_p, _p, _p, _p, _p]
        L.zro_p2g_apic.restype = _i64
        L.zro_p2g_apic.argtypes = [_p]
        L.zro_backward_euler.restype = _i
        L.zro_backward_euler.argtypes = [_p, _d, ctypes.POINTER(ZroImplicitOpts), ctypes.POINTER(ZroImplicitStats)]
        L.zro_g2p_apic.restype = _i
        L.zro_g2p_apic.argtypes = [_p, _d]
        L.zro_step_implicit.restype = _i
        L.zro_step_implicit.argtypes = [_p, _d, ctypes.POINTER(ZroImplicitOpts), ctypes.POINTER(ZroImplicitStats)]
        L.zro_implicit_multiply.argtypes = [_p, _p, _p]
        L.zro_implicit_residual.argtypes = [_p, _p, _p]
        L.zro_implicit_energy.restype = _d
        L.zro_implicit_energy.argtypes = [_p, _p]
        L.zro_num_node_slots.restype = _i64
        L.zro_num_node_slots.argtypes = [_p]
        L.zro_get_implicit.argtypes = [_p, _p, _p, _p]
        L.zro_implicit_diagonal.argtypes = [_p, _p]
        for f in ("zro_stvk_hencky_psi",):
            getattr(L, f).restype = _d
            getattr(L, f).argtypes = [_d, _d, _p]
        for f in ("zro_stvk_hencky_first_piola", "zro_stvk_hencky_kirchhoff"):
            getattr(L, f).argtypes = [_d, _d, _p, _p]
        L.zro_nhb_psi.restype = _d
        L.zro_nhb_psi.argtypes = [_d, _d, _d, _p]
        L.zro_nhb_psi_pos.restype = _d
        L.zro_nhb_psi_pos.argtypes = [_d, _d, _p]
        L.zro_nhb_kirchhoff.argtypes = [_d, _d, _d, _p, _p]
        L.zro_von_mises_project.restype = _i
        L.zro_von_mises_project.argtypes = [_d, _d, _d, _d, _p, _p, _p]
        L.zro_drucker_prager_alpha.restype = _d
        L.zro_drucker_prager_alpha.argtypes = [_d]
        L.zro_drucker_prager_project.restype = _i
        L.zro_drucker_prager_project.argtypes = [_d, _d, _d, _d, _d, _i, _p, _p]
        L.zro_set_viscoelastic.argtypes = [_p, _p, _p, _p, _d, _d]
        L.zro_get_viscoelastic.argtypes = [_p, _p]
        L.zro_visco_relax.argtypes = [_d, _d, _d, _d, _d, _p, _p]
        L.zro_admm_visco_map.argtypes = [_d, _d, _d, _d, _d, _p, _p, _p]
        L.zro_cam_clay_M.restype = _d
        L.zro_cam_clay_M.argtypes = [_d, _i]
        L.zro_cam_clay_project.restype = _i
        L.zro_cam_clay_project.argtypes = [_d, _d, _d, _d, _d, _p, _p]
        L.zro_set_materials.argtypes = [_p, _p, _p, _p, _p, _p]
        L.zro_set_phase_field.argtypes = [_p] + [_p] * 8 + [ctypes.POINTER(ZroPfOpts)]
        L.zro_get_phase_field.argtypes = [_p, _p, _p, _p]
        L.zro_phase_field_solve.argtypes = [_p, _d, ctypes.POINTER(ZroPfStats)]
        L.zro_phase_field_system.argtypes = [_p, _p, _p, _p]
        L.zro_phase_field_multiply.argtypes = [_p, _p, _p]
        L.zro_fcr_kirchhoff_adjoint.argtypes = [_d, _d, _p, _p, _p, _p, _p]
        L.zro_adjoint_step.restype = _i
        L.zro_adjoint_step.argtypes = [_p, _d] + [_p] * 6
        L.zro_step_admm.restype = _i
        L.zro_step_admm.argtypes = [_p, _d, ctypes.POINTER(ZroAdmmOpts), ctypes.POINTER(ZroAdmmStats)]
        L.zro_get_admm.argtypes = [_p, _p, _p, _p]
        L.zro_sym3_eigenvalues.argtypes = [_p, _p]
        L.zro_sym3_max_eigenvalue.restype = _d
        L.zro_sym3_max_eigenvalue.argtypes = [_p]
        L.zro_get_materials.argtypes = [_p, _p, _p, _p]
        L.zro_eval_max_particle_speed.argtypes = [_p, _p]
        L.zro_calculate_dt.restype = _d
        L.zro_calculate_dt.argtypes = [_p, _d]
        L.zro_next_dt.restype = _d
        L.zro_next_dt.argtypes = [_d] * 6
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.size == n, (a.shape, n)
    return a


# ---------------------------------------------------------------- functions
def linear_offset(i, j, k):
    return int(lib().zro_linear_offset(int(i), int(j), int(k)))


def linear_to_coord(off):
    out = np.zeros(3, dtype=np.int32)
    lib().zro_linear_to_coord(ctypes.c_uint64(off), _ptr(out))
    return tuple(int(v) for v in out)


def packed_add(a, b):
    return int(lib().zro_packed_add(ctypes.c_uint64(a), ctypes.c_uint64(b)))


def bspline_weights(x):
    base = np.zeros(1, dtype=np.int32)
    w = np.zeros(3)
    dw = np.zeros(3)
    lib().zro_bspline_weights(float(x), _ptr(base), _ptr(w), _ptr(dw))
    return int(base[0]), w, dw


def _cm(M):
    """3x3 numpy matrix -> column-major 9-vector."""
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64).T.reshape(9))


def _from_cm(a):
    return np.asarray(a, dtype=np.float64).reshape(3, 3).T.copy()


def svd3(F):
    U = np.zeros(9)
    s = np.zeros(3)
    V = np.zeros(9)
    n = lib().zro_svd3(_ptr(_cm(F)), _ptr(U), _ptr(s), _ptr(V))
    return _from_cm(U), s, _from_cm(V), n


def polar3(F):
    R = np.zeros(9)
    S = np.zeros(9)
    lib().zro_polar3(_ptr(_cm(F)), _ptr(R), _ptr(S))
    return _from_cm(R), _from_cm(S)


def lame(E, nu):
    mu = np.zeros(1)
    lam = np.zeros(1)
    lib().zro_lame(float(E), float(nu), _ptr(mu), _ptr(lam))
    return float(mu[0]), float(lam[0])


def stvk_hencky_psi(mu, lam, F):
    return float(lib().zro_stvk_hencky_psi(mu, lam, _ptr(_cm(F))))


def stvk_hencky_first_piola(mu, lam, F):
    P = np.zeros(9)
    lib().zro_stvk_hencky_first_piola(mu, lam, _ptr(_cm(F)), _ptr(P))
    return _from_cm(P)


def stvk_hencky_kirchhoff(mu, lam, F):
    t = np.zeros(9)
    lib().zro_stvk_hencky_kirchhoff(mu, lam, _ptr(_cm(F)), _ptr(t))
    return _from_cm(t)


def nhb_psi(mu, lam, g, F):
    return float(lib().zro_nhb_psi(mu, lam, g, _ptr(_cm(F))))


def nhb_psi_pos(mu, lam, F):
    return float(lib().zro_nhb_psi_pos(mu, lam, _ptr(_cm(F))))


def nhb_kirchhoff(mu, lam, g, F):
    t = np.zeros(9)
    lib().zro_nhb_kirchhoff(mu, lam, g, _ptr(_cm(F)), _ptr(t))
    return _from_cm(t)


def von_mises_project(mu, lam, xi, fail_stress, yield_stress, F):
    """Returns (F_new, yield_stress_new, broken, changed)."""
    y = np.array([yield_stress], dtype=np.float64)
    f = _cm(F)
    b = np.zeros(1, dtype=np.int32)
    ch = lib().zro_von_mises_project(mu, lam, xi, fail_stress, _ptr(y), _ptr(f), _ptr(b))
    return _from_cm(f), float(y[0]), bool(b[0]), bool(ch)


def drucker_prager_alpha(friction_angle_deg):
    return float(lib().zro_drucker_prager_alpha(float(friction_angle_deg)))


def drucker_prager_project(mu, lam, alpha, beta, cohesion, volume_correction, logJp, F):
    """Returns (F_new, logJp_new)."""
    j = np.array([logJp], dtype=np.float64)
    f = _cm(F)
    lib().zro_drucker_prager_project(mu, lam, alpha, beta, cohesion, int(bool(volume_correction)), _ptr(j), _ptr(f))
    return _from_cm(f), float(j[0])


def visco_relax(mu, lam, viscosity_d, viscosity_v, dt, gradV, Fneq):
    """evolveStrain of FElasticNonequilibratedBasedMpmForceHelper on one Fneq."""
    f = _cm(Fneq)
    lib().zro_visco_relax(float(mu), float(lam), float(viscosity_d), float(viscosity_v), float(dt), _ptr(_cm(gradV)),
                          _ptr(f))
    return _from_cm(f)


def admm_visco_map(mu, lam, viscosity_d, viscosity_v, dt, z):
    """The ADMM local step's viscoelastic return map: (relaxed singular values,
    the derivative the reference forms, row-major)."""
    z = np.ascontiguousarray(z, dtype=np.float64)
    zp, D = np.zeros(3), np.zeros((3, 3))
    lib().zro_admm_visco_map(float(mu), float(lam), float(viscosity_d), float(viscosity_v), float(dt), _ptr(z),
                             _ptr(zp), _ptr(D))
    return zp, D


def cam_clay_M(friction_angle_deg, dim=3):
    return float(lib().zro_cam_clay_M(float(friction_angle_deg), int(dim)))


def cam_clay_project(mu, lam, M, beta, xi, logJp, F):
    """Returns (F_new, logJp_new, moved)."""
    j = np.array([logJp], dtype=np.float64)
    f = _cm(F)
    moved = lib().zro_cam_clay_project(mu, lam, M, beta, xi, _ptr(j), _ptr(f))
    return _from_cm(f), float(j[0]), bool(moved)


def fcr_kirchhoff_adjoint(mu, lam, F, taubar, Fbar=None):
    """(Fbar + (dtau/dF)^T taubar, mubar, lambar) of the fixed-corotated tau."""
    fb = _cm(np.zeros((3, 3)) if Fbar is None else Fbar)
    mb = np.zeros(1)
    lb = np.zeros(1)
    lib().zro_fcr_kirchhoff_adjoint(float(mu), float(lam), _ptr(_cm(F)), _ptr(_cm(taubar)), _ptr(fb), _ptr(mb),
                                    _ptr(lb))
    return _from_cm(fb), float(mb[0]), float(lb[0])


def sym3_eigenvalues(A):
    lam = np.zeros(3)
    lib().zro_sym3_eigenvalues(_ptr(_cm(A)), _ptr(lam))
    return lam


def sym3_max_eigenvalue(A):
    return float(lib().zro_sym3_max_eigenvalue(_ptr(_cm(A))))


def fcr_psi(mu, lam, F):
    return float(lib().zro_fcr_psi(mu, lam, _ptr(_cm(F))))


def fcr_first_piola(mu, lam, F):
    P = np.zeros(9)
    lib().zro_fcr_first_piola(mu, lam, _ptr(_cm(F)), _ptr(P))
    return _from_cm(P)


def fcr_kirchhoff(mu, lam, F):
    t = np.zeros(9)
    lib().zro_fcr_kirchhoff(mu, lam, _ptr(_cm(F)), _ptr(t))
    return _from_cm(t)


def fcr_first_piola_differential(mu, lam, F, dF):
    dP = np.zeros(9)
    lib().zro_fcr_first_piola_differential(mu, lam, _ptr(_cm(F)), _ptr(_cm(dF)), _ptr(dP))
    return _from_cm(dP)


def next_dt(dt, min_dt, max_dt, max_dt_original, frame_dt, time_since_last_frame):
    """TimeStepping::nextDt (TimeStepping.h:45-61)."""
    return float(lib().zro_next_dt(float(dt), float(min_dt), float(max_dt), float(max_dt_original),
                                   float(frame_dt), float(time_since_last_frame)))


# ---------------------------------------------------------------- simulation
def make_collider(shape, type, origin, normal=(0, 1, 0), radius=0.0, velocity=(0, 0, 0), friction=0.0,
                  translation=(0, 0, 0), rotation=(0, 0, 0, 0), omega=(0, 0, 0), scale=0.0, dscale_dt=0.0):
    """The oracle's collider record (zro_add_collider normalises a
    half-space normal, as HalfSpace does, AnalyticLevelSet.cpp:262-266;
    zro_fake_separate and zro_collider_frame take it as given)."""
    return ZroCollider(int(shape), int(type), float(friction), (_d * 3)(*origin),
                       (_d * 3)(*normal), float(radius), (_d * 3)(*velocity), (_d * 3)(*translation),
                       (_d * 4)(*rotation), (_d * 3)(*omega), float(scale), float(dscale_dt))


def project_sigma(mu, lam, plastic, pparam, pstate, sig):
    """projectSigma of the ADMM local step (PlasticityApplier.cpp:319-346,
    526-547)."""
    q = np.ascontiguousarray(pparam, dtype=np.float64)
    sv = np.ascontiguousarray(sig, dtype=np.float64)
    out = np.zeros(3)
    lib().zro_project_sigma(float(mu), float(lam), int(plastic), _ptr(q), float(pstate), _ptr(sv), _ptr(out))
    return out


def project_sigma_and_derivative(mu, lam, plastic, pparam, pstate, sig):
    """projectSigmaAndDerivative (PlasticityApplier.cpp:389-423, 580-606):
    (projected sigma, d/d sigma as a 3x3 matrix)."""
    q = np.ascontiguousarray(pparam, dtype=np.float64)
    sv = np.ascontiguousarray(sig, dtype=np.float64)
    zp, D = np.zeros(3), np.zeros(9)
    lib().zro_project_sigma_and_derivative(float(mu), float(lam), int(plastic), _ptr(q), float(pstate), _ptr(sv),
                                           _ptr(zp), _ptr(D))
    return zp, D.reshape(3, 3)


def collider_frame(t, **kw):
    """(b, R, s) of a collider at time t: the state updateState(t) sets for
    constant rates (translation velocity, angular velocity, ds/dt)."""
    c = make_collider(**kw)
    b = np.zeros(3); R = np.zeros(9); s = np.zeros(1)
    lib().zro_collider_frame(ctypes.byref(c), float(t), _ptr(b), _ptr(R), _ptr(s))
    return b, R.reshape(3, 3), float(s[0])


def fake_separate(colliders, x, v, t=0.0):
    """multiObjectCollisionWithFakeSeparate on one node: (any, v, P)."""
    arr = (ZroCollider * max(len(colliders), 1))(*[make_collider(**c) for c in colliders])
    xv = np.ascontiguousarray(x, dtype=np.float64)
    vv = np.array(v, dtype=np.float64)
    P = np.zeros(9)
    hit = lib().zro_fake_separate(arr, len(colliders), float(t), _ptr(xv), _ptr(vv), _ptr(P))
    return bool(hit), vv, P.reshape(3, 3)


class OracleSim:
    """FP64 CPU MPM simulation (explicit MLS step), restating
    MpmSimulationBase<double,3> (Lib/MPM/MpmSimulationBase.cpp)."""

    def __init__(self, dx, gravity=(0.0, -9.8, 0.0), cfl=0.6, nthreads=1):
        g = np.asarray(gravity, dtype=np.float64)
        self._g = g
        self.h = lib().zro_create(float(dx), _ptr(g), float(cfl), int(nthreads))
        self.dx = dx
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None):
            lib().zro_destroy(self.h)
            self.h = None

    def set_particles(self, X, V, C, F, mass, vol, mu, lam):
        n = len(mass)
        self.n = n
        arrs = [_f64(X, 3 * n), _f64(V, 3 * n), _f64(C, 9 * n), _f64(F, 9 * n),
                _f64(mass, n), _f64(vol, n), _f64(mu, n), _f64(lam, n)]
        lib().zro_set_particles(self.h, n, *[_ptr(a) for a in arrs])

    def set_materials(self, model=None, plastic=None, pparam=None, pstate=None, g=None):
        n = self.n
        m = None if model is None else np.ascontiguousarray(np.broadcast_to(model, (n,)), dtype=np.int32)
        p = None if plastic is None else np.ascontiguousarray(np.broadcast_to(plastic, (n,)), dtype=np.int32)
        q = None if pparam is None else _f64(np.broadcast_to(np.asarray(pparam, dtype=np.float64).reshape(-1, 4),
                                                             (n, 4)), 4 * n)
        st = None if pstate is None else _f64(np.broadcast_to(pstate, (n,)), n)
        gg = None if g is None else _f64(np.broadcast_to(g, (n,)), n)
        self._mat_keep = (m, p, q, st, gg)
        lib().zro_set_materials(self.h, _ptr(m), _ptr(p), _ptr(q), _ptr(st), _ptr(gg))

    def set_phase_field(self, opts, c=None, H=None, l0=None, one_over_sigma_c=None, pf_Fp=None, H_max=None,
                        residual_phase=None, allow_damage=None):
        n = self.n

        def f(a):
            return None if a is None else _f64(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,)), n)
        arrs = [f(c), f(H), f(l0), f(one_over_sigma_c), f(pf_Fp), f(H_max), f(residual_phase)]
        al = None if allow_damage is None else np.ascontiguousarray(np.broadcast_to(allow_damage, (n,)),
                                                                     dtype=np.int32)
        self._pf_keep = (arrs, al, opts)
        lib().zro_set_phase_field(self.h, *[_ptr(a) for a in arrs], _ptr(al), ctypes.byref(opts))

    def get_phase_field(self):
        n = self.n
        c, H, fp = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().zro_get_phase_field(self.h, _ptr(c), _ptr(H), _ptr(fp))
        return dict(c=c, H=H, pf_Fp=fp)

    def phase_field_solve(self, dt):
        st = ZroPfStats()
        lib().zro_phase_field_solve(self.h, float(dt), ctypes.byref(st))
        return st

    def phase_field_system(self):
        ns = self.num_node_slots()
        rhs, diag, x = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        lib().zro_phase_field_system(self.h, _ptr(rhs), _ptr(diag), _ptr(x))
        return rhs, diag, x

    def phase_field_multiply(self, x):
        x = _f64(x, self.num_node_slots())
        b = np.zeros_like(x)
        lib().zro_phase_field_multiply(self.h, _ptr(x), _ptr(b))
        return b

    def adjoint_step(self, dt, aX, aV, aC, aF, amu, alam):
        """Reverse of one explicit step from the sim's current (input) state;
        aX (n,3), aV (n,3), aC (n,9), aF (n,9) are updated in place from the
        output adjoint to the input adjoint; amu, alam (n,) accumulate."""
        for a in (aX, aV, aC, aF, amu, alam):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        rc = lib().zro_adjoint_step(self.h, float(dt), _ptr(aX), _ptr(aV), _ptr(aC), _ptr(aF), _ptr(amu),
                                    _ptr(alam))
        if rc != 0:
            raise RuntimeError("adjoint: a collider branch without a derivative was hit")

    def step_admm(self, dt, opts=None):
        st = ZroAdmmStats()
        o = opts if opts is not None else admm_opts()
        flags = int(lib().zro_step_admm(self.h, float(dt), ctypes.byref(o), ctypes.byref(st)))
        return flags, st

    def get_admm(self):
        n = self.n
        F, y, om = np.zeros(9 * n), np.zeros(9 * n), np.zeros(n)
        lib().zro_get_admm(self.h, _ptr(F), _ptr(y), _ptr(om))
        return dict(F=F.reshape(n, 9), y=y.reshape(n, 9), omega=om)

    def get_materials(self):
        n = self.n
        st = np.zeros(n)
        b = np.zeros(n, dtype=np.int32)
        g = np.zeros(n)
        lib().zro_get_materials(self.h, _ptr(st), _ptr(b), _ptr(g))
        return dict(pstate=st, broken=b, g=g)

    def add_collider(self, **kw):
        c = make_collider(**kw)
        assert lib().zro_add_collider(self.h, ctypes.byref(c)) == 0

    def set_time(self, t):
        lib().zro_set_time(self.h, float(t))

    def get_particles(self):
        n = self.n
        X = np.zeros(3 * n); V = np.zeros(3 * n); C = np.zeros(9 * n); F = np.zeros(9 * n)
        lib().zro_get_particles(self.h, _ptr(X), _ptr(V), _ptr(C), _ptr(F))
        return X.reshape(n, 3), V.reshape(n, 3), C.reshape(n, 9), F.reshape(n, 9)

    def sort_and_pollute(self):
        lib().zro_sort_and_pollute(self.h)

    def sort_result(self):
        n = self.n
        sorter = np.zeros(n, dtype=np.uint64)
        order = np.zeros(n, dtype=np.int32)
        base = np.zeros(n, dtype=np.uint64)
        lib().zro_get_sort(self.h, _ptr(sorter), _ptr(order), _ptr(base))
        ng = lib().zro_num_groups(self.h)
        first = np.zeros(ng, dtype=np.int64)
        last = np.zeros(ng, dtype=np.int64)
        blk = np.zeros(ng, dtype=np.uint64)
        lib().zro_get_groups(self.h, _ptr(first), _ptr(last), _ptr(blk))
        npg = lib().zro_num_pages(self.h)
        pages = np.zeros(npg, dtype=np.uint64)
        lib().zro_get_pages(self.h, _ptr(pages))
        return dict(sorter=sorter, order=order, base_offset=base, group_first=first,
                    group_last=last, block_offset=blk, pages=pages)

    def update_stress(self):
        lib().zro_update_stress(self.h)

    def p2g(self):
        return int(lib().zro_p2g_mls(self.h))

    def grid_update(self, dt):
        lib().zro_grid_update(self.h, float(dt))

    def g2p(self, dt):
        return int(lib().zro_g2p_mls(self.h, float(dt)))

    def step(self, dt):
        return int(lib().zro_step_explicit(self.h, float(dt)))

    # ---- viscoelastic branch (FElasticNonequilibratedBasedMpmForceHelper.h)
    def set_viscoelastic(self, Fneq, mu, lam, viscosity_d, viscosity_v):
        n = self.n
        self._fneq = None if Fneq is None else np.ascontiguousarray(np.asarray(Fneq, dtype=np.float64).reshape(9 * n))
        self._mun = np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,)))
        self._lamn = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (n,)))
        lib().zro_set_viscoelastic(self.h, _ptr(self._fneq), _ptr(self._mun), _ptr(self._lamn), float(viscosity_d),
                                   float(viscosity_v))

    def get_viscoelastic(self):
        out = np.zeros(9 * self.n)
        lib().zro_get_viscoelastic(self.h, _ptr(out))
        return out.reshape(self.n, 9)

    # ---- adaptive time step (MpmSimulationBase.cpp:766-790,1143-1170)
    def eval_max_particle_speed(self):
        """(max |v|, max corner, min corner) as evalMaxParticleSpeed."""
        out = np.zeros(7)
        lib().zro_eval_max_particle_speed(self.h, _ptr(out))
        return out[0], out[1:4].copy(), out[4:7].copy()

    def calculate_dt(self, max_dt):
        return float(lib().zro_calculate_dt(self.h, float(max_dt)))

    # ---- implicit backward Euler
    def p2g_apic(self):
        return int(lib().zro_p2g_apic(self.h))

    def backward_euler(self, dt, opts=None):
        st = ZroImplicitStats()
        o = opts if opts is not None else implicit_opts()
        lib().zro_backward_euler(self.h, float(dt), ctypes.byref(o), ctypes.byref(st))
        return st

    def g2p_apic(self, dt):
        return int(lib().zro_g2p_apic(self.h, float(dt)))

    def step_implicit(self, dt, opts=None):
        st = ZroImplicitStats()
        o = opts if opts is not None else implicit_opts()
        flags = int(lib().zro_step_implicit(self.h, float(dt), ctypes.byref(o), ctypes.byref(st)))
        return flags, st

    def num_node_slots(self):
        return int(lib().zro_num_node_slots(self.h))

    def implicit_multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        b = np.zeros_like(x)
        lib().zro_implicit_multiply(self.h, _ptr(x), _ptr(b))
        return b.reshape(-1, 3)

    def implicit_residual(self, dv):
        dv = np.ascontiguousarray(dv, dtype=np.float64).reshape(-1)
        r = np.zeros_like(dv)
        lib().zro_implicit_residual(self.h, _ptr(dv), _ptr(r))
        return r.reshape(-1, 3)

    def implicit_diagonal(self):
        k = self.num_node_slots()
        d = np.zeros(3 * k)
        lib().zro_implicit_diagonal(self.h, _ptr(d))
        return d.reshape(k, 3)

    def implicit_energy(self, dv):
        dv = np.ascontiguousarray(dv, dtype=np.float64).reshape(-1)
        return float(lib().zro_implicit_energy(self.h, _ptr(dv)))

    def implicit_state(self):
        k = self.num_node_slots()
        dv = np.zeros(3 * k); vn = np.zeros(3 * k)
        col = np.zeros(k, dtype=np.int8)
        lib().zro_get_implicit(self.h, _ptr(dv), _ptr(vn), _ptr(col))
        return dict(dv=dv.reshape(k, 3), vn=vn.reshape(k, 3), collided=col.astype(bool))

    def grid(self):
        npg = lib().zro_num_pages(self.h)
        k = npg * 32
        m = np.zeros(k); v = np.zeros(3 * k); nv = np.zeros(3 * k)
        idx = np.zeros(k, dtype=np.int64)
        lib().zro_get_grid(self.h, _ptr(m), _ptr(v), _ptr(nv), _ptr(idx))
        return dict(m=m, v=v.reshape(k, 3), new_v=nv.reshape(k, 3), idx=idx)
