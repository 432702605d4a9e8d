"""Colliders in a moving frame x = R s X + b(t) (AnalyticCollisionObject,
CollisionObject.h:81-172, CollisionObject.cpp:158-194,279-296,340-405) in
the oracle, pinned by what the geometry fixes:

- a sticky sphere that spins (omega) and grows (ds/dt) around b imposes the
  rigid-plus-dilation field v = omega x (x - b) + (ds/dt / s)(x - b) on every
  node inside it; quadratic B-splines reproduce a linear field, so particles
  whose stencils lie inside take exactly that velocity after one step;
- a half-space and a sphere written in a rotated, scaled, translated material
  frame (normal R^T n, origin R^T (p - b) / s, radius r / s) are the same
  world objects: the steps match the identity-frame ones to rounding (a
  transposed R, a normal left in material space or a signed distance not
  scaled by s breaks this);
- evalMaxSpeed of a spinning sphere is its speed at a kept box corner.
"""
import numpy as np
import pytest

import oracle
from _helpers import perturbed

DT = 1e-3


def _quat(axis, angle):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    return (np.cos(angle / 2), *(np.sin(angle / 2) * a))


def _rot(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


CFL = 0.6


def _sim(s, colliders, gravity=(0.0, -9.8, 0.0)):
    o = oracle.OracleSim(dx=s["dx"], gravity=gravity, cfl=CFL)
    o.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    for c in colliders:
        o.add_collider(**c)
    return o


def test_spinning_growing_sticky_sphere_imposes_its_velocity_field():
    s = perturbed(31, n_cells=4, F_noise=0.0, C_noise=0.0)
    X0 = s["X"].copy()
    c = X0.mean(axis=0)
    radius = np.sqrt(3) * np.abs(X0 - c).max() + 4 * s["dx"]
    q = _quat((1.0, 2.0, -0.5), 0.7)
    R, sc = _rot(q), 1.3
    b = c + np.array([0.01, -0.02, 0.005])
    omega, dsdt = np.array([1.5, -2.0, 0.7]), 0.4
    sphere = dict(shape=1, type=1, origin=R.T @ (c - b) / sc, radius=radius / sc, translation=b, rotation=q,
                  omega=omega, scale=sc, dscale_dt=dsdt)
    o = _sim(s, [sphere])
    o.step(DT)
    X, V, C, F = o.get_particles()
    want = np.cross(omega, X0 - b) + (dsdt / sc) * (X0 - b)
    assert np.abs(V - want).max() < 1e-12 * np.abs(want).max()
    # the affine velocity's gradient is the field's: [omega]x + (ds/dt / s) I
    W = np.array([[0.0, -omega[2], omega[1]], [omega[2], 0.0, -omega[0]], [-omega[1], omega[0], 0.0]])
    G = W + (dsdt / sc) * np.eye(3)
    assert np.abs(C.reshape(-1, 3, 3, order="F") - G).max() < 1e-9 * np.abs(G).max()


def _frames(s):
    """A slip floor with friction and a separating sphere, in the identity
    frame and in a rotated, scaled, translated one."""
    dx = s["dx"]
    p0 = np.array([0.0, s["X"][:, 1].min() + 0.3 * dx, 0.0])
    n0 = np.array([0.2, 1.0, -0.1])
    n0 = n0 / np.linalg.norm(n0)
    cw = s["X"].mean(axis=0) + np.array([0.6, 0.2, 0.1]) * np.abs(s["X"] - s["X"].mean(axis=0)).max()
    rw = 0.5 * np.abs(s["X"] - s["X"].mean(axis=0)).max()
    plain = [dict(shape=0, type=2, origin=p0, normal=n0, friction=0.3),
             dict(shape=1, type=3, origin=cw, radius=rw, friction=0.2)]
    q = _quat((-0.3, 1.0, 0.8), 1.1)
    R = _rot(q)
    b, sc = np.array([0.03, -0.01, 0.02]), 0.7
    framed = [dict(shape=0, type=2, origin=R.T @ (p0 - b) / sc, normal=R.T @ n0, friction=0.3, translation=b,
                   rotation=q, scale=sc),
              dict(shape=1, type=3, origin=R.T @ (cw - b) / sc, radius=rw / sc, friction=0.2, translation=b,
                   rotation=q, scale=sc)]
    return plain, framed


def test_rotated_scaled_frame_gives_the_same_world_objects():
    s = perturbed(32, n_cells=5, shuffle=False, F_noise=0.02, C_noise=0.2)
    s["V"][:, 1] -= 1.0
    s["V"][:, 0] += 0.5
    plain, framed = _frames(s)
    a, b = _sim(s, plain), _sim(s, framed)
    for _ in range(6):
        a.step(DT)
        b.step(DT)
    Xa, Va, _, Fa = a.get_particles()
    Xb, Vb, _, Fb = b.get_particles()
    assert not np.allclose(Va, s["V"])  # the objects acted
    assert np.abs(Xa - Xb).max() < 1e-12
    assert np.abs(Va - Vb).max() < 1e-9 * np.abs(Va).max()
    assert np.abs(Fa - Fb).max() < 1e-9


def test_eval_max_speed_of_a_spinning_sphere():
    s = perturbed(33, n_cells=4)
    s["V"][:] = 0.0
    c = s["X"].mean(axis=0)
    omega = np.array([0.0, 0.0, 3.0])
    sphere = dict(shape=1, type=1, origin=(0.0, 0.0, 0.0), radius=0.01, translation=c, omega=omega)
    o = _sim(s, [sphere], gravity=(0.0, 0.0, 0.0))
    dt = o.calculate_dt(1.0)
    # the corners kept are the particle box's (widened by 4 dx) and the
    # sphere box's; the fastest is a particle-box corner
    sp, hi, lo = o.eval_max_particle_speed()
    hi, lo = hi + 4 * s["dx"], lo - 4 * s["dx"]
    corners = np.array([[lo[d] if (i >> d) & 1 else hi[d] for d in range(3)] for i in range(8)])
    speed = np.linalg.norm(np.cross(omega, corners - c), axis=1).max()
    assert dt == pytest.approx(CFL * s["dx"] / speed, rel=1e-14)


def _skew(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def test_collider_frame_follows_constant_rates():
    """updateState(t) for constant rates: b(t) = b0 + v t, s(t) = s0 + ds/dt t
    and R(t) = exp(t [omega]x) R0 (the rotation the reference's scene
    callbacks write as (cos(theta t/2), 0, sin(theta t/2), 0) for omega =
    (0, theta, 0), AdmmInit3D.h:298-313)."""
    from scipy.linalg import expm
    q0 = _quat((0.3, -1.0, 0.4), 0.9)
    omega = np.array([0.7, -1.3, 2.1])
    kw = dict(shape=1, type=1, origin=(0.0, 0.0, 0.0), radius=0.1, translation=(0.1, 0.2, 0.3),
              velocity=(1.0, -2.0, 0.5), rotation=q0, omega=omega, scale=1.4, dscale_dt=-0.2)
    for t in (0.0, 1e-3, 0.37, 2.5):
        b, R, s = oracle.collider_frame(t, **kw)
        assert np.allclose(b, np.array([0.1, 0.2, 0.3]) + t * np.array([1.0, -2.0, 0.5]), rtol=0, atol=1e-15)
        assert s == pytest.approx(1.4 - 0.2 * t, rel=1e-15)
        want = expm(t * _skew(omega)) @ _rot(q0)
        assert np.abs(R - want).max() < 1e-14
    # the reference scene's form, and no rotation without omega
    theta = 24 * 5 / 180 * np.pi
    for t in (0.1, 1.3):
        _, R, _ = oracle.collider_frame(t, shape=1, type=1, origin=(0, 0, 0), omega=(0.0, theta, 0.0))
        assert np.abs(R - _rot((np.cos(theta * t / 2), 0.0, np.sin(theta * t / 2), 0.0))).max() < 1e-15
    _, R, s = oracle.collider_frame(3.0, shape=1, type=1, origin=(0, 0, 0), rotation=q0)
    assert np.abs(R - _rot(q0)).max() < 1e-16
    assert s == 1.0


def test_growing_spinning_sphere_over_several_steps():
    """A sticky sphere that translates, spins and grows: in step k (from t_k
    to t_k + dt) every node inside takes the object's velocity at t_k,
    omega x (x - b(t_k)) + (ds/dt / s(t_k))(x - b(t_k)) + db/dt, and the
    B-spline G2P reproduces that affine field at the particles' positions at
    t_k.  With s held at its time-0 value the dilation rate would be
    ds/dt / s0, 10% off after the steps below."""
    s = perturbed(34, n_cells=4, F_noise=0.0, C_noise=0.0)
    X0 = s["X"].copy()
    c = X0.mean(axis=0)
    radius = np.sqrt(3) * np.abs(X0 - c).max() + 6 * s["dx"]
    omega, dsdt, vb, s0 = np.array([2.0, -1.0, 3.0]), 40.0, np.array([0.05, -0.02, 0.01]), 1.0
    sphere = dict(shape=1, type=1, origin=(0.0, 0.0, 0.0), radius=radius, translation=c, velocity=vb,
                  omega=omega, scale=s0, dscale_dt=dsdt)
    o = _sim(s, [sphere], gravity=(0.0, 0.0, 0.0))
    K = 5
    for _ in range(K - 1):
        o.step(DT)
    Xk = o.get_particles()[0].copy()
    o.step(DT)
    V = o.get_particles()[1]
    tk = (K - 1) * DT
    bk, sk = c + vb * tk, s0 + dsdt * tk
    want = np.cross(omega, Xk - bk) + (dsdt / sk) * (Xk - bk) + vb
    assert np.abs(V - want).max() < 1e-11 * np.abs(want).max()
    stale = np.cross(omega, Xk - bk) + (dsdt / s0) * (Xk - bk) + vb
    assert np.abs(stale - want).max() > 1e-3 * np.abs(want).max()


def _sep(**kw):
    d = dict(shape=1, type=3, origin=(0.0, 0.0, 0.0), radius=1.0)
    d.update(kw)
    return d


def test_fake_separate_single_node_closed_forms():
    """multiObjectCollisionWithFakeSeparate (CollisionObject.cpp:108-160) on
    one node inside a separating unit sphere, outward normal n = x/|x|:
    approaching (v.n < 0) projects the normal part away and keeps the node in
    the projected set (P = I - n n^T); separating (v.n > 0) leaves v and the
    node alone; v.n = 0 stays in contact, also when the zero carries
    rounding noise; friction shortens the tangential part by mu |v.n|;
    everything is relative to the object's own velocity."""
    x = np.array([0.3, -0.4, 0.5])
    n = x / np.linalg.norm(x)
    t1 = np.cross(n, [1.0, 0.0, 0.0]); t1 /= np.linalg.norm(t1)
    vt = 0.7 * t1
    I = np.eye(3)
    hit, v, P = oracle.fake_separate([_sep()], x, vt - 2.0 * n)
    assert hit and np.abs(v - vt).max() < 1e-15 and np.abs(P - (I - np.outer(n, n))).max() < 1e-15
    hit, v, P = oracle.fake_separate([_sep()], x, vt + 2.0 * n)
    assert not hit and np.array_equal(v, vt + 2.0 * n) and np.array_equal(P, I)
    for noise in (0.0, 3e-17, -3e-17):
        hit, v, P = oracle.fake_separate([_sep()], x, vt + noise * n)
        assert hit and np.abs(P - (I - np.outer(n, n))).max() < 1e-15
    # Coulomb friction: |v_t| -> |v_t| - mu |v.n|, or zero when that is negative
    hit, v, _ = oracle.fake_separate([_sep(friction=0.2)], x, vt - 2.0 * n)
    assert hit and np.abs(v - vt * (1 - 0.2 * 2.0 / 0.7)).max() < 1e-15
    hit, v, _ = oracle.fake_separate([_sep(friction=0.5)], x, vt - 2.0 * n)
    assert hit and np.abs(v).max() < 1e-15
    # a moving object: the test and the projection act on v - v_obj
    vo = np.array([0.4, 1.0, -0.3])
    hit, v, _ = oracle.fake_separate([_sep(velocity=vo)], x, vo + vt - 2.0 * n)
    assert hit and np.abs(v - (vo + vt)).max() < 1e-15
    hit, v, _ = oracle.fake_separate([_sep(velocity=vo)], x, vo + vt + 2.0 * n)
    assert not hit
    # a sticky object wins outright; two contacts give an orthonormal basis
    hit, v, P = oracle.fake_separate([_sep(), dict(shape=1, type=1, origin=(0, 0, 0), radius=1.0)], x,
                                     vt - 2.0 * n)
    assert hit and np.array_equal(P, np.zeros((3, 3))) and np.abs(v).max() == 0.0
    m = np.array([0.0, 1.0, 0.0])
    floor = dict(shape=0, type=3, origin=(0.0, 0.0, 0.0), normal=tuple(m))
    hit, v, P = oracle.fake_separate([_sep(), floor], x, vt - 2.0 * n - 1.0 * m)
    k2 = m - (m @ n) * n
    k2 /= np.linalg.norm(k2)
    assert hit and np.abs(P - (I - np.outer(n, n) - np.outer(k2, k2))).max() < 1e-14


def test_admm_step_with_a_separating_sphere_is_well_conditioned():
    """The ADMM step re-tests the separation of nodes the previous iteration
    projected, whose relative normal velocity is zero in exact arithmetic.
    With that test taken against a rounding bound, a 1e-15 change of the
    input velocities changes the step's result at the 1e-14 level (taken
    against zero, as the reference writes it, by 1-13%), and the separating
    sphere still acts."""
    def run(scale, separating=True):
        s = perturbed(45, n_cells=5, shuffle=False, F_noise=0.02, C_noise=0.1, E=1e5)
        s["V"] = s["V"] * scale
        lo, hi, c = s["X"].min(axis=0), s["X"].max(axis=0), s["X"].mean(axis=0)
        r = 0.3 * (hi[1] - lo[1])
        cols = [dict(shape=0, type=1, origin=(0.0, lo[1] - 0.5 * s["dx"], 0.0), normal=(0.0, 1.0, 0.0)),
                dict(shape=1, type=1, origin=(0.0, 0.0, 0.0), radius=r, translation=(lo[0], c[1], c[2]))]
        if separating:
            cols.append(_frames(s)[0][1])
        o = _sim(s, cols)
        o.set_materials(model=1, plastic=0, pparam=(0, 0, 0, 0), pstate=0.0)
        counts = [o.step_admm(1e-3, oracle.admm_opts(admm_iterations=20))[1].minres_iterations for _ in range(2)]
        return counts, o.get_particles()[1]
    ca, Va = run(1.0)
    cb, Vb = run(1.0 + 1e-15)
    _, Vn = run(1.0, separating=False)
    assert ca == cb
    assert np.abs(Va - Vb).max() < 1e-12 * np.abs(Va).max()
    assert np.abs(Va - Vn).max() > 1e-6 * np.abs(Va).max()
