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
