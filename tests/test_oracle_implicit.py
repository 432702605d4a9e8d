"""Pins the oracle's implicit backward-Euler step (CPU only).

Following the reference's own validation method, runDiffTest
(Lib/Ziran/Sim/DiffTest.h:55-136):
- the residual is the negative gradient of the objective energy (central FD);
- the matrix-free operator is the negative derivative of the residual
  (central FD), i.e. the Hessian-vector product;
plus closed forms:
- the operator is symmetric and positive definite near rest;
- with zero stiffness an implicit step is free fall: V = g dt;
- Newton's residual decreases and the solve meets its tolerance.
"""
import numpy as np
import pytest

import oracle
from ziran2019_amd import scenes


def _setup(seed=3, colliders=(), stiffness=1.0, newton=1):
    s = scenes.cube_scene_small(seed=seed, ncells=4, E=2e4 * stiffness)
    rng = np.random.default_rng(seed)
    n = len(s["mass"])
    F = np.tile(np.eye(3), (n, 1, 1)) + 0.04 * rng.normal(size=(n, 3, 3))
    s["F"] = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9))
    s["C"] = 0.2 * rng.normal(size=(n, 9))
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0), nthreads=2)
    sim.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    for c in colliders:
        sim.add_collider(**c)
    sim.sort_and_pollute()
    sim.p2g_apic()
    st = sim.backward_euler(2e-3, oracle.implicit_opts(newton_max_iterations=newton))
    active = sim.grid()["idx"] >= 0
    return sim, s, active, st


def _rand_field(rng, active):
    x = rng.normal(size=(len(active), 3))
    x[~active] = 0.0
    return x


def test_residual_is_negative_energy_gradient():
    sim, s, active, _ = _setup()
    rng = np.random.default_rng(0)
    dv = 0.01 * _rand_field(rng, active)
    r = sim.implicit_residual(dv)
    e = _rand_field(rng, active)
    errs = []
    for i in range(1, 8):
        h = 2.0 ** -i * 1e-3
        fd = (sim.implicit_energy(dv + h * e) - sim.implicit_energy(dv - h * e)) / (2 * h)
        errs.append(abs(fd + np.sum(r * e)) / abs(np.sum(r * e)))
    assert min(errs) < 1e-7


def test_operator_is_residual_derivative_and_symmetric():
    sim, s, active, _ = _setup(seed=4)
    rng = np.random.default_rng(1)
    dv = 0.01 * _rand_field(rng, active)
    x = _rand_field(rng, active)
    y = _rand_field(rng, active)
    sim.implicit_residual(dv)  # sets F(dv)
    Ax = sim.implicit_multiply(x)
    Ay = sim.implicit_multiply(y)
    assert abs(np.sum(y * Ax) - np.sum(x * Ay)) < 1e-10 * abs(np.sum(y * Ax))
    assert np.sum(x * Ax) > 0
    h = 1e-6
    fd = (sim.implicit_residual(dv + h * x) - sim.implicit_residual(dv - h * x)) / (2 * h)
    assert np.abs(fd + Ax).max() < 1e-6 * np.abs(Ax).max()


def test_free_fall_without_stiffness():
    s = scenes.cube_scene_small(seed=5, ncells=4)
    n = len(s["mass"])
    s["V"] = np.zeros((n, 3))
    s["mu"] = np.zeros(n)
    s["lam"] = np.zeros(n)
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0))
    sim.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    dt = 2e-3
    flags, st = sim.step_implicit(dt)
    X, V, C, F = sim.get_particles()
    assert np.allclose(V, np.array([0, -9.8 * dt, 0])[None, :], atol=1e-14)
    assert st.converged == 1 and st.newton_iterations == 0


def test_newton_converges_with_tight_cg():
    opts = oracle.implicit_opts(newton_max_iterations=6, newton_tolerance=1e-12, cg_max_iterations=400,
                                cg_tolerance=1e30, cg_relative_tolerance=1e-10)
    s = scenes.cube_scene_small(seed=6, ncells=4, E=2e4)
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0), nthreads=2)
    n = len(s["mass"])
    rng = np.random.default_rng(6)
    F = np.tile(np.eye(3), (n, 1, 1)) + 0.05 * rng.normal(size=(n, 3, 3))
    sim.set_particles(s["X"], s["V"], s["C"], np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9)),
                      s["mass"], s["vol"], s["mu"], s["lam"])
    sim.sort_and_pollute()
    sim.p2g_apic()
    st = sim.backward_euler(2e-3, opts)
    norms = [st.residual_norm[i] for i in range(st.num_residuals)]
    assert st.converged == 1
    assert norms[-1] < 1e-12
    # quadratic convergence once close
    assert norms[2] < 1e-3 * norms[1]


def test_sticky_projection_zeroes_collided_rows():
    s = scenes.cube_scene_small(seed=7, ncells=4)
    y0 = s["X"][:, 1].min() + 1.5 * s["dx"]
    sim, s2, active, _ = _setup(seed=7, colliders=[dict(shape=0, type=1, origin=(0, y0, 0), normal=(0, 1, 0))])
    st = sim.implicit_state()
    assert st["collided"].any()
    rng = np.random.default_rng(2)
    dv = st["dv"]
    r = sim.implicit_residual(dv)
    assert np.all(r[st["collided"]] == 0.0)
    # collided sticky nodes keep v + dv = 0 (object velocity)
    g = sim.grid()
    vnew = g["v"] + dv
    assert np.abs(vnew[st["collided"] & active]).max() < 1e-14


def test_diagonal_matches_unit_vector_products():
    """The Jacobi diagonal equals e^T A e for unit vectors (brute force)."""
    sim, s, active, _ = _setup(seed=8)
    rng = np.random.default_rng(5)
    dv = 0.01 * _rand_field(rng, active)
    sim.implicit_residual(dv)
    diag = sim.implicit_diagonal()
    idx = np.nonzero(active)[0]
    for k in rng.choice(idx, size=12, replace=False):
        for d in range(3):
            e = np.zeros((len(active), 3))
            e[k, d] = 1.0
            assert abs(sim.implicit_multiply(e)[k, d] - diag[k, d]) <= 1e-10 * abs(diag[k, d])


def test_jacobi_and_mass_preconditioners_reach_the_same_step():
    kw = dict(newton_max_iterations=4, newton_tolerance=1e-10, cg_max_iterations=500, cg_tolerance=1e30,
              cg_relative_tolerance=1e-10)
    s = scenes.cube_scene_small(seed=9, ncells=4, E=2e4)
    n = len(s["mass"])
    rng = np.random.default_rng(9)
    F = np.tile(np.eye(3), (n, 1, 1)) + 0.05 * rng.normal(size=(n, 3, 3))
    mu = s["mu"] * np.exp(rng.normal(size=n))  # strong stiffness contrast
    its, dvs = [], []
    for pre in (0, 1):
        sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0), nthreads=2)
        sim.set_particles(s["X"], s["V"], s["C"], np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9)),
                          s["mass"], s["vol"], mu, s["lam"])
        sim.sort_and_pollute()
        sim.p2g_apic()
        st = sim.backward_euler(2e-3, oracle.implicit_opts(preconditioner=pre, **kw))
        assert st.converged == 1
        its.append(st.cg_iterations)
        dvs.append(sim.implicit_state()["dv"])
    assert np.abs(dvs[0] - dvs[1]).max() < 1e-9 * np.abs(dvs[0]).max()


def _implicit_sim(seed, ncells=4, E=2e4):
    s = scenes.cube_scene_small(seed=seed, ncells=ncells, E=E)
    n = len(s["mass"])
    rng = np.random.default_rng(seed)
    F = np.tile(np.eye(3), (n, 1, 1)) + 0.05 * rng.normal(size=(n, 3, 3))
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0), nthreads=2)
    sim.set_particles(s["X"], s["V"], s["C"], np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9)),
                      s["mass"], s["vol"], s["mu"], s["lam"])
    sim.sort_and_pollute()
    sim.p2g_apic()
    return sim


@pytest.mark.parametrize("pre", [0, 1], ids=["mass", "jacobi"])
def test_minres_newton_reaches_the_cg_step(pre):
    """Minres (the reference's default linear solver, BackwardEuler.h:55-62)
    and ConjugateGradient drive Newton to the same step."""
    kw = dict(newton_max_iterations=6, newton_tolerance=1e-11, cg_max_iterations=600, cg_tolerance=1e30,
              cg_relative_tolerance=1e-11, preconditioner=pre)
    dvs = []
    for solver in (0, 1):
        sim = _implicit_sim(11)
        st = sim.backward_euler(2e-3, oracle.implicit_opts(solver=solver, **kw))
        assert st.converged == 1
        dvs.append(sim.implicit_state()["dv"])
    assert np.abs(dvs[0] - dvs[1]).max() < 1e-9 * np.abs(dvs[0]).max()


def test_minres_minimises_the_preconditioned_residual():
    """After k iterations from x = 0, Minres's step minimises |r0 - A x| in
    the norm of the mass preconditioner over the Krylov space, so that norm
    never grows with k and is never above CG's after as many iterations."""
    base = _implicit_sim(12)
    base.backward_euler(2e-3, oracle.implicit_opts(newton_max_iterations=0))
    dv0 = base.implicit_state()["dv"]
    r0 = base.implicit_residual(dv0)
    m = base.grid()["m"]
    active = base.grid()["idx"] >= 0

    def res_after(solver, k):
        sim = _implicit_sim(12)
        sim.backward_euler(2e-3, oracle.implicit_opts(newton_max_iterations=1, cg_max_iterations=k,
                                                      cg_tolerance=0.0, cg_relative_tolerance=1e-300,
                                                      solver=solver))
        x = sim.implicit_state()["dv"] - dv0
        rr = (r0 - base.implicit_multiply(x))[active]
        return np.sqrt(np.sum(rr * rr / m[active, None]))
    mr = [res_after(1, k) for k in (1, 2, 4, 8, 16, 32)]
    cg = [res_after(0, k) for k in (1, 2, 4, 8, 16, 32)]
    floor = 1e-12 * mr[0]  # both reach rounding level by 32 iterations
    assert all(b <= a * (1 + 1e-12) + floor for a, b in zip(mr, mr[1:]))
    assert all(a <= b * (1 + 1e-9) + floor for a, b in zip(mr, cg))
    assert all(a < b for a, b in zip(mr[:4], cg[:4]))
    assert mr[-1] < 1e-12 * mr[0]
