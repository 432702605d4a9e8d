"""Pins the oracle's adjoint of the explicit MLS step (no counterpart in the
reference; the differentiable form of the step) against FP64 central finite
differences of the oracle's own forward step:

- the fixed-corotated Kirchhoff stress adjoint (F, mu, lambda directions);
- one step: <adjoint, direction> = d/dh L(step(state + h direction)) for a
  random linear functional L of the output, in x, v, C, F, mu and lambda;
- a multi-step rollout onto a sticky ground with the loss
  |x_com(T) - x*|^2: gradients in per-particle E and the initial velocity.
"""
import numpy as np
import pytest

import oracle
from ziran2019_amd import scenes

DT = 2e-4


def _rot_dir(rng, shape):
    d = rng.normal(size=shape)
    return d / np.linalg.norm(d)


def test_stress_adjoint_matches_fd():
    rng = np.random.default_rng(0)
    mu, lam = oracle.lame(1e4, 0.3)
    for _ in range(5):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        tb = rng.normal(size=(3, 3))
        Fb, mb, lb = oracle.fcr_kirchhoff_adjoint(mu, lam, F, tb)
        E = rng.normal(size=(3, 3))
        h = 1e-6
        fd = np.sum(tb * (oracle.fcr_kirchhoff(mu, lam, F + h * E) - oracle.fcr_kirchhoff(mu, lam, F - h * E))) / (2 * h)
        assert abs(fd - np.sum(Fb * E)) < 1e-7 * max(abs(fd), np.abs(Fb).max())
        fdm = np.sum(tb * (oracle.fcr_kirchhoff(mu + 1e-3, lam, F) - oracle.fcr_kirchhoff(mu - 1e-3, lam, F))) / 2e-3
        fdl = np.sum(tb * (oracle.fcr_kirchhoff(mu, lam + 1e-3, F) - oracle.fcr_kirchhoff(mu, lam - 1e-3, F))) / 2e-3
        assert abs(fdm - mb) < 1e-8 * max(abs(mb), 1.0)
        assert abs(fdl - lb) < 1e-8 * max(abs(lb), 1.0)


def _state(s):
    return dict(X=s["X"].copy(), V=s["V"].copy(), C=s["C"].copy(), F=s["F"].copy())


def _sim(s, st, mu, lam, colliders, time=0.0):
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0.0, -9.8, 0.0), nthreads=4)
    sim.set_particles(st["X"], st["V"], st["C"], st["F"], s["mass"], s["vol"], mu, lam)
    for c in colliders:
        sim.add_collider(**c)
    sim.set_time(time)
    return sim


def _step(s, st, mu, lam, colliders, time=0.0):
    sim = _sim(s, st, mu, lam, colliders, time)
    sim.step(DT)
    X, V, C, F = sim.get_particles()
    return dict(X=X, V=V, C=C, F=F)


def _scene(seed=3, ncells=4):
    s = scenes.cube_scene_small(seed=seed, ncells=ncells, E=1e4)
    n = len(s["mass"])
    rng = np.random.default_rng(seed)
    F = np.tile(np.eye(3), (n, 1, 1)) + 0.03 * rng.normal(size=(n, 3, 3))
    s["F"] = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9))
    s["C"] = 0.5 * rng.normal(size=(n, 9))
    s["V"][:, 1] -= 1.0
    y0 = float(s["X"][:, 1].min() + 0.3 * s["dx"])
    cols = [dict(shape=0, type=1, origin=(0.0, y0, 0.0), normal=(0.0, 1.0, 0.0))]
    mu = s["mu"] * np.exp(0.3 * rng.normal(size=n))
    lam = s["lam"] * np.exp(0.3 * rng.normal(size=n))
    return s, cols, mu, lam


@pytest.mark.parametrize("key", ["X", "V", "C", "F", "mu", "lam"])
def test_one_step_adjoint_matches_fd(key):
    s, cols, mu, lam = _scene()
    n = len(s["mass"])
    st = _state(s)
    rng = np.random.default_rng(1)
    w = {k: rng.normal(size=st[k].shape) for k in st}

    def L(out):
        return sum(np.sum(w[k] * out[k]) for k in out)

    aX, aV, aC, aF = (np.ascontiguousarray(w[k].copy()) for k in ("X", "V", "C", "F"))
    amu, alam = np.zeros(n), np.zeros(n)
    sim = _sim(s, st, mu, lam, cols)
    sim.adjoint_step(DT, aX, aV, aC, aF, amu, alam)
    grads = dict(X=aX, V=aV, C=aC, F=aF, mu=amu, lam=alam)
    base = dict(st, mu=mu, lam=lam)
    d = _rot_dir(rng, base[key].shape) * (np.abs(base[key]).max() + 1.0)
    h = 1e-7 * s["dx"] / np.abs(d).max() if key == "X" else 1e-6
    vals = []
    for sign in (1, -1):
        p = {k: v.copy() for k, v in base.items()}
        p[key] = p[key] + sign * h * d
        vals.append(L(_step(s, p, p["mu"], p["lam"], cols)))
    fd = (vals[0] - vals[1]) / (2 * h)
    ad = np.sum(grads[key] * d)
    assert abs(fd - ad) <= 2e-6 * max(abs(ad), abs(fd)), (fd, ad)


def _rollout_loss(s, cols, E, V0, nu, steps, target):
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    st = dict(X=s["X"].copy(), V=V0.copy(), C=s["C"].copy(), F=s["F"].copy())
    states = [st]
    for k in range(steps):
        st = _step(s, st, mu, lam, cols, time=k * DT)
        states.append(st)
    m = s["mass"]
    com = (m[:, None] * st["X"]).sum(0) / m.sum()
    return float(np.sum((com - target) ** 2)), states, mu, lam, com


def test_rollout_gradients_match_fd():
    s, cols, _, _ = _scene(seed=5)
    n = len(s["mass"])
    nu = 0.3
    rng = np.random.default_rng(5)
    E = 1e4 * np.exp(0.4 * rng.normal(size=n))
    V0 = s["V"].copy()
    steps = 12
    target = s["X"].mean(0) + np.array([0.01, 0.02, -0.01])
    loss, states, mu, lam, com = _rollout_loss(s, cols, E, V0, nu, steps, target)
    m = s["mass"]
    aX = np.ascontiguousarray(2 * (com - target)[None, :] * (m / m.sum())[:, None])
    aV, aC, aF = np.zeros((n, 3)), np.zeros((n, 9)), np.zeros((n, 9))
    amu, alam = np.zeros(n), np.zeros(n)
    for k in range(steps - 1, -1, -1):
        sim = _sim(s, states[k], mu, lam, cols, time=k * DT)
        sim.adjoint_step(DT, aX, aV, aC, aF, amu, alam)
    gE = amu / (2 * (1 + nu)) + alam * nu / ((1 + nu) * (1 - 2 * nu))
    gV = aV
    for grad, which in ((gE, "E"), (gV, "V")):
        d = rng.normal(size=grad.shape)
        h = 1e-3 * (E.mean() if which == "E" else 1.0) * 1e-3
        vals = []
        for sign in (1, -1):
            Ep = E + sign * h * d if which == "E" else E
            Vp = V0 + sign * h * d if which == "V" else V0
            vals.append(_rollout_loss(s, cols, Ep, Vp, nu, steps, target)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = np.sum(grad * d)
        assert abs(fd - ad) <= 1e-4 * abs(fd), (which, fd, ad)
    assert np.abs(gE).max() > 0


def _moving_colliders(s):
    """A frictional slip ground that rotates and translates, and a slipping
    sphere that spins and grows, both cutting through the cube's nodes."""
    X = s["X"]
    lo, hi, c = X.min(axis=0), X.max(axis=0), X.mean(axis=0)
    th = 0.2
    ground = dict(shape=0, type=2, friction=0.4, origin=(0.0, 0.0, 0.0), normal=(0.0, 1.0, 0.0),
                  translation=(0.0, float(lo[1] + 0.6 * s["dx"]), 0.0), velocity=(0.05, 0.1, 0.0),
                  rotation=(np.cos(th / 2), 0.0, 0.0, np.sin(th / 2)), omega=(0.0, 0.0, 0.5))
    sphere = dict(shape=1, type=2, origin=(0.0, 0.0, 0.0), radius=float(0.35 * (hi[0] - lo[0])),
                  translation=(float(hi[0]), float(c[1]), float(c[2])), velocity=(-0.2, 0.0, 0.0),
                  rotation=(1.0, 0.0, 0.0, 0.0), omega=(3.0, -2.0, 1.0), scale=1.0, dscale_dt=0.5)
    return [ground, sphere]


@pytest.mark.parametrize("key", ["X", "V", "F", "mu"])
def test_one_step_adjoint_with_moving_frictional_colliders(key):
    """The collider Jacobian for objects in moving frames (rotation, spin,
    scaling, translation) and Coulomb friction, against central differences
    of the forward step; the step starts at t = 0.3 so that every frame has
    moved."""
    s, _, mu, lam = _scene(seed=7)
    cols = _moving_colliders(s)
    n = len(s["mass"])
    st = _state(s)
    rng = np.random.default_rng(2)
    w = {k: rng.normal(size=st[k].shape) for k in st}
    t0 = 0.3

    def L(out):
        return sum(np.sum(w[k] * out[k]) for k in out)

    aX, aV, aC, aF = (np.ascontiguousarray(w[k].copy()) for k in ("X", "V", "C", "F"))
    amu, alam = np.zeros(n), np.zeros(n)
    sim = _sim(s, st, mu, lam, cols, time=t0)
    sim.adjoint_step(DT, aX, aV, aC, aF, amu, alam)
    grads = dict(X=aX, V=aV, C=aC, F=aF, mu=amu, lam=alam)
    base = dict(st, mu=mu, lam=lam)
    d = _rot_dir(rng, base[key].shape) * (np.abs(base[key]).max() + 1.0)
    h = 1e-7 * s["dx"] / np.abs(d).max() if key == "X" else 1e-6
    vals = []
    for sign in (1, -1):
        p = {k: v.copy() for k, v in base.items()}
        p[key] = p[key] + sign * h * d
        vals.append(L(_step(s, p, p["mu"], p["lam"], cols, time=t0)))
    fd = (vals[0] - vals[1]) / (2 * h)
    ad = np.sum(grads[key] * d)
    assert abs(fd - ad) <= 2e-6 * max(abs(ad), abs(fd)), (fd, ad)


def test_moving_colliders_touch_the_nodes():
    """The scene of the test above exercises the branches it claims: nodes
    inside the rotating ground (sliding and sticking under friction) and
    inside the spinning sphere."""
    s, _, mu, lam = _scene(seed=7)
    cols = _moving_colliders(s)
    sim = _sim(s, _state(s), mu, lam, cols, time=0.3)
    sim.sort_and_pollute()
    sim.update_stress()
    sim.p2g()
    g = sim.grid()
    idx = np.nonzero(g["idx"] >= 0)[0]
    info = sim.sort_result()
    pos = []
    for p, base in enumerate(info["pages"]):
        for e in range(32):
            pos.append(oracle.linear_to_coord(int(base) + (e << 7)))
    pos = np.array(pos, dtype=float) * s["dx"]
    inside = [0, 0]
    for k, c in enumerate(cols):
        b, R, sc = oracle.collider_frame(0.3, **c)
        Xm = (pos[idx] - b) @ R / sc
        if c["shape"] == 0:
            inside[k] = int(np.sum(Xm @ np.array(c["normal"]) <= 0.0))
        else:
            inside[k] = int(np.sum(np.linalg.norm(Xm - np.array(c["origin"]), axis=1) < c["radius"]))
    assert inside[0] > 10 and inside[1] > 10, inside


def test_friction_enters_the_gradient():
    """Coulomb friction on the rotating ground changes the step's adjoint (its
    sliding branch is exercised in the scene of the tests above)."""
    s, _, mu, lam = _scene(seed=7)
    n = len(s["mass"])
    out = []
    for friction in (0.4, 0.0):
        cols = _moving_colliders(s)
        cols[0]["friction"] = friction
        rng = np.random.default_rng(2)
        aX, aV, aC, aF = (np.ascontiguousarray(rng.normal(size=sh)) for sh in ((n, 3), (n, 3), (n, 9), (n, 9)))
        amu, alam = np.zeros(n), np.zeros(n)
        _sim(s, _state(s), mu, lam, cols, time=0.3).adjoint_step(DT, aX, aV, aC, aF, amu, alam)
        out.append(aV)
    assert np.abs(out[0] - out[1]).max() > 1e-3 * np.abs(out[1]).max()
