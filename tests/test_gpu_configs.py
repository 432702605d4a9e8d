"""GPU parity on each benchmark configuration's own seeded scene (the
generators in ziran2019_amd/scenes.py, shrunk to sizes the FP64 oracle runs
in seconds; the generator and seeds are the full workload's, so the shrunk
scene is the first part of the full one):

- implicit boxes (backward Euler, Newton + CG): the first 4 of the 64 boxes,
  both preconditioners; the plain FP64 path takes the oracle's Newton and CG
  iteration counts, the FP32-bulk form stays within 1e-5 of the FP64 oracle
  relative to the displacement;
- elastoplastic ADMM slabs: a von Mises slab beside a Drucker-Prager slab,
  3 steps of 20 ADMM iterations; the FP64 path takes the oracle's MINRES
  iteration counts, the FP32-bulk global step stays within the bounds
  written in the test;
- phase-field fracture: the Voronoi-toughness block struck by the moving
  sticky sphere at 5 m/s, dx = 1/320, dt = 5e-4;
- differentiable rollout: the lognormal-modulus cube, 64 steps, checkpoints
  every 8, gradients against the CPU adjoint oracle and FP64 central
  differences of the GPU forward (<= 1e-4).
"""
import numpy as np
import pytest

import oracle
from ziran2019_amd import scenes
from ziran2019_amd.engine import ZR_FP32_REFINED, MpmEngine, implicit_opts, admm_opts, phase_field_opts
from _helpers import rel_err

pytestmark = pytest.mark.gpu

THREADS = 16


def _oracle(s):
    o = oracle.OracleSim(dx=s["dx"], gravity=s["gravity"], nthreads=THREADS)
    o.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    for c in s["colliders"]:
        o.add_collider(**c)
    return o


def _engine(s):
    e = MpmEngine(dx=s["dx"], gravity=s["gravity"])
    e.upload(s)
    e.set_colliders(s["colliders"])
    return e


IMPLICIT = dict(newton_max_iterations=3, newton_tolerance=1e-5, cg_max_iterations=500, cg_tolerance=1e30,
                cg_relative_tolerance=1e-6)


@pytest.mark.parametrize("preconditioner", [0, 1], ids=["mass", "jacobi"])
def test_implicit_boxes(preconditioner):
    s = scenes.boxes_scene(nboxes=4)
    print("implicit boxes:", len(s["mass"]), "particles, preconditioner", preconditioner, flush=True)
    ora = _oracle(s)
    e64, e32 = _engine(s), _engine(s)
    dt = 2e-3
    for step in range(2):
        _, so = ora.step_implicit(dt, oracle.implicit_opts(preconditioner=preconditioner, **IMPLICIT))
        s64 = e64.advanceOneTimeStepImplicit(dt, implicit_opts(precision=0, preconditioner=preconditioner,
                                                               **IMPLICIT))
        s32 = e32.advanceOneTimeStepImplicit(dt, implicit_opts(precision=1, preconditioner=preconditioner,
                                                               **IMPLICIT))
        print("step", step, "newton", s64.newton_iterations, so.newton_iterations, "cg", s64.cg_iterations,
              so.cg_iterations, "fp32 cg", s32.cg_iterations, flush=True)
        assert s64.newton_iterations == so.newton_iterations
        assert s64.cg_iterations == so.cg_iterations
    X, V, C, F = ora.get_particles()
    o64, o32 = e64.download(), e32.download()
    assert rel_err(o64["X"], X) < 1e-11
    assert rel_err(o64["V"], V) < 1e-8
    assert rel_err(o64["F"], F) < 1e-9
    disp = np.abs(X - s["X"]).max()
    err32 = np.abs(o32["X"] - X).max() / disp
    print("FP32-bulk position error relative to the displacement:", err32, flush=True)
    assert err32 < 1e-5
    assert rel_err(o32["V"], V) < 1e-5
    e64.close()
    e32.close()


def test_admm_slabs():
    s = scenes.slabs_scene(cells=(8, 6, 8), gap=4)
    m = s["materials"]
    print("admm slabs:", len(s["mass"]), "particles", flush=True)
    ora, eng, e32 = _oracle(s), _engine(s), _engine(s)
    ora.set_materials(model=m["model"], plastic=m["plasticity"], pparam=m["plastic_param"],
                      pstate=m["plastic_state"])
    eng.set_materials(**m)
    e32.set_materials(**m)
    for step in range(3):
        _, so = ora.step_admm(1e-3, oracle.admm_opts(admm_iterations=20))
        sg = eng.advanceOneTimeStepAdmm(1e-3, admm_opts(admm_iterations=20))
        s32 = e32.advanceOneTimeStepAdmm(1e-3, admm_opts(admm_iterations=20, precision=ZR_FP32_REFINED))
        print("step", step, "minres", sg.minres_iterations, so.minres_iterations, "fp32", s32.minres_iterations,
              flush=True)
        assert sg.minres_iterations == so.minres_iterations
    X, V, C, F = ora.get_particles()
    out = eng.download()
    assert rel_err(out["X"], X) < 1e-10
    assert rel_err(out["V"], V) < 1e-8
    assert rel_err(out["F"], F) < 1e-9
    pg, po = eng.get_materials()["plastic_state"], ora.get_materials()["pstate"]
    dp = s["slab"] == 1
    assert rel_err(pg, po) < 1e-8
    assert np.abs(po[dp]).max() > 0.0  # the sand slab flowed
    # the FP32-bulk global step (measured on MI355X: 2e-8 of the displacement,
    # V 3e-8, F 6e-10 relative, with the oracle's MINRES counts); every MINRES
    # solve stops at half its start residual, so FP32 rounding that moved a
    # stopping test by one iteration would change that solve's dv by a
    # fraction of its correction
    o32 = e32.download()
    disp = np.abs(X - s["X"]).max()
    ex, ev, ef = (np.abs(o32["X"] - X).max() / disp, rel_err(o32["V"], V), rel_err(o32["F"], F))
    print("FP32-bulk ADMM: X error / displacement", ex, "V", ev, "F", ef, flush=True)
    assert ex < 1e-6 and ev < 1e-6 and ef < 1e-7
    eng.close()
    e32.close()


def test_fracture_sphere():
    s = scenes.fracture_scene(cells=(12, 8, 10), nsites=20, sphere_radius=0.015)
    n = len(s["mass"])
    print("fracture:", n, "particles", flush=True)
    ora, eng = _oracle(s), _engine(s)
    kw = dict(l0=s["l0"], one_over_sigma_c=s["one_over_sigma_c"], residual_phase=s["residual_phase"])
    eng.set_phase_field(phase_field_opts(), **kw)
    ora.set_phase_field(oracle.pf_opts(), **kw)
    dt = 5e-4
    for step in range(8):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
        print("step", step, "cg", eng.phase_field_last_stats().cg_iterations, flush=True)
    X, V, C, F = ora.get_particles()
    out = eng.download()
    assert rel_err(out["X"], X) < 1e-10
    assert rel_err(out["V"], V) < 1e-8
    assert rel_err(out["F"], F) < 1e-9
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    assert rel_err(pg["c"], po["c"]) < 1e-9
    assert rel_err(pg["H"], po["H"]) < 1e-9
    assert np.abs(V[:, 1]).max() > 1.0  # the sphere struck the block
    eng.close()


def test_adjoint_rollout_64_steps():
    from ziran2019_amd.autodiff import Rollout, lame_from_E
    s = scenes.adjoint_scene(ncells=8)
    n = len(s["mass"])
    print("adjoint rollout:", n, "particles, 64 steps, checkpoints every 8", flush=True)
    nu, dt, steps = 0.3, 2e-4, 64
    E, V0 = s["E"].copy(), s["V"].copy()
    target = s["X"].mean(0) + np.array([0.01, 0.02, -0.01])
    m = s["mass"]

    def loss_and_grad(X):
        com = (m[:, None] * X).sum(0) / m.sum()
        return float(np.sum((com - target) ** 2)), 2 * (com - target)[None, :] * (m / m.sum())[:, None]

    prob = Rollout(s, steps, dt, nu=nu, checkpoint_every=8)
    X = prob.forward(E, V0)
    loss, gX = loss_and_grad(X)
    gE, gV = prob.backward(gX)
    # the CPU adjoint oracle over the same 64 steps
    mu, lam = lame_from_E(E, nu)
    states = []
    st = dict(X=s["X"].copy(), V=V0.copy(), C=s["C"].copy(), F=s["F"].copy())
    for k in range(steps):
        states.append(st)
        sim = oracle.OracleSim(dx=s["dx"], gravity=s["gravity"], nthreads=THREADS)
        sim.set_particles(st["X"], st["V"], st["C"], st["F"], m, s["vol"], mu, lam)
        sim.add_collider(**s["colliders"][0])
        sim.set_time(k * dt)
        sim.step(dt)
        Xo, Vo, Co, Fo = sim.get_particles()
        st = dict(X=Xo, V=Vo, C=Co, F=Fo)
    assert rel_err(X, st["X"]) < 1e-10
    aX = np.ascontiguousarray(gX)
    aV, aC, aF = np.zeros((n, 3)), np.zeros((n, 9)), np.zeros((n, 9))
    amu, alam = np.zeros(n), np.zeros(n)
    for k in range(steps - 1, -1, -1):
        st = states[k]
        sim = oracle.OracleSim(dx=s["dx"], gravity=s["gravity"], nthreads=THREADS)
        sim.set_particles(st["X"], st["V"], st["C"], st["F"], m, s["vol"], mu, lam)
        sim.add_collider(**s["colliders"][0])
        sim.set_time(k * dt)
        sim.adjoint_step(dt, aX, aV, aC, aF, amu, alam)
    gE_o = amu / (2 * (1 + nu)) + alam * nu / ((1 + nu) * (1 - 2 * nu))
    print("adjoint vs oracle: E", rel_err(gE, gE_o), "V0", rel_err(gV, aV), flush=True)
    assert rel_err(gE, gE_o) < 1e-8
    assert rel_err(gV, aV) < 1e-8
    rng = np.random.default_rng(64)
    # the moduli move the centre of mass by ~1e-11 m per Pa: a step of
    # 1e-3 E keeps the loss difference far above its rounding noise (the
    # central difference then agrees with the adjoint to ~1e-7; at 1e-6 E
    # the noise alone is ~5e-4 relative, the same on the CPU oracle)
    for grad, which in ((gE, "E"), (gV, "V")):
        d = rng.normal(size=grad.shape)
        h = 1e-3 * np.abs(E).mean() if which == "E" else 1e-4
        vals = []
        for sign in (1, -1):
            Ep = E + sign * h * d if which == "E" else E
            Vp = V0 + sign * h * d if which == "V" else V0
            vals.append(loss_and_grad(prob.forward(Ep, Vp))[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = np.sum(grad * d)
        print(which, "central difference", fd, "adjoint", ad, flush=True)
        assert abs(fd - ad) <= 1e-4 * abs(fd)
    prob.close()
