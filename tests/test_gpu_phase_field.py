"""GPU parity of the CD-MPM phase-field damage solve and the naive Rankine
damage (FractureSimulation.h:209-554) against the FP64 oracle: one solve on
the same grid (same CG iteration count, c / H / pf_Fp to 1e-10), and several
explicit steps with the damage solve inside the step (1e-9)."""
import numpy as np
import pytest

import oracle
from ziran2019_amd import scenes
from ziran2019_amd.engine import ZR_DAMAGE_PHASE_FIELD_RANKINE, nhb_one_over_sigma_c, phase_field_opts
from _helpers import make_pair, rel_err

pytestmark = pytest.mark.gpu


def _scene(seed, ncells=6, stretch=0.12, E=500.0, nu=0.4):
    s = scenes.cube_scene_small(seed=seed, ncells=ncells, E=E)
    n = len(s["mass"])
    rng = np.random.default_rng(seed)
    F = np.tile(np.eye(3), (n, 1, 1)) * (1 + stretch) + 0.03 * rng.normal(size=(n, 3, 3))
    s["F"] = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9))
    mu, lam = scenes.lame(E, nu)
    s["mu"], s["lam"] = np.full(n, mu), np.full(n, lam)
    osc = nhb_one_over_sigma_c(0.05, mu, lam) * np.exp(0.3 * rng.normal(size=n))
    return s, osc


def _pair(s, osc, M, naive=False, sigma_F=0.0, colliders=(), pf_Fp=None, tol=0.0, c=None, delete_threshold=0.0,
          grid_log2=None):
    """tol = 0 leaves CG to its relative tolerance (1e-12 |r0|), so that the
    solve runs many iterations; the reference's absolute 1e-6 stops it after
    one or two at these particle volumes."""
    eng, ora = make_pair(s, colliders=colliders, grid_log2=grid_log2)
    l0 = 0.5 * s["dx"]
    allow = (np.arange(len(s["mass"])) % 7 != 0).astype(np.int32)
    kw = dict(l0=l0, one_over_sigma_c=osc, allow_damage=allow, pf_Fp=pf_Fp, c=c)
    o = dict(parabolic_M=M, naive=naive, sigma_F=sigma_F, cg_tolerance=tol, delete_threshold=delete_threshold)
    eng.set_phase_field(phase_field_opts(**o), **kw)
    ora.set_phase_field(oracle.pf_opts(**o), **kw)
    return eng, ora


@pytest.mark.parametrize("M", [0.0, 1.0, 15.0])
def test_single_solve_parity(M):
    s, osc = _scene(31)
    pf_Fp = 1.0 + 2.0 * np.random.default_rng(31).uniform(size=len(s["mass"]))
    eng, ora = _pair(s, osc, M, pf_Fp=pf_Fp)
    dt = 1e-4
    eng.sortParticlesAndPolluteGrid()
    eng.particlesToGrid()
    st_g = eng.phase_field_solve(dt)
    ora.sort_and_pollute()
    ora.update_stress()
    ora.p2g()
    st_o = ora.phase_field_solve(dt)
    # the elliptic form (M = 0) needs many iterations; with M > 0 the 1/dt
    # lumped term dominates and CG meets 1e-12 |r0| in one or two
    assert st_o.cg_iterations > (5 if M == 0.0 else 0)
    assert st_g.cg_iterations == st_o.cg_iterations
    assert st_g.residual_norm < 1e-10 and st_o.residual_norm < 1e-10
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    assert rel_err(pg["c"], po["c"]) < 1e-10
    assert rel_err(pg["H"], po["H"]) < 1e-12
    assert rel_err(pg["pf_Fp"], po["pf_Fp"]) < 1e-12
    assert pg["c"].min() < 1.0
    eng.close()


@pytest.mark.parametrize("M", [0.0, 15.0])
def test_steps_with_damage_parity(M):
    s, osc = _scene(32, stretch=0.15)
    s["V"][:, 1] -= 0.5
    y0 = float(s["X"][:, 1].min() - 0.5 * s["dx"])
    cols = [dict(shape=0, type=1, origin=(0, y0, 0), normal=(0, 1, 0))]
    eng, ora = _pair(s, osc, M, colliders=cols)
    dt = 1e-4
    for _ in range(5):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
    out = eng.download()
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-9
    assert rel_err(out["V"], V) < 1e-9
    assert rel_err(out["F"], F) < 1e-9
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    assert rel_err(pg["c"], po["c"]) < 1e-9
    assert np.all(pg["c"] <= 1.0) and pg["c"].min() >= 0.0
    assert rel_err(eng.get_materials()["degradation"], ora.get_materials()["g"]) < 1e-9
    eng.close()


def test_naive_rankine_parity():
    s, osc = _scene(33, stretch=0.1)
    eng, ora = _pair(s, osc, 0.0, naive=True, sigma_F=20.0)
    for _ in range(3):
        eng.advanceOneTimeStep(1e-4)
        ora.step(1e-4)
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    assert rel_err(pg["c"], po["c"]) < 1e-10
    assert pg["c"].min() < 1.0
    out = eng.download()
    assert rel_err(out["X"], ora.get_particles()[0]) < 1e-10
    eng.close()


def test_phase_field_then_rankine_steps_parity():
    """ZR_DAMAGE_PHASE_FIELD_RANKINE (the phase-field solve, then the Rankine
    cut of its c) inside explicit steps, against the oracle."""
    s, osc = _scene(34, stretch=0.15)
    eng, ora = _pair(s, osc, 15.0, naive=ZR_DAMAGE_PHASE_FIELD_RANKINE, sigma_F=40.0)
    dt = 1e-4
    for _ in range(4):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
    out = eng.download()
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-9
    assert rel_err(out["V"], V) < 1e-9
    assert rel_err(out["F"], F) < 1e-9
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    assert rel_err(pg["c"], po["c"]) < 1e-9
    assert pg["c"].min() < 1.0
    assert rel_err(eng.get_materials()["degradation"], ora.get_materials()["g"]) < 1e-9
    eng.close()


def test_delete_particles_parity():
    """deleteParticles at the start of every explicit step: particles whose c
    is below the threshold are parked at (4, 4, 4) (a grid that reaches it:
    2^9 - 8 nodes at dx = 1/64), against the oracle over several steps."""
    s, osc = _scene(35, stretch=0.15)
    n = len(s["mass"])
    c0 = np.ones(n)
    c0[::11] = 1e-3
    eng, ora = _pair(s, osc, 15.0, c=c0, delete_threshold=0.01, grid_log2=9)
    dt = 1e-4
    for _ in range(3):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
    out = eng.download()
    X, V, C, F = ora.get_particles()
    parked = np.abs(X - 4.0).max(axis=1) < 1e-6
    assert parked.sum() == len(c0[::11])
    assert rel_err(out["X"], X) < 1e-12
    assert rel_err(out["V"], V) < 1e-9
    assert rel_err(out["F"], F) < 1e-9
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    assert rel_err(pg["c"], po["c"]) < 1e-9
    eng.close()


def test_phase_field_frame_output():
    """The fracture project's frame output (FractureSimulation.h:111-144):
    position, pf = c and sigmaC = 1 / one_over_sigma_c (float-rounded) per
    particle, read back from the .bgeo file."""
    import io
    from ziran2019_amd import partio
    s, osc = _scene(40)
    eng, ora = _pair(s, osc, 0.0)
    eng.advanceOneTimeStep(1e-4)
    b = io.BytesIO()
    eng.write_phase_field_partio(b)
    b.seek(0)
    X, attrs = partio.read_bgeo(b)
    out, pf = eng.download(), eng.get_phase_field()
    osc = np.broadcast_to(np.asarray(osc, dtype=np.float64), (len(X),))
    assert np.array_equal(X, out["X"].astype(np.float32))
    assert np.array_equal(attrs["pf"][:, 0], pf["c"].astype(np.float32))
    want = (1.0 / osc.astype(np.float32).astype(np.float64)).astype(np.float32)
    assert np.array_equal(attrs["sigmaC"][:, 0], want)
    eng.close()
