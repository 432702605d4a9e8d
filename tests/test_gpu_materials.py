"""GPU parity of the StvkWithHencky / NeoHookeanBorden models and the von
Mises / Drucker-Prager / Cam-Clay return maps (explicit MLS steps) against the FP64
oracle on the same seeded particles.  Tolerances: 1e-11 relative after one
step (summation order only), 1e-9 after several."""
import numpy as np
import pytest

from ziran2019_amd.engine import (ZR_MODEL_NEO_HOOKEAN_BORDEN, ZR_MODEL_STVK_HENCKY, ZR_PLASTIC_CAM_CLAY,
                                  ZR_PLASTIC_DRUCKER_PRAGER, ZR_PLASTIC_VON_MISES, ZrError, cam_clay_M,
                                  drucker_prager_alpha, implicit_opts)
from _helpers import make_pair, perturbed, rel_err

pytestmark = pytest.mark.gpu


def _ground(s, depth=-0.5):
    return [dict(shape=0, type=1, origin=(0.0, float(s["X"][:, 1].min() + depth * s["dx"]), 0.0),
                 normal=(0.0, 1.0, 0.0))]


def _run(s, mats, steps, dt=1e-4):
    eng, ora = make_pair(s, colliders=_ground(s))
    eng.set_materials(**mats)
    ora.set_materials(model=mats.get("model"), plastic=mats.get("plasticity"), pparam=mats.get("plastic_param"),
                      pstate=mats.get("plastic_state"), g=mats.get("degradation"))
    for _ in range(steps):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
    out = eng.download()
    X, V, C, F = ora.get_particles()
    gm, om = eng.get_materials(), ora.get_materials()
    eng.close()
    return out, (X, V, C, F), gm, om


def _check(out, ref, tol):
    X, V, C, F = ref
    assert rel_err(out["X"], X) < tol
    assert rel_err(out["V"], V) < tol
    assert rel_err(out["C"], C) < 10 * tol
    assert rel_err(out["F"], F) < tol


@pytest.mark.parametrize("steps", [1, 10])
def test_stvk_hencky_elastic(steps):
    s = perturbed(21, shuffle=False, E=1e5)
    s["V"][:, 1] -= 0.5
    out, ref, _, _ = _run(s, dict(model=ZR_MODEL_STVK_HENCKY), steps)
    _check(out, ref, 1e-11 if steps == 1 else 1e-9)


@pytest.mark.parametrize("xi", [0.0, 2e4])
def test_von_mises_return_mapping(xi):
    s = perturbed(22, shuffle=False, E=1e5)
    s["V"][:, 1] -= 0.5
    mats = dict(model=ZR_MODEL_STVK_HENCKY, plasticity=ZR_PLASTIC_VON_MISES, plastic_param=(xi, 1e30, 0.0, 0.0),
                plastic_state=200.0)
    out, ref, gm, om = _run(s, mats, 8)
    _check(out, ref, 1e-9)
    assert rel_err(gm["plastic_state"], om["pstate"]) < 1e-9
    if xi > 0:
        assert gm["plastic_state"].min() < 200.0  # hardening happened


def test_von_mises_fail_stress_flags():
    s = perturbed(23, shuffle=False, E=1e5)
    mats = dict(model=ZR_MODEL_STVK_HENCKY, plasticity=ZR_PLASTIC_VON_MISES, plastic_param=(0.0, 50.0, 0.0, 0.0),
                plastic_state=200.0)
    out, ref, gm, om = _run(s, mats, 1)
    _check(out, ref, 1e-11)
    assert np.array_equal(gm["broken"], om["broken"])
    assert gm["broken"].any()


@pytest.mark.parametrize("cohesion,beta", [(0.0, 1.0), (0.01, 0.5)])
def test_drucker_prager_sand(cohesion, beta):
    s = perturbed(24, shuffle=False, E=1e5, F_noise=0.03)
    s["V"][:, 1] -= 1.0
    a = drucker_prager_alpha(30.0)
    mats = dict(model=ZR_MODEL_STVK_HENCKY, plasticity=ZR_PLASTIC_DRUCKER_PRAGER,
                plastic_param=(a, beta, cohesion, 1.0), plastic_state=0.0)
    out, ref, gm, om = _run(s, mats, 8)
    _check(out, ref, 1e-9)
    scale = max(np.abs(om["pstate"]).max(), 1e-12)
    assert np.abs(gm["plastic_state"] - om["pstate"]).max() < 1e-9 * scale


@pytest.mark.parametrize("logJp,beta", [(np.log(0.98), 0.5), (0.0, 2.0)])
def test_cam_clay_snow(logJp, beta):
    # NeoHookeanBorden + NonAssociativeCamClay as in the reference's snow
    # demos (friction angle 45, xi 3): compressed against the ground
    s = perturbed(27, shuffle=False, E=1e4, F_noise=0.05)
    s["V"][:, 1] -= 1.0
    mats = dict(model=ZR_MODEL_NEO_HOOKEAN_BORDEN, plasticity=ZR_PLASTIC_CAM_CLAY,
                plastic_param=(cam_clay_M(45.0), beta, 3.0, 0.0), plastic_state=logJp)
    out, ref, gm, om = _run(s, mats, 8)
    _check(out, ref, 1e-9)
    scale = max(np.abs(om["pstate"]).max(), 1e-12)
    assert np.abs(gm["plastic_state"] - om["pstate"]).max() < 1e-9 * scale
    assert np.any(gm["plastic_state"] != logJp)  # the return map moved states


def test_cam_clay_needs_neo_hookean_borden():
    s = perturbed(28, shuffle=False)
    eng, _ = make_pair(s)
    with pytest.raises(ZrError):
        eng.set_materials(model=ZR_MODEL_STVK_HENCKY, plasticity=ZR_PLASTIC_CAM_CLAY,
                          plastic_param=(cam_clay_M(45.0), 0.5, 3.0, 0.0), plastic_state=0.0)
    eng.close()


def test_neo_hookean_borden_with_degradation():
    s = perturbed(25, shuffle=False, E=500.0, F_noise=0.1)
    n = len(s["mass"])
    g = np.random.default_rng(25).uniform(1e-3, 1.0, size=n)
    out, ref, gm, _ = _run(s, dict(model=ZR_MODEL_NEO_HOOKEAN_BORDEN, degradation=g), 5)
    _check(out, ref, 1e-9)
    assert np.array_equal(gm["degradation"], g)


def test_mixed_materials_in_one_scene():
    s = perturbed(26, shuffle=True, E=1e5, F_noise=0.03)
    n = len(s["mass"])
    rng = np.random.default_rng(26)
    model = rng.integers(0, 3, size=n).astype(np.int32)
    plastic = np.where(model == ZR_MODEL_STVK_HENCKY, rng.integers(0, 3, size=n), 0).astype(np.int32)
    pp = np.zeros((n, 4))
    pp[plastic == ZR_PLASTIC_VON_MISES] = (1e3, 1e30, 0.0, 0.0)
    pp[plastic == ZR_PLASTIC_DRUCKER_PRAGER] = (drucker_prager_alpha(30.0), 1.0, 0.0, 1.0)
    st = np.where(plastic == ZR_PLASTIC_VON_MISES, 300.0, 0.0)
    mats = dict(model=model, plasticity=plastic, plastic_param=pp, plastic_state=st,
                degradation=rng.uniform(0.2, 1.0, size=n))
    out, ref, gm, om = _run(s, mats, 5)
    _check(out, ref, 1e-9)


def test_implicit_rejects_other_models():
    s = perturbed(27, shuffle=False)
    eng, _ = make_pair(s)
    eng.set_materials(model=ZR_MODEL_STVK_HENCKY)
    with pytest.raises(ZrError):
        eng.advanceOneTimeStepImplicit(1e-3, implicit_opts())
    eng.close()


@pytest.mark.parametrize("viscosities", [(5.0, 8.0), (0.2, 0.5)])
def test_viscoelastic_branch(viscosities):
    # FElasticNonequilibratedBasedMpmForceHelper on top of the particles' own
    # model: a pre-strained branch that relaxes while the body falls
    s = perturbed(29, shuffle=False, E=1e4, F_noise=0.03)
    s["V"][:, 1] -= 0.5
    n = len(s["mass"])
    rng = np.random.default_rng(29)
    Fn = np.tile(np.eye(3), (n, 1, 1)) + 0.05 * rng.normal(size=(n, 3, 3))
    Fneq = np.ascontiguousarray(Fn.transpose(0, 2, 1).reshape(n, 9))
    mu_n = np.full(n, 2e3) * np.exp(0.2 * rng.normal(size=n))
    lam_n = np.full(n, 3e3)
    eng, ora = make_pair(s, colliders=_ground(s))
    eng.set_viscoelastic(Fneq, mu_n, lam_n, *viscosities)
    ora.set_viscoelastic(Fneq, mu_n, lam_n, *viscosities)
    for _ in range(8):
        eng.advanceOneTimeStep(1e-4)
        ora.step(1e-4)
    out = eng.download()
    _check(out, ora.get_particles(), 1e-9)
    assert rel_err(eng.get_viscoelastic(), ora.get_viscoelastic()) < 1e-9
    with pytest.raises(ZrError):  # the implicit step has no viscoelastic branch
        eng.advanceOneTimeStepImplicit(1e-3, implicit_opts())
    eng.set_viscoelastic(None, None, None, 1.0, 1.0)  # off again
    eng.advanceOneTimeStep(1e-4)
    eng.close()
