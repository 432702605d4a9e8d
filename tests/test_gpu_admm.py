"""GPU parity of the ADMM step (AdmmSimulationFoam.h:163-818) against the FP64
oracle: StvkWithHencky particles with no / von Mises / Drucker-Prager
plasticity on a sticky ground, several steps (warm starts, dual rescaling);
ADMM variables, particle state and MINRES iteration counts; and the
viscoelastic branch (enable_visco) with its relaxed Fe_Nonequilibrated."""
import numpy as np
import pytest

import oracle
from ziran2019_amd.engine import (ZR_FP32_REFINED, ZR_MODEL_STVK_HENCKY, ZR_PLASTIC_DRUCKER_PRAGER,
                                  ZR_PLASTIC_VON_MISES, ZrError, admm_opts, drucker_prager_alpha)
from _helpers import make_pair, perturbed, rel_err

pytestmark = pytest.mark.gpu


def _materials(n, rng, kind):
    if kind == "elastic":
        return dict(model=ZR_MODEL_STVK_HENCKY, plasticity=0, plastic_param=(0, 0, 0, 0), plastic_state=0.0)
    if kind == "vm":
        return dict(model=ZR_MODEL_STVK_HENCKY, plasticity=ZR_PLASTIC_VON_MISES, plastic_param=(0.0, 1e30, 0, 0),
                    plastic_state=200.0)
    if kind == "dp":
        return dict(model=ZR_MODEL_STVK_HENCKY, plasticity=ZR_PLASTIC_DRUCKER_PRAGER,
                    plastic_param=(drucker_prager_alpha(30.0), 1.0, 0.0, 1.0), plastic_state=0.0)
    plastic = rng.integers(0, 3, size=n).astype(np.int32)
    pp = np.zeros((n, 4))
    pp[plastic == ZR_PLASTIC_VON_MISES] = (0.0, 1e30, 0.0, 0.0)
    pp[plastic == ZR_PLASTIC_DRUCKER_PRAGER] = (drucker_prager_alpha(30.0), 1.0, 0.0, 1.0)
    st = np.where(plastic == ZR_PLASTIC_VON_MISES, 200.0, 0.0)
    return dict(model=ZR_MODEL_STVK_HENCKY, plasticity=plastic, plastic_param=pp, plastic_state=st)


@pytest.mark.parametrize("kind", ["elastic", "vm", "dp", "mixed"])
def test_admm_steps_parity(kind):
    s = perturbed(41, n_cells=5, shuffle=False, F_noise=0.03, C_noise=0.1, E=1e5)
    s["V"][:, 1] -= 0.5
    n = len(s["mass"])
    y0 = float(s["X"][:, 1].min() + 0.5 * s["dx"])
    cols = [dict(shape=0, type=1, origin=(0.0, y0, 0.0), normal=(0.0, 1.0, 0.0))]
    eng, ora = make_pair(s, colliders=cols)
    m = _materials(n, np.random.default_rng(41), kind)
    eng.set_materials(**m)
    ora.set_materials(model=m["model"], plastic=m["plasticity"], pparam=m["plastic_param"], pstate=m["plastic_state"])
    dt = 1e-3
    og = admm_opts(admm_iterations=20)
    oo = oracle.admm_opts(admm_iterations=20)
    for step in range(3):
        sg = eng.advanceOneTimeStepAdmm(dt, og)
        _, so = ora.step_admm(dt, oo)
        assert sg.minres_iterations == so.minres_iterations, step
    a_g, a_o = eng.get_admm_state(), ora.get_admm()
    assert rel_err(a_g["omega"], a_o["omega"]) < 1e-10
    assert rel_err(a_g["F"], a_o["F"]) < 1e-9
    assert rel_err(a_g["y"], a_o["y"]) < 1e-8
    out = eng.download()
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-10
    assert rel_err(out["V"], V) < 1e-8
    assert rel_err(out["F"], F) < 1e-9
    if kind != "elastic":
        pg, po = eng.get_materials()["plastic_state"], ora.get_materials()["pstate"]
        worst = int(np.argmax(np.abs(pg - po)))
        print(kind, "plastic state: worst particle", worst, pg[worst], po[worst], flush=True)
        assert rel_err(pg, po) < 1e-8
    eng.close()


def _visco_pair(seed):
    s = perturbed(seed, n_cells=5, shuffle=False, F_noise=0.03, C_noise=0.1, E=1e5)
    s["V"][:, 1] -= 0.5
    n = len(s["mass"])
    rng = np.random.default_rng(seed)
    Fv = np.tile(np.eye(3), (n, 1, 1)) + 0.04 * rng.normal(size=(n, 3, 3))
    fneq = np.ascontiguousarray(Fv.transpose(0, 2, 1).reshape(n, 9))
    mu2, lam2 = oracle.lame(5e4, 0.3)
    y0 = float(s["X"][:, 1].min() + 0.5 * s["dx"])
    cols = [dict(shape=0, type=1, origin=(0.0, y0, 0.0), normal=(0.0, 1.0, 0.0))]
    eng, ora = make_pair(s, colliders=cols)
    return s, eng, ora, fneq, mu2, lam2


def test_admm_viscoelastic_parity():
    """AdmmSimulationFoam.h's enable_visco branch: the second F per particle
    from Fe_Nonequilibrated through initStep, the local step with the
    relaxation map, the global operator, rhs and preconditioner, and uStep;
    then gridToParticles' relaxation of Fneq with the APIC grad v.  Three
    steps of 20 ADMM iterations against the oracle."""
    s, eng, ora, fneq, mu2, lam2 = _visco_pair(43)
    eng.set_materials(**_materials(len(s["mass"]), None, "elastic"))
    ora.set_materials(model=ZR_MODEL_STVK_HENCKY)
    eng.set_viscoelastic(fneq, mu2, lam2, 20.0, 80.0)
    ora.set_viscoelastic(fneq, mu2, lam2, 20.0, 80.0)
    dt = 1e-3
    for step in range(3):
        sg = eng.advanceOneTimeStepAdmm(dt, admm_opts(admm_iterations=20))
        _, so = ora.step_admm(dt, oracle.admm_opts(admm_iterations=20))
        print("visco admm step", step, "minres", sg.minres_iterations, so.minres_iterations, flush=True)
        assert sg.minres_iterations == so.minres_iterations, step
    a_g, a_o = eng.get_admm_state(), ora.get_admm()
    assert rel_err(a_g["omega"], a_o["omega"]) < 1e-10
    assert rel_err(a_g["F"], a_o["F"]) < 1e-9
    out = eng.download()
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-10
    assert rel_err(out["V"], V) < 1e-8
    assert rel_err(out["F"], F) < 1e-9
    fg, fo = eng.get_viscoelastic(), ora.get_viscoelastic()
    assert rel_err(fg, fo) < 1e-9
    assert np.abs(fo - fneq).max() > 1e-3  # the branch moved and relaxed
    eng.close()


def test_admm_viscoelastic_fp32_global_step():
    """The FP32-bulk global step with the viscoelastic branch (the FP32
    operator adds the FFn term): three steps of 20 ADMM iterations against
    the FP64 oracle, within the bounds written here (measured on MI355X:
    1e-8 of the displacement, V 2e-8, F and Fneq 5e-9 relative)."""
    s, eng, ora, fneq, mu2, lam2 = _visco_pair(45)
    eng.set_materials(**_materials(len(s["mass"]), None, "elastic"))
    ora.set_materials(model=ZR_MODEL_STVK_HENCKY)
    eng.set_viscoelastic(fneq, mu2, lam2, 20.0, 80.0)
    ora.set_viscoelastic(fneq, mu2, lam2, 20.0, 80.0)
    dt = 1e-3
    for step in range(3):
        sg = eng.advanceOneTimeStepAdmm(dt, admm_opts(admm_iterations=20, precision=ZR_FP32_REFINED))
        _, so = ora.step_admm(dt, oracle.admm_opts(admm_iterations=20))
        print("fp32 visco admm step", step, "minres", sg.minres_iterations, so.minres_iterations, flush=True)
    X, V, C, F = ora.get_particles()
    out = eng.download()
    disp = np.abs(X - s["X"]).max()
    ex, ev, ef = (np.abs(out["X"] - X).max() / disp, rel_err(out["V"], V), rel_err(out["F"], F))
    efn = rel_err(eng.get_viscoelastic(), ora.get_viscoelastic())
    print("FP32-bulk visco ADMM: X error / displacement", ex, "V", ev, "F", ef, "Fneq", efn, flush=True)
    assert ex < 1e-6 and ev < 1e-6 and ef < 1e-7 and efn < 1e-7
    eng.close()


def test_admm_rejects_unknown_precision():
    s, eng, ora, fneq, mu2, lam2 = _visco_pair(44)
    eng.set_materials(**_materials(len(s["mass"]), None, "elastic"))
    with pytest.raises(ZrError):
        eng.advanceOneTimeStepAdmm(1e-3, admm_opts(admm_iterations=2, precision=7))
    eng.close()


def test_admm_viscoelastic_refuses_plasticity():
    s, eng, ora, fneq, mu2, lam2 = _visco_pair(44)
    eng.set_materials(**_materials(len(s["mass"]), None, "vm"))
    eng.set_viscoelastic(fneq, mu2, lam2, 20.0, 80.0)
    with pytest.raises(ZrError):
        eng.advanceOneTimeStepAdmm(1e-3, admm_opts(admm_iterations=2))
    eng.close()
