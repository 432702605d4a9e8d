"""GPU parity of colliders in a moving frame x = R s X + b(t) (rotation,
angular velocity, scaling, translation; CollisionObject.cpp:158-194,
279-296,340-405) against the oracle: explicit MLS steps, the implicit step
(collision projection with world-space normals), calculateDt with a spinning
sphere, and the adjoint's refusal of frames it has no derivative for."""
import numpy as np
import pytest

import oracle
from ziran2019_amd.engine import ZrError, implicit_opts
from _helpers import make_pair, perturbed, rel_err
from test_oracle_collider_frame import _frames, _quat

pytestmark = pytest.mark.gpu


def _spinner(s, typ=1):
    c = s["X"].mean(axis=0)
    ext = np.abs(s["X"] - c).max()
    return dict(shape=1, type=typ, origin=(0.0, 0.0, 0.0), radius=0.6 * ext / 1.2,
                translation=c + np.array([0.4, 0.0, 0.0]) * ext, rotation=_quat((0.0, 1.0, 1.0), 0.4),
                omega=(-2.0 * np.pi * 0.1 * 20, 3.0, 0.5), scale=1.2, dscale_dt=0.3, velocity=(0.0, -0.1, 0.0))


@pytest.mark.parametrize("steps", [1, 6])
def test_explicit_steps_with_framed_colliders(steps):
    s = perturbed(41, n_cells=5, shuffle=False, F_noise=0.02, C_noise=0.2)
    s["V"][:, 1] -= 1.0
    s["V"][:, 0] += 0.5
    plain, framed = _frames(s)
    cols = framed + [_spinner(s)]
    print("framed colliders, explicit steps", steps, flush=True)
    eng, ora = make_pair(s, colliders=cols)
    for _ in range(steps):
        eng.advanceOneTimeStep(1e-3)
        ora.step(1e-3)
    out = eng.download()
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-12
    assert rel_err(out["V"], V) < 1e-9
    assert rel_err(out["F"], F) < 1e-10


def test_implicit_step_with_framed_colliders():
    s = perturbed(42, n_cells=5, shuffle=False, F_noise=0.02, C_noise=0.1, E=2e4)
    s["V"][:, 1] -= 0.5
    plain, framed = _frames(s)
    cols = [framed[0], _spinner(s, typ=2)]
    print("framed colliders, implicit step", flush=True)
    eng, ora = make_pair(s, colliders=cols)
    for _ in range(2):
        st_g = eng.advanceOneTimeStepImplicit(2e-3, implicit_opts())
        flags, st_o = ora.step_implicit(2e-3, oracle.implicit_opts())
        assert st_g.cg_iterations == st_o.cg_iterations
    out = eng.download()
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-11
    assert rel_err(out["V"], V) < 1e-8


def test_calculate_dt_with_a_spinning_sphere_bit_exact():
    s = perturbed(43, n_cells=4)
    s["V"] *= 0.01
    eng, ora = make_pair(s, colliders=[_spinner(s)])
    assert eng.calculateDt(1.0) == ora.calculate_dt(1.0)
    eng.close()


def test_adjoint_refuses_a_spinning_collider():
    from ziran2019_amd.autodiff import Rollout
    s = perturbed(44, n_cells=3)
    n = len(s["mass"])
    prob = Rollout(s, 2, 1e-4, colliders=[_spinner(s)])
    X = prob.forward(np.full(n, 1e4), np.zeros((n, 3)))
    with pytest.raises(ZrError):
        prob.backward(np.ones_like(X))
    prob.close()
