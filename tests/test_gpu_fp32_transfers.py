"""The FP32 form of the explicit transfers (zr_config.precision =
ZR_FP32_REFINED): the MLS P2G's per-particle terms in FP32 (weights, offsets,
affine and stress products; the NeoHookeanBorden stress itself), the node
sums, grid update, G2P, F update, damage and particle state in FP64.
Checked against the FP64 oracle on the phase-field fracture scene
(NeoHookeanBorden with the moving sticky sphere) and on a fixed-corotated
cube (its stress evaluated in FP64 and rounded), within the bounds written
in each test; and the configuration's precision field is validated."""
import numpy as np
import pytest

import oracle
from ziran2019_amd import scenes
from ziran2019_amd.engine import ZR_FP32_REFINED, MpmEngine, ZrError, phase_field_opts
from _helpers import perturbed, rel_err

pytestmark = pytest.mark.gpu


def _engine(s, precision):
    e = MpmEngine(dx=s["dx"], gravity=s["gravity"], precision=precision)
    e.upload(s)
    e.set_colliders(s["colliders"])
    return e


def test_fp32_transfers_fracture():
    s = scenes.fracture_scene(cells=(12, 8, 10), nsites=20, sphere_radius=0.015)
    n = len(s["mass"])
    print("fracture, FP32 transfers:", n, "particles", flush=True)
    ora = oracle.OracleSim(dx=s["dx"], gravity=s["gravity"], nthreads=16)
    ora.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    for c in s["colliders"]:
        ora.add_collider(**c)
    eng = _engine(s, ZR_FP32_REFINED)
    kw = dict(l0=s["l0"], one_over_sigma_c=s["one_over_sigma_c"], residual_phase=s["residual_phase"])
    eng.set_phase_field(phase_field_opts(), **kw)
    ora.set_phase_field(oracle.pf_opts(), **kw)
    dt = 5e-4
    for step in range(8):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
    X, V, C, F = ora.get_particles()
    out = eng.download()
    pg, po = eng.get_phase_field(), ora.get_phase_field()
    disp = np.abs(X - s["X"]).max()
    ex = np.abs(out["X"] - X).max() / disp
    ev, ef, ec, eh = rel_err(out["V"], V), rel_err(out["F"], F), rel_err(pg["c"], po["c"]), rel_err(pg["H"], po["H"])
    print("FP32 transfers, fracture: X error / displacement", ex, "V", ev, "F", ef, "c", ec, "H", eh, flush=True)
    assert np.abs(V[:, 1]).max() > 1.0  # the sphere struck the block
    assert ex < 1e-4 and ev < 1e-4 and ef < 1e-6 and ec < 1e-5 and eh < 1e-4
    eng.close()


def test_fp32_transfers_fixed_corotated_cube():
    s = perturbed(61, n_cells=6, F_noise=0.03, C_noise=0.3)
    cols = [dict(shape=0, type=1, origin=(0.0, float(s["X"][:, 1].min()) + 0.5 * s["dx"], 0.0),
                 normal=(0.0, 1.0, 0.0))]
    s["gravity"] = (0.0, -9.8, 0.0)
    s["colliders"] = cols
    ora = oracle.OracleSim(dx=s["dx"], gravity=s["gravity"], nthreads=8)
    ora.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    for c in cols:
        ora.add_collider(**c)
    eng = _engine(s, ZR_FP32_REFINED)
    dt = 1e-4
    for _ in range(10):
        eng.advanceOneTimeStep(dt)
        ora.step(dt)
    X, V, C, F = ora.get_particles()
    out = eng.download()
    disp = np.abs(X - s["X"]).max()
    ex, ev, ec, ef = (np.abs(out["X"] - X).max() / disp, rel_err(out["V"], V), rel_err(out["C"], C),
                      rel_err(out["F"], F))
    print("FP32 transfers, cube: X error / displacement", ex, "V", ev, "C", ec, "F", ef, flush=True)
    assert ex < 1e-5 and ev < 1e-5 and ec < 1e-4 and ef < 1e-7
    eng.close()


def test_unknown_transfer_precision_is_rejected():
    with pytest.raises(ZrError):
        MpmEngine(dx=1.0 / 64, precision=7)
