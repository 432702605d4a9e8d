"""The viscoelastic branch on the CPU: the oracle's evolveStrain of
FElasticNonequilibratedBasedMpmForceHelper (FElasticNonequilibratedBasedMpmForceHelper.h:62-87)
against what its Hencky-strain relaxation fixes: no relaxation without
viscosity, the isotropic closed form, the deviatoric decay 1 / (1 + dt alpha),
rotation invariance, and the reference's volumetric term built from the first
principal log-strain (Eigen's trace() of a 3x1 vector)."""
import numpy as np
import pytest

import oracle

MU, LAM, DT = 2e3, 3e3, 1e-3


def _rot(seed):
    R = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))[0]
    return R * np.sign(np.linalg.det(R))


def _coeffs(vd, vv):
    alpha = 2 * MU / vd
    beta = 2 * (2 * MU + 3 * LAM) / (9 * vv) - 2 * MU / (3 * vd)
    return alpha, beta


def test_no_viscosity_is_the_elastic_update():
    rng = np.random.default_rng(1)
    G = 0.3 * rng.normal(size=(3, 3))
    F = np.eye(3) + 0.1 * rng.normal(size=(3, 3))
    out = oracle.visco_relax(MU, LAM, 1e300, 1e300, DT, G, F)
    want = F + DT * G @ F
    assert np.allclose(out, want, rtol=1e-13, atol=1e-15)


def test_isotropic_stretch_closed_form():
    vd, vv = 5.0, 7.0
    alpha, beta = _coeffs(vd, vv)
    e = 0.02
    R = _rot(2)
    F = R * np.exp(e)
    out = oracle.visco_relax(MU, LAM, vd, vv, DT, np.zeros((3, 3)), F)
    e_new = e * (1 - DT * beta / (1 + DT * (alpha + 3 * beta))) / (1 + DT * alpha)
    assert np.allclose(out, R * np.exp(e_new), rtol=1e-13, atol=1e-15)


def test_principal_strains_and_rotation_invariance():
    vd, vv = 2.0, 3.0
    alpha, beta = _coeffs(vd, vv)
    eps = np.array([0.05, -0.02, 0.01])
    U, V = _rot(3), _rot(4)
    F = U @ np.diag(np.exp(eps)) @ V.T
    out = oracle.visco_relax(MU, LAM, vd, vv, DT, np.zeros((3, 3)), F)
    # the volumetric term uses the first principal log-strain; the SVD sorts
    # the singular values, so that is the largest one
    t = eps.max()
    want_eps = (eps - DT * beta / (1 + DT * (alpha + 3 * beta)) * t) / (1 + DT * alpha)
    assert np.allclose(out, U @ np.diag(np.exp(want_eps)) @ V.T, rtol=1e-12, atol=1e-15)
    # differences of principal log-strains decay by 1 / (1 + dt alpha)
    s_new = np.log(np.linalg.svd(out, compute_uv=False))
    s_old = np.sort(eps)[::-1]
    assert np.allclose(s_new[0] - s_new[1], (s_old[0] - s_old[1]) / (1 + DT * alpha), rtol=1e-10)
    # rotating the particle rotates the result
    Q = _rot(5)
    out_q = oracle.visco_relax(MU, LAM, vd, vv, DT, np.zeros((3, 3)), Q @ F)
    assert np.allclose(out_q, Q @ out, rtol=1e-12, atol=1e-15)


def test_oracle_step_adds_the_branch_stress():
    from _helpers import perturbed
    s = perturbed(4, shuffle=False)
    n = len(s["mass"])
    o1 = oracle.OracleSim(dx=s["dx"], nthreads=4)
    o1.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    o2 = oracle.OracleSim(dx=s["dx"], nthreads=4)
    o2.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    # Fneq = identity: no branch stress, so the first step is the plain one
    o2.set_viscoelastic(None, np.full(n, MU), np.full(n, LAM), 10.0, 10.0)
    o1.step(1e-4)
    o2.step(1e-4)
    assert np.array_equal(o1.get_particles()[0], o2.get_particles()[0])
    # after it Fneq has moved with the flow and relaxed
    assert not np.allclose(o2.get_viscoelastic(), np.tile(np.eye(3).reshape(9), (n, 1)))
