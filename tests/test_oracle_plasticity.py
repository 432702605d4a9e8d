"""Pins the oracle's StvkWithHencky and NeoHookeanBorden models and the von
Mises / Drucker-Prager return maps (CPU only).

- first Piola stresses are the FP64 central-difference gradient of psi
  (StvkWithHencky.cpp:114-143; NeoHookeanBorden.cpp:38-86 with P = tau F^-T);
- return maps are checked by what the mathematics fixes, not by retyping the
  update: the projected strain lies on the yield surface (von Mises:
  |dev log sigma| = tau_y / 2 mu; Drucker-Prager: on the cone or at the tip),
  the projection is coaxial with F (same singular vectors), idempotent, keeps
  volume where the model says so, and leaves states inside the surface alone
  (PlasticityApplier.cpp:249-316, 443-501).
"""
import numpy as np
import pytest

import oracle


def _rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _F(rng, sig):
    return _rot(rng) @ np.diag(sig) @ _rot(rng).T


def _fd_grad(psi, F, h=1e-6):
    G = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            E = np.zeros((3, 3))
            E[i, j] = h
            G[i, j] = (psi(F + E) - psi(F - E)) / (2 * h)
    return G


def _dev_log_norm(F):
    s = np.linalg.svd(F, compute_uv=False)
    e = np.log(s)
    return np.linalg.norm(e - e.mean()), e.sum()


def _coaxial(F0, F1):
    """F1 F0^-1 is symmetric (same singular vectors, possibly rescaled)."""
    A = F1 @ np.linalg.inv(F0)
    return np.abs(A - A.T).max() <= 1e-10 * np.abs(A).max()


@pytest.mark.parametrize("seed", range(4))
def test_stvk_hencky_piola_is_energy_gradient(seed):
    rng = np.random.default_rng(seed)
    mu, lam = oracle.lame(1e5, 0.3)
    F = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
    if np.linalg.det(F) < 0.2:
        F = F + np.eye(3)
    P = oracle.stvk_hencky_first_piola(mu, lam, F)
    G = _fd_grad(lambda X: oracle.stvk_hencky_psi(mu, lam, X), F)
    assert np.abs(G - P).max() < 1e-7 * np.abs(P).max()
    tau = oracle.stvk_hencky_kirchhoff(mu, lam, F)
    assert np.abs(tau - P @ F.T).max() < 1e-12 * np.abs(tau).max()
    assert np.abs(tau - tau.T).max() < 1e-9 * np.abs(tau).max()


def test_stvk_hencky_rest_and_rotation():
    rng = np.random.default_rng(7)
    mu, lam = oracle.lame(1e4, 0.3)
    R = _rot(rng)
    assert np.abs(oracle.stvk_hencky_first_piola(mu, lam, R)).max() < 1e-10
    assert abs(oracle.stvk_hencky_psi(mu, lam, R)) < 1e-20
    F = _F(rng, [1.2, 0.9, 0.8])
    assert abs(oracle.stvk_hencky_psi(mu, lam, R @ F) - oracle.stvk_hencky_psi(mu, lam, F)) < 1e-9


@pytest.mark.parametrize("J_side,g", [(1.2, 1.0), (0.8, 1.0), (1.2, 0.3), (0.8, 0.3)])
def test_nhb_kirchhoff_is_energy_gradient(J_side, g):
    rng = np.random.default_rng(int(J_side * 10 + g * 10))
    mu, lam = oracle.lame(500.0, 0.4)
    sig = np.array([1.1, 0.95, 1.0])
    sig = sig * (J_side / np.prod(sig)) ** (1.0 / 3.0)
    F = _F(rng, sig)
    tau = oracle.nhb_kirchhoff(mu, lam, g, F)
    P = tau @ np.linalg.inv(F).T
    G = _fd_grad(lambda X: oracle.nhb_psi(mu, lam, g, X), F)
    assert np.abs(G - P).max() < 1e-7 * np.abs(P).max()
    # psi(g=1) = psi_pos + psi_neg, and psi_pos carries the volumetric part only in tension
    pos = oracle.nhb_psi_pos(mu, lam, F)
    assert abs(oracle.nhb_psi(mu, lam, g, F) - (g * pos + (oracle.nhb_psi(mu, lam, 0.0, F)))) < 1e-12 * abs(pos)


def test_nhb_rest_state_and_kappa_literal():
    mu, lam = oracle.lame(500.0, 0.4)
    assert np.abs(oracle.nhb_kirchhoff(mu, lam, 1.0, np.eye(3))).max() < 1e-12
    # pure dilation J: tau = J * kappa/2 (J - 1/J) I with kappa = 0.666666 mu + lambda
    J = 1.1
    tau = oracle.nhb_kirchhoff(mu, lam, 1.0, np.eye(3) * J ** (1.0 / 3.0))
    kappa = 0.666666 * mu + lam
    assert np.allclose(tau, np.eye(3) * J * kappa / 2 * (J - 1 / J), rtol=1e-13, atol=1e-12)


@pytest.mark.parametrize("seed", range(5))
def test_von_mises_projects_onto_yield_surface(seed):
    rng = np.random.default_rng(seed)
    mu, lam = oracle.lame(1e5, 0.3)
    ty = 200.0
    F = _F(rng, np.exp(rng.normal(scale=0.05, size=3)))
    n0, tr0 = _dev_log_norm(F)
    assert n0 > ty / (2 * mu)  # starts outside
    F1, ty1, broken, changed = oracle.von_mises_project(mu, lam, 0.0, 1e30, ty, F)
    assert changed and not broken and ty1 == ty
    n1, tr1 = _dev_log_norm(F1)
    assert abs(n1 - ty / (2 * mu)) < 1e-12
    assert abs(tr1 - tr0) < 1e-12  # volume preserving
    assert _coaxial(F, F1)
    F2, _, _, changed2 = oracle.von_mises_project(mu, lam, 0.0, 1e30, ty, F1)
    assert np.abs(F2 - F1).max() < 1e-13 or not changed2


def test_von_mises_inside_hardening_and_failure():
    rng = np.random.default_rng(11)
    mu, lam = oracle.lame(1e5, 0.3)
    F = _F(rng, [1.0005, 0.9998, 0.9999])
    F1, ty1, _, changed = oracle.von_mises_project(mu, lam, 0.0, 1e30, 200.0, F)
    assert not changed and np.array_equal(F1, F) and ty1 == 200.0
    F = _F(rng, [1.05, 0.97, 0.99])
    n0, _ = _dev_log_norm(F)
    xi = 1e4
    F1, ty1, _, _ = oracle.von_mises_project(mu, lam, xi, 1e30, 200.0, F)
    dg = n0 - 200.0 / (2 * mu)
    assert abs(ty1 - (200.0 - xi * dg)) < 1e-9
    assert abs(_dev_log_norm(F1)[0] - 200.0 / (2 * mu)) < 1e-12  # projected with the pre-hardening yield stress
    _, _, broken, _ = oracle.von_mises_project(mu, lam, 0.0, 1.0, 200.0, F)
    assert broken


def test_drucker_prager_alpha():
    s = np.sin(30.0 / 180.0 * np.pi)
    assert abs(oracle.drucker_prager_alpha(30.0) - np.sqrt(2 / 3) * 2 * s / (3 - s)) < 1e-9


def _dp(F, logJp=0.0, beta=1.0, cohesion=0.0, vc=True):
    mu, lam = oracle.lame(1e5, 0.3)
    a = oracle.drucker_prager_alpha(30.0)
    F1, j1 = oracle.drucker_prager_project(mu, lam, a, beta, cohesion, vc, logJp, F)
    return F1, j1, mu, lam, a


def test_drucker_prager_tension_goes_to_tip():
    rng = np.random.default_rng(3)
    F = _F(rng, [1.05, 1.02, 0.99])  # tr log sigma > 0
    F1, j1, *_ = _dp(F, beta=0.5)
    assert np.abs(F1.T @ F1 - np.eye(3)).max() < 1e-13  # a rotation: the tip with zero cohesion
    assert abs(j1 - 0.5 * np.log(np.linalg.svd(F, compute_uv=False)).sum()) < 1e-13
    F1, j1, *_ = _dp(F, beta=0.5, vc=False)
    assert j1 == 0.0


@pytest.mark.parametrize("seed", range(4))
def test_drucker_prager_compression_projects_onto_cone(seed):
    rng = np.random.default_rng(100 + seed)
    sig = np.exp(np.array([0.08, -0.06, -0.07]) + 0.002 * rng.normal(size=3))
    F = _F(rng, sig)
    n0, tr0 = _dev_log_norm(F)
    assert tr0 < 0
    F1, j1, mu, lam, a = _dp(F, logJp=0.0)
    n1, tr1 = _dev_log_norm(F1)
    assert j1 == 0.0
    assert abs(tr1 - tr0) < 1e-13
    # on the cone: |dev| + (3 lambda + 2 mu)/(2 mu) tr alpha = 0
    assert abs(n1 + (3 * lam + 2 * mu) / (2 * mu) * tr1 * a) < 1e-12
    assert _coaxial(F, F1)
    F2, _, *_ = _dp(F1)
    assert np.abs(F2 - F1).max() < 1e-12


def test_drucker_prager_inside_cone_unchanged():
    rng = np.random.default_rng(9)
    F = _F(rng, np.exp([-0.01, -0.011, -0.012]))
    F1, j1, *_ = _dp(F, logJp=-0.001)
    assert np.abs(F1 - F).max() < 1e-14 and j1 == 0.0


def test_oracle_step_applies_plasticity():
    from ziran2019_amd import scenes
    s = scenes.cube_scene_small(seed=3, ncells=4, E=1e5)
    n = len(s["mass"])
    mu, lam = oracle.lame(1e5, 0.3)
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0), nthreads=2)
    rng = np.random.default_rng(0)
    F = np.tile(np.eye(3), (n, 1, 1)) + 0.05 * rng.normal(size=(n, 3, 3))
    Fcm = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9))
    sim.set_particles(s["X"], s["V"], s["C"], Fcm, s["mass"], s["vol"], np.full(n, mu), np.full(n, lam))
    ty = 200.0
    sim.set_materials(model=1, plastic=1, pparam=(0.0, 1e30, 0.0, 0.0), pstate=ty)
    sim.step(1e-4)
    _, _, _, Fo = sim.get_particles()
    for p in range(0, n, 7):
        Fp = Fo[p].reshape(3, 3).T
        assert _dev_log_norm(Fp)[0] <= ty / (2 * mu) + 1e-12


# ---- NonAssociativeCamClay (PlasticityApplier.cpp:10-19, 36-135) on NeoHookeanBorden
def _cc_setup():
    mu, lam = 3e3, 4e3
    kappa = 0.666666 * mu + lam  # NeoHookeanBorden.cpp:20
    return mu, lam, kappa


def test_cam_clay_M_closed_form():
    # M = sqrt(2/3) 2 s / (3 - s) * 3 / sqrt(2/3) = 6 s / (3 - s) in 3-D
    for phi in (15.0, 30.0, 45.0, 60.0):
        s = np.sin(phi / 180.0 * 3.141592653)
        assert oracle.cam_clay_M(phi) == pytest.approx(6.0 * s / (3.0 - s), rel=1e-15)
        assert engine_cam_clay_M(phi) == pytest.approx(oracle.cam_clay_M(phi), rel=1e-15)


def engine_cam_clay_M(phi):
    from ziran2019_amd.engine import cam_clay_M
    return cam_clay_M(phi)


@pytest.mark.parametrize("J", [0.9, 1.2])
def test_cam_clay_tips(J):
    # logJp = 0: p0 = kappa 1e-5, so compression (p > p0) and tension
    # (p < -beta p0) both land on a tip: a volumetric Se with the tip's Je
    mu, lam, kappa = _cc_setup()
    beta, xi = 0.5, 3.0
    R = np.linalg.qr(np.random.default_rng(1).normal(size=(3, 3)))[0]
    R *= np.sign(np.linalg.det(R))
    F = R @ np.diag([J ** (1 / 3)] * 3)
    Fn, lj, moved = oracle.cam_clay_project(mu, lam, oracle.cam_clay_M(45.0), beta, xi, 0.0, F)
    p0 = kappa * 1e-5
    Je_new = np.sqrt(-2 * p0 / kappa + 1) if J < 1 else np.sqrt(2 * beta * p0 / kappa + 1)
    assert moved
    sv = np.linalg.svd(Fn, compute_uv=False)
    assert np.allclose(sv, Je_new ** (1 / 3), rtol=1e-13)
    assert lj == pytest.approx(np.log(J / Je_new), rel=1e-12)


def test_cam_clay_inside_keeps_state():
    mu, lam, kappa = _cc_setup()
    F = np.diag([1.01, 0.995, 1.0])
    Fn, lj, moved = oracle.cam_clay_project(mu, lam, oracle.cam_clay_M(45.0), 0.5, 3.0, -0.5, F)
    assert not moved
    assert lj == -0.5
    assert np.allclose(Fn, F, rtol=0, atol=1e-14)


def test_cam_clay_projection_onto_surface():
    mu, lam, kappa = _cc_setup()
    M, beta, xi, lj0 = oracle.cam_clay_M(45.0), 0.5, 3.0, -0.05
    rng = np.random.default_rng(3)
    U = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    V = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    U *= np.sign(np.linalg.det(U))
    V *= np.sign(np.linalg.det(V))
    sig = np.array([1.3, 0.85, 0.93])
    F = U @ np.diag(sig) @ V.T
    Fn, lj, moved = oracle.cam_clay_project(mu, lam, M, beta, xi, lj0, F)
    assert moved
    # trial quantities from the closed forms
    p0 = kappa * (1e-5 + np.sinh(xi * max(-lj0, 0.0)))
    Je = sig.prod()
    s2 = sig ** 2
    s_hat = mu * Je ** (-2 / 3) * (s2 - s2.mean())
    p = -kappa / 2 * (Je - 1 / Je) * Je
    a = (1 + 2 * beta) * 3 / 2
    assert -beta * p0 < p < p0
    assert a * s_hat @ s_hat + M * M * (p + beta * p0) * (p - p0) > 1e-4  # outside
    # the return keeps the mean of Se^2 and puts the deviatoric part at the
    # surface's radius k along the trial direction (with the trial Je)
    k = np.sqrt(-M * M * (p + beta * p0) * (p - p0) / a)
    sn = np.sort(np.linalg.svd(Fn, compute_uv=False) ** 2)
    want = np.sort(k / mu * Je ** (2 / 3) * s_hat / np.linalg.norm(s_hat) + s2.mean())
    assert np.allclose(sn, want, rtol=1e-12)
    # hardening: the logJp step is the tip of the line from the surface's
    # centre (pc, 0) through the trial point (p, q) where it meets the surface
    pc = (1 - beta) * p0 / 2
    q = np.sqrt(1.5) * np.linalg.norm(s_hat)
    Je_fake = Je / np.exp(lj - lj0)
    p_fake = kappa / 2 * (1 - Je_fake ** 2)
    q_fake = q * (p_fake - pc) / (p - pc)
    y_fake = (1 + 2 * beta) * q_fake ** 2 + M * M * (p_fake + beta * p0) * (p_fake - p0)
    assert abs(y_fake) < 1e-9 * (M * M * p0 * p0)
