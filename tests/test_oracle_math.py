"""Pins the oracle's B-splines, 3x3 SVD and fixed-corotated model (CPU only).

- quadratic B-spline: partition of unity, zero first moment, second moment
  1/4 in index space (hence D^-1 = 4/dx^2, MpmSimulationBase.cpp:114-116),
  derivative by central finite differences;
- SVD: singular values against LAPACK (numpy.linalg.svd) with the
  reference's sign/order convention (ImplicitQRSVD.h:257-345), reconstruction,
  orthogonality and det(U) = det(V) = +1, special cases (diagonal, rotation,
  repeated, rank-deficient, reflection);
- fixed-corotated: P = dpsi/dF and dP by FP64 central finite differences,
  rotation invariance, zero stress at F = R, tau = P F^T.
"""
import numpy as np
import pytest

import oracle


def test_bspline_moments(oracle_lib):
    for x in np.linspace(3.0, 4.0, 101):
        base, w, dw = oracle.bspline_weights(x)
        nodes = base + np.arange(3)
        assert base == int(np.floor(x - 0.5))
        assert abs(w.sum() - 1.0) < 1e-15
        assert abs(np.dot(w, nodes - x)) < 1e-14
        assert abs(np.dot(w, (nodes - x) ** 2) - 0.25) < 1e-14
        assert abs(dw.sum()) < 1e-14
        h = 1e-6
        # derivative of w_i(x) wrt x at fixed base node
        bp, wp, _ = oracle.bspline_weights(x + h)
        bm, wm, _ = oracle.bspline_weights(x - h)
        if bp == base and bm == base:
            fd = (wp - wm) / (2 * h)
            assert np.allclose(fd, dw, atol=1e-8)


def _check_svd(F):
    U, s, V, _ = oracle.svd3(F)
    nrm = max(np.linalg.norm(F), 1e-300)
    # the reference deflates below tol = 1024 eps max(|B|/2, 1), an absolute
    # floor for small matrices (ImplicitQRSVD.h:381,526-533)
    tol = 4 * 1024 * np.finfo(float).eps * max(nrm, 1.0)
    assert np.linalg.norm(U @ np.diag(s) @ V.T - F) <= tol
    assert np.allclose(U.T @ U, np.eye(3), atol=1e-13)
    assert np.allclose(V.T @ V, np.eye(3), atol=1e-13)
    assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12
    ref = np.linalg.svd(F, compute_uv=False)
    assert np.allclose(np.sort(np.abs(s))[::-1], ref, atol=tol)
    assert s[0] >= 0 and s[1] >= 0
    assert abs(s[0]) >= abs(s[1]) - 1e-15 and abs(s[1]) >= abs(s[2]) - 1e-15
    d = np.linalg.det(F)
    if abs(d) > 1e-10 * max(nrm, 1.0) ** 3:
        assert np.sign(s[2]) == np.sign(d)
    return U, s, V


def test_svd_random(oracle_lib):
    rng = np.random.default_rng(11)
    for _ in range(2000):
        F = rng.normal(size=(3, 3)) * rng.choice([1e-3, 1.0, 1e3])
        _check_svd(F)


def test_svd_near_identity_and_degenerate(oracle_lib):
    rng = np.random.default_rng(12)
    for _ in range(500):
        F = np.eye(3) + 1e-6 * rng.normal(size=(3, 3))
        _check_svd(F)
    cases = [np.eye(3), np.diag([3.0, 2.0, 1.0]), np.diag([1.0, 2.0, 3.0]), np.diag([2.0, 2.0, 2.0]),
             np.diag([1.0, 1.0, -1.0]), np.diag([-1.0, 1.0, 1.0]), np.zeros((3, 3)),
             np.outer([1.0, 2.0, 3.0], [0.5, -1.0, 2.0]), np.diag([5.0, 0.0, 0.0]), np.diag([1.0, 1e-20, 1.0])]
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    cases += [Rz, Rz @ np.diag([2.0, 2.0, 0.5]), -Rz]
    for F in cases:
        _check_svd(F)


def test_svd_sorted_diagonal_is_identity_factors(oracle_lib):
    U, s, V, n = oracle.svd3(np.diag([3.0, 2.0, 1.0]))
    assert np.allclose(s, [3, 2, 1])
    assert np.allclose(U, np.eye(3)) and np.allclose(V, np.eye(3))


def test_polar(oracle_lib):
    rng = np.random.default_rng(5)
    for _ in range(200):
        F = np.eye(3) + 0.3 * rng.normal(size=(3, 3))
        R, S = oracle.polar3(F)
        assert np.allclose(R @ S, F, atol=1e-13)
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-13)
        assert np.allclose(S, S.T, atol=1e-13)


def _rand_F(rng, scale=0.3):
    F = np.eye(3) + scale * rng.normal(size=(3, 3))
    if np.linalg.det(F) < 0.2:
        F = np.eye(3) + 0.1 * rng.normal(size=(3, 3))
    return F


def test_lame(oracle_lib):
    mu, lam = oracle.lame(1e4, 0.3)
    assert abs(mu - 1e4 / 2.6) < 1e-9
    assert abs(lam - 1e4 * 0.3 / (1.3 * 0.4)) < 1e-9


def test_fcr_piola_is_energy_gradient(oracle_lib):
    rng = np.random.default_rng(7)
    mu, lam = oracle.lame(1e4, 0.3)
    for _ in range(40):
        F = _rand_F(rng)
        P = oracle.fcr_first_piola(mu, lam, F)
        h = 1e-6
        fd = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                E = np.zeros((3, 3)); E[i, j] = h
                fd[i, j] = (oracle.fcr_psi(mu, lam, F + E) - oracle.fcr_psi(mu, lam, F - E)) / (2 * h)
        assert np.allclose(P, fd, rtol=1e-6, atol=1e-6 * np.abs(P).max())


def test_fcr_differential_by_fd(oracle_lib):
    rng = np.random.default_rng(8)
    mu, lam = oracle.lame(1e4, 0.3)
    for _ in range(40):
        F = _rand_F(rng)
        dF = rng.normal(size=(3, 3))
        dP = oracle.fcr_first_piola_differential(mu, lam, F, dF)
        h = 1e-6
        fd = (oracle.fcr_first_piola(mu, lam, F + h * dF) - oracle.fcr_first_piola(mu, lam, F - h * dF)) / (2 * h)
        assert np.allclose(dP, fd, rtol=1e-6, atol=1e-6 * np.abs(dP).max())


def test_fcr_invariances(oracle_lib):
    rng = np.random.default_rng(9)
    mu, lam = oracle.lame(5e4, 0.3)
    for _ in range(20):
        F = _rand_F(rng)
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] *= -1
        assert abs(oracle.fcr_psi(mu, lam, q @ F) - oracle.fcr_psi(mu, lam, F)) < 1e-9 * max(1, oracle.fcr_psi(mu, lam, F))
        # zero stress at a pure rotation
        assert np.abs(oracle.fcr_first_piola(mu, lam, q)).max() < 1e-9
        P = oracle.fcr_first_piola(mu, lam, F)
        tau = oracle.fcr_kirchhoff(mu, lam, F)
        assert np.allclose(tau, P @ F.T, rtol=1e-12, atol=1e-10 * np.abs(tau).max())
        assert np.allclose(tau, tau.T, atol=1e-9 * np.abs(tau).max())


def test_fcr_hand_values(oracle_lib):
    """Fixed-corotated psi = mu |F - R|^2 + lambda/2 (J - 1)^2 and
    P = 2 mu (F - R) + lambda (J - 1) J F^-T (Stomakhin et al. 2012) at
    diagonal F, worked by hand: R = I for positive entries; for an inverted F
    the rotation comes from the sign-adjusted SVD (only the smallest singular
    value negative), so diag(1.2, 0.9, -1.1) has sigma = (1.2, 1.1, -0.9)."""
    mu, lam = 2.0, 3.0
    F = np.diag([1.2, 0.9, 1.1])
    # mu (0.2^2 + 0.1^2 + 0.1^2) + 1.5 (1.188 - 1)^2 = 0.12 + 0.053016
    assert oracle.fcr_psi(mu, lam, F) == pytest.approx(0.173016, rel=1e-14)
    P = oracle.fcr_first_piola(mu, lam, F)
    # 2 mu (F_ii - 1) + lambda (J - 1) J / F_ii with J = 1.188
    want = np.diag([0.8 + 3 * 0.188 * 1.188 / 1.2, -0.4 + 3 * 0.188 * 1.188 / 0.9,
                    0.4 + 3 * 0.188 * 1.188 / 1.1])
    assert np.abs(P - want).max() < 1e-14
    # inverted: mu (0.2^2 + 0.1^2 + 1.9^2) + 1.5 (-1.188 - 1)^2 = 7.32 + 7.181016
    Fi = np.diag([1.2, 0.9, -1.1])
    assert oracle.fcr_psi(mu, lam, Fi) == pytest.approx(14.501016, rel=1e-14)
    # the same energy for any rotation applied on either side
    rng = np.random.default_rng(10)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    assert oracle.fcr_psi(mu, lam, q @ Fi @ q.T) == pytest.approx(14.501016, rel=1e-12)
