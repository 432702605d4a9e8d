"""Pins the oracle's CD-MPM phase-field damage and Rankine (naive) damage
(Projects/fracture/FractureSimulation.h:209-554, PhaseField.h:70-76) on CPU.

- the lumped phase-field operator equals a dense matrix assembled here from
  independently computed quadratic B-spline weights (brute force);
- it is symmetric positive definite and reproduces constants exactly
  (sum_j grad w_jp = 0), so an undamaged, uniform block keeps c = 1;
- the CG result satisfies the system to the solver's tolerance;
- damage is monotone (c never grows, stays in [0, 1]), the history H never
  decreases and g = c^2 (1 - k) + k;
- the naive form cuts c to sigma_F / sigma_max, with sigma_max checked
  against numpy's symmetric eigensolver.
"""
import numpy as np
import pytest

import oracle
from ziran2019_amd import scenes


def _bspline(x, dx):
    """Textbook quadratic B-spline weights and gradients of one coordinate."""
    y = x / dx
    base = int(np.floor(y - 0.5))
    d = y - base
    w = np.array([0.5 * (1.5 - d) ** 2, 0.75 - (d - 1.0) ** 2, 0.5 * (d - 0.5) ** 2])
    dw = np.array([d - 1.5, -2.0 * (d - 1.0), d - 0.5]) / dx
    return base, w, dw


def _scene(seed=1, ncells=3, stretch=0.05):
    s = scenes.cube_scene_small(seed=seed, ncells=ncells, E=500.0)
    n = len(s["mass"])
    rng = np.random.default_rng(seed)
    F = np.tile(np.eye(3), (n, 1, 1)) * (1 + stretch) + 0.02 * rng.normal(size=(n, 3, 3))
    s["F"] = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9))
    s["mu"], s["lam"] = scenes.lame(500.0, 0.4)
    s["mu"] = np.full(n, s["mu"])
    s["lam"] = np.full(n, s["lam"])
    return s


def _sim(s, opts, **pf):
    sim = oracle.OracleSim(dx=s["dx"], gravity=(0, -9.8, 0), nthreads=2)
    sim.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    sim.set_phase_field(opts, **pf)
    return sim


def _node_coords(sim):
    sim.sort_and_pollute()
    info = sim.sort_result()
    coords = {}
    for p, base in enumerate(info["pages"]):
        for e in range(32):
            off = int(base) + (e << 7)
            coords[tuple(oracle.linear_to_coord(off))] = p * 32 + e
    return coords


@pytest.mark.parametrize("M", [0.0, 15.0])
def test_operator_matches_dense_assembly(M):
    s = _scene(2)
    n = len(s["mass"])
    rng = np.random.default_rng(0)
    l0 = 0.5 * s["dx"]
    pf_Fp = 1.0 + rng.uniform(0, 3, size=n)
    dt = 1e-3
    # H above any psi+ keeps pf_Fp fixed through the solve's history update
    sim = _sim(s, oracle.pf_opts(parabolic_M=M), l0=l0, pf_Fp=pf_Fp, one_over_sigma_c=1e-3, H=1e300)
    coords = _node_coords(sim)
    sim.update_stress()
    sim.p2g()
    sim.phase_field_solve(dt)
    active = sim.grid()["idx"] >= 0
    ns = len(active)
    A = np.zeros((ns, ns))
    Fm = s["F"].reshape(n, 3, 3).transpose(0, 2, 1)
    gscale = M if M > 0 else 1.0
    for p in range(n):
        vol = s["vol"][p] * np.linalg.det(Fm[p])
        lump = vol * (pf_Fp[p] * M + 1.0 / dt) if M > 0 else vol * pf_Fp[p]
        bw = [_bspline(s["X"][p, d], s["dx"]) for d in range(3)]
        ks, ws, gs = [], [], []
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    k = coords[(bw[0][0] + a, bw[1][0] + b, bw[2][0] + c)]
                    w = bw[0][1][a] * bw[1][1][b] * bw[2][1][c]
                    g = np.array([bw[0][2][a] * bw[1][1][b] * bw[2][1][c], bw[0][1][a] * bw[1][2][b] * bw[2][1][c],
                                  bw[0][1][a] * bw[1][1][b] * bw[2][2][c]])
                    if active[k]:
                        ks.append(k)
                        ws.append(w)
                        gs.append(g)
        G = np.array(gs)
        kap = vol * 4 * l0 * l0 * gscale
        for i, ki in enumerate(ks):
            A[ki, ki] += lump * ws[i]
            for j, kj in enumerate(ks):
                A[ki, kj] += kap * G[i] @ G[j]
    x = rng.normal(size=ns)
    x[~active] = 0.0
    b = sim.phase_field_multiply(x)
    assert np.abs(b - A @ x).max() < 1e-12 * np.abs(A @ x).max()
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
    Aa = A[np.ix_(active, active)]
    assert np.linalg.eigvalsh(Aa).min() > 0
    # constants: A 1 = lumped diagonal
    rhs, diag, xs = sim.phase_field_system()
    one = active.astype(float)
    assert np.abs(sim.phase_field_multiply(one) - diag).max() < 1e-12 * np.abs(diag).max()
    # CG solution satisfies the system (absolute 1e-6 or 1e-12 relative, with slack for the final step)
    res = np.linalg.norm((A @ xs - rhs)[active])
    assert res < max(1e-6, 1e-12 * np.linalg.norm(rhs)) * 1.0001


def test_undamaged_uniform_block_keeps_c():
    s = _scene(3, stretch=0.0)
    n = len(s["mass"])
    s["F"] = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    sim = _sim(s, oracle.pf_opts(parabolic_M=15.0), l0=0.5 * s["dx"], one_over_sigma_c=1e-3)
    for _ in range(3):
        sim.step(1e-4)
    pf = sim.get_phase_field()
    assert pf["c"].min() > 1 - 1e-8
    assert np.all(pf["c"] <= 1.0)


@pytest.mark.parametrize("M", [0.0, 1.0, 15.0])
def test_damage_monotone_and_history(M):
    s = _scene(4, stretch=0.15)
    n = len(s["mass"])
    osc = 1.0 / 2.0 * np.exp(0.3 * np.random.default_rng(4).normal(size=n))
    sim = _sim(s, oracle.pf_opts(parabolic_M=M), l0=0.5 * s["dx"], one_over_sigma_c=osc)
    c_prev, H_prev = np.ones(n), np.zeros(n)
    for _ in range(4):
        sim.step(1e-4)
        pf = sim.get_phase_field()
        assert np.all(pf["c"] <= c_prev + 1e-15) and pf["c"].min() >= 0.0
        assert np.all(pf["H"] >= H_prev)
        assert np.all(pf["pf_Fp"] >= 1.0)
        g = sim.get_materials()["g"]
        assert np.allclose(g, pf["c"] ** 2 * (1 - 0.001) + 0.001, rtol=1e-15, atol=0)
        c_prev, H_prev = pf["c"], pf["H"]
    assert c_prev.min() < 1.0  # it does damage under this load


def test_history_update_closed_form():
    s = _scene(5, stretch=0.1)
    n = len(s["mass"])
    l0 = 0.5 * s["dx"]
    sim = _sim(s, oracle.pf_opts(parabolic_M=1.0), l0=l0, one_over_sigma_c=0.25)
    sim.sort_and_pollute()
    sim.update_stress()
    sim.p2g()
    sim.phase_field_solve(1e-4)
    pf = sim.get_phase_field()
    F = s["F"].reshape(n, 3, 3).transpose(0, 2, 1)
    for p in range(0, n, 11):
        psi = oracle.nhb_psi_pos(s["mu"][p], s["lam"][p], F[p])
        assert pf["H"][p] == pytest.approx(max(psi, 0.0), rel=1e-15)
        assert pf["pf_Fp"][p] == pytest.approx(4 * l0 * (1 - 0.001) * pf["H"][p] * 0.25 + 1, rel=1e-14)


def test_sym3_max_eigenvalue():
    rng = np.random.default_rng(8)
    cases = [np.diag([1.0, 1.0, 1.0]), np.diag([2.0, 2.0, -1.0]), np.zeros((3, 3)), -np.eye(3),
             np.diag([3.0, 1.0, 2.0])]
    for _ in range(20):
        A = rng.normal(size=(3, 3))
        cases.append(A + A.T)
    for A in cases:
        ref = np.linalg.eigvalsh(A).max()
        assert abs(oracle.sym3_max_eigenvalue(A) - ref) < 1e-13 * max(1.0, np.abs(A).max())


def test_naive_rankine_cut():
    s = _scene(6, stretch=0.1)
    n = len(s["mass"])
    sigma_F = 20.0
    sim = _sim(s, oracle.pf_opts(naive=True, sigma_F=sigma_F))
    sim.sort_and_pollute()
    sim.update_stress()
    sim.p2g()
    sim.phase_field_solve(1e-4)
    c = sim.get_phase_field()["c"]
    F = s["F"].reshape(n, 3, 3).transpose(0, 2, 1)
    for p in range(n):
        J = np.linalg.det(F[p])
        cauchy = oracle.nhb_kirchhoff(s["mu"][p], s["lam"][p], 1.0, F[p]) / J
        smax = np.linalg.eigvalsh(0.5 * (cauchy + cauchy.T)).max()
        expect = min(1.0, sigma_F / smax) if smax > sigma_F else 1.0
        assert c[p] == pytest.approx(expect, rel=1e-12)
    assert c.min() < 1.0


def _lattice_bar(nx, nyz, dx, lo=8):
    """8 particles per cell on a regular lattice (cell centre +- dx/4) over
    cells [lo, lo + nx) x [lo, lo + nyz)^2; unit density, identity F."""
    o = (np.arange(2) + 0.5) * 0.5
    cells = np.stack(np.meshgrid(np.arange(nx), np.arange(nyz), np.arange(nyz), indexing="ij"), -1).reshape(-1, 3)
    sub = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)
    X = ((cells[:, None, :] + lo + sub[None, :, :]) * dx).reshape(-1, 3)
    n = len(X)
    vol = np.full(n, dx ** 3 / 8)
    mu, lam = scenes.lame(500.0, 0.4)
    return dict(X=X, V=np.zeros((n, 3)), C=np.zeros((n, 9)), F=np.tile(np.eye(3).reshape(9), (n, 1)),
                mass=vol.copy(), vol=vol, mu=np.full(n, mu), lam=np.full(n, lam), dx=dx)


@pytest.mark.parametrize("M", [0.0, 8.0])
def test_uniform_block_has_the_closed_form_solution(M):
    """With uniform c0 and pf_Fp = f the lumped system (FractureSimulation.h:
    312-437) is solved by a constant, because the gradient term annihilates
    constants: x = (M + c0/dt) / (M f + 1/dt) for M > 0, x = 1/f for M = 0;
    phaseG2P then gives c = min(c0, x) (the increment form for M > 0)."""
    s = _lattice_bar(6, 6, 1.0 / 64)
    dt, c0, f = 1e-3, 0.9, 3.0
    opts = oracle.pf_opts(parabolic_M=M, cg_max_iterations=500, cg_tolerance=1e-300, cg_relative=1e-15)
    sim = _sim(s, opts, c=c0, l0=2 * s["dx"], pf_Fp=f, one_over_sigma_c=1e-3, H=1e300)
    sim.sort_and_pollute()
    sim.update_stress()
    sim.p2g()
    sim.phase_field_solve(dt)
    c = sim.get_phase_field()["c"]
    x = (M + c0 / dt) / (M * f + 1.0 / dt) if M > 0 else 1.0 / f
    assert np.abs(c - min(c0, x)).max() < 1e-10
    assert x < c0  # the solve lowers c


def test_damage_front_decays_over_2_l0():
    """Next to a fully damaged region (pf_Fp = 1e4) an intact one (pf_Fp = 1)
    approaches c = 1 as 1 - c ~ exp(-x / (2 l0)): the 1D form of
    f c - 4 l0^2 c'' = 1 (CD-MPM's phase-field equation, the coefficient
    vol 4 l0^2 of FractureSimulation.h:360).  A coefficient of 2 l0^2 or
    8 l0^2 changes the length by a factor sqrt(2)."""
    dx = 1.0 / 128
    s = _lattice_bar(72, 4, dx)
    l0 = 3 * dx
    xmid = (8 + 24) * dx
    f = np.where(s["X"][:, 0] < xmid, 1e4, 1.0)
    opts = oracle.pf_opts(parabolic_M=0.0, cg_max_iterations=2000, cg_tolerance=1e-300, cg_relative=1e-14)
    sim = _sim(s, opts, l0=l0, pf_Fp=f, one_over_sigma_c=1e-3, H=1e300)
    sim.sort_and_pollute()
    sim.update_stress()
    sim.p2g()
    sim.phase_field_solve(1e-3)
    c = sim.get_phase_field()["c"]
    x = s["X"][:, 0] - xmid
    sel = (x > 3 * dx) & (x < 20 * dx)
    slope = np.polyfit(x[sel], np.log(1.0 - c[sel]), 1)[0]
    assert -1.0 / slope == pytest.approx(2 * l0, rel=0.08)
    assert c[s["X"][:, 0] < xmid - 3 * dx].max() < 1e-3


def test_phase_field_then_rankine_rule():
    """ZR_DAMAGE_PHASE_FIELD_RANKINE: the phase-field solve, then the Rankine
    cut of its result: c = min(c_pf, sigma_F / sigma_max) where sigma_max
    (numpy's eigensolver on the undegraded Cauchy stress) exceeds sigma_F;
    g follows the final c."""
    s = _scene(7, stretch=0.1)
    n = len(s["mass"])
    F = s["F"].reshape(n, 3, 3).transpose(0, 2, 1)
    smax = np.empty(n)
    for p in range(n):
        cauchy = oracle.nhb_kirchhoff(s["mu"][p], s["lam"][p], 1.0, F[p]) / np.linalg.det(F[p])
        smax[p] = np.linalg.eigvalsh(0.5 * (cauchy + cauchy.T)).max()
    sigma_F = float(np.median(smax))  # about half the particles are cut
    pf = dict(l0=0.5 * s["dx"], one_over_sigma_c=np.full(n, 3.0))
    out = {}
    for rule in (0, 2):
        sim = _sim(s, oracle.pf_opts(naive=rule, sigma_F=sigma_F), **pf)
        sim.sort_and_pollute()
        sim.update_stress()
        sim.p2g()
        sim.phase_field_solve(1e-4)
        out[rule] = sim.get_phase_field()["c"]
    cut = 0
    for p in range(n):
        expect = min(out[0][p], sigma_F / smax[p]) if smax[p] > sigma_F else out[0][p]
        cut += expect < out[0][p]
        assert out[2][p] == pytest.approx(expect, rel=1e-12)
    assert 0 < cut < n
    assert out[0].min() < 1.0  # the phase field itself damaged some particles


def test_delete_particles_parks_at_rest():
    """deleteParticles (FractureSimulation.h:162-207) at the start of the
    step: a particle whose c is below the threshold is parked at (4, 4, 4)
    with v = 0 and F = I, its C kept.  Alone there, one MLS step moves it as
    a free particle: the grid velocity is C (x_i - x_p) + g dt, whose
    weighted first moment vanishes and whose weighted second moment gives C
    back (sum_i w_i (x_i - x_p)(x_i - x_p)^T = dx^2 / 4 I), so v = g dt,
    x = (4, 4, 4) + g dt^2, C unchanged and F = I + dt C."""
    s = _scene(8, ncells=3)
    n = len(s["mass"])
    rng = np.random.default_rng(8)
    s["C"] = np.ascontiguousarray(0.3 * rng.normal(size=(n, 9)))
    c0 = np.ones(n)
    gone = [3, 17]
    c0[gone] = 1e-3
    sim = _sim(s, oracle.pf_opts(delete_threshold=0.01), c=c0, l0=0.5 * s["dx"], one_over_sigma_c=np.full(n, 3.0))
    dt = 1e-4
    sim.step(dt)
    X, V, C, F = sim.get_particles()
    g = np.array([0.0, -9.8, 0.0])
    Cg = s["C"][gone].reshape(-1, 3, 3).transpose(0, 2, 1)
    # the parked particles share one position, so their grid sees the
    # mass-weighted mean of their C
    m = s["mass"][gone]
    Cm = (m[:, None, None] * Cg).sum(axis=0) / m.sum()
    for k, p in enumerate(gone):
        assert np.allclose(X[p], 4.0 + g * dt * dt, rtol=0, atol=1e-14)
        assert np.allclose(V[p], g * dt, rtol=0, atol=1e-13)
        assert np.allclose(C[p].reshape(3, 3).T, Cm, rtol=0, atol=1e-12)
        assert np.allclose(F[p].reshape(3, 3).T, np.eye(3) + dt * Cm, rtol=0, atol=1e-14)
    keep = np.setdiff1d(np.arange(n), gone)
    assert np.abs(X[keep] - s["X"][keep]).max() < 1e-3
