"""Seeded synthetic particle scenes for the MPM time step.

The reference samples with Poisson disks inside level sets
(Lib/MPM/MpmInitializationHelper.h:187-202) using std::mt19937, whose
uniform_real_distribution is implementation-defined.  The workloads here use a
counter-based splitmix64 stream instead, so the same seed gives the same bits
on every machine.  Particle attributes follow what the reference creates per
particle: X, V, C (APIC affine matrix), F, mass, element measure V0, and a
per-particle constitutive model (mu, lambda) -- MpmParticleHandleBase.cpp:457,
FBasedMpmForceHelper.h:17-164.

Particles are emitted in the engine's sort order (page Morton order, then
cell, then index) so that the storage order of the first step is coherent.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(counter, seed):
    """k-th output of the splitmix64 generator seeded with ``seed``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(counter, dtype=np.uint64) + np.uint64(1)) * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(counter, seed):
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (splitmix64(counter, seed) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def lame(E, nu):
    """Lame parameters, NeoHookeanBorden::setLameParameters
    (Lib/Ziran/Physics/ConstitutiveModel/NeoHookeanBorden.cpp:16-21)."""
    E = np.asarray(E, dtype=np.float64)
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = E / (2.0 * (1.0 + nu))
    return mu, lam


# ---- SPGrid page/cell code of a node (GridState<double,3> masks) ---------
def _spread3(v):
    v = np.asarray(v, dtype=np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x030000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def cell_code(bx, by, bz):
    """Linear_Offset(b) >> 7 for base nodes with 0 <= b < 1024
    (SPGrid_Mask.h:150-174 with the GridState<double,3> masks)."""
    bx = np.asarray(bx, dtype=np.int64).astype(np.uint64)
    by = np.asarray(by, dtype=np.int64).astype(np.uint64)
    bz = np.asarray(bz, dtype=np.int64).astype(np.uint64)
    page = _spread3(bx >> np.uint64(1)) | (_spread3(bz >> np.uint64(2)) << np.uint64(1)) | \
        (_spread3(by >> np.uint64(2)) << np.uint64(2))
    e = ((bx & np.uint64(1)) << np.uint64(4)) | ((by & np.uint64(3)) << np.uint64(2)) | (bz & np.uint64(3))
    return (page << np.uint64(5)) | e


def base_nodes(X, dx):
    """int_floor(X * (1/dx) - 0.5) per axis (MpmSimulationBase.cpp:1046-1049)."""
    y = np.asarray(X, dtype=np.float64) * (1.0 / dx) - 0.5
    t = np.trunc(y)
    return (t - (t > y)).astype(np.int64)


def morton_order(X, dx):
    b = base_nodes(X, dx)
    code = cell_code(b[:, 0], b[:, 1], b[:, 2])
    return np.lexsort((np.arange(len(code)), code))


def _particles_in_cells(cells, ppc, dx, seed, stream):
    """``ppc`` particles per cell, each uniform in its cell."""
    ncell = len(cells)
    k = np.arange(ncell * ppc * 3, dtype=np.uint64) + np.uint64(stream) * np.uint64(1 << 40)
    u = uniform01(k, seed).reshape(ncell, ppc, 3)
    X = (cells[:, None, :].astype(np.float64) + u) * dx
    return X.reshape(-1, 3)


def _finish(X, dx, rho, ppc, mu, lam, order=True):
    n = len(X)
    if order:
        perm = morton_order(X, dx)
        X = X[perm]
        mu = np.broadcast_to(mu, (n,))[perm] if np.ndim(mu) else mu
        lam = np.broadcast_to(lam, (n,))[perm] if np.ndim(lam) else lam
    vol = dx ** 3 / ppc
    F = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    return dict(
        X=np.ascontiguousarray(X),
        V=np.zeros((n, 3)),
        C=np.zeros((n, 9)),
        F=F,
        mass=np.full(n, rho * vol),
        vol=np.full(n, vol),
        mu=np.ascontiguousarray(np.broadcast_to(np.asarray(mu, dtype=np.float64), (n,))),
        lam=np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (n,))),
    )


def cube_scene(dx=1.0 / 128, lo_cell=51, ncells=26, ppc=8, E=1e4, nu=0.3, rho=1000.0, seed=20190701):
    """Plain explicit MLS-MPM workload: a cube of ``ncells``^3 cells starting at
    cell index ``lo_cell`` (26^3 cells at dx=1/128 from cell 51: the side-0.2
    cube centred at 0.5), ``ppc`` particles per cell jittered uniformly in-cell,
    fixed-corotated material.  Default size: 26^3 * 8 = 140,608 particles."""
    r = np.arange(lo_cell, lo_cell + ncells)
    cells = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    X = _particles_in_cells(cells, ppc, dx, seed, 0)
    mu, lam = lame(E, nu)
    out = _finish(X, dx, rho, ppc, mu, lam)
    out.update(dx=dx, gravity=(0.0, -9.8, 0.0),
               colliders=[dict(shape=0, type=1, origin=(0.0, 0.1, 0.0), normal=(0.0, 1.0, 0.0))])
    return out


def slabs_scene(dx=1.0 / 256, cells=(64, 40, 50), gap=8, ppc=8, E=1e4, nu=0.3, rho=1000.0, yield_stress=200.0,
                friction_angle=30.0, seed=20190703):
    """Elastoplastic ADMM workload: two slabs of ``cells`` cells side by side
    in x (64 x 40 x 50 = 128,000 cells each, 2,048,000 particles at 8 ppc
    and dx = 1/256), ``gap`` cells apart, on a sticky ground, StvkWithHencky;
    the first slab von Mises (yield stress ``yield_stress`` Pa, no hardening),
    the second Drucker-Prager sand (``friction_angle`` degrees, beta 1, no
    cohesion, volume correction).  Materials come as the ``materials`` entry,
    the keyword arguments of MpmEngine.set_materials."""
    cells = np.asarray(cells)
    ground = 0.1
    lo0 = np.array([int(round(0.5 / dx)) - cells[0] - gap // 2, int(np.ceil(ground / dx)),
                    int(round(0.5 / dx)) - cells[2] // 2])
    los = [lo0, lo0 + np.array([cells[0] + gap, 0, 0])]
    Xs = []
    for k, lo in enumerate(los):
        r = [np.arange(lo[d], lo[d] + cells[d]) for d in range(3)]
        cidx = np.stack(np.meshgrid(*r, indexing="ij"), axis=-1).reshape(-1, 3)
        Xs.append(_particles_in_cells(cidx, ppc, dx, seed, k))
    X = np.concatenate(Xs)
    slab = np.concatenate([np.zeros(len(Xs[0]), np.int32), np.ones(len(Xs[1]), np.int32)])
    perm = morton_order(X, dx)
    X, slab = X[perm], slab[perm]
    mu, lam = lame(E, nu)
    out = _finish(X, dx, rho, ppc, mu, lam, order=False)
    n = len(X)
    s_phi = np.sin(friction_angle / 180.0 * 3.141592653)
    alpha = np.sqrt(2.0 / 3.0) * 2.0 * s_phi / (3.0 - s_phi)  # PlasticityApplier.cpp:249-260
    pp = np.zeros((n, 4))
    pp[slab == 0] = (0.0, 1e30, 0.0, 0.0)
    pp[slab == 1] = (alpha, 1.0, 0.0, 1.0)
    out.update(dx=dx, gravity=(0.0, -9.8, 0.0), slab=slab,
               materials=dict(model=1, plasticity=np.where(slab == 0, 1, 2).astype(np.int32), plastic_param=pp,
                              plastic_state=np.where(slab == 0, yield_stress, 0.0)),
               colliders=[dict(shape=0, type=1, origin=(0.0, ground, 0.0), normal=(0.0, 1.0, 0.0))])
    return out


def adjoint_scene(dx=1.0 / 128, ncells=32, lo_cell=48, ppc=8, logE_mu=np.log(1e4), logE_sigma=0.4, nu=0.3,
                  rho=1000.0, seed=20190705):
    """Differentiable explicit workload: a cube of ``ncells``^3 cells (32^3 x 8
    = 262,144 particles at dx = 1/128) standing on a sticky ground (half a
    cell into its bottom layer, so that the ground's reaction and with it the
    centre of mass depend on the moduli) with per-particle Young's modulus
    lognormal(``logE_mu``, ``logE_sigma``) from the seed; the ``E`` entry holds
    the moduli (mu / lam are set from them with ``nu``)."""
    r = np.arange(lo_cell, lo_cell + ncells)
    cells = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    X = _particles_in_cells(cells, ppc, dx, seed, 0)
    perm = morton_order(X, dx)
    X = X[perm]
    n = len(X)
    E = np.exp(logE_mu + logE_sigma * _normal01(np.arange(n, dtype=np.uint64) + np.uint64(1 << 46), seed))
    mu, lam = lame(E, nu)
    out = _finish(X, dx, rho, ppc, mu, lam, order=False)
    out.update(dx=dx, gravity=(0.0, -9.8, 0.0), E=E,
               colliders=[dict(shape=0, type=1, origin=(0.0, (lo_cell + 0.5) * dx, 0.0), normal=(0.0, 1.0, 0.0))])
    return out


def cube_scene_small(seed=7, ncells=6, lo_cell=20, dx=1.0 / 64, E=1e4, v0=(0.3, -0.2, 0.1), spin=2.0):
    """A small cube with a nonzero initial velocity field (translation plus a
    rigid spin) used by parity tests."""
    s = cube_scene(dx=dx, lo_cell=lo_cell, ncells=ncells, ppc=8, E=E, seed=seed)
    X = s["X"]
    c = X.mean(axis=0)
    r = X - c
    w = np.array([0.0, spin, 0.5 * spin])
    s["V"] = np.asarray(v0)[None, :] + np.cross(w[None, :], r)
    return s


def _normal01(counter, seed):
    """Standard normals by Box-Muller on the splitmix64 stream."""
    c = np.asarray(counter, dtype=np.uint64)
    u1 = uniform01(c * np.uint64(2), seed)
    u2 = uniform01(c * np.uint64(2) + np.uint64(1), seed)
    u1 = np.maximum(u1, 2.0 ** -53)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def boxes_scene(dx=1.0 / 256, nboxes=64, side_lo=0.025, side_hi=0.075, ppc=8, logE_mu=np.log(5e4),
                logE_sigma=0.5, nu=0.3, rho=1000.0, seed=20190702, v_scale=0.1, grad_scale=2.0, gap=8):
    """Implicit workload: ``nboxes`` seeded boxes with sides U[side_lo,
    side_hi] (cell-aligned) at least ``gap`` cells apart, Young's modulus
    lognormal(logE_mu, logE_sigma) per box, ``ppc`` particles per cell.  Each
    box starts with a seeded affine velocity v = v0 + L (x - c) (APIC C = L)
    so that every implicit step has a non-trivial elastic solve; v0 and the
    gaps keep the boxes apart and every particle well inside the one-cell
    CFL bound for tens of steps at dt = 2e-3.  Defaults give about 1M
    particles at dx = 1/256 with 8 ppc."""
    rng_ctr = 0
    boxes = []
    lo_cell = int(np.ceil(0.12 / dx))
    hi_cell = int(np.floor(0.88 / dx))
    attempt = 0
    while len(boxes) < nboxes:
        u = uniform01(np.arange(rng_ctr, rng_ctr + 6, dtype=np.uint64), seed)
        rng_ctr += 6
        attempt += 1
        if attempt > 100000:
            raise RuntimeError("could not place the boxes")
        side = np.round((side_lo + (side_hi - side_lo) * u[:3]) / dx).astype(np.int64)
        side = np.maximum(side, 2)
        lo = (lo_cell + np.floor(u[3:] * (hi_cell - lo_cell - side))).astype(np.int64)
        hi = lo + side
        if any(np.all(lo < b[1] + gap) and np.all(b[0] < hi + gap) for b in boxes):
            continue
        boxes.append((lo, hi))
    Xs, mus, lams, Vs, Cs = [], [], [], [], []
    for b, (lo, hi) in enumerate(boxes):
        r = [np.arange(lo[d], hi[d]) for d in range(3)]
        cells = np.stack(np.meshgrid(*r, indexing="ij"), axis=-1).reshape(-1, 3)
        X = _particles_in_cells(cells, ppc, dx, seed, 1 + b)
        z = _normal01(np.arange(1 + 16 * b, 1 + 16 * b + 13, dtype=np.uint64), seed + 1)
        E = np.exp(logE_mu + logE_sigma * z[0])
        mu, lam = lame(E, nu)
        v0 = v_scale * z[1:4]
        L = grad_scale * z[4:13].reshape(3, 3)
        c = (lo + hi) * 0.5 * dx
        n = len(X)
        Xs.append(X)
        mus.append(np.full(n, mu))
        lams.append(np.full(n, lam))
        Vs.append(v0[None, :] + (X - c[None, :]) @ L.T)
        Cs.append(np.tile(L.T.reshape(1, 9), (n, 1)))  # column-major of L
    X = np.concatenate(Xs)
    perm = morton_order(X, dx)
    out = _finish(X[perm], dx, rho, ppc, np.concatenate(mus)[perm], np.concatenate(lams)[perm], order=False)
    out["V"] = np.ascontiguousarray(np.concatenate(Vs)[perm])
    out["C"] = np.ascontiguousarray(np.concatenate(Cs)[perm])
    out.update(dx=dx, gravity=(0.0, -9.8, 0.0), boxes=boxes,
               colliders=[dict(shape=0, type=1, origin=(0.0, 0.1, 0.0), normal=(0.0, 1.0, 0.0))])
    return out


def fracture_scene(dx=1.0 / 320, cells=(100, 75, 50), ppc=8, E=500.0, nu=0.4, rho=1000.0, nsites=200,
                   toughness_sigma=0.3, percentage=0.05, sphere_radius=0.05, sphere_speed=5.0, seed=20190704):
    """CD-MPM fracture workload: a NeoHookeanBorden block of ``cells`` cells
    (100 x 75 x 50 = 375,000 cells, 3,000,000 particles at 8 ppc and
    dx = 1/320) resting on a sticky ground, struck from above by a sticky
    sphere moving down at ``sphere_speed``.  The block is split into
    ``nsites`` seeded Voronoi cells; each cell scales the fracture toughness
    by a lognormal(0, ``toughness_sigma``) factor (one_over_sigma_c divided by
    it).  Base toughness: PhaseField::Get_Sigma_C_Based_On_Max_Deformation at
    ``percentage`` (PhaseField.h:59-68); l0 = dx/2, residual phase 1e-3
    (MpmParticleHandleBase.cpp:105-114)."""
    cells = np.asarray(cells)
    ground = 0.1
    lo = np.array([int(round(0.5 / dx)) - cells[0] // 2, int(np.ceil(ground / dx)), int(round(0.5 / dx)) - cells[2] // 2])
    r = [np.arange(lo[d], lo[d] + cells[d]) for d in range(3)]
    cell_idx = np.stack(np.meshgrid(*r, indexing="ij"), axis=-1).reshape(-1, 3)
    X = _particles_in_cells(cell_idx, ppc, dx, seed, 0)
    mu, lam = lame(E, nu)
    perm = morton_order(X, dx)
    X = X[perm]
    n = len(X)
    out = _finish(X, dx, rho, ppc, mu, lam, order=False)
    # Voronoi sites and their toughness factors
    box_lo, box_hi = lo * dx, (lo + cells) * dx
    us = uniform01(np.arange(3 * nsites, dtype=np.uint64) + np.uint64(1 << 44), seed).reshape(nsites, 3)
    sites = box_lo + us * (box_hi - box_lo)
    factor = np.exp(toughness_sigma * _normal01(np.arange(nsites, dtype=np.uint64) + np.uint64(1 << 45), seed))
    owner = np.empty(n, dtype=np.int64)
    chunk = 1 << 16
    for s0 in range(0, n, chunk):
        d2 = ((X[s0:s0 + chunk, None, :] - sites[None, :, :]) ** 2).sum(axis=2)
        owner[s0:s0 + chunk] = d2.argmin(axis=1)
    J = (1.0 + percentage) ** 3
    kappa = 0.666666 * mu + lam
    base = 1.0 / (kappa / 2.0 * ((J * J - 1.0) / 2.0 - np.log(J)))
    top = box_hi[1]
    centre = (0.5 * (box_lo[0] + box_hi[0]), top + sphere_radius + 2 * dx, 0.5 * (box_lo[2] + box_hi[2]))
    out.update(dx=dx, gravity=(0.0, -9.8, 0.0), owner=owner, sites=sites,
               one_over_sigma_c=base / factor[owner], l0=0.5 * dx, residual_phase=1e-3,
               colliders=[dict(shape=0, type=1, origin=(0.0, ground, 0.0), normal=(0.0, 1.0, 0.0)),
                          dict(shape=1, type=1, origin=centre, radius=sphere_radius,
                               velocity=(0.0, -sphere_speed, 0.0))])
    return out
