"""GPU parity of the adaptive time step: evalMaxParticleSpeed and calculateDt
(MpmSimulationBase.cpp:766-790,1143-1170) bit-exact against the oracle, and
frames of CFL substeps (SimulationBase::advanceOneFrame with
TimeStepping::nextDt, SimulationBase.h:288-321, TimeStepping.h:45-80) through
the Python binding and the C++ host against the oracle's frame loop."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from ziran2019_amd import build as zbuild
from ziran2019_amd import scenes
from ziran2019_amd.engine import MpmEngine, RestartException, TimeStepping, ZrError
from _helpers import perturbed, rel_err

pytestmark = pytest.mark.gpu

FLT_MAX = float(np.finfo(np.float32).max)


def _pair(s, cfl=0.6, colliders=()):
    eng = MpmEngine(dx=s["dx"], cfl=cfl)
    eng.upload(s)
    eng.set_colliders(list(colliders))
    ora = oracle.OracleSim(dx=s["dx"], cfl=cfl, nthreads=8)
    ora.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    for c in colliders:
        ora.add_collider(**c)
    return eng, ora


def _oracle_frame(ora, st):
    dts = []
    while True:
        dt = oracle.next_dt(ora.calculate_dt(st.max_dt), st.min_dt, st.max_dt, st.max_dt_original, st.frame_dt,
                            st.time_since_last_frame)
        assert ora.step(dt) & 1 == 0  # bit 0: moved more than a cell (a restart)
        dts.append(dt)
        if st.advance(dt) == TimeStepping.FRAME:
            return dts


@pytest.mark.parametrize("seed", [1, 2])
def test_max_speed_bit_exact(seed):
    s = perturbed(seed)
    s["V"] = np.random.default_rng(seed).normal(size=s["V"].shape)
    print("max speed, seed", seed, "particles", len(s["mass"]), flush=True)
    eng, ora = _pair(s)
    sp, hi, lo = eng.evalMaxParticleSpeed()
    osp, ohi, olo = ora.eval_max_particle_speed()
    assert sp == osp
    assert np.array_equal(hi, ohi) and np.array_equal(lo, olo)
    eng.close()


def test_max_speed_of_no_particles():
    s = perturbed(1)
    for k in ("X", "V", "C", "F", "mass", "vol", "mu", "lam"):
        s[k] = np.asarray(s[k])[:0]
    eng = MpmEngine(dx=s["dx"])
    sp, hi, lo = eng.evalMaxParticleSpeed()
    assert sp == 0.0 and np.all(hi == -FLT_MAX) and np.all(lo == FLT_MAX)
    assert eng.calculateDt(0.125) == 0.125
    eng.close()


def test_calculate_dt_bit_exact():
    s = perturbed(5)
    s["V"] = 0.3 * np.random.default_rng(5).normal(size=s["V"].shape)
    eng, ora = _pair(s, cfl=0.45)
    assert eng.calculateDt(1.0) == ora.calculate_dt(1.0)
    eng.close()
    # a moving collider faster than every particle sets the step
    fast = [dict(shape=0, type=1, origin=(0.0, -1.0, 0.0), normal=(0.0, 1.0, 0.0), velocity=(0.0, 0.0, 40.0))]
    eng, ora = _pair(s, cfl=0.45, colliders=fast)
    assert eng.calculateDt(1.0) == ora.calculate_dt(1.0) == 0.45 * s["dx"] / 40.0
    eng.close()
    # nothing moves: max_dt
    s["V"][:] = 0.0
    eng, ora = _pair(s)
    assert eng.calculateDt(0.03125) == ora.calculate_dt(0.03125) == 0.03125
    eng.close()


def test_advance_one_frame_matches_oracle():
    s = perturbed(6, shuffle=False)
    s["V"][:, 1] -= 2.0
    y0 = float(s["X"][:, 1].min() - 1.5 * s["dx"])
    ground = [dict(shape=0, type=1, origin=(0.0, y0, 0.0), normal=(0.0, 1.0, 0.0))]
    # cfl 0.1: substeps of 0.1 dx / ~2.3 m/s ~ 7e-4 s at dx = 1/64, several per frame
    eng, ora = _pair(s, cfl=0.1, colliders=ground)
    frame_dt = 2e-3
    st_e, st_o = TimeStepping(frame_dt=frame_dt), TimeStepping(frame_dt=frame_dt)
    for f in range(2):
        print("frame", f, flush=True)
        dts = eng.advanceOneFrame(st_e)
        odts = _oracle_frame(ora, st_o)
        assert len(dts) == len(odts) >= 3
        assert np.allclose(dts, odts, rtol=1e-9, atol=0)
        assert st_e.time == pytest.approx((f + 1) * frame_dt, rel=1e-14)
    X, V, C, F = ora.get_particles()
    out = eng.download()
    assert rel_err(out["X"], X) < 1e-10
    assert rel_err(out["V"], V) < 1e-9
    eng.close()


def test_host_frames_match_oracle(tmp_path):
    from test_gpu_host import _read, _write
    exe = zbuild.HOST_EXE
    if not os.path.exists(exe):
        raise ZrError(f"{exe} is missing: build it with `python -m ziran2019_amd.build`")
    s = perturbed(8, shuffle=False)
    s["V"][:, 1] -= 2.0
    y0 = float(s["X"][:, 1].min() - 1.5 * s["dx"])
    frame_dt = 2e-3
    src, dst = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write(src, s)
    cmd = [exe, src, dst, repr(s["dx"]), "2", repr(-frame_dt), "1", repr(y0), "0", "0.1"]
    print("host frames:", " ".join(cmd[3:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("frame")]
    assert len(lines) == 2
    ora = oracle.OracleSim(dx=s["dx"], cfl=0.1, nthreads=8)
    ora.set_particles(s["X"], s["V"], s["C"], s["F"], s["mass"], s["vol"], s["mu"], s["lam"])
    ora.add_collider(shape=0, type=1, origin=(0.0, y0, 0.0), normal=(0.0, 1.0, 0.0))
    st = TimeStepping(frame_dt=frame_dt)
    for f in range(2):
        odts = _oracle_frame(ora, st)
        assert int(lines[f][3]) == len(odts)
        assert float(lines[f][5]) == pytest.approx(st.time, rel=1e-14)
    out = _read(dst)
    X, V, C, F = ora.get_particles()
    assert rel_err(out["X"], X) < 1e-10
    assert rel_err(out["V"], V) < 1e-9


def test_rollout_stops_after_the_restart_step():
    """zr_rollout_forward reads each step's CFL flags with the next step's
    sort: a rollout stops after the step that asked for a restart, with the
    state that stepping one call at a time leaves (bit for bit), and a
    rollout that ends before that step runs to its end."""
    s = scenes.cube_scene_small(seed=5, ncells=4, v0=(0.0, 0.0, 0.0), spin=0.0)
    dx, dt, g = s["dx"], 1e-3, 9.8
    # a uniform fall whose step 5 (the speed dx/dt + 0.5 g dt) moves more than a cell
    s["V"][:, 1] = -(dx / dt - 4.5 * g * dt)

    def engine():
        e = MpmEngine(dx=dx, gravity=(0.0, -g, 0.0))
        e.upload(s)
        e.set_colliders([])
        return e

    a = engine()
    states, restart = [], None
    for k in range(1, 10):
        try:
            a.advanceOneTimeStep(dt)
        except RestartException:
            restart = k
            break
        states.append(a.download())
    print("one call per step: restart at step", restart, flush=True)
    assert restart == 5
    last = a.download()

    b = engine()
    with pytest.raises(RestartException):
        b.rollout_forward(9, dt, 3)
    got = b.download()
    for key in ("X", "V", "C", "F"):
        assert np.array_equal(got[key], last[key]), key

    c = engine()
    c.rollout_forward(4, dt, 3)
    got = c.download()
    for key in ("X", "V", "C", "F"):
        assert np.array_equal(got[key], states[3][key]), key
    for e in (a, b, c):
        e.close()
