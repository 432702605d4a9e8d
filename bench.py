"""Benchmark of the MPM time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload implicit-boxes|explicit-cube|admm-slabs|
                                                                fracture-sphere|adjoint-cube]

One step = MpmSimulationBase::advanceOneTimeStep on the whole particle set,
inputs resident in HBM:
  implicit-boxes (default; the metric's workload): 64 seeded boxes, ~0.98M
    particles at dx = 1/256, dt = 2e-3, backward Euler with Newton (tol 1e-5,
    <= 3 iterations) and mass-preconditioned CG (relative tolerance 1e-6)
  explicit-cube: 140,608 particles, dx = 1/128, dt = 1e-4, explicit MLS
Prints one JSON line with particle-steps/s, the roofline position of the
dominant kernel (HIP-event kernel durations over the timed region) and a CPU
baseline (the FP64 C oracle on the host cores, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s mixed-FP32 vs FP64 at fixed error; P2G/G2P HBM GB/s"
HBM_PEAK_GBS = 8000.0
F64 = 8
F32 = 4


def algorithmic_bytes(kernel, n, g):
    """Required HBM bytes per launch, each stream counted once (n particles,
    g active grid nodes with m != 0, T the value size, 4 B sorted index):

      P2G (MLS, stress fused from F): n (26T + 4) + g 7T
            reads x3 v3 C9 F9 m1 V0 1, writes m, mv3, f3 per node
      P2G (APIC, implicit path, no stress): n (16T + 4) + g 4T
            reads x3 v3 C9 m1, writes m, mv3 per node
      G2P (+F update): n (36T + 4) + g 3T
            reads x3 F9 and v*3 per node, writes x3 v3 C9 F9
      operator apply (fixed-corotated): n (34T + 4) + g 7T
            per particle x3 Fn9 U9 Sigma3 V9 V0 1, per node x3 in, m, b3 out
      Newton state: n (15 FP64 + 4 + 26T + 4) + g 6 FP64
            reads x3 Fn9 mu lambda V0 and (vn + dv) per node, writes the
            operator data (26 values + cell)
    """
    if kernel == "p2g":
        return n * (26 * F64 + 4) + g * 7 * F64
    if kernel == "p2g_apic":
        return n * (16 * F64 + 4) + g * 4 * F64
    if kernel == "g2p":
        return n * (36 * F64 + 4) + g * 3 * F64
    if kernel == "hvp":
        return n * (34 * F32 + 4) + g * 7 * F32
    if kernel == "hvp64":
        return n * (34 * F64 + 4) + g * 7 * F64
    if kernel == "state":
        return n * (15 * F64 + 4 + 26 * F32 + 4) + g * 6 * F64
    raise ValueError(kernel)


# HBM bytes per launch measured with rocprofv3 FETCH_SIZE (x 2.0, from a
# 256 MiB copy) and WRITE_SIZE (x 1.0) on one MI355X for these workloads
# (profiles/hbm_traffic_counters.md); None where no counter run exists
TRAFFIC = {
    ("implicit-boxes", "hvp"): 139.0e6,
    ("implicit-boxes", "hvp64"): 275.4e6,
    ("explicit-cube", "p2g"): 40.6e6,
    ("explicit-cube", "g2p"): 49.5e6,
}

IMPLICIT_OPTS = dict(newton_max_iterations=3, newton_tolerance=1e-5, cg_max_iterations=500, cg_tolerance=1e30,
                     cg_relative_tolerance=1e-6)
# the reference's default lumped mass (the diagonal of the operator's mass
# part) or JacobiPreconditioner on the whole operator diagonal; the metric
# is quoted with the first, the second is measured beside it
PRECONDITIONERS = dict(jacobi=1, mass=0)
# what holds each kernel below its HBM roofline, from its rocprofv3 SQ counters
# (profiles/hvp_issue_counters.md)
LIMITER = dict(hvp="latency and VALU/LDS issue (waves wait 64% of cycles, LDS bank conflicts 0.60 of LDS-active cycles; HBM traffic 0.96x the algorithmic bytes)")
# steps after the timed region whose kernel launches are timed with HIP events
SAMPLE_STEPS = 3


def host_cpu():
    """(CPU model from lscpu, online cores, cores this process may use): the
    threads a baseline uses are the scheduler's share, capped by
    OMP_NUM_THREADS where the host sets it."""
    import subprocess
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    online = os.cpu_count() or 1
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else online
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return model, online, share


def _box_subset(scene, idx):
    """The particles of boxes ``idx`` of the implicit-boxes scene."""
    import numpy as np
    cells = np.floor(scene["X"] / scene["dx"]).astype(np.int64)
    keep = np.zeros(len(cells), dtype=bool)
    for i in idx:
        lo, hi = scene["boxes"][i]
        keep |= np.all((cells >= lo) & (cells < hi), axis=1)
    n = len(cells)
    return {k: (np.ascontiguousarray(v[keep]) if isinstance(v, np.ndarray) and len(v) == n else v)
            for k, v in scene.items()}


def _median_implicit_steps(sub, colliders, dt, threads, opts, timed):
    """One untimed implicit oracle step, then ``timed`` steps: (median
    seconds per step, Newton and CG iterations of each timed step)."""
    import statistics

    import oracle
    sim = oracle.OracleSim(dx=sub["dx"], gravity=sub["gravity"], nthreads=threads)
    sim.set_particles(sub["X"], sub["V"], sub["C"], sub["F"], sub["mass"], sub["vol"], sub["mu"], sub["lam"])
    for c in colliders:
        sim.add_collider(**c)
    sim.step_implicit(dt, oracle.implicit_opts(**opts))  # first step excluded
    times, its = [], []
    for _ in range(timed):
        t0 = time.perf_counter()
        _, st = sim.step_implicit(dt, oracle.implicit_opts(**opts))
        times.append(time.perf_counter() - t0)
        its.append((st.newton_iterations, st.cg_iterations))
    return statistics.median(times), its


def cpu_baseline_implicit(scene, dt, opts, multi_boxes=(0, 1, 2, 3), single_boxes=(3,), timed=5):
    """The FP64 C oracle (the reference's loop structure, OpenMP) on the
    host: the median of ``timed`` implicit steps after one untimed step, on
    all the cores of this process's share (the particles of the first four
    boxes) and on one core (the particles of box 3, 3584 particles); the
    same solver settings as the GPU run.  The per-step work varies with the
    Newton and CG counts, which are reported beside the times."""
    model, online, share = host_cpu()
    multi = _box_subset(scene, multi_boxes)
    single = _box_subset(scene, single_boxes)
    tm, its_m = _median_implicit_steps(multi, scene["colliders"], dt, share, opts, timed)
    ts, its_s = _median_implicit_steps(single, scene["colliders"], dt, 1, opts, timed)
    nm, ns = len(multi["mass"]), len(single["mass"])
    return dict(value=nm / tm, unit="particle-steps/s", cores=share, kind="port",
                sample=f"FP64 C oracle (oracle/zro_mpm.c, OpenMP, {share} threads) implicit steps on the {nm} "
                       f"particles of boxes {list(multi_boxes)} of the 64, median of {timed} steps after one "
                       f"untimed step ({tm:.2f} s per step; Newton/CG per step {its_m}), same solver settings",
                one_core=dict(value=ns / ts, cores=1,
                              sample=f"the {ns} particles of box {list(single_boxes)}, median of {timed} steps "
                                     f"after one untimed step ({ts:.2f} s per step; Newton/CG per step {its_s})"),
                cpu_model=model, online_cores=online)


def cpu_baseline(scene, dt, seconds):
    """The FP64 C oracle on the same explicit workload: one untimed step,
    then 5 runs of equal step counts sized so that all runs take about
    ``seconds``; the median run's rate, on the process's core share and on
    one core."""
    import statistics

    import oracle
    model, online, share = host_cpu()
    n = len(scene["mass"])
    out = {}
    for threads in (share, 1):
        sim = oracle.OracleSim(dx=scene["dx"], gravity=scene["gravity"], nthreads=threads)
        sim.set_particles(scene["X"], scene["V"], scene["C"], scene["F"], scene["mass"], scene["vol"], scene["mu"],
                          scene["lam"])
        for c in scene["colliders"]:
            sim.add_collider(**c)
        t0 = time.perf_counter()
        sim.step(dt)  # first step excluded
        first = time.perf_counter() - t0
        k = max(1, int(seconds / 2 / 5 / max(first, 1e-3)))
        runs = []
        for _ in range(5):
            t0 = time.perf_counter()
            for _ in range(k):
                sim.step(dt)
            runs.append(time.perf_counter() - t0)
        out[threads] = (n * k / statistics.median(runs), k)
    (v_all, k_all), (v_one, k_one) = out[share], out[1]
    return dict(value=v_all, unit="particle-steps/s", cores=share, kind="port",
                sample=f"FP64 C oracle (oracle/zro_mpm.c, OpenMP) on the same {n}-particle workload, median of 5 "
                       f"runs of {k_all} steps after one untimed step",
                one_core=dict(value=v_one, cores=1, sample=f"median of 5 runs of {k_one} steps, one thread"),
                cpu_model=model, online_cores=online)


def run_explicit(args):
    import numpy as np
    from ziran2019_amd import scenes
    from ziran2019_amd.engine import MpmEngine

    scene = scenes.cube_scene()
    dt = 1e-4
    n = len(scene["mass"])
    eng = MpmEngine(dx=scene["dx"], gravity=scene["gravity"])
    eng.upload(scene)
    eng.set_colliders(scene["colliders"])
    for _ in range(args.warmup):
        eng.advanceOneTimeStep(dt)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.advanceOneTimeStep(dt)
    t1 = time.perf_counter()
    # per-kernel durations from a separate pass with event markers
    eng.profile(True)
    for _ in range(50):
        eng.advanceOneTimeStep(dt)
    kt = {k: eng.kernel_time(k) for k in ("p2g", "grid", "g2p", "sort")}
    eng.profile(False)
    eng.sortParticlesAndPolluteGrid()
    eng.particlesToGrid()
    g_nodes = int(np.count_nonzero(eng.dump("grid_m")))
    eng.close()
    workload = ("explicit MLS-MPM, fixed-corotated cube 26^3 cells x 8 ppc, dx=1/128, dt=1e-4, sticky ground, "
                "FP64 plain path")
    base = None if args.no_cpu_baseline else cpu_baseline(scene, dt, args.cpu_seconds)
    return dict(n=n, dt=dt, dx=scene["dx"], elapsed=t1 - t0, kt=kt, g_nodes=g_nodes, roof_kernels=("p2g", "g2p"),
                bw_kernels=dict(p2g="p2g", g2p="g2p"), workload=workload, base=base, extra={}, dtype="f64")


def _implicit_run(scene, opts, steps, warmup, sample_steps):
    """warmup steps, the timed steps (no event markers), then
    ``sample_steps`` more steps with a HIP event pair around every kernel
    launch for the per-kernel durations (a separate pass, so that the
    markers stay out of the timed region)."""
    from ziran2019_amd.engine import MpmEngine
    eng = MpmEngine(dx=scene["dx"], gravity=scene["gravity"])
    eng.upload(scene)
    eng.set_colliders(scene["colliders"])
    for _ in range(warmup):
        eng.advanceOneTimeStepImplicit(2e-3, opts)
    newton = cg = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        st = eng.advanceOneTimeStepImplicit(2e-3, opts)
        newton += st.newton_iterations
        cg += st.cg_iterations
    t1 = time.perf_counter()
    kt = None
    if sample_steps:
        eng.profile(True, every=1)
        for _ in range(sample_steps):
            eng.advanceOneTimeStepImplicit(2e-3, opts)
        kt = {k: eng.kernel_time(k) for k in ("p2g", "grid", "g2p", "sort", "state", "hvp")}
        eng.profile(False)
    final = eng.download()
    return eng, dict(elapsed=t1 - t0, kt=kt, newton=newton / steps, cg=cg / steps, final=final)


def run_implicit(args):
    import numpy as np
    from ziran2019_amd import scenes
    from ziran2019_amd.engine import implicit_opts

    scene = scenes.boxes_scene()
    dt = 2e-3
    n = len(scene["mass"])
    # plain FP64 path, then the FP32-bulk / FP64-refined form from the same state
    opts = dict(IMPLICIT_OPTS, preconditioner=PRECONDITIONERS[args.preconditioner])
    eng64, r64 = _implicit_run(scene, implicit_opts(**opts), args.steps, args.warmup, SAMPLE_STEPS)
    eng64.close()
    eng, r32 = _implicit_run(scene, implicit_opts(precision=1, **opts), args.steps, args.warmup, SAMPLE_STEPS)
    eng.sortParticlesAndPolluteGrid()
    eng.particlesToGridApic()
    g_nodes = int(np.count_nonzero(eng.dump("grid_m")))
    eng.close()
    X64, X32 = r64["final"]["X"], r32["final"]["X"]
    disp = np.abs(X64 - scene["X"]).max()
    err = dict(max_abs_position_diff=float(np.abs(X32 - X64).max()),
               rel_to_displacement=float(np.abs(X32 - X64).max() / disp),
               rel_to_extent=float(np.abs(X32 - X64).max() / np.abs(X64).max()),
               velocity_rel=float(np.abs(r32["final"]["V"] - r64["final"]["V"]).max() /
                                  np.abs(r64["final"]["V"]).max()),
               steps=args.steps + args.warmup)
    workload = (f"implicit backward Euler, 64 seeded fixed-corotated boxes (E lognormal(ln 5e4, 0.5)), {n} particles "
                f"at 8 ppc, dx=1/256, dt=2e-3, Newton tol 1e-5 (<=3 its) + {args.preconditioner}-preconditioned CG "
                "rel tol 1e-6; "
                "FP32 operator and Krylov vectors inside the FP64 Newton residual")
    base = None if args.no_cpu_baseline else cpu_baseline_implicit(scene, dt, opts)
    fp64_value = n * args.steps / r64["elapsed"]
    extra = dict(newton_iterations_per_step=r32["newton"], cg_iterations_per_step=r32["cg"],
                 fp64_plain=dict(value=fp64_value, ms_per_step=r64["elapsed"] / args.steps * 1e3,
                                 newton_iterations_per_step=r64["newton"], cg_iterations_per_step=r64["cg"],
                                 kernels={k: dict(avg_us=round(v[0] * 1e3 / max(v[1], 1), 2), launches=v[1])
                                          for k, v in r64["kt"].items()}),
                 speedup_vs_fp64=(n * args.steps / r32["elapsed"]) / fp64_value,
                 fp32_vs_fp64_error=err)
    if args.preconditioner == "mass":
        # the same steps with JacobiPreconditioner on the operator diagonal
        jopts = dict(IMPLICIT_OPTS, preconditioner=PRECONDITIONERS["jacobi"])
        e, j64 = _implicit_run(scene, implicit_opts(**jopts), args.steps, args.warmup, 0)
        e.close()
        e, j32 = _implicit_run(scene, implicit_opts(precision=1, **jopts), args.steps, args.warmup, 0)
        e.close()
        extra["jacobi_operator_diagonal"] = dict(
            value=n * args.steps / j32["elapsed"], ms_per_step=j32["elapsed"] / args.steps * 1e3,
            cg_iterations_per_step=j32["cg"], newton_iterations_per_step=j32["newton"],
            fp64_value=n * args.steps / j64["elapsed"], fp64_cg_iterations_per_step=j64["cg"],
            speedup_vs_fp64=j64["elapsed"] / j32["elapsed"])
    return dict(n=n, dt=dt, dx=scene["dx"], elapsed=r32["elapsed"], kt=r32["kt"], g_nodes=g_nodes,
                roof_kernels=("hvp", "state"),
                bw_kernels=dict(hvp="hvp", state="state", p2g="p2g_apic", g2p="g2p"),
                workload=workload, base=base, extra=extra, dtype="f32-bulk/f64-refined")


def copy_calibration(nbytes=256 << 20, launches=5):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of
    its dispatches against ``nbytes`` give the counters' scale factor."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_bytes_each_way=nbytes, launches=launches)))


def run_row_workload(args):
    """The other rows at their configuration sizes (one JSON line each, not
    the metric): admm-slabs (ADMM step, 2.05M particles in a von Mises and a
    Drucker-Prager slab, 20 ADMM iterations, dt 1e-3), fracture-sphere (the
    explicit step with the CD-MPM phase-field solve, 3M NeoHookeanBorden
    particles struck by a sphere at 5 m/s, dt 5e-4) and adjoint-cube (a
    64-step explicit rollout of 262,144 particles at dt 2e-4 with per-particle
    E, forward with a checkpoint every 8 steps plus the adjoint back to E and
    V0; one "step" here is one rollout)."""
    import numpy as np
    from ziran2019_amd import scenes
    from ziran2019_amd.engine import ZR_MODEL_NEO_HOOKEAN_BORDEN, MpmEngine, admm_opts, phase_field_opts
    wl = args.workload
    extra = {}
    eng = None
    if wl == "adjoint-cube":
        from ziran2019_amd.autodiff import Rollout
        s = scenes.adjoint_scene()
        dt, nsteps = 2e-4, 64
        prob = Rollout(s, nsteps, dt, checkpoint_every=8)
        rng = np.random.default_rng(5)
        V0 = 0.1 * rng.normal(size=(len(s["mass"]), 3))
        target = s["X"].mean(axis=0) + np.array([0.05, 0.0, 0.0])

        def host_step():
            # the torch-facing driver: state and parameters uploaded, X
            # downloaded for the loss, the adjoint of E and V0 read back
            X = prob.forward(s["E"], V0)
            com = X.mean(axis=0)
            gX = np.tile(2.0 * (com - target) / len(X), (len(X), 1))
            gE, gV = prob.backward(gX)
            extra["loss"] = float(((com - target) ** 2).sum())
            extra["grad_E_norm"] = float(np.linalg.norm(gE))
            return gX

        # the timed rollout starts from device-resident inputs: the initial
        # state restored from a device checkpoint, the two rollout calls and
        # the loss seed's upload (n x 3 doubles; every rollout reaches the
        # same X, so the seed is the same); nothing is read back
        gX = host_step()
        peng = prob.eng

        def step():
            peng.checkpoint_restore(0)  # the forward's first checkpoint: the initial state
            peng.set_time(0.0)
            peng.rollout_forward(nsteps, dt, prob.every, 0)
            peng.adjoint_set(X=gX)
            peng.rollout_backward(nsteps, dt, prob.every, 0, prob.nseg)
        workload = ("differentiable explicit MLS-MPM rollout, 262144 particles at dx=1/128, 64 steps dt=2e-4, "
                    "E lognormal(ln 1e4, 0.4), checkpoints every 8 steps, adjoint to E and V0; inputs resident "
                    "on the device (the loss seed uploaded per rollout); value counts the 64 forward steps of "
                    "each rollout")
        n, per = len(s["mass"]), nsteps
    else:
        if wl == "admm-slabs":
            s = scenes.slabs_scene()
            dt = 1e-3
        else:
            s = scenes.fracture_scene()
            dt = 5e-4
        n = len(s["mass"])
        fp32_transfers = wl == "fracture-sphere" and args.transfer_precision == "fp32"
        eng = MpmEngine(dx=s["dx"], gravity=s["gravity"], precision=1 if fp32_transfers else 0)
        eng.upload(s)
        eng.set_colliders(s["colliders"])
        if wl == "admm-slabs":
            eng.set_materials(**s["materials"])
            fp32 = args.admm_precision == "fp32-refined"
            o = admm_opts(admm_iterations=20, precision=1 if fp32 else 0)
            minres = []

            def step():
                minres.append(eng.advanceOneTimeStepAdmm(dt, o).minres_iterations)
            workload = ("ADMM step (StvkWithHencky, von Mises slab with yield stress 200 Pa beside a Drucker-Prager "
                        "slab at 30 deg), 2048000 particles at 8 ppc, dx=1/256, dt=1e-3, 20 ADMM iterations, "
                        + ("FP32-bulk global step" if fp32 else "FP64"))
            extra["minres_iterations"] = minres
        else:
            eng.set_materials(model=ZR_MODEL_NEO_HOOKEAN_BORDEN)
            eng.set_phase_field(phase_field_opts(), l0=s["l0"], one_over_sigma_c=s["one_over_sigma_c"],
                                residual_phase=s["residual_phase"])

            def step():
                eng.advanceOneTimeStep(dt)
            workload = ("explicit MLS step with the CD-MPM phase-field solve (CG), 3000000 NeoHookeanBorden "
                        "particles in a Voronoi-toughened block, sticky sphere at 5 m/s, dx=1/320, dt=5e-4, "
                        + ("FP32 P2G terms" if fp32_transfers else "FP64"))
        per = 1
    for _ in range(args.warmup):
        step()
    if eng is not None:
        eng.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    t1 = time.perf_counter()
    kernels = {}
    if eng is not None:
        for k in ("p2g", "grid", "g2p", "sort"):
            ms, launches = eng.kernel_time(k)
            kernels[k] = dict(avg_us=round(ms * 1e3 / max(launches, 1), 2), launches=launches)
        eng.close()
    el = t1 - t0
    if wl == "adjoint-cube":
        # the resident rollouts' gradient of E (read after their timing), to
        # set beside the driver's
        a = peng.adjoint_get()
        extra["grad_E_norm_resident"] = float(np.linalg.norm(
            a["mu"] / (2.0 * (1.0 + prob.nu)) + a["lam"] * prob.nu / ((1.0 + prob.nu) * (1.0 - 2.0 * prob.nu))))
        # the same rollout through the torch-facing driver, host I/O included
        t2 = time.perf_counter()
        for _ in range(args.steps):
            host_step()
        extra["ms_per_step_with_host_io"] = (time.perf_counter() - t2) / args.steps * 1e3
        prob.close()
    if "minres_iterations" in extra:
        extra["minres_iterations"] = extra["minres_iterations"][-args.steps:]
    print(json.dumps(dict(workload=wl, config=workload, n_particles=n, steps=args.steps, warmup=args.warmup,
                          ms_per_step=el / args.steps * 1e3, value=n * per * args.steps / el,
                          unit="particle-steps/s", kernels=kernels, **extra)))


def calculate_dt_timing(calls):
    """calculateDt (evalMaxParticleSpeed: |v| and the position bounds, one
    reduction over the particles) on the implicit-boxes particles, ``calls``
    synchronous calls; kernel durations come from a rocprofv3 trace of this
    command, the wall time per call includes the host round trip."""
    from ziran2019_amd import scenes
    from ziran2019_amd.engine import MpmEngine
    scene = scenes.boxes_scene()
    n = len(scene["mass"])
    eng = MpmEngine(dx=scene["dx"], gravity=scene["gravity"])
    eng.upload(scene)
    eng.set_colliders(scene["colliders"])
    dt = eng.calculateDt(2e-3)
    t0 = time.perf_counter()
    for _ in range(calls):
        dt = eng.calculateDt(2e-3)
    t1 = time.perf_counter()
    eng.close()
    print(json.dumps(dict(workload="calculateDt on the implicit-boxes particles", n_particles=n, calls=calls,
                          dt=dt, wall_us_per_call=(t1 - t0) / calls * 1e6, algorithmic_bytes=n * 6 * F64)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="implicit-boxes",
                    choices=["implicit-boxes", "explicit-cube", "admm-slabs", "fracture-sphere", "adjoint-cube"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--preconditioner", choices=sorted(PRECONDITIONERS), default="mass",
                    help="CG preconditioner of the implicit-boxes workload")
    ap.add_argument("--admm-precision", choices=["fp64", "fp32-refined"], default="fp64",
                    help="admm-slabs: the global step in FP64 or in the FP32-bulk form")
    ap.add_argument("--transfer-precision", choices=["fp64", "fp32"], default="fp64",
                    help="fracture-sphere: the P2G's per-particle terms in FP64 or FP32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 256 MiB (5 launches) to scale HBM counters")
    ap.add_argument("--calculate-dt", type=int, default=0, metavar="CALLS",
                    help="only time CALLS calls of calculateDt on the implicit-boxes particles")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.calculate_dt > 0:
        calculate_dt_timing(args.calculate_dt)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("admm-slabs", "fracture-sphere", "adjoint-cube"):
        args.steps = 3 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        run_row_workload(args)
        return
    implicit = args.workload == "implicit-boxes"
    if args.steps is None:
        args.steps = 10 if implicit else 1000
    if args.warmup is None:
        args.warmup = 2 if implicit else 20
    r = run_implicit(args) if implicit else run_explicit(args)
    n, kt, g_nodes = r["n"], r["kt"], r["g_nodes"]

    kern = {}
    for k, kind in r["bw_kernels"].items():
        ms, launches = kt[k]
        avg_s = ms / 1e3 / max(launches, 1)
        b = algorithmic_bytes(kind, n, g_nodes)
        kern[k] = dict(avg_us=avg_s * 1e6, bytes=b, gbs=b / avg_s / 1e9)
    dom = max(r["roof_kernels"], key=lambda k: kern[k]["avg_us"] * kt[k][1])
    roof = dict(bound="hbm", limiter=LIMITER.get(dom), kernel=dom, achieved=round(kern[dom]["gbs"], 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(kern[dom]["gbs"] / HBM_PEAK_GBS, 4), traffic=TRAFFIC.get((args.workload, dom)),
                bytes_per_launch=kern[dom]["bytes"], avg_launch_us=round(kern[dom]["avg_us"], 2))
    kernels = {k: dict(avg_us=round(v[0] * 1e3 / max(v[1], 1), 2), launches=v[1], total_ms=round(v[0], 3))
               for k, v in kt.items()}
    for k in kern:
        kernels[k]["gbs"] = round(kern[k]["gbs"], 2)
        kernels[k]["hbm_frac"] = round(kern[k]["gbs"] / HBM_PEAK_GBS, 4)

    elapsed = r["elapsed"]
    out = dict(
        metric=METRIC,
        value=n * args.steps / elapsed,
        unit="particle-steps/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=elapsed / args.steps * 1e3,
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype=r["dtype"],
        data="synthetic",
        config=dict(workload=r["workload"], n_particles=n, active_nodes=g_nodes, dx=r["dx"], dt=r["dt"],
                    **r["extra"]),
        roofline=roof,
        kernels=kernels,
        cpu_baseline=r["base"],
    )
    print(json.dumps(out))


if __name__ == "__main__":
    main()
