"""Benchmark of the MPM time step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload explicit-cube]

One step = MpmSimulationBase::advanceOneTimeStep on the whole particle set
(sort + page allocation, P2G with fused stress, grid update, G2P with F
update), inputs resident in HBM.  Prints one JSON line with particle-steps/s,
the roofline position of the dominant kernel (HIP-event kernel durations over
the timed region) and a CPU baseline (the FP64 C oracle on the host cores).
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


def algorithmic_bytes(kernel, n, g):
    """Required HBM bytes per launch.

    n particles, g active grid nodes (m != 0), FP64 (8 B) values, 4 B index:
      P2G : reads x3 v3 C9 F9 m1 V0 1 mu1 lambda1 (28 values) + index,
            writes m, mv3, f3 per node (7 values)
      G2P : reads x3 F9 + index and v*3 per node, writes x3 v3 C9 F9
    """
    if kernel == "p2g":
        return n * (28 * F64 + 4) + g * 7 * F64
    if kernel == "g2p":
        return n * (36 * F64 + 4) + g * 3 * F64
    raise ValueError(kernel)


def cpu_baseline(scene, dt, seconds, threads):
    import oracle
    sim = oracle.OracleSim(dx=scene["dx"], gravity=scene["gravity"], nthreads=threads)
    sim.set_particles(scene["X"], scene["V"], scene["C"], scene["F"], scene["mass"], scene["vol"], scene["mu"],
                      scene["lam"])
    for c in scene["colliders"]:
        sim.add_collider(**c)
    sim.step(dt)  # first step excluded
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.step(dt)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 200:
            break
    n = len(scene["mass"])
    return dict(value=n * steps / el, unit="particle-steps/s", cores=threads, kind="port",
                sample=f"FP64 C oracle (oracle/zro_mpm.c, OpenMP) on the same {n}-particle workload, "
                       f"{steps} steps after one untimed step, {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="explicit-cube", choices=["explicit-cube"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see DESIGN.md for how the step would shard)", file=sys.stderr)
        sys.exit(2)

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
    eng.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.advanceOneTimeStep(dt)
    t1 = time.perf_counter()
    kt = {k: eng.kernel_time(k) for k in ("p2g", "grid", "g2p", "sort")}
    eng.profile(False)
    elapsed = t1 - t0
    # active grid nodes of the final state
    eng.sortParticlesAndPolluteGrid()
    eng.particlesToGrid()
    g_nodes = int(np.count_nonzero(eng.dump("grid_m")))
    eng.close()

    kern = {}
    for k in ("p2g", "g2p"):
        ms, launches = kt[k]
        avg_s = ms / 1e3 / max(launches, 1)
        b = algorithmic_bytes(k, n, g_nodes)
        kern[k] = dict(avg_us=avg_s * 1e6, bytes=b, gbs=b / avg_s / 1e9)
    dom = max(kern, key=lambda k: kern[k]["avg_us"])
    roof = dict(bound="hbm", kernel=dom, achieved=round(kern[dom]["gbs"], 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(kern[dom]["gbs"] / HBM_PEAK_GBS, 4), traffic=None,
                bytes_per_launch=kern[dom]["bytes"], avg_launch_us=round(kern[dom]["avg_us"], 2))
    kernels = {k: dict(avg_us=round(v[0] * 1e3 / max(v[1], 1), 2), launches=v[1]) for k, v in kt.items()}
    for k in ("p2g", "g2p"):
        kernels[k]["gbs"] = round(kern[k]["gbs"], 2)

    base = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        base = cpu_baseline(scene, dt, args.cpu_seconds, threads)

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
        dtype="f64",
        data="synthetic",
        config=dict(workload="explicit MLS-MPM, fixed-corotated cube 26^3 cells x 8 ppc, dx=1/128, dt=1e-4, "
                             "sticky ground, FP64 plain path",
                    n_particles=n, active_nodes=g_nodes, dx=scene["dx"], dt=dt),
        roofline=roof,
        kernels=kernels,
        cpu_baseline=base,
    )
    print(json.dumps(out))


if __name__ == "__main__":
    main()
