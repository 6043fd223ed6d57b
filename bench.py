#!/usr/bin/env python3
"""Benchmark of the MLS-MPM step on one MI355X.

Workload (one GPU): 3D snow, four 64^3 cubes (1,048,576 particles) of side 0.2 at seeded
centres with one seeded velocity per cube, 128^3 grid (dx = 1/128), FP32, dt = 1e-4.  A "step"
is one MPMSimulation<3>::advance() (P2G -> grid update -> G2P) over all particles, with the
particle state resident in HBM for the whole timed region.

Prints one JSON line: particle-steps/s, the roofline of the dominant kernel (algorithmic bytes
per launch / its average launch time measured with HIP events on the engine's stream) and the
CPU baseline (the restated reference step, serial FP32, on a bounded sample of the same
workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/sec fwd and fwd+bwd (2D/3D, FP32) vs plain GPU & host CPU path"

# Algorithmic (minimum) HBM bytes per particle per launch, SoA FP32 (3D snow):
#   P2G reads x, v, F, C, Jp, mass, volume = 3+3+9+9+3 floats = 108 B, writes 16 B per active node
#   G2P reads x, F, Jp = 13 floats = 52 B, writes x, v, C, F, Jp = 25 floats = 100 B,
#       reads 12 B (3 velocity floats) per active node
#   grid update reads and writes 4 floats per active node = 32 B
P2G_BYTES_PER_PARTICLE = {(3, 0): 108, (3, 1): 104, (3, 2): 104, (2, 0): 60, (2, 1): 56, (2, 2): 56}
G2P_BYTES_PER_PARTICLE = {(3, 0): 152, (3, 1): 144, (3, 2): 144, (2, 0): 80, (2, 1): 72, (2, 2): 72}
P2G_BYTES_PER_NODE = {2: 12, 3: 16}
G2P_BYTES_PER_NODE = {2: 8, 3: 12}
GRID_BYTES_PER_NODE = {2: 24, 3: 32}

# Adjoint kernels of a jelly step, algorithmic (minimum) HBM bytes, SoA FP32, d = 2 / 3:
#   G2P adjoint reads x (d), F (d^2) and the cotangent of x, v, F, C, Jp (2d + 2d^2 + 1), writes
#       the partial cotangents of x, F, Jp (d + d^2 + 1); per node with mass it reads the updated
#       grid velocity (d) and writes the grid cotangent (d)
#   grid adjoint, per node with mass: reads the pre-update grid {momentum, mass} (d + 1) and the
#       grid cotangent (d), writes the cotangent of {momentum, mass} (d + 1)
#   P2G adjoint reads x, v, F, C, mass, volume (2d + 2d^2 + 2) and the partial cotangents (d + d^2 + 1),
#       writes the cotangent of x, v, F, C, Jp (2d + 2d^2 + 1) and dL/dE (1); per node with mass it
#       reads the cotangent of {momentum, mass} (d + 1)
def adjoint_bytes(d: int):
    ng = 2 * d + 2 * d * d + 1
    return {
        "g2p_adjoint": (4 * (d + d * d + ng + d + d * d + 1), 4 * 2 * d),
        "grid_adjoint": (0, 4 * (d + 1 + d + d + 1)),
        "p2g_adjoint": (4 * (2 * d + 2 * d * d + 2 + d + d * d + 1 + ng + 1), 4 * (d + 1)),
    }


def load_traffic(kernel: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present, with the state
    it was measured on: (bytes, {"g_active": nodes with mass then, "alg_bytes": the algorithmic
    bytes of that launch, "ratio": bytes / alg_bytes, "source": the summary}).  The counters come
    from their own profiled run, not from this one: the active-node count of the 3D snow workload
    grows with the step count, so the ratio is the comparable figure."""
    path = os.path.join(HERE, "profiles", f"traffic_{kernel}_four_cubes_3d.json")
    if not os.path.exists(path):
        return None, None
    try:
        with open(path) as f:
            rec = json.load(f)
    except Exception:
        return None, None
    b = rec.get("bytes_per_launch")
    state = {k: rec.get(k) for k in ("g_active", "alg_bytes", "steps_before", "source")}
    if b and rec.get("alg_bytes"):
        state["ratio"] = b / rec["alg_bytes"]
    return b, state


def cpu_baseline(sc, steps: int):
    import numpy as np  # noqa: F401

    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O

    O.build()
    p = sc.particles
    st = O.State(x=p.x, v=p.v, F=p.F, C=p.C, Jp=p.Jp, mass=p.mass, volume=p.volume)
    secs = O.time_steps(st, int(sc.material), sc.res, sc.dt, sc.E, sc.nu, sc.gravity, warmup=1, steps=steps)
    # every core this process may run on (the GPU box allots 16 to one GPU's job and says so in
    # OMP_NUM_THREADS); nproc and the affinity mask are recorded beside the count
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = nproc
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(usable, omp) if omp > 0 else usable
    msteps = 2 * steps
    psecs = O.time_steps(st, int(sc.material), sc.res, sc.dt, sc.E, sc.nu, sc.gravity, warmup=1, steps=msteps,
                         threads=threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": p.n * steps / secs,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"same workload ({p.n} particles, res {sc.res}): 1 untimed + {steps} timed steps of the "
                  f"restated serial reference step (oracle/, FP32, -O3 -ffast-math), {secs:.1f} s",
        "cpu_model": model,
        "host": {"nproc": nproc, "affinity_cpus": usable, "OMP_NUM_THREADS": omp or None},
        "multicore": {
            "value": p.n * msteps / psecs,
            "unit": "particle-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"1 untimed + {msteps} timed steps of the same step on {threads} OpenMP threads (each thread adds "
                      f"the contributions to its own band of grid rows, in particle order: bitwise the serial step), {psecs:.1f} s",
        },
    }


def two_cubes_2d():
    """The reference example's 2D snow scene (src/example.cpp:29-46: two 50x50 cubes, 5,000
    particles, res 64) for its 4,000 steps in one advance: tiny, so bounded by per-step latency
    rather than by HBM.  The headline figure takes the autotuner's choice (the tiled step's launch
    shapes and the compact step's particles per thread); the tiled step and each compact shape are
    reported beside it."""
    from nuclearmpm_amd import MPMSimulation, scenes

    sc = scenes.two_cubes_2d()

    def timed(**kw):
        with MPMSimulation(sc.particles, sc.material, res=sc.res, **kw) as sim:
            sim.advance(20)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.advance(sc.steps)
            sim.synchronize()
            return time.perf_counter() - t0, sim.tuning()

    t, tuning = timed(autotune=True)
    variants = {"tiled": timed(compact=0)[0]}
    for ppt in (1, 2, 4):
        variants[f"compact_ppt{ppt}"] = timed(compact=ppt)[0]
    return {"workload": "two_cubes_2d: 2D snow, 2 x 50x50 particles, res 64, dt 1e-4 (src/example.cpp)",
            "particles": sc.particles.n, "steps": sc.steps, "value": sc.particles.n * sc.steps / t,
            "unit": "particle-steps/s", "us_per_step": 1e6 * t / sc.steps, "tuning": tuning,
            "us_per_step_by_step_kind": {k: 1e6 * v / sc.steps for k, v in variants.items()}}


def batched_vs_sequential(steps: int):
    """2D batched workload: 256 seeded scenes (1-4 cubes each, liquid / snow / jelly, res 128) in
    one launch sequence per step for the configured 2,000 steps, against the same scenes advanced
    one plain simulation at a time; particle-steps/s of each and the ratio.  A scene whose particle
    leaves its grid stops at that step on its own (the reference throws out of that scene's
    advance()); it is reported, and only the steps the scenes actually took are counted."""
    from nuclearmpm_amd import MPMBatch, MPMSimulation, scenes

    specs = scenes.batched_2d(seed=1, scenes=256)
    nps = [s.particles.n for s in specs]
    n = sum(nps)
    warm = 5
    with MPMBatch(specs) as b:
        b.advance(warm)
        b.synchronize()
        t0 = time.perf_counter()
        b.advance(steps)
        status = b.scene_status()  # waits for the steps; a stopped scene is not an error here
        tb = time.perf_counter() - t0
    live = 0
    for s, m in enumerate(nps):
        st = status.get(s)
        taken = steps if st is None else min(steps, max(0, st["step"] - 1 - warm))
        live += m * taken
    seq_steps = 50
    ts = 0.0
    for s in specs:
        with MPMSimulation(s.particles, s.material, res=s.res, dt=s.dt, E=s.E, nu=s.nu, gravity=s.gravity) as sim:
            sim.advance(2)
            sim.synchronize()
            t0 = time.perf_counter()
            sim.advance(seq_steps)
            sim.synchronize()
            ts += time.perf_counter() - t0
    batched = live / tb
    sequential = n * seq_steps / ts
    return {
        "workload": "batched_2d: 256 scenes (seed 1), 1-4 cubes each, side resolution U{20..80}, "
                    "materials uniform over liquid/snow/jelly, res 128, dt 1e-4",
        "particles": n,
        "batched": {"value": batched, "unit": "particle-steps/s", "steps": steps,
                    "stopped_scenes": {str(s): {"step": st["step"], "particles": nps[s],
                                                "material": int(specs[s].material)} for s, st in status.items()},
                    "particle_steps_counted": live},
        "sequential": {"value": sequential, "unit": "particle-steps/s", "steps_per_scene": seq_steps},
        "speedup": batched / sequential,
    }


def fwd_bwd(which: str, svd: str = "exact"):
    """Forward + backward of a differentiable jelly rollout with checkpoints (torch.autograd over
    the engine's adjoint kernels); particle-steps/s counts each simulated step once.  `svd`: the
    engine's SVD mode (the accurate mode takes the polar factor of the forward and of the adjoint
    from the FP64 Newton iteration)."""
    import numpy as np
    import torch

    from nuclearmpm_amd import MPMBatch, MPMSimulation, autodiff, scenes
    from nuclearmpm_amd._lib import lib

    if which == "jelly_2d_diff":
        specs = [scenes.jelly_2d_diff()]
        # the forward steps take the autotuner's step (the compact step, for this small scene); the
        # adjoint's recomputed steps are tiled (the adjoint kernels walk the tile sort)
        sim = MPMSimulation(specs[0].particles, specs[0].material, res=specs[0].res, E=specs[0].E, autotune=True,
                            svd=svd)
        desc = "jelly_2d_diff: 140x140 particles at [0.3,0.5]^2, res 128, 512 steps, checkpoint every 64"
    else:
        specs = scenes.jelly_3d_batched(seed=2, scenes=8)
        sim = MPMBatch(specs, svd=svd)
        desc = "jelly_3d_batched: 8 scenes x 64^3 particles, res 128, 256 steps, checkpoint every 32"
    steps, every = specs[0].steps, int(specs[0].extra["checkpoint_every"])
    d = specs[0].particles.dim
    nps = [s.particles.n for s in specs]
    n = sum(nps)
    targets = torch.tensor(np.stack([s.extra["target"] for s in specs]), dtype=torch.float64, device="cuda")
    x0 = torch.cat([autodiff.pack(*(torch.from_numpy(getattr(s.particles, k)).cuda() for k in ("x", "v", "F", "C")))
                    for s in specs], dim=1)
    E = torch.tensor([s.E for s in specs], dtype=torch.float64)

    def run(record="none"):
        s0 = x0.clone().requires_grad_(True)
        e = E.clone().requires_grad_(True)
        out = autodiff.rollout(sim, s0, e, steps, every, record=record)
        xs = autodiff.unpack(out, d)[0]
        loss = 0.0
        start = 0
        for s, m in enumerate(nps):
            loss = loss + ((xs[start:start + m].double().mean(0) - targets[s]) ** 2).sum()
            start += m
        loss.backward()
        torch.cuda.synchronize()
        return float(loss), s0.grad, e.grad

    run()  # warm-up (allocations, tape)
    t0 = time.perf_counter()
    loss, g, ge = run()
    t = time.perf_counter() - t0
    t0 = time.perf_counter()
    with torch.no_grad():
        autodiff.rollout(sim, x0, E, steps, every)
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    # the same rollout with its segments recorded in the forward pass (record="auto": as many as fit
    # in half the free device memory), so that the backward pass does not recompute them
    run("auto")  # warm-up (the tapes' allocations)
    plan = autodiff._record_plan(lib(), sim._handle(), steps, every, "auto", x0.device)
    rec_bytes = sum(int(lib().mpm_record_bytes(sim._handle(), min(every, steps - i * every), m))
                    for i, m in enumerate(plan) if m is not None)
    t0 = time.perf_counter()
    _, g_rec, ge_rec = run("auto")
    t_rec = time.perf_counter() - t0
    recorded = {
        "value": n * steps / t_rec,
        "segments": len(plan),
        "recorded_with_grids": sum(m == 1 for m in plan),
        "recorded_without_grids": sum(m == 0 for m in plan),
        "tape_GB": rec_bytes / 1e9,
        "gradients_bitwise_equal": bool(torch.equal(g_rec, g)) and bool(torch.equal(ge_rec, ge)),
        # a single-scene engine's plain forward may take the compact step, whose fixed-point sums
        # round differently from the tiled step's that recorded segments (and recomputations) take
        "grad_max_rel_diff": float((g_rec - g).abs().max() / g.abs().max().clamp_min(1e-30)),
        "gradE_max_rel_diff": float((ge_rec - ge).abs().max() / ge.abs().max().clamp_min(1e-30)),
    }
    del g_rec, ge_rec
    # a separate pass with HIP events per kernel: the adjoint kernels' rooflines
    sim.set_timing(True)
    sim.reset_kernel_times()
    run()
    sim.synchronize()
    g_active = sim.active_nodes()
    kernels = {}
    for name, (per_particle, per_node) in adjoint_bytes(d).items():
        avg_ms, launches = sim.kernel_time(name)
        nbytes = per_particle * n + per_node * g_active
        gbs = nbytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        kernels[name] = {"avg_ms": avg_ms, "launches": launches, "alg_bytes": nbytes, "achieved_GBs": gbs,
                         "frac": gbs / HBM_PEAK_GBS}
    tuning = sim.tuning()
    sim.close()
    return {
        "workload": desc,
        "particles": n,
        "steps": steps,
        "value": n * steps / t,
        "unit": "particle-steps/s (forward + backward)",
        "forward_only": n * steps / tf,
        "recorded": recorded,
        "loss": loss,
        "grad_finite": bool(torch.isfinite(g).all()) and bool(torch.isfinite(ge).all()),
        "adjoint_kernels": kernels,
        "g_active_end": g_active,
        "tuning": tuning,
        "svd": SVD_MODES[svd],
    }


def scene_generation():
    """The 3D snow workload's 1,048,576-particle state built on the device (lattices, velocities,
    F, C, Jp: a 25 x n float image, 104.9 MB written) against the host scene generator plus the
    upload of the same image."""
    import numpy as np
    import torch

    from nuclearmpm_amd import generators, scenes

    sc = scenes.four_cubes_3d()
    generators.scene_image(sc)  # warm the kernels up
    dev_ms, kern_ms = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        img = generators.scene_image(sc)
        e1.record()
        torch.cuda.synchronize()
        dev_ms.append(1e3 * (time.perf_counter() - t0))
        kern_ms.append(e0.elapsed_time(e1))
    host_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        h = scenes.four_cubes_3d().particles
        n = h.x.shape[0]
        host = np.concatenate([h.x.T, h.v.T, h.F.reshape(n, -1).T, h.C.reshape(n, -1).T, h.Jp[None]])
        up = torch.as_tensor(host, device="cuda")
        torch.cuda.synchronize()
        host_ms.append(1e3 * (time.perf_counter() - t0))
    same = bool(torch.equal(img, up))
    nbytes = img.numel() * 4
    return {"particles": int(img.shape[1]), "image_bytes": nbytes, "device_ms": float(np.median(dev_ms)),
            "device_stream_ms": float(np.median(kern_ms)),
            "device_GBs": nbytes / (float(np.median(kern_ms)) * 1e-3) / 1e9,
            "host_plus_upload_ms": float(np.median(host_ms)), "bitwise_equal": same}


LEGS = {
    "four_cubes_3d_accurate": lambda: accurate_step(),
    "two_cubes_2d": lambda: two_cubes_2d(),
    "batched": lambda: batched_vs_sequential(2000),
    "fwd_bwd_2d": lambda: fwd_bwd("jelly_2d_diff"),
    "fwd_bwd_3d": lambda: fwd_bwd("jelly_3d_batched"),
    "scene_gen": lambda: scene_generation(),
}


SVD_MODES = {"exact": "nclr_svd order (bitwise the reference's FP32)",
             "fast": "fast FP32 Jacobi of A^T A (not bitwise the reference)",
             "accurate": "FP64 Newton polar factor and FP64 Jacobi SVD (closer to the FP64 step than the "
                         "reference's FP32 arithmetic)"}


def measure_step(sc, svd: str, steps: int, warmup: int, tuning_arg: str = "auto", graph: bool = False):
    """The plain step on scene `sc`: `warmup` untimed steps, the launch shape (autotuned on the
    warmed-up state or given as CHUNKxTHREADS), `steps` timed steps, then a separate pass with HIP
    events per kernel (so the events do not perturb the end-to-end number).  Returns (particle-steps/s,
    ms per step, tuning, per-kernel records, nodes with mass, whether every position is finite)."""
    from nuclearmpm_amd import MPMSimulation

    p = sc.particles
    n = p.n
    dim, mat = p.dim, int(sc.material)
    sim = MPMSimulation(p, sc.material, res=sc.res, dt=sc.dt, E=sc.E, nu=sc.nu, gravity=sc.gravity,
                        graph=graph, svd=svd)
    sim.advance(warmup)
    sim.synchronize()
    if tuning_arg == "auto":
        tuning = sim.autotune(5)  # untimed; restores the state and the step count
    else:
        chunk, threads = (int(v) for v in tuning_arg.lower().split("x"))
        sim.set_tuning(chunk, threads)
        tuning = sim.tuning()
    sim.advance(2)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.advance(steps)
    sim.synchronize()
    t1 = time.perf_counter()
    total = t1 - t0
    ms_per_step = 1e3 * total / steps
    value = n * steps / total

    ksteps = max(10, min(steps, 200))
    sim.set_timing(True)
    sim.reset_kernel_times()
    sim.advance(ksteps)
    sim.synchronize()
    kt = {k: sim.kernel_time(k)[0] for k in ("p2g", "grid", "g2p")}
    sim.set_timing(False)
    g_active = sim.active_nodes()
    finite = bool(__import__("numpy").isfinite(sim.particles().x).all())
    sim.close()

    p2g_bytes = P2G_BYTES_PER_PARTICLE[(dim, mat)] * n + P2G_BYTES_PER_NODE[dim] * g_active
    g2p_bytes = G2P_BYTES_PER_PARTICLE[(dim, mat)] * n + G2P_BYTES_PER_NODE[dim] * g_active
    grid_bytes = GRID_BYTES_PER_NODE[dim] * g_active
    per_kernel = {
        "p2g": {"avg_ms": kt["p2g"], "alg_bytes": p2g_bytes},
        "grid": {"avg_ms": kt["grid"], "alg_bytes": grid_bytes},
        "g2p": {"avg_ms": kt["g2p"], "alg_bytes": g2p_bytes},
    }
    # P2G accumulates every particle's 3^d x (d + 1) quantised contributions with 64-bit integer
    # atomics in LDS (no HBM atomics: each work item stores its own slot)
    lds_atomics = (3 ** dim) * (dim + 1) * n
    per_kernel["p2g"]["lds_atomic_u64_per_launch"] = lds_atomics
    per_kernel["p2g"]["lds_atomics_per_s"] = lds_atomics / (kt["p2g"] * 1e-3) if kt["p2g"] > 0 else 0.0
    for v in per_kernel.values():
        v["achieved_GBs"] = v["alg_bytes"] / (v["avg_ms"] * 1e-3) / 1e9 if v["avg_ms"] > 0 else 0.0
        v["frac"] = v["achieved_GBs"] / HBM_PEAK_GBS
    return value, ms_per_step, tuning, per_kernel, g_active, finite


def accurate_step():
    """The headline workload (1M-particle 3D snow, 1000 timed steps after 20 warm-up steps,
    autotuned) in the accurate SVD mode: per-particle results closer to the FP64 step than the
    reference's FP32 arithmetic, not bitwise the reference's SVD."""
    from nuclearmpm_amd import scenes

    value, ms, tuning, per_kernel, g_active, finite = measure_step(scenes.four_cubes_3d(), "accurate", 1000, 20)
    return {"workload": "four_cubes_3d, svd accurate", "value": value, "unit": "particle-steps/s",
            "ms_per_step": ms, "svd": SVD_MODES["accurate"],
            "tuning": {"p2g_chunk": int(tuning["p2g_chunk"]), "g2p_threads": int(tuning["g2p_threads"])},
            "kernels": per_kernel, "g_active": g_active, "state_finite": finite}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and forward+backward legs")
    ap.add_argument("--graph", action="store_true",
                    help="replay each step as a captured hipGraph (measured slower than direct launches on this "
                         "ROCm stack for every workload here, so off by default)")
    ap.add_argument("--fast-svd", action="store_true", help="engine option fast_svd (not bitwise the reference)")
    ap.add_argument("--svd", choices=["exact", "fast", "accurate"], default=None,
                    help="SVD mode of the step: exact (nclr_svd's FP32 operation order, bitwise the reference's "
                         "arithmetic), fast (FP32 Jacobi of A^T A) or accurate (FP64 Newton polar factor and FP64 "
                         "Jacobi SVD)")
    ap.add_argument("--tuning", default="auto",
                    help="P2G particles per work item and G2P workgroup size as 'CHUNKxTHREADS' (e.g. 512x256), "
                         "or 'auto' to let the engine's autotuner time every shape on the warmed-up state")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 5 device copies of 256 MiB (the reference kernel for HBM traffic counters)")
    ap.add_argument("--leg", choices=sorted(LEGS),
                    help="run only this secondary leg and print its JSON record (for profiling it alone); "
                         "the headline line is not printed")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (batched scenes would "
              "shard by scene without communication). Run with --gpus 1.", file=sys.stderr)
        return 2
    if args.leg:
        print(json.dumps({args.leg: LEGS[args.leg]()}), flush=True)
        return 0

    from nuclearmpm_amd import scenes

    sc = scenes.four_cubes_3d()
    n = sc.particles.n
    dim = sc.particles.dim

    if args.copy_calibration:
        import torch

        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    svd = args.svd or ("fast" if args.fast_svd else "exact")
    value, ms_per_step, tuning, per_kernel, g_active, finite = measure_step(
        sc, svd, args.steps, args.warmup, args.tuning, args.graph)
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_ms"])
    achieved = per_kernel[dom]["achieved_GBs"]
    traffic, traffic_state = load_traffic(dom)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator)",
        "config": {
            "workload": "four_cubes_3d: 3D snow plain step, 4 cubes x 64^3 = 1,048,576 particles, side 0.2, "
                        "res 128 (dx=1/128), dt 1e-4, E 1e4, nu 0.2, gravity -100",
            "particles": n,
            "res": sc.res,
            "dim": dim,
            "material": "snow",
            "mode": "fwd",
            "hipgraph": args.graph,
            "tuning": {"p2g_chunk": int(tuning["p2g_chunk"]), "g2p_threads": int(tuning["g2p_threads"]),
                       "chosen_by": "autotuner" if args.tuning == "auto" else "flag"},
            "svd": SVD_MODES[svd],
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_measured_on": traffic_state,
        },
        "kernels": per_kernel,
        "g_active": g_active,
        "state_finite": finite,
    }
    if not args.no_extra:
        for name, fn in LEGS.items():
            try:
                out[name] = fn()
            except Exception as e:  # report the failing leg without losing the headline number
                out[name] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
