"""Benchmark: tet-steps/s of the per-timestep soft-body solve on one MI355X.

Workload (default): the ~2e5-tet octopus-like body (ellipsoid mantle + 8
tapered arms, Neo-Hookean with a log-normal stiffness field) on a ground
plane with IPC barrier (dhat = 1e-3) and friction (mu = 0.3), dt = 1/100,
12 Newton iterations per step, block-Jacobi PCG to relative residual 1e-6
(FP64 residual), line search on.  Synthetic mesh, generated in-process.

One "step" = RefinementSolver.step on the whole body.  value = tets * steps /
wall time of the timed steps (inputs resident in HBM).  The roofline object
is for the dominant kernel, the persistent PCG (k_pcg): algorithmic bytes per
CG iteration = 40 B per nonzero 3x3 block (36 B values + 4 B column) + 24 B
per vertex (gather and publish of the FP32 direction), times the iterations
each launch ran, divided by the launch's duration from HIP events on the
launching stream.  cpu_baseline = the FP64 C oracle (oracle/mfem_oracle.c),
one step of the same scene from its initial state, one thread.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "warp-mfem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="octopus", choices=["octopus", "beam"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--pcg", default="standard", choices=["standard", "pipelined"])
    ap.add_argument("--stream-rows", action="store_true", help="stream matrix rows instead of keeping them in VGPRs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the solve shards by body or by vertex partition; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mfem_amd import scene
    from mfem_amd.solver import RefinementSolver, model_state

    if args.workload == "octopus":
        m, dt = scene.octopus_scene(), 0.01
        wl = "octopus_like_204k_tets_ground_contact"
    else:
        m, dt = scene.beam_scene(), 1.0 / 60.0
        wl = "cantilever_beam_1920_tets"
    kw = {k: v for k, v in m.params.items() if k != "iterations"}
    solver = RefinementSolver(m, m.params["iterations"], max_tets=m.tet_count, enable_refinement=False,
                              device="cuda:0", pcg_pipelined=args.pcg == "pipelined",
                              pcg_resident_rows=not args.stream_rows, **kw)
    state = model_state(m, "cuda:0")
    out = state.clone()
    print(f"[bench] {wl}: {m.tet_count} tets, {m.particle_count} verts, warmup {args.warmup}", flush=True)
    for k in range(args.warmup):
        solver.step(state, out, None, None, dt)
        state, out = out, state
    torch.cuda.synchronize()
    if solver.status() != 0:
        raise SystemExit("device error during warmup")

    timers: dict = {}
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.step_instrumented(state, out, dt, timers)
        hist.append(solver.buffers["stats"].clone())
        state, out = out, state
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"[bench] timed {args.steps} steps in {wall:.3f} s", flush=True)
    if solver.status() != 0:
        raise SystemExit("device error during the timed steps")

    from mfem_amd import _ffi

    st = np.stack([h.cpu().numpy().reshape(_ffi.MFEM_MAX_STATS_ITERS, _ffi.MFEM_STATS_PER_ITER)[: solver.iterations]
                   for h in hist])  # (K, iters, stats)
    khz = solver.lib.mfem_wall_clock_khz()
    ticks = st[:, :, 8:14].sum(axis=(0, 1))
    n_it = max(1.0, float(st[:, :, 0].sum()))
    pcg_phase_us = {name: float(ticks[k] / khz * 1e3 / n_it) for k, name in
                    enumerate(["spmv", "reduce1", "update", "reduce2", "checks", "prologue"])}
    cg_iters = st[:, :, 0]
    relres = st[:, :, 1]
    assert np.all(relres <= 1e-6), "a PCG solve missed rel-res 1e-6"
    phase_ms = {k: sum(a.elapsed_time(b) for a, b in v) for k, v in timers.items()}
    pcg_ms = np.array([a.elapsed_time(b) for a, b in timers["global_solve"]])
    nnz = int(len(solver.pattern.cols))
    V, T = m.particle_count, m.tet_count
    bytes_per_iter = 40.0 * nnz + 24.0 * V
    iters_flat = cg_iters.reshape(-1)
    active = iters_flat > 0
    achieved = float(np.sum(iters_flat[active] * bytes_per_iter) / np.sum(pcg_ms[active] * 1e-3) / 1e9) if active.any() else 0.0
    traffic = None

    value = T * args.steps / wall
    result = {
        "metric": "tet-steps/s at CG rel-res 1e-6, learned-precond vs plain PCG, 1 MI355X",
        "value": value,
        "unit": "tet-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl,
            "tets": T,
            "verts": V,
            "newton_iterations": solver.iterations,
            "dt": dt,
            "preconditioner": "block_jacobi",
            "pcg": args.pcg,
            "learned_preconditioner": False,
            "reductions": "f64",
            "cg_iters_mean": float(cg_iters.mean()),
            "cg_iters_per_step": float(cg_iters.sum(axis=1).mean()),
            "cg_relres_max": float(relres.max()),
            "cg_iters_last_step": cg_iters[-1].astype(int).tolist(),
            "cg_restarts_last_step": st[-1, :, 7].astype(int).tolist(),
            "ls_trials_last_step": st[-1, :, 2].astype(int).tolist(),
            "pcg_phase_us_per_iter": pcg_phase_us,
            "phase_ms_per_step": {k: v / args.steps for k, v in phase_ms.items()},
        },
        "roofline": {
            "kernel": "k_pcg",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_cg_iter": bytes_per_iter,
            "us_per_cg_iter": float(np.sum(pcg_ms[active]) * 1e3 / max(1.0, np.sum(iters_flat[active]))),
        },
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(m, dt)
    print(json.dumps(result), flush=True)


def cpu_baseline(m, dt):
    from oracle import Oracle

    o = Oracle(m, 64)
    t0 = time.perf_counter()
    o.step(dt)
    sec = time.perf_counter() - t0
    return {
        "value": m.tet_count / sec,
        "unit": "tet-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"1 step ({int(o.param('iterations'))} Newton iterations) of the same scene from its initial "
                  f"state, FP64 C oracle, 1 thread, {sec:.1f} s",
    }


if __name__ == "__main__":
    main()
