"""Benchmark: tet-steps/s of the per-timestep soft-body solve on one MI355X,
learned preconditioning (network-weighted coarse correction on top of block
Jacobi, with a learned Galerkin warm start) vs plain block-Jacobi PCG.

Workload (default): the ~2e5-tet octopus-like body (ellipsoid mantle + 8
tapered arms, Neo-Hookean with a log-normal stiffness field) on a ground
plane with IPC barrier (dhat = 1e-3) and friction (mu = 0.3), dt = 1/100,
12 Newton iterations per step, block-Jacobi PCG to relative residual 1e-6
(FP64 residual), line search on.  Synthetic mesh, generated in-process.

Order: W warmup steps (plain PCG); ``--coarse-steps`` plain steps whose
Newton systems are logged, the coarse-weight network (per-vertex smoothing
weights of the PCG's workgroup-local rigid-mode correction) trained on them on
the GPU (BF16 MFMA forward/backward, Adam) on the energy-norm error left by
five preconditioned CG iterations, and kept only if two validation steps take
fewer CG iterations than with the network off; ``--train-steps`` plain steps
whose solves are logged for the warm-start MLP (a Galerkin-start candidate),
trained the same way and kept only with a 5 % margin; plain steps up to
``--start-step`` (default 50 of the 100-step scene); the autotuner's choice of
PCG variant on whole steps; then K timed steps with plain PCG (the reference's
start x~ - x and block Jacobi), K with the learned path (Galerkin start +
block Jacobi + the network-weighted rigid-mode correction: the value), and as
ablations K with the networks off (unsmoothed rigid modes, no learned guess),
K with a hand-tuned uniform weight in place of the network, and K with block
Jacobi alone under the Galerkin start, all from the same state and solver
state.  One "step" = RefinementSolver.step on the whole body.  value = tets *
K / wall time of the learned leg (inputs resident in HBM); the plain leg's
rate is reported beside it (config.plain_value).

The roofline object is for the dominant kernel, the persistent PCG (k_pcg):
algorithmic bytes per CG iteration = 40 B per nonzero 3x3 block (36 B values
+ 4 B column) + 24 B per vertex (gather and publish of the FP32 direction),
times the iterations each launch ran, divided by the launch's duration from
HIP events on the launching stream.  The mlp object is the MFMA figure of the
network's forward kernel: 2 * (16*64 + 64*64 + 64*3) FLOP per vertex over its
event-timed duration, against the dense BF16 peak.  cpu_baseline = the C
oracle (oracle/mfem_oracle.c, -O3 -march=x86-64-v3, OpenMP over vertices and
tets), one step from the state the timed legs start from, FP32 and FP64, on
all host threads the process may use and on one; value = the faster
all-threads leg.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--start-step S] [--train-steps T] [--epochs E]
                       [--no-learned] [--workload octopus|beam|batch]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "warp-mfem_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
_OBJECTIVES = {"energy": "log e^T H e per solve, e = dx0 - dx* (energy-norm error of the guess)",
               "residual": "log |g - H dx0|^2 per solve (initial residual)"}
BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="octopus", choices=["octopus", "beam", "batch"],
                    help="octopus: the ~2e5-tet body (default); beam: the 1,920-tet cantilever; batch: 64 seeded "
                         "soft bodies (~1.2e6 tets) in one batched solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-no-single-thread", action="store_true",
                    help="time the CPU baseline on all threads only (skip the one-thread legs)")
    ap.add_argument("--pcg", default="standard", choices=["standard", "pipelined"])
    ap.add_argument("--stream-rows", action="store_true", help="stream matrix rows instead of keeping them in VGPRs")
    ap.add_argument("--train-steps", type=int, default=6, help="plain steps logged as training solves")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--objective", default="energy", choices=sorted(_OBJECTIVES),
                    help="training objective of the warm-start network")
    ap.add_argument("--local-sweeps", type=int, default=None,
                    help="workgroup-local block-Jacobi sweeps added to the preconditioner of the accelerated legs "
                         "(even; 0 = block Jacobi as the reference; default: the autotuner picks 0, 2 or 4)")
    ap.add_argument("--local-omega", type=float, default=0.5)
    ap.add_argument("--no-learned", action="store_true", help="time plain PCG only")
    ap.add_argument("--start-step", type=int, default=50,
                    help="step of the 100-step scene the timed legs start at: after warmup and the logged training "
                         "steps, plain steps advance the body to it (the settled, contact-dense phase)")
    ap.add_argument("--coarse-steps", type=int, default=3, help="plain steps logged for the coarse-weight network")
    ap.add_argument("--coarse-epochs", type=int, default=10)
    ap.add_argument("--coarse-lr", type=float, default=2e-2)
    ap.add_argument("--coarse-omega", type=float, default=0.0,
                    help="uniform smoothing weight of the coarse basis when the network is off")
    ap.add_argument("--uniform-omega", type=float, default=0.75,
                    help="the hand-tuned uniform weight of the ablation that replaces the network")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph legs")
    ap.add_argument("--dump-iters", default=None,
                    help="write the CG iteration count of every PCG launch, in launch order, to this JSON file "
                         "(to divide per-dispatch hardware counters by iterations)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default PCG variant choice")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run five 1 GiB device-to-device copies (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the solve shards by body or by vertex partition; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 4, "copy_probe_bytes_written": dst.numel() * 4,
                          "copies": 5}), flush=True)
        return

    if args.workload == "batch":
        return run_batch(args)

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
    T_tets = m.tet_count
    state = model_state(m, "cuda:0")
    out = state.clone()
    print(f"[bench] {wl}: {m.tet_count} tets, {m.particle_count} verts, warmup {args.warmup}", flush=True)
    if args.dump_iters:
        solver.launch_log = []  # (variant, CG iterations) of every PCG launch, in launch order
    for k in range(args.warmup):
        solver.step(state, out, None, None, dt)
        state, out = out, state
    torch.cuda.synchronize()
    if solver.status() != 0:
        raise SystemExit("device error during warmup")

    from mfem_amd import _ffi
    from mfem_amd.coarse import CoarseWeightNet, collect_coarse
    from mfem_amd.learned import WarmStartNet, collect_solves

    learned = not args.no_learned
    train_info = None
    coarse_info = None
    net = None
    cnet = None
    if learned:
        # the coarse-weight network (smoothing of the PCG's workgroup-local
        # rigid-mode basis): systems logged from plain steps, trained on the
        # GPU, kept only if whole steps take fewer CG iterations than with
        # the network off (w = 0)
        t0 = time.perf_counter()
        csol = collect_coarse(solver, state, dt, steps=args.coarse_steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        cnet = CoarseWeightNet("cuda:0", seed=1)
        chist = cnet.fit(csol, solver.pattern, epochs=args.coarse_epochs, lr=args.coarse_lr, seed=0)
        del csol
        cval = cnet.validate_cg(solver, state, dt, steps=2)
        cw = cnet_weights_summary(solver, cnet, state, dt)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        coarse_info = {"systems": args.coarse_steps * solver.iterations, "epochs": args.coarse_epochs,
                       "objective": "log(e_5^T H e_5 / e^T H e): energy-norm error left by 5 CG iterations "
                                    "preconditioned by M(w) on H y = H e, e = dx* - (x~ - x)",
                       "collect_s": t1 - t0, "train_s": t2 - t1, "loss_network_off": chist["off"],
                       "loss_per_epoch": [round(v, 5) for v in chist["loss"]], "best_epoch": chist["best_epoch"],
                       "loss_uniform_weight": chist.get("uniform"), "validation": cval, "weights": cw}
        print(f"[bench] coarse-weight network: loss {chist['loss'][0]:.5g} -> {min(chist['loss']):.5g} (kept epoch "
              f"{chist['best_epoch']}); validation CG iterations {cval['cg_iters']}, kept {cval['kept']}", flush=True)
        # the warm start's network (Galerkin start candidate)
        t0 = time.perf_counter()
        solves = collect_solves(solver, state, dt, steps=args.train_steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        net = WarmStartNet("cuda:0")
        hist = net.fit_solves(solves, solver, args.objective, epochs=args.epochs, lr=1e-4, seed=0)
        del solves
        solver.set_coarse_net(cnet)
        # its launches (features, MLP, blend) cost ~5 % of a step: kept only
        # if it saves at least that share of the CG iterations
        val = net.validate_cg(solver, state, dt, steps=2, margin=0.05)
        solver.set_coarse_net(None)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        train_info = {"solves": args.train_steps * solver.iterations,
                      "pairs": args.train_steps * solver.iterations * m.particle_count, "train_steps": args.train_steps,
                      "epochs": args.epochs, "objective": _OBJECTIVES[args.objective], "collect_s": t1 - t0,
                      "train_s": t2 - t1, "loss_prev_newton_step": hist["base"],
                      "loss_per_epoch": [round(v, 4) for v in hist["loss"]], "best_epoch": hist["best_epoch"],
                      "validation": val}
        print(f"[bench] warm start trained in {t2 - t1:.2f} s, objective {hist['loss'][0]:.4g} -> "
              f"{min(hist['loss']):.4g} (kept epoch {hist['best_epoch']}); validation CG iterations "
              f"{val['cg_iters']}, kept {val['kept']}", flush=True)
        if val["kept"] != "trained":
            net = None  # a zero correction: no warm-start network launches (the Galerkin start without it)

    # plain steps up to the timed window (the networks were trained on the
    # early steps: the window measures how they carry to later ones)
    steps_done = args.warmup + ((args.coarse_steps + args.train_steps) if learned else 0)
    solver.set_warm_start(None)
    solver.set_coarse_net(None)
    solver.set_coarse_correction(False)
    solver._S.pcg_flags |= _ffi.PCG_PIPELINED
    while steps_done < args.start_step:
        solver.step(state, out, None, None, dt)
        state, out = out, state
        steps_done += 1
    torch.cuda.synchronize()
    if solver.status() != 0:
        raise SystemExit("device error while advancing to the timed window")
    print(f"[bench] timed window: steps {steps_done} to {steps_done + args.steps - 1}", flush=True)

    khz = solver.lib.mfem_wall_clock_khz()

    def configure(ws_net, coarse_net, coarse_on, omega):
        if ws_net is not None:
            solver.set_warm_start(ws_net)
        else:
            solver.set_galerkin_start(learned)  # the Galerkin start without the learned guess
        solver.set_local_preconditioner(0, args.local_omega)
        solver.set_coarse_net(None)
        solver.set_coarse_correction(coarse_on, omega=omega)
        if coarse_net is not None:
            solver.set_coarse_net(coarse_net)
        if not coarse_on:
            solver._S.pcg_flags |= _ffi.PCG_PIPELINED

    tune = None
    if learned and not args.no_autotune:
        configure(net, cnet if coarse_info["validation"]["kept"] == "trained" else None, True, args.coarse_omega)
        tune = solver.autotune(state, dt)
        c = tune["chosen"]
        print(f"[bench] autotuned PCG: variant {c['variant']} ({c['chunks_per_wave']} chunk(s)/wave, "
              f"{c['resident_slots']} resident slots, {c['mode']}), {c['ms_per_step']:.2f} ms/step, "
              f"{c['us_per_iter']:.2f} us/iter", flush=True)
    tuned_variant = tune["chosen"]["variant"] if tune else -1

    def timed_leg(s0, ws_net, coarse_net, coarse_on, omega=0.0, ws=True, mode="eager"):
        """K steps from s0 and the same solver state.  mode "eager": one
        ``solver.step`` (one mfem_step, no host work between the kernels) per
        step; "graph": replays of two HIP graphs of the step (a -> b, b -> a;
        ``capture_step``); "instrumented": the same steps stage by stage with
        events between the phases (each event record costs ~10 us of device
        idle), for the phase and per-launch timings."""
        st = s0.clone()
        o = st.clone()
        solver.restore(snap0)
        if ws:
            configure(ws_net, coarse_net, coarse_on, omega)
        else:
            solver.set_warm_start(None)
            solver.set_coarse_net(None)
            solver.set_coarse_correction(False)
            solver._S.pcg_flags |= _ffi.PCG_PIPELINED
        solver._S.pcg_variant = tuned_variant if coarse_on else -1
        graphs = None
        if mode == "graph":
            graphs = (solver.capture_step(st, o, dt), solver.capture_step(o, st, dt))
            st.assign(s0)
            solver.restore(snap0)
        timers: dict = {}
        hist = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            if mode == "instrumented":
                solver.step_instrumented(st, o, dt, timers)
            elif mode == "graph":
                graphs[k % 2].replay()
            else:
                solver.step(st, o, None, None, dt)
            hist.append(solver.buffers["stats"].clone())
            st, o = o, st
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if solver.status() != 0:
            raise SystemExit("device error during the timed steps")
        stats = np.stack([h.cpu().numpy().reshape(_ffi.MFEM_MAX_STATS_ITERS, _ffi.MFEM_STATS_PER_ITER)
                          [: solver.iterations] for h in hist])  # (K, iters, stats)
        assert np.all(stats[:, :, 1] <= 1e-6), "a PCG solve missed rel-res 1e-6"
        return wall, stats, timers, st

    s0 = state.clone()
    snap0 = solver.snapshot()  # every leg starts from the same solver state
    wall_p, st_p, timers_p, end_p = timed_leg(s0, None, None, False, ws=False)
    print(f"[bench] plain: {args.steps} steps in {wall_p:.3f} s", flush=True)
    abl = {}
    if learned:
        cn = cnet if coarse_info["validation"]["kept"] == "trained" else None
        wall, st, timers, end_l = timed_leg(s0, net, cn, True, args.coarse_omega)
        print(f"[bench] learned (warm start + learned coarse correction): {args.steps} steps in {wall:.3f} s",
              flush=True)
        pos_diff = float((end_l.particle_q - end_p.particle_q).abs().max().item())
        # ablations, from the same state and solver state: the networks off
        # (Galerkin start without the learned guess, unsmoothed rigid modes,
        # w = 0); a hand-tuned uniform weight in place of the network; the
        # coarse correction off (block Jacobi under the Galerkin start)
        for name, kwl in (("networks_off", dict(ws_net=None, coarse_net=None, coarse_on=True, omega=0.0)),
                          ("uniform_weight", dict(ws_net=net, coarse_net=None, coarse_on=True,
                                                  omega=args.uniform_omega)),
                          ("block_jacobi", dict(ws_net=net, coarse_net=None, coarse_on=False))):
            w_, s_, _, _ = timed_leg(s0, **kwl)
            abl[name] = {"value": T_tets * args.steps / w_, "ms_per_step": w_ * 1e3 / args.steps,
                         "cg_iters_per_step": float(s_[:, :, 0].sum(axis=1).mean())}
            print(f"[bench] ablation {name}: {args.steps} steps in {w_:.3f} s, "
                  f"{abl[name]['cg_iters_per_step']:.0f} CG iterations per step", flush=True)
    else:
        wall, st = wall_p, st_p
    # the plain and the value configuration replayed as HIP graphs; the
    # value is the faster mode, compared with the plain leg in the same mode
    if args.no_graph:
        wall_pg, st_pg, wall_g, st_g = float("inf"), st_p, float("inf"), st
    else:
        wall_pg, st_pg, _, _ = timed_leg(s0, None, None, False, ws=False, mode="graph")
        if learned:
            wall_g, st_g, _, _ = timed_leg(s0, net, cn, True, args.coarse_omega, mode="graph")
        else:
            wall_g, st_g = wall_pg, st_pg
    print(f"[bench] HIP graph replays: plain {wall_pg:.3f} s, value configuration {wall_g:.3f} s", flush=True)
    eager = {"value_ms_per_step": wall * 1e3 / args.steps, "plain_ms_per_step": wall_p * 1e3 / args.steps}
    graph = None if args.no_graph else {"value_ms_per_step": wall_g * 1e3 / args.steps,
                                         "plain_ms_per_step": wall_pg * 1e3 / args.steps}
    wall_e = wall
    step_mode = "graph" if wall_g < wall else "eager"
    if step_mode == "graph":
        wall, wall_p, st = wall_g, wall_pg, st_g
    # the same steps again, stage by stage with events: phase and PCG launch
    # timings of the value leg and of the plain leg
    if learned:
        wall_i, st_i, timers, _ = timed_leg(s0, net, cn, True, args.coarse_omega, mode="instrumented")
    else:
        wall_i, st_i, timers, _ = timed_leg(s0, None, None, False, ws=False, mode="instrumented")
    wall_pi, _, timers_p, _ = timed_leg(s0, None, None, False, ws=False, mode="instrumented")
    same_iters = bool(np.array_equal(st_i[:, :, 0], st[:, :, 0]))

    ticks = st[:, :, 8:14].sum(axis=(0, 1))
    n_it = max(1.0, float(st[:, :, 0].sum()))
    pcg_phase_us = {name: float(ticks[k] / khz * 1e3 / n_it) for k, name in
                    enumerate(["spmv", "reduce1", "update", "reduce2", "checks", "prologue"])}
    cg_iters = st[:, :, 0]
    relres = st[:, :, 1]
    phase_ms = {k: sum(a.elapsed_time(b) for a, b in v) for k, v in timers.items()}
    pcg_ms = np.array([a.elapsed_time(b) for a, b in timers["global_solve"]])
    pcg_ms_p = np.array([a.elapsed_time(b) for a, b in timers_p["global_solve"]])
    nnz = int(len(solver.pattern.cols))
    V, T = m.particle_count, m.tet_count
    bytes_per_iter = 40.0 * nnz + 24.0 * V
    iters_flat = cg_iters.reshape(-1)
    active = iters_flat > 0
    achieved = float(np.sum(iters_flat[active] * bytes_per_iter) / np.sum(pcg_ms[active] * 1e-3) / 1e9) if active.any() else 0.0
    # HBM bytes per launch: the per-CG-iteration FETCH_SIZE / WRITE_SIZE of
    # k_pcg measured by rocprofv3 on this workload (calibrated against a
    # 1 GiB device copy) times this run's mean iterations per launch
    traffic = None
    traffic_file = os.path.join(ROOT, "profiles", "hbm_traffic_octopus_204k.json")
    if args.workload == "octopus" and os.path.exists(traffic_file):
        with open(traffic_file) as f:
            per_iter = json.load(f)["k_pcg"]["hbm_bytes_per_cg_iter"]
        traffic = float(per_iter * np.mean(iters_flat[active])) if active.any() else None

    value = T * args.steps / wall
    value_plain = T * args.steps / wall_p
    config = {
        "workload": wl,
        "tets": T,
        "verts": V,
        "newton_iterations": solver.iterations,
        "dt": dt,
        "timed_steps": [steps_done, steps_done + args.steps - 1],
        "preconditioner": {"plain": "block_jacobi",
                           "value": "block_jacobi + workgroup-local rigid-mode correction, learned smoothing weights"
                           if learned else "block_jacobi"},
        "pcg": tune["chosen"]["mode"] if tune else "pipelined",
        "learned_warm_start": bool(learned and net is not None),
        "reductions": "f64",
        "plain_value": value_plain,
        "plain_ms_per_step": wall_p * 1e3 / args.steps,
        "step_mode": step_mode,
        "eager": eager,
        "hip_graph": graph,
        "instrumented_ms_per_step": {"value_leg": wall_i * 1e3 / args.steps, "plain_leg": wall_pi * 1e3 / args.steps,
                                     "same_cg_iterations": same_iters},
        "learned_speedup": value / value_plain,
        "cg_iters_per_step": float(cg_iters.sum(axis=1).mean()),
        "plain_cg_iters_per_step": float(st_p[:, :, 0].sum(axis=1).mean()),
        "cg_iters_mean": float(cg_iters.mean()),
        "cg_relres_max": float(relres.max()),
        "cg_iters_last_step": cg_iters[-1].astype(int).tolist(),
        "cg_restarts_last_step": st[-1, :, 7].astype(int).tolist(),
        "ls_trials_last_step": st[-1, :, 2].astype(int).tolist(),
        "pcg_phase_us_per_iter": pcg_phase_us,
        "phase_ms_per_step": {k: v / args.steps for k, v in phase_ms.items()},
        "plain_us_per_cg_iter": float(np.sum(pcg_ms_p) * 1e3 / max(1.0, float(st_p[:, :, 0].sum()))),
    }
    if tune is not None:
        config["pcg_autotune"] = {
            "chosen": tune["chosen"]["variant"],
            "table": [{k: r[k] for k in ("variant", "chunks_per_wave", "resident_slots", "mode", "workgroups",
                                         "ms_per_step", "us_per_iter", "cg_iters")}
                      for r in tune["table"]]}
    if learned:
        config["pcg_start_frac"] = {"x_tilde_minus_x": float((st[:, :, 14] == 0).mean()),
                                    "galerkin": float((st[:, :, 14] == 3).mean())}
        # |g - H x0| / |g - H (x~ - x)| of the Galerkin start
        config["start_residual_ratio_median"] = float(np.median(st[:, :, 15]))
        config["max_abs_position_diff_vs_plain"] = pos_diff
        config["training"] = train_info
        config["coarse_weight_network"] = coarse_info
        config["ablations"] = abl
        # what the networks add over the same method without them
        # (ablations run eagerly: compared with the eager value leg)
        value_e = T * args.steps / wall_e
        config["learned_correction_gain"] = value_e / abl["networks_off"]["value"]
        config["learned_vs_uniform_weight"] = value_e / abl["uniform_weight"]["value"]
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
        "config": config,
        "roofline": {
            "kernel": "k_pcg",
            "bound": "latency",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_cg_iter": bytes_per_iter,
            "traffic_note": "achieved = algorithmic bytes (40 B per nonzero block + 24 B per vertex per CG iteration) "
                            "over the launch's event time; traffic = measured HBM bytes per launch (FETCH_SIZE x 2.0 "
                            "+ WRITE_SIZE x 1.0, profiles/hbm_traffic_octopus_204k.json), below the algorithmic bytes "
                            "because the matrix rows stay in VGPRs for the whole solve: an iteration moves only the "
                            "published vector (staged per workgroup in LDS); the kernel is bound by grid-reduction "
                            "and memory latency, not HBM bandwidth",
            "hbm_frac_measured": (float(traffic / (np.mean(pcg_ms[active]) * 1e-3) / 1e9 / HBM_PEAK_GBS)
                                  if traffic else None),
            "us_per_cg_iter": float(np.sum(pcg_ms[active]) * 1e3 / max(1.0, np.sum(iters_flat[active]))),
            "us_per_launch": float(np.mean(pcg_ms[active]) * 1e3) if active.any() else None,
            "cg_iters_per_launch": float(np.mean(iters_flat[active])) if active.any() else None,
        },
    }
    result["roofline_assembly"] = assembly_roofline(phase_ms, args.steps * solver.iterations, T, V, nnz,
                                                    args.workload == "octopus")
    result["roofline_spmv"] = spmv_roofline(solver, V, nnz)
    if learned:
        result["mlp"] = mlp_mfma(solver, cnet, V)
    if args.dump_iters:
        dump_launch_log(args.dump_iters, solver, m, nnz)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(m, dt, s0, snap0, single_thread=not args.cpu_no_single_thread)
    print(json.dumps(result), flush=True)


def run_batch(args):
    """The batch workload: 64 seeded soft bodies (boxes / ellipsoids, 2e3 to
    5e4 tets each, ARAP / Neo-Hookean, E log-uniform in [1e4, 1e6] Pa) over
    the plane z = 0, dt = 1/100, 12 Newton iterations, every body's PCG to
    rel-res 1e-6 on its own, in one persistent PCG launch and one
    line-search launch per Newton iteration.  Legs as for one body (plain,
    learned after training on logged batch solves, autotuned), plus the
    learned leg replayed as HIP graphs of a whole step (value = the faster
    of the eager and graph learned legs).  Roofline: k_pcg algorithmic bytes
    per launch = sum over bodies of that body's CG iterations x (40 B per
    nonzero block + 24 B per vertex), over the launch's event-timed
    duration."""
    import numpy as np
    import torch

    from mfem_amd import _ffi, scene
    from mfem_amd.batch import BatchSolver
    from mfem_amd.learned import WarmStartNet, collect_solves

    dt = 0.01
    t0 = time.perf_counter()
    models = scene.soft_body_batch()
    params = {k: v for k, v in models[0].params.items() if k not in ("iterations", "energy")}
    bs = BatchSolver(models, models[0].params["iterations"], device="cuda:0", **params)
    solver = bs.solver
    nb = bs.n_bodies
    T = bs.tet_count
    Vr = sum(m.particle_count for m in models)
    wl = f"soft_body_batch_{nb}_bodies_{T // 1000}k_tets"
    print(f"[bench] {wl}: {T} tets, {Vr} verts ({solver.active_particle_count} rows, body alignment "
          f"{256 * bs.body_wg_align} rows), built in {time.perf_counter() - t0:.1f} s, warmup {args.warmup}", flush=True)
    state, out = bs.make_state(), bs.make_state()
    for k in range(args.warmup):
        bs.step(state, out, dt)
        state, out = out, state
    torch.cuda.synchronize()
    if bs.status() != 0:
        raise SystemExit("device error during warmup")

    def body_stats():
        return solver.buffers["stats"].view(nb, _ffi.MFEM_MAX_STATS_ITERS, _ffi.MFEM_STATS_PER_ITER)[:, : solver.iterations]

    learned = not args.no_learned
    net, train_info = None, None
    if learned:
        t0 = time.perf_counter()
        solves = collect_solves(solver, state, dt, steps=args.train_steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        net = WarmStartNet("cuda:0")
        hist = net.fit_solves(solves, solver, args.objective, epochs=args.epochs, lr=1e-4, seed=0)
        del solves
        val = net.validate_cg(solver, state, dt, steps=2)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        train_info = {"solves": args.train_steps * solver.iterations,
                      "pairs": args.train_steps * solver.iterations * Vr, "train_steps": args.train_steps,
                      "epochs": args.epochs, "objective": _OBJECTIVES[args.objective], "collect_s": t1 - t0,
                      "train_s": t2 - t1, "loss_prev_newton_step": hist["base"],
                      "loss_per_epoch": [round(v, 4) for v in hist["loss"]], "best_epoch": hist["best_epoch"],
                      "validation": val}
        print(f"[bench] trained in {t2 - t1:.2f} s (kept epoch {hist['best_epoch']}); validation CG iterations "
              f"{val['cg_iters']}, kept {val['kept']}", flush=True)
    # the learned legs' start: the network when validation kept it, else the
    # Galerkin start alone (its output layer would be zero: the features,
    # MLP and blend launches would add nothing)
    value_net = None if not learned else (net if train_info["validation"]["kept"] == "trained" else "galerkin")

    def attach(use_net):
        if isinstance(use_net, str):
            solver.set_galerkin_start(True)
        else:
            solver.set_warm_start(use_net)

    tune = None
    if not args.no_autotune:
        attach(value_net)
        solver.set_local_preconditioner(0, args.local_omega)
        tune = solver.autotune(state, dt, local_sweeps=_sweep_options(args))
        c = tune["chosen"]
        args.local_sweeps = c["local_sweeps"]
        print(f"[bench] autotuned PCG: variant {c['variant']} ({c['chunks_per_wave']} chunk(s)/wave, "
              f"{c['resident_slots']} resident slots, {c['mode']}), {c['local_sweeps']} local sweeps, "
              f"{c['ms_per_step']:.2f} ms/step", flush=True)
    if args.local_sweeps is None:
        args.local_sweeps = 0

    def timed_leg(s0, use_net, sweeps=0, instrumented=False):
        """K eager steps (``solver.step``), or with ``instrumented`` the same
        steps stage by stage with events between the phases."""
        st, o = s0.clone(), s0.clone()
        solver.restore(snap0)
        attach(use_net)
        solver.set_local_preconditioner(sweeps, args.local_omega)
        timers: dict = {}
        hist = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            if instrumented:
                solver.step_instrumented(st, o, dt, timers)
            else:
                solver.step(st, o, None, None, dt)
            hist.append(body_stats().clone())
            st, o = o, st
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if bs.status() != 0:
            raise SystemExit("device error during the timed steps")
        stats = torch.stack(hist).cpu().numpy()  # (K, bodies, iters, stats)
        assert np.all(stats[..., 1] <= 1e-6), "a body's PCG solve missed rel-res 1e-6"
        return wall, stats, timers, st

    def graph_leg(s0, use_net, sweeps=0):
        attach(use_net)
        solver.set_local_preconditioner(sweeps, args.local_omega)
        a, b = s0.clone(), s0.clone()
        solver.restore(snap0)
        g_ab = solver.capture_step(a, b, dt)
        g_ba = solver.capture_step(b, a, dt)
        a.assign(s0)
        solver.restore(snap0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            (g_ab if k % 2 == 0 else g_ba).replay()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if bs.status() != 0:
            raise SystemExit("device error during the graph replays")
        assert np.all(body_stats().cpu().numpy()[..., 1] <= 1e-6)
        return wall

    s0 = state.clone()
    snap0 = solver.snapshot()  # every leg starts from the same solver state
    wall_p, st_p, _, end_p = timed_leg(s0, None)
    print(f"[bench] plain: {args.steps} steps in {wall_p:.3f} s", flush=True)
    if learned:
        wall_e, st, timers, end_l = timed_leg(s0, value_net, args.local_sweeps)
        print(f"[bench] learned + local preconditioner: {args.steps} steps in {wall_e:.3f} s", flush=True)
        wall_j, st_j, _, _ = timed_leg(s0, value_net, 0)
        print(f"[bench] learned, block Jacobi: {args.steps} steps in {wall_j:.3f} s", flush=True)
    else:
        wall_e, st, timers = wall_p, st_p, _
    wall_g = graph_leg(s0, value_net, args.local_sweeps)
    print(f"[bench] learned, HIP graph replays: {args.steps} steps in {wall_g:.3f} s", flush=True)
    wall = min(wall_e, wall_g)
    # phase and PCG launch timings from an instrumented repeat of the value leg
    _, _, timers, _ = timed_leg(s0, value_net, args.local_sweeps, instrumented=True)

    # per-body algorithmic bytes per CG iteration
    rp = solver.pattern.rowptr
    bv = bs.body_vert
    bytes_body = np.array([40.0 * (rp[bv[i] + models[i].particle_count] - rp[bv[i]]) + 24.0 * models[i].particle_count
                           for i in range(nb)])
    pcg_ms = np.array([x.elapsed_time(y) for x, y in timers["global_solve"]])  # (K * iters)
    it_body = st[..., 0]  # (K, bodies, iters)
    launch_bytes = np.einsum("kbi,b->ki", it_body, bytes_body).reshape(-1)
    launch_iters = it_body.max(axis=1).reshape(-1)
    active = launch_iters > 0
    achieved = float(launch_bytes[active].sum() / (pcg_ms[active].sum() * 1e-3) / 1e9) if active.any() else 0.0
    per_step_body = it_body.sum(axis=2)  # (K, bodies)
    value = T * args.steps / wall
    value_plain = T * args.steps / wall_p
    config = {
        "workload": wl,
        "bodies": nb,
        "tets": T,
        "verts": Vr,
        "rows": int(solver.active_particle_count),
        "body_alignment_rows": 256 * bs.body_wg_align,
        "tets_per_body_min_max": [int(min(m.tet_count for m in models)), int(max(m.tet_count for m in models))],
        "neohookean_bodies": int(sum(m.params["energy"] == "neohookean" for m in models)),
        "newton_iterations": solver.iterations,
        "dt": dt,
        "preconditioner": {"plain": "block_jacobi",
                           "value": f"block_jacobi + {args.local_sweeps} workgroup-local sweeps (omega "
                                    f"{args.local_omega})" if args.local_sweeps else "block_jacobi"},
        "learned_warm_start": learned,
        "warm_start": (None if not learned else
                       "network" if value_net is net else "galerkin_start (network not kept by validation)"),
        "reductions": "f64, per body",
        "launches_per_newton_iteration": {"k_pcg": 1, "k_linesearch": 1},
        "eager_ms_per_step": wall_e * 1e3 / args.steps,
        "graph_ms_per_step": wall_g * 1e3 / args.steps,
        "plain_value": value_plain,
        "plain_ms_per_step": wall_p * 1e3 / args.steps,
        "learned_speedup": value / value_plain,
        "cg_iters_per_step_launch": float(st[..., 0].max(axis=1).sum(axis=1).mean()),
        "plain_cg_iters_per_step_launch": float(st_p[..., 0].max(axis=1).sum(axis=1).mean()),
        "body_cg_iters_per_step_min_median_max": [float(per_step_body.mean(axis=0).min()),
                                                  float(np.median(per_step_body.mean(axis=0))),
                                                  float(per_step_body.mean(axis=0).max())],
        "cg_relres_max": float(st[..., 1].max()),
        "phase_ms_per_step": {k: sum(a.elapsed_time(b) for a, b in v) / args.steps for k, v in timers.items()},
        "ls_trials_per_newton_iteration": {"max_over_bodies_mean": float(st[..., 2].max(axis=1).mean()),
                                           "mean_over_bodies": float(st[..., 2].mean())},
        "hardest_bodies": [{"tets": models[i].tet_count, "energy": models[i].params["energy"],
                            "E": float(_young(models[i])), "cg_iters_per_step": float(per_step_body[:, i].mean())}
                           for i in np.argsort(-per_step_body.mean(axis=0))[:4]],
    }
    if tune is not None:
        config["pcg_autotune"] = {
            "chosen": tune["chosen"]["variant"],
            "table": [{k: r[k] for k in ("variant", "chunks_per_wave", "resident_slots", "mode", "workgroups",
                                         "local_sweeps", "ms_per_step", "us_per_iter", "cg_iters")}
                      for r in tune["table"]]}
    if learned:
        config["training"] = train_info
        config["pcg_start_frac"] = {"x_tilde_minus_x": float((st[..., 14] == 0).mean()),
                                    "galerkin": float((st[..., 14] == 3).mean())}
        config["ablation_block_jacobi"] = {
            "value": T * args.steps / wall_j, "ms_per_step": wall_j * 1e3 / args.steps,
            "cg_iters_per_step_launch": float(st_j[..., 0].max(axis=1).sum(axis=1).mean())}
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
        "config": config,
        "roofline": {
            "kernel": "k_pcg",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "us_per_launch_iter": float(pcg_ms[active].sum() * 1e3 / max(1.0, launch_iters[active].sum())),
        },
    }
    if learned:
        result["mlp"] = mlp_mfma(solver, net, solver.active_particle_count)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline_bodies(models, dt, single_thread=not args.cpu_no_single_thread)
    print(json.dumps(result), flush=True)


def _young(m):
    mu, lam = m.tet_materials[0, 0], m.tet_materials[0, 1]
    return mu * (3 * lam + 2 * mu) / (lam + mu)


def cpu_baseline_bodies(models, dt, single_thread: bool = True, budget_tets: int = 150_000):
    """The C oracle (-O3 -march=x86-64-v3) on the batch, one step of every
    body from its initial state, bodies spread over all host threads the
    process may use (one single-threaded oracle per body and thread: the
    bodies are independent, ctypes drops the GIL), FP32 and FP64; and one
    thread on the first bodies up to ``budget_tets`` tets.  ``value`` is the
    faster all-threads rate."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import Oracle

    threads, model = _cpu_info()

    def one(m, bits):
        o = Oracle(m, bits, threads=1)  # set in the worker thread: it owns its OpenMP settings
        o.step(dt)
        return m.tet_count

    def run(bits, bodies, nthreads):
        t0 = time.perf_counter()
        if nthreads == 1:
            tets = sum(one(m, bits) for m in bodies)
        else:
            # largest bodies first so the pool's tail is short
            order = sorted(bodies, key=lambda m: -m.tet_count)
            with ThreadPoolExecutor(max_workers=nthreads) as ex:
                tets = sum(ex.map(lambda m: one(m, bits), order))
        sec = time.perf_counter() - t0
        return {"bits": bits, "threads": nthreads, "bodies": len(bodies), "tets": tets, "value": tets / sec,
                "s": sec}

    legs = [run(32, models, threads), run(64, models, threads)]
    best = max(legs, key=lambda r: r["value"])
    if single_thread:
        sample, tets = [], 0
        for m in models:
            if sample and tets + m.tet_count > budget_tets:
                break
            sample.append(m)
            tets += m.tet_count
        legs += [run(32, sample, 1), run(64, sample, 1)]
    return {"value": best["value"], "unit": "tet-steps/s", "cores": threads, "kind": "port", "cpu": model,
            "sample": f"1 step (12 Newton iterations) of all {len(models)} bodies ({best['tets']} tets) from their "
                      f"initial states, one body per thread on {threads} threads, C oracle -O3 "
                      f"-march=x86-64-v3, FP{best['bits']}, {best['s']:.1f} s",
            "legs": legs}


def _sweep_options(args):
    return (0, 2, 4) if args.local_sweeps is None else (args.local_sweeps,)


def cnet_weights_summary(solver, cnet, state, dt):
    """Statistics of the network's smoothing weights on the first Newton
    system of a step from ``state`` (the solver state is restored)."""
    import torch

    snap = solver.snapshot()
    try:
        st = state.clone()
        solver.bind(st)
        solver.begin_step(dt)
        solver.update_system(dt)
        solver.assemble(dt)
        solver.set_coarse_net(cnet)
        solver.coarse_weights_from_net(dt)
        torch.cuda.synchronize()
        w = solver.buffers["coarse_w_net"][: solver.active_particle_count].double()
        q = torch.quantile(w, torch.tensor([0.05, 0.5, 0.95], dtype=w.dtype, device=w.device)).tolist()
        return {"mean": float(w.mean()), "std": float(w.std()), "p05": q[0], "median": q[1], "p95": q[2]}
    finally:
        solver.set_coarse_net(None)
        solver.restore(snap)


FP32_VALU_PEAK_TFLOPS = 157.3  # published (packed FP32); 78.6 without packed instructions


def assembly_roofline(phase_ms, n_newton, T, V, B, octopus):
    """Element + assembly per Newton iteration: k_element with its vertex part (the
    update_system phase) and k_assemble (the assemble phase, right-hand side included), event
    timed over the timed leg.  Algorithmic bytes = 84 T + 40 V + 36 B (tets,
    DmInv, s, Lame pair per tet; x, x~, mass and g per vertex; 36 B per
    nonzero block of H).  The VALU figure, when the counter summary of this
    workload is present, is the measured FP32 FLOP count of those kernels
    (64 lanes x (ADD + MUL + TRANS + 2 FMA) instructions) over the same time."""
    us = (phase_ms["update_system"] + phase_ms["assemble"]) * 1e3 / max(1, n_newton)
    algo = 84.0 * T + 40.0 * V + 36.0 * B
    gbs = algo / (us * 1e-6) / 1e9
    out = {"kernels": ["k_element (tets and vertices)", "k_assemble"], "bound": "valu", "achieved": gbs,
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "us_per_newton_iter": us,
           "bytes_per_newton_iter": algo}
    f = os.path.join(ROOT, "profiles", "valu_element_assembly_octopus_204k.json")
    if octopus and os.path.exists(f):
        with open(f) as fh:
            prof = json.load(fh)
        flop = float(prof["flop_per_newton_iter"])
        tf = flop / (us * 1e-6) / 1e12
        out.update({"valu_tflops": tf, "valu_peak_tflops": FP32_VALU_PEAK_TFLOPS,
                    "valu_frac": tf / FP32_VALU_PEAK_TFLOPS, "flop_per_newton_iter": flop,
                    "traffic": prof.get("hbm_bytes_per_newton_iter")})
    return out


def spmv_roofline(solver, V, B):
    """The stand-alone SELL-64 SpMV (k_spmv, mfem_spmv) on the last assembled
    system, event timed on the launching stream over graph replays (and, for
    comparison, over back-to-back host launches): algorithmic bytes = 40 B +
    4 (V + 1) + 24 V (block values and column, row pointer, x and y)."""
    import torch

    from mfem_amd import _ffi

    x = torch.randn(solver.max_particles, 3, device=solver.device)
    y = torch.empty_like(x)
    S = solver._S
    stream = solver._stream

    def launch():
        _ffi.call("mfem_spmv", ctypes.byref(S), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                  stream())

    for _ in range(3):
        launch()
    n = 50
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(n):
        launch()
    b.record()
    torch.cuda.synchronize()
    us_eager = a.elapsed_time(b) * 1e3 / n
    # the same launches replayed from a HIP graph: no host launch cost
    # between them (the kernel's own duration plus the device's dispatch)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(n):
            launch()
    g.replay()
    torch.cuda.synchronize()
    a.record()
    for _ in range(4):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    us = a.elapsed_time(b) * 1e3 / (4 * n)
    algo = 40.0 * B + 4.0 * (V + 1) + 24.0 * V
    gbs = algo / (us * 1e-6) / 1e9
    return {"kernel": "k_spmv", "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "us_per_launch": us, "us_per_launch_eager": us_eager,
            "bytes_per_launch": algo, "timing": "HIP events around 4 replays of a graph of 50 launches"}


def mlp_mfma(solver, net, V):
    """Event-timed forward pass of the per-vertex MLP over the body's
    features (the coarse-weight network's launch of every step)."""
    import torch

    X = solver.buffers["cw_x"][:V] if "cw_x" in solver.buffers else solver.buffers["ws_x"][:V]
    for _ in range(3):
        net.forward(X)
    n = 20
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(n):
        net.forward(X)
    b.record()
    torch.cuda.synchronize()
    us = a.elapsed_time(b) * 1e3 / n
    flop = 2.0 * V * (16 * 64 + 64 * 64 + 64 * 3)
    tflops = flop / (us * 1e-6) / 1e12
    return {"kernel": "k_mlp_forward", "bound": "mfma", "achieved": tflops, "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": tflops / BF16_PEAK_TFLOPS, "us_per_launch": us, "flop_per_launch": flop}


def dump_launch_log(path, solver, m, nnz):
    """Every PCG launch of the run in launch order (kernel name, CG
    iterations), to divide per-dispatch hardware counters by iterations."""
    modes = {"standard": 0, "pipelined": 1, "standard_lowreg": 2, "pipelined_coarse": 3}
    names = {}
    for v in solver.pcg_variants():
        names[v["variant"]] = ("k_pcg_lowreg" if v["mode"] == "standard_lowreg" else
                               f"k_pcg<{v['chunks_per_wave']}, {v['resident_slots']}, {modes[v['mode']]}>")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"launches": [[names.get(v, str(v)), n] for v, n in solver.launch_log], "tets": m.tet_count,
                   "verts": m.particle_count, "nnz_blocks": nnz}, f)


def _cpu_info():
    """Host threads this process may use (affinity mask, capped by
    OMP_NUM_THREADS when that is set) and the CPU model name."""
    n = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, n), model


def cpu_baseline(m, dt, state=None, snap=None, steps: int = 1, single_thread: bool = True):
    """The C oracle (oracle/mfem_oracle.c) built -O3 -march=x86-64-v3 with
    OpenMP, timed on ``steps`` steps from the state the timed GPU legs start
    from (positions, velocities, stretch and multipliers copied from the
    device), in FP32 and FP64 on all host threads this process may use, and
    at one thread.  ``value`` is the faster all-threads leg."""
    from oracle import Oracle

    threads, model = _cpu_info()

    def run(bits, nthreads):
        o = Oracle(m, bits, threads=nthreads)
        if state is not None:
            n = m.particle_count
            o.buf("x")[:] = state.particle_q[:n].double().cpu().numpy().ravel()
            o.buf("v")[:] = state.particle_qd[:n].double().cpu().numpy().ravel()
            o.buf("s")[:] = snap["s"][: m.tet_count].double().cpu().numpy().ravel()
            o.buf("lam")[:] = snap["lam"][: m.tet_count].double().cpu().numpy().ravel()
        t0 = time.perf_counter()
        cg = 0
        for _ in range(steps):
            o.step(dt)
            cg += int(o.stats()[:, 0].sum())
        sec = time.perf_counter() - t0
        return {"bits": bits, "threads": o.threads, "value": m.tet_count * steps / sec, "s_per_step": sec / steps,
                "cg_iters_per_step": cg / steps}

    legs = [run(32, threads), run(64, threads)]
    best = max(legs, key=lambda r: r["value"])
    if single_thread:
        legs += [run(32, 1), run(64, 1)]
    return {
        "value": best["value"],
        "unit": "tet-steps/s",
        "cores": best["threads"],
        "kind": "port",
        "cpu": model,
        "sample": f"{steps} step(s) ({int(m.params['iterations'])} Newton iterations, CG to rel-res 1e-6) from the "
                  f"state the timed GPU legs start from, C oracle -O3 -march=x86-64-v3 + OpenMP, "
                  f"FP{best['bits']}, {best['threads']} threads, {best['s_per_step'] * steps:.1f} s",
        "legs": legs,
    }


def _release():
    """Frees the device memory held by torch's caching allocator (the
    solvers are gone with main's frame) before the interpreter tears down
    HIP, so that no device free runs after the runtime or a profiler's tool
    has finalised."""
    import gc

    import torch

    if torch.cuda.is_initialized():
        torch.cuda.synchronize()
        gc.collect()
        torch.cuda.empty_cache()
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
    _release()
