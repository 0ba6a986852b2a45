"""Benchmark: forward+adjoint mass-spring rollout throughput on one MI355X.

Workload (one "step" of this harness = one full training rollout of the
batch): the 3D quadruped (N=54 vertices, S=323 springs, 4 actuator groups,
MLP 179-64-4 tanh) with 1024 target tasks, FP32, horizon 1000 advances,
checkpoints every 50 steps recomputed in the backward pass, windowed COM
loss over the last 500 states, gradients of every MLP weight.  Inputs and
weights are synthetic (seeded), resident in HBM before the timed region.

Printed as ONE JSON line (rank 0):
  value            batched persistent robot-steps/s (forward + adjoint)
  plain            one-robot-per-launch form (reference launch structure,
                   per-step kernels + tape replay) on a sample of robots
  roofline         HBM roofline of the dominant kernel (adjoint rollout)
  cpu_baseline     C/OpenMP restatement (oracle/cpu_rollout.c), FP32, 32
                   tasks x 1000 steps with the same checkpointing and one
                   Adam step, all cores and one thread; the FP64 torch
                   oracle's figure beside it
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "robot-steps/s fwd+bwd batched vs plain; MPM particle-steps/s; p99 latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3          # dense FP32 matrix (MFMA) peak, MI355X
UNEVEN_DASHPOT = 0.03

# HBM bytes moved per launch, measured with rocprofv3 counters on the default
# workloads (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each from its own
# run, scaled by what the same counters read for a 256 MiB device copy:
# FETCH_SIZE x 2.00, WRITE_SIZE x 1.00).  The commands and tables are in
# profiles/ms_rollout_quadruped1024_kernel_stats.md and
# profiles/mpm_quadruped64_8robots_kernel_stats.md.  Reported only when the
# run's configuration is the measured one.
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_KIB = {
    # (kernel, robots, horizon, checkpoint interval): (FETCH_SIZE, WRITE_SIZE)
    ("k_wave_bwd_pipe", 1024, 1000, 50): (3029464.6, 4624975.7),
    ("k_wave_fwd", 1024, 1000, 50): (1250.6, 27251.4),
}
# MLS-MPM, 8 robots x 27 648 particles on 64^3 grids, checkpoints every 100
# steps, replayed segments: FETCH_SIZE and WRITE_SIZE (KiB) of one step of
# every forward, replay and adjoint kernel, summed
MPM_STEP_KIB = {(8, 100): (63832.0, 98505.0)}
MPM_MEASURED_SCATTER = {"scatter": 0, "box_limit": 512, "tile_nodes": 2048, "unroll": 3}


def measured_bytes(fetch_kib, write_kib):
    return (fetch_kib * FETCH_SCALE + write_kib * WRITE_SCALE) * 1024.0


def algorithmic_bytes(N, S, I, H, O, K, R_d, dtype_bytes=4, stored=False):
    """Bytes per robot-step by kernel (state round-trip model): forward reads
    and writes x,v (48N) plus a checkpoint every K steps (24N/K) plus one pass
    over springs (20 B) and weights; the adjoint recomputes the forward (48N)
    and runs the adjoint (read x,v, read/write xbar,vbar: 72N) with two passes
    over springs and weights.  With stored rows the forward also writes each
    step's adjoint rows (x, v, per-vertex impulse sums, network input, hidden
    values and derivatives, outputs and derivatives, COM: 9N + I + 2H + 2O + 6
    reals) and the adjoint reads them instead of recomputing.  Weight/topology
    bytes amortise over the R_d robots sharing a design."""
    P = I * H + H * O + H + O
    s = dtype_bytes / 4.0
    rows = 4 * (9 * N + I + 2 * H + 2 * O + 6)
    fwd = (48 * N + 24 * N / K + (20 * S + 4 * P) / R_d + (rows if stored else 0)) * s
    bwd = ((72 * N + rows if stored else 120 * N) + (40 * S + 8 * P) / R_d) * s
    return fwd, bwd


def mpm_bytes(g_per_p, K, f_full=1.0, f_replay=0.0, save_horizon=False):
    """MLS-MPM bytes per particle-step, forward + adjoint (state round-trip
    model; G = active grid nodes, 16 B per node record): forward
    196 + 56 G/P, plus 32 G/P when it saves each step's grid (momentum/mass
    and velocity records) for the backward; adjoint 292 + 112 G/P.  Segments
    recomputed in the backward are charged for the work they do: the
    fraction f_full in full (the forward again plus the checkpoint read,
    96 / K), the fraction f_replay as the replay (saved velocities read and
    restored, 32 G/P; G2P with the F update: x, C, F read, x, v, C, F
    written, velocities gathered: 180 + 16 G/P; the checkpoint read)."""
    fwd = 196 + 56 * g_per_p + (32 * g_per_p if save_horizon else 0.0)
    return (fwd + 292 + 112 * g_per_p + f_full * (196 + 56 * g_per_p + 96 / K)
            + f_replay * (180 + 48 * g_per_p + 96 / K))


def mlp_ab(B, I, H, O, reps=50):
    """The batched MLP's forward + adjoint (dl_mlp_forward + dl_mlp_backward,
    FP32, the [B x I] x [I x H] layer-1 GEMM of the per-step form) in the
    library's default mix (MFMA forward and weight gradients, VALU input
    adjoint), every stage on the matrix cores, and on the VALU kernels, timed
    with HIP events on the launching stream.  FLOPs per pair: 6 (I H + H O) B
    (forward 2, adjoint 4: input gradient and weight gradient)."""
    import torch
    from diffloco import _native as nat
    g = torch.Generator(device="cpu").manual_seed(5)
    mk = lambda *s_: torch.randn(*s_, generator=g).cuda()
    inp, og = mk(B, I), mk(B, O)
    W = [mk(H, I) / I ** 0.5, 0.1 * mk(H), mk(O, H) / H ** 0.5, 0.1 * mk(O)]
    hid, out = torch.zeros(B, H, device="cuda"), torch.zeros(B, O, device="cuda")
    G = [torch.zeros_like(w) for w in W]
    ing = torch.zeros_like(inp)
    work = torch.zeros(B * (2 * H + O), device="cuda")
    P = nat.ptr
    flops = 6.0 * (I * H + H * O) * B
    res = {"shape": [B, I, H, O], "flops_per_fwd_bwd": flops}
    st = torch.cuda.current_stream()
    for name, on in (("mfma", 1), ("mfma_all_stages", 2), ("valu", 0)):
        nat.call("dl_set_option", nat.OPT_MLP_MFMA, on)

        def pair():
            nat.call("dl_mlp_forward", nat.F32, nat.ACT_TANH, B, I, H, O, P(inp), *map(P, W),
                     P(hid), P(out), nat.stream())
            nat.call("dl_mlp_backward", nat.F32, nat.ACT_TANH, B, I, H, O, P(inp), *map(P, W),
                     P(og), P(ing), *map(P, G), P(work), nat.stream())
        for _ in range(5):
            pair()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            pair()
        e1.record(st)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        res[name] = {"us_per_fwd_bwd": us, "tflops": flops / (us * 1e-6) / 1e12,
                     "frac_of_fp32_peak": flops / (us * 1e-6) / 1e12 / FP32_PEAK_TFLOPS}
    nat.call("dl_set_option", nat.OPT_MLP_MFMA, 1)
    res["mfma_speedup"] = res["valu"]["us_per_fwd_bwd"] / res["mfma"]["us_per_fwd_bwd"]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--robots", type=int, default=1024)
    ap.add_argument("--horizon", type=int, default=1000)
    ap.add_argument("--ckpt", type=int, default=50)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--variant", type=int, default=3,
                    help="mass-spring kernel variant (0 auto, 3 checkpoints every --ckpt steps "
                         "and pipelined segment recompute, 4 every step's rows kept by the "
                         "forward)")
    ap.add_argument("--plain-robots", type=int, default=2)
    ap.add_argument("--skip-per-step", action="store_true",
                    help="skip the per-step batched form (reference launch structure at the "
                         "full batch: per-step kernels over all tasks, MLP GEMMs on MFMA)")
    ap.add_argument("--skip-mlp", action="store_true", help="skip the MLP MFMA / VALU A/B")
    ap.add_argument("--cpu-robots", type=int, default=32,
                    help="robots of the FP64 torch oracle's CPU figure")
    ap.add_argument("--cpu-tasks", type=int, default=32,
                    help="tasks of the C/OpenMP FP32 CPU baseline")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-mpm", action="store_true")
    ap.add_argument("--skip-alt", action="store_true",
                    help="skip the comparison runs of the other memory form (stored rows / "
                         "states against checkpoints and segment recompute)")
    ap.add_argument("--mpm-robots", type=int, default=8)
    ap.add_argument("--mpm-horizon", type=int, default=2000)
    ap.add_argument("--mpm-ckpt", type=int, default=100)
    ap.add_argument("--mpm-steps", type=int, default=1)
    ap.add_argument("--mpm-tune", type=int, default=1,
                    help="1: the autotuner picks the MPM scatter configuration; 0: "
                         "--mpm-scatter / --mpm-tile with box limit 512, unroll 3")
    ap.add_argument("--mpm-scatter", type=int, default=0,
                    help="0: global-atomic commits (staged run sums, LDS tile fallback), "
                         "1: tiles stored and gathered by the grid pass")
    ap.add_argument("--mpm-tile", type=int, default=2048, help="LDS scatter tile, nodes")
    ap.add_argument("--mpm-store-states", type=int, default=0,
                    help="1: the MPM forward keeps every step's particle state (adjoint without "
                         "replay); 0: checkpoints every --mpm-ckpt steps")
    ap.add_argument("--skip-uneven", action="store_true")
    ap.add_argument("--uneven-robots", type=int, default=512)
    ap.add_argument("--uneven-iters", type=int, default=3)
    ap.add_argument("--uneven-stored", type=int, default=0,
                    help="1: uneven forwards keep every step's adjoint rows; 0: checkpoints "
                         "every 50 steps")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--serve-rates", default="50,200,1000,2000",
                    help="Poisson arrival rates (per second) of the serving sweep")
    ap.add_argument("--serve-requests", type=int, default=120,
                    help="requests per sweep point (the same robots at every rate)")
    ap.add_argument("--serve-p99-bound", type=float, default=2.0,
                    help="p99 latency bound (s) the sweep reports the highest rate for")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--skip-serial", action="store_true")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 4 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not supported "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibrate:
        import torch
        n = args.copy_calibrate << 18          # float32 elements in MIB mebibytes
        src = torch.ones(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": n * 4, "copies": 4}))
        return

    import numpy as np
    import torch
    from __graft_entry__ import load_package
    dl = load_package()
    from diffloco.rollout import MassSpringBatch, init_weights
    from diffloco.solver_mass_spring import _physics
    from diffloco.taichi_env import sensor_layout

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    R, T, K, W = args.robots, args.horizon, args.ckpt, 500
    design = dl.build_quadruped()
    qc = dl.quadruped_config(batch_size=R, steps=T, window=W).get_config()
    lay = sensor_layout(qc)
    N, S, O, H = design.n_vert, design.n_spring, design.n_act, 64
    I = qc["nn"]["n_input_states"]
    rng = np.random.default_rng(1)
    tv = np.stack([rng.uniform(-0.08, 0.08, R), np.zeros(R), rng.uniform(-0.08, 0.08, R)], 1)
    th = rng.uniform(0.1, 0.25, R)
    tg = np.concatenate([tv, th[:, None], np.zeros((R, 1))], 1)
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    batch = MassSpringBatch([design], [0] * R, T, _physics(qc), lay, tg, [x0] * R, act="tanh",
                            ckpt_every=K, dtype=torch.float32, block=args.block, loss_window=W,
                            variant=args.variant)
    w = init_weights(I, H, O, seed=0)
    batch.set_weights(0, *w)
    lg = np.full(R, 1.0 / R)

    def one_step():
        batch.reset()
        batch.forward()
        batch.backward(loss_grad=lg)

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    fwd_ms, bwd_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.reset()
        ev[0].record()
        batch.forward()
        ev[1].record()
        ev[2].record()
        batch.backward(loss_grad=lg)
        ev[3].record()
        torch.cuda.synchronize()
        fwd_ms.append(ev[0].elapsed_time(ev[1]))
        bwd_ms.append(ev[2].elapsed_time(ev[3]))
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    robot_steps = R * T
    value = robot_steps * args.steps / wall
    stored = batch.variant == 4
    f_b, b_b = algorithmic_bytes(N, S, I, H, O, batch.K, R, stored=stored)
    bwd_avg_ms = float(np.mean(bwd_ms))
    fwd_avg_ms = float(np.mean(fwd_ms))
    achieved_bwd = b_b * robot_steps / (bwd_avg_ms * 1e-3) / 1e9
    achieved_fwd = f_b * robot_steps / (fwd_avg_ms * 1e-3) / 1e9
    # which kernels the launches used: variant 0 picks the wave kernels with the
    # pipelined adjoint when the robot fits them (as here), else the block kernels
    wave = batch.variant in (2, 3, 4) or (batch.variant == 0 and batch.wave_eligible())
    kname = "k_wave" if wave else "k_rollout"
    kname_bwd = kname + ("_bwd_pipe" if wave and batch.variant != 2 else "_bwd")
    loss = float(batch.loss.double().mean().item())
    gnorm = float(batch.grad_out.double().norm().item())

    # ---- the same rollout in the other memory form: checkpoints every K steps
    # with each segment recomputed in the adjoint (variant 3) against every
    # step's rows kept by the forward (variant 4)
    alt = None
    alt_variant = {3: 4, 4: 3}.get(batch.variant)
    if alt_variant is not None and not args.skip_alt:
        g_main = batch.grad_out.double().clone()
        sc_main = batch.scratch.numel() * batch.scratch.element_size()
        cbat = MassSpringBatch([design], [0] * R, T, _physics(qc), lay, tg, [x0] * R, act="tanh",
                               ckpt_every=K, dtype=torch.float32, block=args.block,
                               loss_window=W, variant=alt_variant)
        cbat.set_weights(0, *w)
        cbat.reset()
        cbat.forward()
        cbat.backward(loss_grad=lg)
        torch.cuda.synchronize()
        cf, cbk = [], []
        t1 = time.perf_counter()
        for _ in range(args.steps):
            cbat.reset()
            ev[0].record()
            cbat.forward()
            ev[1].record()
            cbat.backward(loss_grad=lg)
            ev[3].record()
            torch.cuda.synchronize()
            cf.append(ev[0].elapsed_time(ev[1]))
            cbk.append(ev[1].elapsed_time(ev[3]))
        cwall = time.perf_counter() - t1
        gc = cbat.grad_out.double()
        alt = {"variant": alt_variant, "ckpt_every": cbat.K,
               "value": robot_steps * args.steps / cwall,
               "ms_forward": float(np.mean(cf)), "ms_backward": float(np.mean(cbk)),
               "scratch_gb": cbat.scratch.numel() * cbat.scratch.element_size() / 1e9,
               "scratch_gb_main": sc_main / 1e9,
               "grad_rel_diff_vs_main": float((gc - g_main).norm() / g_main.norm())}
        del cbat
        torch.cuda.empty_cache()

    # ---- plain form: per-step kernels, one robot per launch
    plain = None
    if not args.skip_plain and args.plain_robots > 0:
        qc1 = dl.quadruped_config(batch_size=1, steps=T, window=W)
        tr = dl.DiffPhyTrainer(qc1)
        tr.taichi_env.setup_robot()
        tr.nn.set_weights(*[x.float() for x in w])
        tr.simulate_task(T, tv[:1], th[:1])  # warm-up
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for r in range(args.plain_robots):
            tr.simulate_task(T, tv[r:r + 1], th[r:r + 1])
        torch.cuda.synchronize()
        pw = time.perf_counter() - t1
        pv = args.plain_robots * T / pw
        plain = {"value": pv, "unit": "robot-steps/s", "sample": f"{args.plain_robots} robots x {T} steps, "
                 "per-step kernels (center, sensors, MLP, advance) + tape adjoints, B=1 per launch",
                 "ms_per_robot_rollout": pw * 1e3 / args.plain_robots}

    # ---- per-step batched form: the reference's launch structure (one launch
    # per kernel per step over all tasks, tape replay), the per-step MLP GEMMs
    # [R x I] x [I x H] on the matrix cores; one rollout, MFMA then VALU
    per_step = None
    if not args.skip_per_step:
        from diffloco import _native as nat
        qcb = dl.quadruped_config(batch_size=R, steps=T, window=W)
        trb = dl.DiffPhyTrainer(qcb)
        trb.taichi_env.setup_robot()
        trb.nn.set_weights(*[x.float() for x in w])
        per_step = {"robots": R, "horizon": T,
                    "form": "per-step kernels (center, sensors, MLP, advance) over all tasks + "
                            "tape replay of their adjoints"}
        for name, on in (("mfma", 1), ("valu", 0)):
            nat.call("dl_set_option", nat.OPT_MLP_MFMA, on)
            trb.simulate_task(T, tv, th)        # warm-up
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            trb.simulate_task(T, tv, th)
            torch.cuda.synchronize()
            pw = time.perf_counter() - t1
            per_step[name] = {"value": R * T / pw, "unit": "robot-steps/s",
                              "ms_per_rollout": pw * 1e3}
        nat.call("dl_set_option", nat.OPT_MLP_MFMA, 1)
        del trb
        torch.cuda.empty_cache()
    mlp = None if args.skip_mlp else mlp_ab(R, I, H, O)

    # ---- uneven batch: 512 seeded random voxel robots, size classes, HIP graphs
    uneven = None
    if not args.skip_uneven:
        from diffloco import robot_design as rdm
        from diffloco.uneven import UnevenBatch, n_inputs, tune_schedule
        designs_u, hz = rdm.random_robot_batch(args.uneven_robots, seed=3)
        Ru = len(designs_u)
        rngu = np.random.default_rng(4)
        tgu = np.zeros((Ru, 5))
        tgu[:, 0] = rngu.uniform(-0.08, 0.08, Ru)
        tgu[:, 2] = rngu.uniform(-0.08, 0.08, Ru)
        tgu[:, 3] = rngu.uniform(0.1, 0.25, Ru)
        # dashpot impulse 0.03: the no-dt impulse c (va - vb).d d of the
        # reference integrator diverges at c = 0.12 for vertices with many
        # springs (dense voxel bodies), also in the FP64 oracle
        phys_u = _physics(qc)
        phys_u.dashpot = UNEVEN_DASHPOT

        wu = [init_weights(n_inputs(lay, d.n_vert), 64, d.n_act, seed=1000 + r)
              for r, d in enumerate(designs_u)]

        def load_weights(ub):
            for r in range(Ru):
                ub.set_weights(r, *wu[r])

        def uneven_run(stored):
            # robot-to-launch schedule (size-class boundaries), then each
            # class's launch shape, both by the autotuner
            t_tune = time.perf_counter()
            bounds, sched = tune_schedule(designs_u, hz, phys_u, lay, tgu, set_weights=load_weights,
                                          ckpt_every=50, loss_window=100, stored=stored)
            ub = UnevenBatch(designs_u, hz, phys_u, lay, tgu, ckpt_every=50, loss_window=100,
                             stored=stored, bounds=bounds)
            load_weights(ub)
            ub.tune()
            t_tune = time.perf_counter() - t_tune
            ub.prepare(np.full(Ru, 1.0 / Ru))
            ub.run_iteration()
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            ub.run_iteration()
            torch.cuda.synchronize()
            eager_ms = (time.perf_counter() - t3) * 1e3
            ub.capture()
            ub.replay()
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            for _ in range(args.uneven_iters):
                ub.replay()
            torch.cuda.synchronize()
            graph_ms = (time.perf_counter() - t3) * 1e3 / args.uneven_iters
            ub.prepare(np.full(Ru, 1.0 / Ru))
            per_class = ub.class_times()
            rs = ub.robot_steps
            ubytes = 0.0
            for d, T_r in zip(designs_u, hz):
                I_r = n_inputs(lay, d.n_vert)
                f_r, b_r = algorithmic_bytes(d.n_vert, d.n_spring, I_r, 64, d.n_act,
                                             T_r if ub.stored else 50, 1, stored=ub.stored)
                ubytes += (f_r + b_r) * T_r
            uval = rs / (graph_ms * 1e-3)
            nv = np.array([d.n_vert for d in designs_u])
            out = {"value": uval, "unit": "robot-steps/s", "robots": Ru, "robot_steps": rs,
                      "n_vert": [int(nv.min()), float(nv.mean()), int(nv.max())],
                      "horizon": [int(min(hz)), int(max(hz))], "dtype": "f32",
                      "ms_per_iteration_graph": graph_ms, "ms_per_iteration_eager": eager_ms,
                      "autotune_s": t_tune,
                      "classes": [{"name": bk["name"], "robots": len(bk["ids"]), "launch": bk["config"],
                                   "ms_alone": per_class[bk["name"]]}
                                  for bk in ub.buckets],
                      "schedule": {"bounds": list(bounds),
                                   "candidates_ms": [(t["candidate"]["bounds"], t.get("ms"))
                                                     for t in (sched or [])]},
                      "roofline": {"bound": "hbm", "achieved": ubytes / (graph_ms * 1e-3) / 1e9,
                                   "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": ubytes / (graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "algorithmic_bytes_per_iteration": ubytes, "traffic": None},
                      "dashpot_impulse": UNEVEN_DASHPOT,
                      "nonfinite_robots": int((~torch.isfinite(ub.losses())).sum().item()),
                      "mean_loss": float(ub.losses().double().mean().item()), "stored": ub.stored}
            del ub
            torch.cuda.empty_cache()
            return out

        uneven = uneven_run(bool(args.uneven_stored))
        if not args.skip_alt:
            ua = uneven_run(not uneven["stored"])
            uneven["checkpointed" if uneven["stored"] else "stored_rows"] = {
                k: ua[k] for k in ("value", "stored", "ms_per_iteration_graph", "autotune_s",
                                   "classes", "schedule", "nonfinite_robots", "mean_loss")}

    # ---- MLS-MPM: 8 soft quadrupeds, each on its own 64^3 grid
    mpm = None
    if not args.skip_mpm:
        from diffloco import mpm as dm
        Bm, Tm, Km = args.mpm_robots, args.mpm_horizon, args.mpm_ckpt
        pos, ids, n_act = dm.quadruped3d_particles(cells_per_cube=6, n_grid=64)

        scfg = {"scatter": args.mpm_scatter, "tile_nodes": args.mpm_tile or 2048,
                "box_limit": 512, "unroll": 3}
        t_st, tms = 0.0, []
        if args.mpm_tune:
            # scatter configuration of this body and batch, by the autotuner
            from diffloco.autotune import Autotuner
            t5 = time.perf_counter()
            tuner = Autotuner()
            scfg = dm.tune_scatter(dm.quadruped3d_params(64), pos, ids, n_act, Bm,
                                   dtype=torch.float32, K=Km, tuner=tuner)
            t_st = time.perf_counter() - t5
            tms = sorted(t["ms"] for e in tuner.cache.values() if "timings" in e
                         and any("scatter" in t["candidate"] for t in e["timings"])
                         for t in e["timings"] if "ms" in t)

        def mpm_run(store, save_horizon=True):
            mb = dm.MPMBatch(dm.quadruped3d_params(64), pos, ids, n_act, Bm, Tm, K=Km,
                             window=Tm // 2, dtype=torch.float32, store_states=store,
                             save_horizon=save_horizon, **scfg)
            rngm = np.random.default_rng(7)
            mb.set_targets(np.stack([rngm.uniform(-0.08, 0.08, Bm), np.zeros(Bm),
                                     rngm.uniform(-0.08, 0.08, Bm)], 1),
                           rngm.uniform(0.1, 0.25, Bm))
            mw = init_weights(mb.I, 64, n_act, seed=8)
            mb.set_weights(*mw)
            mb.forward()
            mb.backward()
            torch.cuda.synchronize()
            mb.check()
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            mf, mbk = [], []
            for _ in range(args.mpm_steps):
                e0.record()
                mb.forward()
                e1.record()
                mb.backward()
                e2.record()
                torch.cuda.synchronize()
                mb.check()
                mf.append(e0.elapsed_time(e1))
                mbk.append(e1.elapsed_time(e2))
            mb.seq_backward_ms = None
            if mb.overlap_replay:
                # the same backward with every recompute on the main stream
                mb.overlap_replay, seq = False, []
                for _ in range(args.mpm_steps):
                    mb.forward()
                    e1.record()
                    mb.backward()
                    e2.record()
                    torch.cuda.synchronize()
                    seq.append(e1.elapsed_time(e2))
                mb.overlap_replay = True
                mb.seq_backward_ms = float(np.mean(seq))
            return mb, mf, mbk

        mb, mf, mbk = mpm_run(bool(args.mpm_store_states))
        active = mb.active_nodes()
        P = len(pos)
        ms_rollout = float(np.mean(mf) + np.mean(mbk))
        pps = Bm * P * Tm / (ms_rollout * 1e-3)
        gp = active / P
        f_rep = sum(mb.replay) / mb.nck
        bpp = mpm_bytes(gp, Km, 1.0 - f_rep, 0.0 if mb.store_states else f_rep, mb.save_horizon)
        mpm = {"value": pps, "unit": "particle-steps/s", "robots": Bm, "particles": P,
               "grid": 64, "horizon": Tm, "ckpt_every": Km, "dtype": "f32",
               "scatter": ["staged run sums + per-node gather, global atomics",
                           "tile store + gather"][scfg["scatter"]],
               "scatter_config": scfg, "scatter_tune_s": t_st,
               "scatter_tune_ms_best_median_worst": ([tms[0], tms[len(tms) // 2], tms[-1]]
                                                     if args.mpm_tune and tms else None),
               "tile_nodes": scfg["tile_nodes"], "scatter_windows": len(mb.windows),
               "store_states": mb.store_states, "replayed_segments": int(sum(mb.replay)),
               "resort_every": Km,
               "ms_forward": float(np.mean(mf)), "ms_backward": float(np.mean(mbk)),
               "overlap_replay": mb.overlap_replay,
               "ms_backward_one_stream": mb.seq_backward_ms,
               "active_nodes_per_particle": gp,
               "roofline": {"bound": "hbm", "achieved": bpp * pps / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": bpp * pps / 1e9 / HBM_PEAK_GBS,
                            "algorithmic_bytes_per_particle_step": bpp,
                            "traffic": (measured_bytes(*MPM_STEP_KIB[(Bm, Km)]) / (Bm * P)
                                        if (Bm, Km) in MPM_STEP_KIB and P == 27648
                                        and not mb.store_states and mb.overlap_replay
                                        and scfg == MPM_MEASURED_SCATTER else None),
                            "traffic_unit": "bytes per particle-step, all MPM kernels"},
               "mean_loss": float(mb.loss.item()), "device_gb": mb.nbytes() / 1e9}
        if not args.skip_alt:
            # the other memory form: every step's particle state stored (the
            # adjoint without replay) against checkpoints every K steps
            main_store = mb.store_states
            del mb
            torch.cuda.empty_cache()
            mc, cf, cb = mpm_run(not main_store)
            mpm["checkpointed" if main_store else "stored_states"] = {
                "value": Bm * P * Tm / ((np.mean(cf) + np.mean(cb)) * 1e-3),
                "store_states": mc.store_states, "ckpt_every": Km,
                "ms_forward": float(np.mean(cf)), "ms_backward": float(np.mean(cb)),
                "replayed_segments": int(sum(mc.replay)),
                "mean_loss": float(mc.loss.item()), "device_gb": mc.nbytes() / 1e9}
            del mc
            torch.cuda.empty_cache()
            # checkpoints only (no per-step grid save): every segment
            # recomputed in full in the backward, memory O(T / K)
            mr, rf, rb = mpm_run(False, save_horizon=False)
            rpps = Bm * P * Tm / ((np.mean(rf) + np.mean(rb)) * 1e-3)
            rbpp = mpm_bytes(gp, Km, 1.0, 0.0, False)
            mpm["recompute"] = {
                "value": rpps, "store_states": False, "save_horizon": False, "ckpt_every": Km,
                "ms_forward": float(np.mean(rf)), "ms_backward": float(np.mean(rb)),
                "algorithmic_bytes_per_particle_step": rbpp,
                "frac": rbpp * rpps / 1e9 / HBM_PEAK_GBS,
                "mean_loss": float(mr.loss.item()), "device_gb": mr.nbytes() / 1e9}
            del mr
            torch.cuda.empty_cache()

    # ---- request loop: Poisson arrivals, 70 % random mass-spring robots, 30 % MPM bodies
    serve = None
    if not args.skip_serve:
        from diffloco import serving as sv
        phys_s = _physics(qc)
        phys_s.dashpot = UNEVEN_DASHPOT

        def serve_run(rate, max_active):
            reqs = sv.make_requests(rate, None, seed=4, layout=lay, count=args.serve_requests)
            ex = sv.default_executors(phys_s, lay)
            srv = sv.Server(ex, budget_bytes=int(args.serve_budget_gb * 2 ** 30), chunk=50,
                            max_active=max_active)
            t4 = time.perf_counter()
            done = srv.run(reqs)
            wall_s = time.perf_counter() - t4
            st = sv.latency_stats(done, wall_s)
            st.update({"rate_per_s": rate, "arrival_window_s": max(r.arrival for r in reqs),
                       "wall_s": wall_s, "ticks": srv.ticks,
                       "rejected": len(srv.rejected),
                       "mpm_requests": sum(r.kind == "mpm" for r in done),
                       "grad_requests": sum(r.grad for r in done),
                       "robot_steps_per_s": st["robot_steps"] / wall_s})
            del ex, srv
            torch.cuda.empty_cache()
            return st

        rates = [float(r) for r in args.serve_rates.split(",")]
        sweep = [serve_run(r, None) for r in rates]
        ok = [p["rate_per_s"] for p in sweep if p["p99_s"] <= args.serve_p99_bound]
        serve = {"requests_per_point": args.serve_requests,
                 "hbm_budget_gb": args.serve_budget_gb, "tick_steps": 50,
                 "mix": "70 % random voxel mass-spring robots (horizons 100-1000), 30 % MPM "
                        "bodies of 8k-60k particles on 32^3 / 64^3 grids (horizons 200-2000), "
                        "half of each with gradients",
                 "sweep": sweep, "p99_bound_s": args.serve_p99_bound,
                 "max_rate_within_p99_bound": max(ok) if ok else None,
                 "sustained_req_s": max(p["throughput_req_s"] for p in sweep)}
        if not args.skip_serial:
            # one request at a time at the highest rate (both forms work-bound)
            one = serve_run(rates[-1], 1)
            serve["one_at_a_time"] = one
            serve["throughput_vs_one_at_a_time"] = (sweep[-1]["throughput_req_s"]
                                                    / one["throughput_req_s"])

    # ---- CPU baseline: the C/OpenMP restatement (FP32) and the FP64 torch oracle
    cpu = None
    if not args.skip_cpu:
        from oracle import builders as ob
        from oracle import mass_spring as om
        from oracle.cpu_rollout import CpuRollout
        objects, springs, groups = ob.build_quadruped()
        rob = om.make_robot(objects, springs, groups)
        # the C/OpenMP restatement in FP32 on the small training configuration:
        # 32 tasks x 1000 steps, checkpoints every 50 steps recomputed in the
        # backward, one Adam step per iteration; median of 5 after a warm-up
        Bc, Tc = args.cpu_tasks, T
        cfg = om.TaskConfig(steps=Tc, window=min(W, Tc))
        x0c = np.repeat(om.initial_state(rob, 1)[0].numpy(), Bc, 0)
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)

        def c_iters(nthreads, iters=5):
            cr = CpuRollout(rob, cfg, K=K, dtype="f32")
            cr.set_weights(*[x.float().cpu().numpy() for x in w])
            times = []
            for it in range(iters + 1):
                t2 = time.perf_counter()
                _, _, gcpu = cr.run(x0c, tv[:Bc], th[:Bc], threads=nthreads)
                cr.adam(gcpu)
                if it:
                    times.append(time.perf_counter() - t2)
            return Bc * Tc / float(np.median(times))

        v_all, v_one = c_iters(threads), c_iters(1)
        # the FP64 torch restatement (autograd, no checkpointing) beside it
        Bt = args.cpu_robots
        cfg64 = om.TaskConfig(steps=Tc, window=min(W, Tc))
        w64 = [x.double().cpu().clone().requires_grad_() for x in w]
        x0o, v0o = om.initial_state(rob, Bt)
        t2 = time.perf_counter()
        lo, _, _ = om.rollout_task(rob, cfg64, w64, torch.tensor(tv[:Bt]), torch.tensor(th[:Bt]),
                                   x0o, v0o)
        lo.backward()
        cw = time.perf_counter() - t2
        cpu = {"value": v_all, "unit": "robot-steps/s", "cores": threads, "kind": "port",
               "sample": f"C/OpenMP restatement (oracle/cpu_rollout.c), FP32, {Bc} tasks x {Tc} "
                         f"steps, checkpoints every {K} steps recomputed in the backward, one "
                         f"Adam step per iteration, median of 5 iterations after 1 warm-up",
               "one_thread": v_one,
               "fp64_torch_oracle": {"value": Bt * Tc / cw, "unit": "robot-steps/s",
                                     "cores": torch.get_num_threads(),
                                     "sample": f"FP64 torch-CPU oracle, {Bt} robots x {Tc} "
                                               "steps, forward + autograd backward (no "
                                               "checkpointing)"}}

    key = (R, T, batch.K) if batch.variant == 3 and not stored else None
    mk = {k[0]: v for k, v in MEASURED_KIB.items() if key and k[1:] == key}
    traffic_bwd = measured_bytes(*mk[kname_bwd]) if kname_bwd in mk else None
    traffic_fwd = measured_bytes(*mk[kname + "_fwd"]) if kname + "_fwd" in mk else None
    out = {
        "metric": METRIC, "value": value, "unit": "robot-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded targets, N(0,1/fan_in) weights)",
        "config": {"workload": f"quadruped mass-spring rollout fwd+adjoint, {R} tasks x {T} steps, "
                   f"MLP {I}-{H}-{O} tanh, "
                   + ("every step's adjoint rows kept by the forward (no recompute)" if stored
                      else f"checkpoints every {batch.K} steps, segments recomputed"),
                   "robots": R, "horizon": T, "ckpt_every": batch.K, "variant": batch.variant,
                   "n_vert": N, "n_spring": S, "mlp": [I, H, O], "block": args.block,
                   "form": "batched persistent (one launch forward, one launch adjoint)"},
        "roofline": {"bound": "hbm", "achieved": achieved_bwd, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_bwd / HBM_PEAK_GBS, "traffic": traffic_bwd,
                     "algorithmic_bytes_per_launch": b_b * robot_steps,
                     "kernel": kname_bwd, "algorithmic_bytes_per_robot_step": b_b,
                     "avg_launch_ms": bwd_avg_ms,
                     "forward": {"kernel": kname + "_fwd", "achieved": achieved_fwd,
                                 "avg_launch_ms": fwd_avg_ms, "traffic": traffic_fwd,
                                 "algorithmic_bytes_per_robot_step": f_b}},
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "per_step_batched": per_step,
        "mlp": mlp,
        "cpu_baseline": cpu,
        "mpm": mpm,
        "uneven": uneven,
        "serve": serve,
        "check": {"mean_loss": loss, "grad_norm": gnorm},
        "stored_rows" if alt_variant == 4 else "checkpointed": alt,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
