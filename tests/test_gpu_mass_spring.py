"""HIP mass-spring path against the FP64 oracle (per-step kernels and the
persistent batched rollout).  Tolerances: FP64 kernels agree with the oracle
to rounding (1e-10 relative); FP32 positions/velocities within 1e-5 relative
per step, loss within 1e-4, gradients within 1e-3 (relative, L2)."""

import ctypes
import sys

import numpy as np
import pytest
import torch

import diffloco as dl
from diffloco import _native as nat
from diffloco.rollout import MassSpringBatch, rollout_loss
from diffloco.solver_mass_spring import DeviceTopology, _physics
from diffloco.taichi_env import sensor_layout
from oracle import mass_spring as om

from helpers import (from_soa, max_rel, oracle_quadruped, oracle_reference_robot, perturbed_state,
                     rel_err, task_targets, to_soa)

pytestmark = pytest.mark.gpu
DT = {"f64": torch.float64, "f32": torch.float32}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _phys_ref():
    return om.Physics()


def _phys_quad():
    return om.TaskConfig().physics


def _nat_phys(p):
    return nat.MSPhysics(dt=p.dt, gravity=p.gravity, ground=p.ground, drag=p.drag,
                         dashpot=p.dashpot, friction=p.friction)


CASES = [("ref", 100), ("quad", None)]


def _design(kind):
    if kind == "ref":
        return dl.build_reference(100), oracle_reference_robot(100), _phys_ref()
    return dl.build_quadruped(), oracle_quadruped(), _phys_quad()


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", ["ref", "quad"])
def test_advance_matches_oracle(kind, prec):
    _say(f"advance {kind} {prec}")
    design, rob, phys = _design(kind)
    B, dt = 6, DT[prec]
    x, v = perturbed_state(rob, B, seed=3, ground=phys.ground)
    act = torch.rand(B, design.n_act, dtype=torch.float64, generator=torch.Generator().manual_seed(4)) * 2 - 1
    topo = DeviceTopology(design, dt, "cuda", 0.1)
    xd, vd, ad = to_soa(x, dt), to_soa(v, dt), act.to("cuda", dt)
    x1, v1 = torch.empty_like(xd), torch.empty_like(vd)
    nat.call("dl_ms_advance", nat.dtype_code(dt), ctypes.byref(topo.c), ctypes.byref(_nat_phys(phys)),
             B, nat.ptr(xd), nat.ptr(vd), nat.ptr(ad), nat.ptr(x1), nat.ptr(v1), nat.stream())
    torch.cuda.synchronize()
    # oracle fed with the same (rounded) inputs
    xo, vo = from_soa(xd), from_soa(vd)
    ex, ev = om.step(xo, vo, ad.double().cpu(), rob, phys)
    tol = 1e-11 if prec == "f64" else 1e-5
    assert max_rel(from_soa(x1), ex, 1.0) < tol
    assert max_rel(from_soa(v1), ev, 1.0) < tol * 10


@pytest.mark.parametrize("kind", ["ref", "quad"])
def test_advance_grad_matches_autograd(kind):
    _say(f"advance_grad {kind}")
    design, rob, phys = _design(kind)
    B, dt = 4, torch.float64
    x, v = perturbed_state(rob, B, seed=5, ground=phys.ground)
    g = torch.Generator().manual_seed(6)
    act = torch.rand(B, design.n_act, dtype=dt, generator=g) * 2 - 1
    xb1 = torch.randn(x.shape, dtype=dt, generator=g)
    vb1 = torch.randn(x.shape, dtype=dt, generator=g)
    topo = DeviceTopology(design, dt, "cuda", 0.1)
    xg, vg = torch.zeros(B, 3, design.n_vert, dtype=dt, device="cuda"), torch.zeros(B, 3, design.n_vert, dtype=dt, device="cuda")
    ag = torch.zeros(B, design.n_act, dtype=dt, device="cuda")
    keep = [to_soa(x), to_soa(v), act.cuda(), to_soa(xb1), to_soa(vb1)]   # alive across the launch
    nat.call("dl_ms_advance_grad", nat.F64, ctypes.byref(topo.c), ctypes.byref(_nat_phys(phys)), B,
             *[nat.ptr(t) for t in keep], nat.ptr(xg), nat.ptr(vg), nat.ptr(ag), nat.stream())
    torch.cuda.synchronize()
    xr, vr, ar = x.clone().requires_grad_(), v.clone().requires_grad_(), act.clone().requires_grad_()
    ex, ev = om.step(xr, vr, ar, rob, phys)
    (ex * xb1).sum().add((ev * vb1).sum()).backward()
    assert rel_err(from_soa(xg), xr.grad) < 1e-10
    assert rel_err(from_soa(vg), vr.grad) < 1e-10
    assert rel_err(ag.cpu(), ar.grad) < 1e-10


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("act", ["sin", "tanh"])
def test_mlp_forward_backward(prec, act):
    _say(f"mlp {prec} {act}")
    dt = DT[prec]
    B, I, H, O = 5, 179, 64, 4
    g = torch.Generator().manual_seed(7)
    inp = torch.randn(B, I, dtype=torch.float64, generator=g)
    W = om.init_weights(I, H, O, seed=8)
    W[1] = 0.1 * torch.randn(H, dtype=torch.float64, generator=g)
    W[3] = 0.1 * torch.randn(O, dtype=torch.float64, generator=g)
    og = torch.randn(B, O, dtype=torch.float64, generator=g)
    Wd = [w.to("cuda", dt).contiguous() for w in W]
    hid = torch.zeros(B, H, dtype=dt, device="cuda")
    out = torch.zeros(B, O, dtype=dt, device="cuda")
    code = nat.ACT_SIN if act == "sin" else nat.ACT_TANH
    ind = inp.to("cuda", dt)
    nat.call("dl_mlp_forward", nat.dtype_code(dt), code, B, I, H, O, nat.ptr(ind),
             *[nat.ptr(w) for w in Wd], nat.ptr(hid), nat.ptr(out), nat.stream())
    Gs = [torch.zeros_like(w) for w in Wd]
    ing = torch.zeros_like(ind)
    work = torch.zeros(B * (2 * H + O), dtype=dt, device="cuda")
    ogd = og.to("cuda", dt)
    nat.call("dl_mlp_backward", nat.dtype_code(dt), code, B, I, H, O, nat.ptr(ind),
             *[nat.ptr(w) for w in Wd], nat.ptr(ogd), nat.ptr(ing),
             *[nat.ptr(x) for x in Gs], nat.ptr(work), nat.stream())
    torch.cuda.synchronize()
    Wr = [w.clone().requires_grad_() for w in W]
    ir = inp.clone().requires_grad_()
    eh, eo = om.mlp_forward(ir, *Wr, act=act)
    (eo * og).sum().backward()
    tol = 1e-12 if prec == "f64" else 2e-5
    assert rel_err(out.cpu(), eo.detach()) < tol
    assert rel_err(hid.cpu(), eh.detach()) < tol
    gtol = 1e-11 if prec == "f64" else 1e-4
    assert rel_err(ing.cpu(), ir.grad) < gtol
    for a, b in zip(Gs, Wr):
        assert rel_err(a.cpu(), b.grad) < gtol


def _mlp_run(B, I, H, O, act, seed, mfma):
    g = torch.Generator().manual_seed(seed)
    inp = torch.randn(B, I, dtype=torch.float64, generator=g)
    W = om.init_weights(I, H, O, seed=seed + 1)
    W[1] = 0.1 * torch.randn(H, dtype=torch.float64, generator=g)
    W[3] = 0.1 * torch.randn(O, dtype=torch.float64, generator=g)
    og = torch.randn(B, O, dtype=torch.float64, generator=g)
    dt = torch.float32
    Wd = [w.to("cuda", dt).contiguous() for w in W]
    ind, ogd = inp.to("cuda", dt), og.to("cuda", dt)
    hid = torch.zeros(B, H, dtype=dt, device="cuda")
    out = torch.zeros(B, O, dtype=dt, device="cuda")
    Gs = [torch.zeros_like(w) for w in Wd]
    ing = torch.full_like(ind, 0.5)                 # the input adjoint accumulates
    work = torch.zeros(B * (2 * H + O), dtype=dt, device="cuda")
    code = nat.ACT_SIN if act == "sin" else nat.ACT_TANH
    nat.call("dl_set_option", nat.OPT_MLP_MFMA, int(mfma))
    try:
        nat.call("dl_mlp_forward", nat.F32, code, B, I, H, O, nat.ptr(ind),
                 *[nat.ptr(w) for w in Wd], nat.ptr(hid), nat.ptr(out), nat.stream())
        nat.call("dl_mlp_backward", nat.F32, code, B, I, H, O, nat.ptr(ind),
                 *[nat.ptr(w) for w in Wd], nat.ptr(ogd), nat.ptr(ing),
                 *[nat.ptr(x) for x in Gs], nat.ptr(work), nat.stream())
        torch.cuda.synchronize()
    finally:
        nat.call("dl_set_option", nat.OPT_MLP_MFMA, 1)
    Wr = [w.clone().requires_grad_() for w in W]
    ir = inp.clone().requires_grad_()
    eh, eo = om.mlp_forward(ir, *Wr, act=act)
    (eo * og).sum().backward()
    return (out, hid, ing - 0.5, Gs), (eo.detach(), eh.detach(), ir.grad, [w.grad for w in Wr])


@pytest.mark.parametrize("B,I,H,O", [(1024, 179, 64, 4), (37, 179, 64, 4), (1, 179, 64, 4),
                                     (300, 38, 64, 4), (70, 240, 48, 16)])
@pytest.mark.parametrize("act", ["tanh", "sin"])
@pytest.mark.parametrize("mode", [1, 2])
def test_mlp_mfma_matches_fp64_and_valu(B, I, H, O, act, mode):
    """The MFMA paths of dl_mlp_forward / dl_mlp_backward (FP32, tiles of 16
    robots; ragged batches, inputs not a multiple of 4, 48 hidden units):
    DL_OPT_MLP_MFMA = 1 (MFMA forward and weight gradients, VALU input
    adjoint) and 2 (every stage on MFMA), against torch FP64 and against the
    VALU kernels (DL_OPT_MLP_MFMA = 0)."""
    _say(f"mlp mfma B={B} I={I} H={H} O={O} {act} mode={mode}")
    got, ref = _mlp_run(B, I, H, O, act, 70, mode)
    valu, _ = _mlp_run(B, I, H, O, act, 70, 0)
    out, hid, ing, G = got
    eo, eh, eig, eG = ref
    assert rel_err(out.cpu(), eo) < 2e-6 and rel_err(hid.cpu(), eh) < 2e-6
    assert rel_err(ing.cpu(), eig) < 2e-5
    for a, b in zip(G, eG):
        assert rel_err(a.cpu(), b) < 2e-5
    for a, b in zip(got[:2] + (got[2],) + tuple(got[3]), valu[:2] + (valu[2],) + tuple(valu[3])):
        assert rel_err(a.cpu(), b.cpu()) < 2e-5


@pytest.mark.parametrize("layout", ["ref", "quad"])
def test_nn_input_and_grad(layout):
    _say(f"nn_input {layout}")
    if layout == "ref":
        cfg = dl.ConfigSim.from_file("tests/golden/sim_quad.json").get_config()
        lay_o = om.SensorLayout()
        rob = oracle_reference_robot()
    else:
        cfg = dl.quadruped_config(batch_size=4).get_config()
        lay_o = om.TaskConfig().layout
        rob = oracle_quadruped()
    lay = sensor_layout(cfg)
    B, N, t = 4, rob.n_vert, 37
    dt = torch.float64
    x, v = perturbed_state(rob, B, seed=9)
    c = (x / N).sum(1)
    g = torch.Generator().manual_seed(10)
    tv = torch.randn(B, 3, dtype=dt, generator=g) * 0.05
    th = torch.rand(B, dtype=dt, generator=g)
    tc = torch.zeros(B, dtype=dt)
    I = lay_o.n_input(N)
    inp = torch.zeros(B, I, dtype=dt, device="cuda")
    keep = [to_soa(x), to_soa(v), c.cuda(), tv.cuda(), th.cuda(), tc.cuda()]
    nat.call("dl_nn_input", nat.F64, ctypes.byref(lay), B, N, t, 0, *[nat.ptr(u) for u in keep],
             nat.ptr(inp), nat.stream())
    torch.cuda.synchronize()
    ref = om.nn_input(t, x, v, c, tv, th, tc, lay_o)
    assert max_rel(inp.cpu(), ref, 1.0) < 1e-12
    # adjoint
    ig = torch.randn(B, I, dtype=dt, generator=g)
    xg = torch.zeros(B, 3, N, dtype=dt, device="cuda")
    vg = torch.zeros_like(xg)
    cg = torch.zeros(B, 3, dtype=dt, device="cuda")
    igd = ig.cuda()
    nat.call("dl_nn_input_grad", nat.F64, ctypes.byref(lay), B, N, nat.ptr(igd), nat.ptr(xg),
             nat.ptr(vg), nat.ptr(cg), nat.stream())
    torch.cuda.synchronize()
    xr, vr, cr = x.clone().requires_grad_(), v.clone().requires_grad_(), c.clone().requires_grad_()
    (om.nn_input(t, xr, vr, cr, tv, th, tc, lay_o) * ig).sum().backward()
    assert rel_err(from_soa(xg), xr.grad) < 1e-12
    assert rel_err(cg.cpu(), cr.grad) < 1e-12
    if layout == "quad":
        assert rel_err(from_soa(vg), vr.grad) < 1e-12


def _ref_small_config(steps, batch, prec):
    import json
    with open("tests/golden/sim_quad.json") as fh:
        raw = json.load(fh)
    raw["process"].update(max_steps=steps, run_period=20, turn_period=60, jump_period=60)
    raw["nn"]["batch_size"] = batch
    raw["precision"] = prec
    return dl.ConfigSim(raw)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_per_step_trainer_matches_oracle(prec):
    """The reference-shaped per-step loop with Tape replay (ref preset)."""
    _say(f"trainer ref {prec}")
    T, B = 130, 3
    config = _ref_small_config(T, B, prec)
    trainer = dl.DiffPhyTrainer(config)
    env = trainer.taichi_env
    env.setup_robot()
    trainer.nn.weights_init(seed=11)
    W = [w.data.double().cpu().clone() for w in (trainer.nn.weights1, trainer.nn.bias1,
                                                  trainer.nn.weights2, trainer.nn.bias2)]
    pool = torch.rand(B * (T // 60 + 1) * 3, dtype=torch.float64,
                      generator=torch.Generator().manual_seed(12))
    trainer.simulate(T, iter=0, max_speed=0.1, max_height=0.1, pool=pool)
    torch.cuda.synchronize()
    loss = env.loss[None]
    # oracle
    rob = oracle_reference_robot()
    cfg = om.RefConfig(steps=T, run_period=20, turn_period=60, jump_period=60,
                       layout=om.SensorLayout(run_period=20))
    tv, th, tc = om.train_targets(pool.numpy(), T, B, 60, 0.1)
    Wr = [w.clone().requires_grad_() for w in W]
    x0, v0 = om.initial_state(rob, B)
    lo, _, _ = om.rollout_ref(rob, cfg, Wr, tv, th, tc, x0, v0)
    lo.backward()
    ltol, gtol = (1e-10, 1e-8) if prec == "f64" else (1e-4, 1e-3)
    assert abs(loss - lo.item()) / abs(lo.item()) < ltol
    G = [trainer.nn.weights1.grad, trainer.nn.bias1.grad, trainer.nn.weights2.grad,
         trainer.nn.bias2.grad]
    for a, b in zip(G, Wr):
        assert rel_err(a.data.cpu(), b.grad) < gtol


def _quad_batch(B, T, W, K, prec, tv, th, store_traj=False, loss_window=None):
    design = dl.build_quadruped()
    qc = dl.quadruped_config(batch_size=B, steps=T, window=W).get_config()
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    tg = np.concatenate([tv, th[:, None], np.zeros((B, 1))], 1)
    return MassSpringBatch([design], [0] * B, T, _physics(qc), sensor_layout(qc), tg, [x0] * B,
                           act="tanh", ckpt_every=K, dtype=DT[prec],
                           loss_window=W if loss_window is None else loss_window,
                           store_traj=store_traj)


def _oracle_task(B, T, W, tv, th, w64):
    rob = oracle_quadruped()
    cfg = om.TaskConfig(steps=T, window=W)
    Wr = [w.clone().requires_grad_() for w in w64]
    x0, v0 = om.initial_state(rob, B)
    lo, lb, _ = om.rollout_task(rob, cfg, Wr, torch.tensor(tv), torch.tensor(th), x0, v0)
    lo.backward()
    return lo.item(), lb.detach(), [w.grad for w in Wr]


@pytest.mark.parametrize("prec,K", [("f64", 16), ("f64", 7), ("f32", 16)])
def test_batched_rollout_matches_oracle(prec, K):
    _say(f"batched quad {prec} K={K}")
    B, T, W = 5, 90, 40
    tv, th = task_targets(B, 13)
    design = dl.build_quadruped()
    I = om.TaskConfig().layout.n_input(design.n_vert)
    w64 = om.init_weights(I, 64, 4, seed=14)
    batch = _quad_batch(B, T, W, K, prec, tv, th)
    wd = [w.to("cuda", DT[prec]).requires_grad_() for w in w64]
    lossv = rollout_loss(batch, [wd])
    lossv.mean().backward()
    torch.cuda.synchronize()
    lo, lb, go = _oracle_task(B, T, W, tv, th, w64)
    ltol, gtol = (1e-10, 1e-8) if prec == "f64" else (1e-4, 1e-3)
    assert abs(lossv.mean().item() - lo) / abs(lo) < ltol
    assert rel_err(lossv.detach().cpu(), lb) < ltol
    for a, b in zip(wd, go):
        assert rel_err(a.grad.cpu(), b) < gtol


def test_batched_chunked_equals_single_launch():
    """Forward in 3 chunks and backward in chunks give the same numbers as one launch."""
    _say("batched chunked")
    B, T, W, K = 3, 70, 30, 10
    tv, th = task_targets(B, 15)
    design = dl.build_quadruped()
    I = om.TaskConfig().layout.n_input(design.n_vert)
    w = [x.to("cuda") for x in om.init_weights(I, 64, 4, seed=16)]
    a = _quad_batch(B, T, W, K, "f64", tv, th)
    b = _quad_batch(B, T, W, K, "f64", tv, th)
    for bt in (a, b):
        bt.set_weights(0, *w)
        bt.reset()
    a.forward()
    a.backward(loss_grad=np.full(B, 1.0 / B))
    for _ in range(3):
        b.forward(max_steps=25)
    while np.any(b.table["t_bwd"] > 0):
        b.backward(loss_grad=np.full(B, 1.0 / B), max_steps=13, reduce=False)
    b.reduce_grads()
    torch.cuda.synchronize()
    # chunking only changes floating-point summation order
    assert rel_err(b.loss.cpu(), a.loss.cpu()) < 1e-14
    assert rel_err(b.grad_out.cpu(), a.grad_out.cpu()) < 1e-12


def test_batched_deterministic():
    _say("batched determinism")
    B, T, W, K = 4, 60, 20, 16
    tv, th = task_targets(B, 17)
    design = dl.build_quadruped()
    I = om.TaskConfig().layout.n_input(design.n_vert)
    w = [x.to("cuda", torch.float32) for x in om.init_weights(I, 64, 4, seed=18)]
    outs = []
    for _ in range(2):
        bt = _quad_batch(B, T, W, K, "f32", tv, th)
        bt.set_weights(0, *w)
        bt.reset()
        bt.forward()
        bt.backward(loss_grad=np.full(B, 1.0 / B))
        torch.cuda.synchronize()
        outs.append((bt.loss.clone(), bt.grad_out.clone()))
    assert torch.equal(outs[0][0], outs[1][0])
    assert torch.equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("K,variant", [(16, 2), (7, 2), (16, 3), (7, 3), (16, 4)])
def test_wave_kernel_matches_oracle(K, variant):
    """Wave-per-robot kernel (FP32, one wavefront per robot) against the FP64 oracle;
    variant 3 runs the adjoint's segment recompute on a second wavefront,
    variant 4 keeps the forward's rows instead of recomputing."""
    _say(f"wave quad K={K} variant={variant}")
    B, T, W = 6, 90, 40
    tv, th = task_targets(B, 21)
    design = dl.build_quadruped()
    I = om.TaskConfig().layout.n_input(design.n_vert)
    w64 = om.init_weights(I, 64, 4, seed=22)
    qc = dl.quadruped_config(batch_size=B, steps=T, window=W).get_config()
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    tg = np.concatenate([tv, th[:, None], np.zeros((B, 1))], 1)
    batch = MassSpringBatch([design], [0] * B, T, _physics(qc), sensor_layout(qc), tg, [x0] * B,
                            act="tanh", ckpt_every=K, dtype=torch.float32, loss_window=W,
                            variant=variant)
    wd = [w.to("cuda", torch.float32).requires_grad_() for w in w64]
    lossv = rollout_loss(batch, [wd])
    lossv.mean().backward()
    torch.cuda.synchronize()
    lo, lb, go = _oracle_task(B, T, W, tv, th, w64)
    assert abs(lossv.mean().item() - lo) / abs(lo) < 1e-4
    assert rel_err(lossv.detach().cpu(), lb) < 1e-4
    for a, b in zip(wd, go):
        assert rel_err(a.grad.cpu(), b) < 1e-3


@pytest.mark.parametrize("budget", [0, 23])
def test_pipelined_adjoint_bit_identical_to_sequential(budget):
    """The adjoint with the segment recompute on a second wavefront per robot
    (variant 3) gives the sequential wave adjoint's (variant 2) gradients and
    adjoint state bit for bit; the one reading the rows the forward kept
    (variant 4, no recompute) gives the same adjoint state bit for bit and the
    same weight gradients up to the grouping of the per-segment sums (it folds
    every 50 steps, not every K): one design, horizons 45..90 mixed inside a
    workgroup (2 to 6 segments of K = 16), a partly filled second workgroup
    (6 robots, 4 per workgroup), whole and step-budgeted backward calls."""
    _say(f"pipelined vs sequential adjoint budget={budget}")
    horizons = [90, 61, 77, 90, 45, 83]
    B, K, W = len(horizons), 16, 40
    tv, th = task_targets(B, 23)
    design = dl.build_quadruped()
    I = om.TaskConfig().layout.n_input(design.n_vert)
    w64 = [w.to("cuda", torch.float32) for w in om.init_weights(I, 64, 4, seed=24)]
    qc = dl.quadruped_config(batch_size=B, steps=max(horizons), window=W).get_config()
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    tg = np.concatenate([tv, th[:, None], np.zeros((B, 1))], 1)
    out = []
    for variant in (2, 3, 4):
        batch = MassSpringBatch([design], [0] * B, horizons, _physics(qc), sensor_layout(qc), tg,
                                [x0] * B, act="tanh", ckpt_every=K, dtype=torch.float32,
                                loss_window=W, variant=variant)
        assert batch.packed4()
        batch.set_weights(0, *w64)
        batch.forward()
        if budget:
            calls = 0
            while batch.table["t_bwd"].max() > 0:
                batch.backward(loss_grad=np.linspace(0.5, 1.5, B), max_steps=budget, reduce=False)
                calls += 1
            assert calls == -(-max(horizons) // budget)
            batch.reduce_grads()
        else:
            batch.backward(loss_grad=np.linspace(0.5, 1.5, B))
        torch.cuda.synchronize()
        out.append(([g.clone() for g in batch.weight_grads(0)], batch.adjoint.clone()))
    ga, aa = out[0]
    assert torch.isfinite(aa).all()
    for v, (gb, ab) in zip((3, 4), out[1:]):
        for x, y in zip(ga, gb):
            assert torch.equal(x, y) if v == 3 else rel_err(y.cpu(), x.cpu()) < 1e-5
        assert torch.equal(aa, ab)


@pytest.mark.parametrize("variant", [2, 3])
def test_wave_recompute_reproduces_forward(variant):
    """The adjoint's segment recompute repeats the forward bit for bit (same
    accumulation orders in layer 1, the spring pass and the incidence
    gathers, with the weights and topology read from LDS instead of
    registers): after the adjoint, the recompute rows of the first segment
    hold exactly the positions the forward stored for steps 0 .. K-1."""
    _say(f"recompute vs forward variant={variant}")
    B, T, W, K = 5, 70, 30, 16
    tv, th = task_targets(B, 25)
    design = dl.build_quadruped()
    N, S, O = design.n_vert, design.n_spring, design.n_act
    I = om.TaskConfig().layout.n_input(N)
    w64 = [w.to("cuda", torch.float32) for w in om.init_weights(I, 64, 4, seed=26)]
    qc = dl.quadruped_config(batch_size=B, steps=T, window=W).get_config()
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    tg = np.concatenate([tv, th[:, None], np.zeros((B, 1))], 1)
    batch = MassSpringBatch([design], [0] * B, T, _physics(qc), sensor_layout(qc), tg, [x0] * B,
                            act="tanh", ckpt_every=K, dtype=torch.float32, loss_window=W,
                            store_traj=True, variant=variant)
    batch.set_weights(0, *w64)
    batch.forward()
    batch.backward(loss_grad=np.full(B, 1.0 / B))
    torch.cuda.synchronize()
    buf_len = K * (9 * N + I + 3 * 64 + 3 * O + 6) + 10 * S + 4   # buffer 0 rows + spill
    for r in range(B):
        o = int(batch.table[r]["traj"])
        fwd = batch.traj[o:o + K * (3 * N + 3)].view(K, 3 * N + 3)[:, :3 * N]
        sc = int(batch.table[r]["scratch"])
        bufs = [batch.scratch[b:b + K * 6 * N].view(K, 6 * N)[:, :3 * N] for b in (sc, sc + buf_len)]
        hits = [torch.equal(fwd, x) for x in bufs]
        assert hits[0] if variant == 2 else any(hits), (r, hits)


@pytest.mark.parametrize("n_cubes", [2, 6])
def test_pipelined_weight_fold_split_by_input_width(n_cubes):
    """The pipelined adjoint (variant 3) splits each segment's weight-gradient
    fold between its two wavefronts by input block (the reversing one takes
    the last block of 64 inputs and the b1, W2, b2 sums).  A chain of 2 cubes
    has 53 inputs (one block: the reversing wavefront folds everything), a
    chain of 6 has 101 (two blocks); both give the sequential adjoint's
    (variant 2) gradients and adjoint state bit for bit.  The robots start
    0.1 above the ground and land after about 80 steps: the horizons run past
    that, or the gradients would be zero (internal forces do not move the
    centre of mass in free fall)."""
    from diffloco import robot_design as rd
    _say(f"pipelined fold split, chain of {n_cubes} cubes")
    design = rd.build_voxels([((i, 0, 0), True, i % 2) for i in range(n_cubes)], n_act=2)
    I = om.TaskConfig().layout.n_input(design.n_vert)
    assert I == {2: 53, 6: 101}[n_cubes]
    horizons = [160, 121, 144, 133]
    B, K, W = len(horizons), 16, 60
    tv, th = task_targets(B, 29)
    w64 = [w.to("cuda", torch.float32) for w in om.init_weights(I, 64, 2, seed=30)]
    qc = dl.quadruped_config(batch_size=B, steps=max(horizons), window=W).get_config()
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    tg = np.concatenate([tv, th[:, None], np.zeros((B, 1))], 1)
    out = []
    for variant in (2, 3):
        batch = MassSpringBatch([design], [0] * B, horizons, _physics(qc), sensor_layout(qc), tg,
                                [x0] * B, act="tanh", ckpt_every=K, dtype=torch.float32,
                                loss_window=W, variant=variant)
        batch.set_weights(0, *w64)
        batch.forward()
        batch.backward(loss_grad=np.linspace(0.5, 1.5, B))
        torch.cuda.synchronize()
        out.append(([g.clone() for g in batch.weight_grads(0)], batch.adjoint.clone()))
    (ga, aa), (gb, ab) = out
    assert all(torch.isfinite(g).all() and g.abs().max() > 0 for g in ga)
    for x, y in zip(ga, gb):
        assert torch.equal(x, y)
    assert torch.equal(aa, ab)


def _uneven(prec, variant, seed=31, block=256, lds_policy=0):
    from diffloco import robot_design as rd
    rng = np.random.default_rng(seed)
    cubes_list, designs = [], []
    for n_cubes, n_groups in ((8, 4), (11, 5), (9, 4)):
        cubes, g = rd.random_voxel_cubes(rng, n_cubes, n_groups)
        cubes_list.append((cubes, g))
        designs.append(rd.build_voxels(cubes, n_act=g))
    horizons = [100, 130, 115]
    W = 50
    qc = dl.quadruped_config(batch_size=3, steps=130, window=W).get_config()
    tv, th = task_targets(3, seed + 1)
    tg = np.concatenate([tv, th[:, None], np.zeros((3, 1))], 1)
    x0 = [d.objects + np.array([0.4, 0.0, 0.0]) for d in designs]
    batch = MassSpringBatch(designs, [0, 1, 2], horizons, _physics(qc), sensor_layout(qc), tg, x0,
                            act="tanh", ckpt_every=16, dtype=DT[prec], loss_window=W,
                            variant=variant, block=block, lds_policy=lds_policy)
    return batch, designs, cubes_list, horizons, W, tv, th


@pytest.mark.parametrize("prec,variant,block,pol", [("f64", 1, 256, 0), ("f32", 1, 256, 0),
                                                    ("f32", 2, 256, 0), ("f32", 1, 512, 4),
                                                    ("f32", 1, 128, 4), ("f64", 5, 256, 0),
                                                    ("f32", 5, 512, 4)])
def test_uneven_batch_matches_oracle(prec, variant, block, pol):
    """Three different random designs (own sizes, MLPs and horizons) in one launch;
    also the block kernel at 512 threads and with everything staged in LDS
    (policy 4), and with the forward keeping every step's rows (variant 5)."""
    _say(f"uneven {prec} variant={variant} block={block} policy={pol}")
    from oracle import builders as ob
    batch, designs, cubes_list, horizons, W, tv, th = _uneven(prec, variant, block=block,
                                                                lds_policy=pol)
    lay = om.TaskConfig().layout
    weights64 = [om.init_weights(lay.n_input(d.n_vert), 64, d.n_act, seed=40 + i)
                 for i, d in enumerate(designs)]
    wd = [[w.to("cuda", DT[prec]).requires_grad_() for w in ws] for ws in weights64]
    lossv = rollout_loss(batch, wd)
    lossv.sum().backward()
    torch.cuda.synchronize()
    ltol, gtol = (1e-10, 1e-8) if prec == "f64" else (1e-4, 1e-3)
    for r, ((cubes, g), T) in enumerate(zip(cubes_list, horizons)):
        objects, springs, groups = ob.build_voxels(cubes)
        rob = om.make_robot(objects, springs, groups)
        cfg = om.TaskConfig(steps=T, window=W)
        Wr = [w.clone().requires_grad_() for w in weights64[r]]
        x0, v0 = om.initial_state(rob, 1)
        lo, _, _ = om.rollout_task(rob, cfg, Wr, torch.tensor(tv[r:r + 1]), torch.tensor(th[r:r + 1]), x0, v0)
        lo.backward()
        assert abs(lossv[r].item() - lo.item()) / abs(lo.item()) < ltol
        for a, b in zip(wd[r], Wr):
            assert rel_err(a.grad.cpu(), b.grad) < gtol


def test_reference_preset_through_batched_trajectory():
    """sim_quad semantics through the persistent kernel: T-1 advances,
    trajectory rows stored, velocity + rotation losses evaluated on them by
    the loss kernels, their adjoint fed back as external trajectory gradients."""
    _say("ref preset batched")
    T, B = 130, 2
    cfg = om.RefConfig(steps=T, run_period=20, turn_period=60, jump_period=60,
                       layout=om.SensorLayout(run_period=20))
    rob = oracle_reference_robot()
    design = dl.build_reference(100)
    pool = torch.rand(B * (T // 60 + 1) * 3, dtype=torch.float64, generator=torch.Generator().manual_seed(50))
    tv, th, tc = om.train_targets(pool.numpy(), T, B, 60, 0.1)
    W = om.init_weights(838, 64, 272, seed=51)
    W = [(torch.rand(w.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(52 + i)) * 2 - 1) * 0.05
         if i in (0, 2) else w for i, w in enumerate(W)]
    Wr = [w.clone().requires_grad_() for w in W]
    x0, v0 = om.initial_state(rob, B)
    lo, _, _ = om.rollout_ref(rob, cfg, Wr, tv, th, tc, x0, v0)
    lo.backward()
    # product: per-step targets [T, 5] per robot
    raw = dl.ConfigSim.from_file("tests/golden/sim_quad.json").get_config()
    raw["process"]["run_period"] = 20
    from diffloco import taichi_env as te
    lay = te.sensor_layout(raw)
    tg = [np.concatenate([tv[:, k].numpy(), th[:, k, None].numpy(), tc[:, k, None].numpy()], 1)
          for k in range(B)]
    xs0 = [design.objects + np.array([0.4, 0.0, 0.0])] * B
    batch = MassSpringBatch([design], [0] * B, T - 1, _physics(raw), lay, tg, xs0, act="sin",
                            ckpt_every=16, dtype=torch.float64, loss_window=0, store_traj=True)
    for i, w in enumerate((W[0], W[1], W[2], W[3])):
        pass
    batch.set_weights(0, *[w.cuda() for w in W])
    batch.reset()
    batch.forward()
    # external losses on the stored trajectory, with the loss kernels
    N = design.n_vert
    xs = torch.zeros(T, B, 3, N, dtype=torch.float64, device="cuda")
    cs = torch.zeros(T, B, 3, dtype=torch.float64, device="cuda")
    for k in range(B):
        xk, ck = batch.trajectory(k)
        xs[:, k] = xk.transpose(1, 2)
        cs[:T - 1, k] = ck[:T - 1]
    tvd, thd = tv.cuda().contiguous(), th.cuda().contiguous()
    loss = torch.zeros(2, dtype=torch.float64, device="cuda")
    lb = torch.zeros(B, dtype=torch.float64, device="cuda")
    work = torch.zeros(B, dtype=torch.float64, device="cuda")
    nat.call("dl_loss_velocity", nat.F64, B, T, T - 1, T, 20, 60, nat.ptr(cs), nat.ptr(tvd),
             nat.ptr(loss[0:1]), nat.ptr(lb), nat.stream())
    nat.call("dl_loss_rotation", nat.F64, B, N, T, T - 1, T, 20, 60, nat.ptr(xs), nat.ptr(cs),
             nat.ptr(thd), nat.ptr(loss[1:2]), nat.ptr(lb), nat.ptr(work), nat.stream())
    xg = torch.zeros_like(xs)
    cg = torch.zeros_like(cs)
    nat.call("dl_loss_velocity_grad", nat.F64, B, T, T - 1, T, 20, 60, nat.ptr(cs), nat.ptr(tvd),
             1.0, nat.ptr(cg), nat.stream())
    nat.call("dl_loss_rotation_grad", nat.F64, B, N, T, T - 1, T, 20, 60, nat.ptr(xs), nat.ptr(cs),
             nat.ptr(thd), 1.0, nat.ptr(xg), nat.ptr(cg), nat.stream())
    torch.cuda.synchronize()
    assert abs(loss.sum().item() - lo.item()) / abs(lo.item()) < 1e-10
    tgrad = torch.zeros_like(batch.traj)
    for k in range(B):
        o = int(batch.table[k]["traj"])
        rows = tgrad[o:o + T * (3 * N + 3)].view(T, 3 * N + 3)
        rows[:, :3 * N] = xg[:, k].reshape(T, 3 * N)
        rows[:, 3 * N:] = cg[:, k]
    batch.traj_grad = tgrad
    batch.backward(loss_grad=np.zeros(B))
    torch.cuda.synchronize()
    G = batch.weight_grads(0)
    for a, b in zip(G, Wr):
        assert rel_err(a.cpu(), b.grad) < 1e-8


def test_batched_gradient_matches_central_differences():
    """Adjoint of the persistent rollout (FP64, quadruped, contact reached)
    against FP64 central differences, eps = 1e-4, on 16 random MLP weights
    (seed 2)."""
    _say("batched gradient vs central differences")
    from diffloco.rollout import init_weights
    B, T, W, K = 2, 150, 60, 16
    design = dl.build_quadruped()
    qc = dl.quadruped_config(batch_size=B, steps=T, window=W, precision="f64").get_config()
    tv, th = task_targets(B, 3)
    tg = np.concatenate([tv, th[:, None], np.zeros((B, 1))], 1)
    x0 = design.objects + np.array([0.4, 0.0, 0.0])
    batch = MassSpringBatch([design], [0] * B, T, _physics(qc), sensor_layout(qc), tg, [x0] * B,
                            act="tanh", ckpt_every=K, dtype=torch.float64, loss_window=W)
    I = batch.n_in[0]
    w = init_weights(I, 64, 4, seed=0, dtype=torch.float64)

    def loss_of(ws):
        batch.set_weights(0, *ws)
        batch.reset()
        batch.forward()
        return batch.loss.double().mean().item()

    loss_of(w)
    batch.backward(loss_grad=np.full(B, 1.0 / B))
    grads = [g.clone().double().cpu() for g in batch.weight_grads(0)]
    gmax = max(g.abs().max().item() for g in grads)
    assert gmax > 0
    rng = np.random.default_rng(2)
    eps = 1e-4
    for _ in range(16):
        pi = int(rng.integers(4))
        j = int(rng.integers(w[pi].numel()))
        wp = [x.clone() for x in w]
        wm = [x.clone() for x in w]
        wp[pi].view(-1)[j] += eps
        wm[pi].view(-1)[j] -= eps
        fd = (loss_of(wp) - loss_of(wm)) / (2 * eps)
        ad = grads[pi].view(-1)[j].item()
        assert abs(fd - ad) <= 1e-4 * abs(ad) + 1e-6 * gmax, (pi, j, fd, ad)


def test_training_iterations_match_oracle():
    """Three training iterations of the reference-shaped loop (rollout under
    the tape, TNS, Adam with b1 = b2 = 0.9 and iteration-based bias
    correction) against the oracle's rollout + autograd + Adam, FP64: the
    loss curve, TNS and final weights (diffphy_trainer.py:277-302,
    nn.py:173-191)."""
    _say("training iterations")
    T, B = 130, 2
    config = _ref_small_config(T, B, "f64")
    trainer = dl.DiffPhyTrainer(config)
    env, nn = trainer.taichi_env, trainer.nn
    env.setup_robot()
    nn.weights_init(seed=31)
    nn.clear_adam()
    W = [w.data.double().cpu().clone() for w in (nn.weights1, nn.bias1, nn.weights2, nn.bias2)]
    ms = [torch.zeros_like(w) for w in W]
    vs = [torch.zeros_like(w) for w in W]
    rob = oracle_reference_robot()
    cfg = om.RefConfig(steps=T, run_period=20, turn_period=60, jump_period=60,
                       layout=om.SensorLayout(run_period=20))
    lr = nn.learning_rate
    for it in range(3):
        pool = torch.rand(B * (T // 60 + 1) * 3, dtype=torch.float64,
                          generator=torch.Generator().manual_seed(40 + it))
        trainer.simulate(T, iter=it, max_speed=0.1, max_height=0.1, pool=pool)
        tns = nn.get_TNS()
        nn.gradient_update(it)
        torch.cuda.synchronize()
        tv, th, tc = om.train_targets(pool.numpy(), T, B, 60, 0.1)
        Wr = [w.clone().requires_grad_() for w in W]
        x0, v0 = om.initial_state(rob, B)
        lo, _, _ = om.rollout_ref(rob, cfg, Wr, tv, th, tc, x0, v0)
        lo.backward()
        tns_o = sum(float((w.grad ** 2).sum()) for w in Wr)
        for w, m, v, wr in zip(W, ms, vs, Wr):
            om.adam_update(w, wr.grad, m, v, it, lr=lr)
        assert abs(env.loss[None] - lo.item()) <= 1e-10 * abs(lo.item()), it
        assert abs(tns - tns_o) <= 1e-8 * tns_o, it
    for a, b in zip((nn.weights1, nn.bias1, nn.weights2, nn.bias2), W):
        assert rel_err(a.data.double().cpu(), b) < 1e-9
