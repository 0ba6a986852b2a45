"""The reference's height, pose, actuation and crawl losses
(taichi_env.py:213-233, 276-293): the HIP loss kernels and their adjoints
against the oracle's restatement and its autograd, and the per-step trainer
with every loss enabled against the oracle rollout."""

import sys

import numpy as np
import pytest
import torch

import diffloco as dl
from diffloco import _native as nat
from oracle import mass_spring as om

from helpers import oracle_reference_robot, rel_err

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _soa_x(x):
    """[T,B,N,3] -> device [T,B,3,N]."""
    return x.transpose(-1, -2).contiguous().cuda()


@pytest.mark.parametrize("written", [15, 11])
def test_loss_kernels_match_oracle(written):
    """Values, per-robot values and adjoints of the four loss kernels (FP64)
    against the oracle, with rows at and after ``written`` unwritten (they
    read as zero)."""
    _say(f"loss kernels written={written}")
    g = torch.Generator().manual_seed(31 + written)
    T, B, N, O, J, steps = 16, 3, 5, 4, 5, 15
    r = lambda *s: torch.randn(*s, dtype=F64, generator=g)
    height = 0.2 + 0.1 * r(T, B)
    upper = 0.1 + 0.1 * r(T, B)
    x = 0.3 * r(T, B, N, 3)
    obj = 0.3 * r(N, 3)
    act = r(T, B, O)
    tv = 0.05 * r(T, B, 3)
    th = torch.where(torch.rand(T, B, generator=g, dtype=F64) < 0.5, 0.0, 0.2).to(F64)
    tc = torch.where(torch.rand(T, B, generator=g, dtype=F64) < 0.5, 0.0, 1.0).to(F64)
    cut = lambda a: a[:written]
    z = lambda *s: torch.zeros(*s, dtype=F64, device="cuda")
    dev = [t.cuda() for t in (height, upper, act, tv, th, tc, obj)]
    hd, ud, ad, tvd, thd, tcd, objd = dev
    xd = _soa_x(x)
    P = nat.ptr
    loss = z(4)
    lbs = [z(B) for _ in range(4)]
    work = z(2 * B)
    mh = torch.full((1,), -1.0, dtype=F64, device="cuda")
    nat.call("dl_loss_height", nat.F64, B, T, written, steps, J, P(hd), P(thd), P(loss[0:1]),
             P(lbs[0]), P(mh), P(work), nat.stream())
    nat.call("dl_loss_pose", nat.F64, B, N, T, written, steps, J, P(xd), P(objd), P(loss[1:2]),
             P(lbs[1]), P(work), nat.stream())
    nat.call("dl_loss_actuation", nat.F64, B, O, T, written, steps, P(ad), P(tvd), P(thd),
             P(loss[2:3]), P(lbs[2]), P(work), nat.stream())
    nat.call("dl_loss_crawl", nat.F64, B, T, written, steps, P(ud), P(tcd), P(loss[3:4]),
             P(lbs[3]), P(work), nat.stream())
    gs = [0.7, 1.3, 0.9, 1.1]
    hg, ug, ag, xg = z(T, B), z(T, B), z(T, B, O), torch.zeros_like(xd)
    nat.call("dl_loss_height_grad", nat.F64, B, T, written, steps, J, P(hd), P(thd), gs[0], P(hg),
             nat.stream())
    nat.call("dl_loss_pose_grad", nat.F64, B, N, T, written, steps, J, P(xd), P(objd), gs[1],
             P(xg), nat.stream())
    nat.call("dl_loss_actuation_grad", nat.F64, B, O, T, written, steps, P(ad), P(tvd), P(thd),
             gs[2], P(ag), nat.stream())
    nat.call("dl_loss_crawl_grad", nat.F64, B, T, written, steps, P(ud), P(tcd), gs[3], P(ug),
             nat.stream())
    torch.cuda.synchronize()
    hr, ur, ar, xr = (cut(a).clone().requires_grad_() for a in (height, upper, act, x))
    eh, ebh, hmax = om.loss_height(hr, th, steps, J)
    ep, ebp = om.loss_pose(xr, obj, steps, J)
    ea, eba = om.loss_actuation(ar, tv, th, steps)
    ec, ebc = om.loss_crawl(ur, tc, steps)
    want = torch.stack([eh, ep, ea, ec]).detach()
    assert torch.allclose(loss.cpu(), want, rtol=1e-13, atol=1e-300)
    for got, exp in zip(lbs, (ebh, ebp, eba, ebc)):
        assert rel_err(got.cpu(), exp.detach()) < 1e-13
    assert want.abs().min() > 0
    assert abs(mh.item() - max(hmax, -1.0)) == 0.0
    (gs[0] * eh + gs[1] * ep + gs[2] * ea + gs[3] * ec).backward()
    for got, exp in ((hg, hr), (ug, ur), (ag, ar)):
        full = torch.zeros_like(got.cpu())
        full[:written] = exp.grad
        assert rel_err(got.cpu(), full) < 1e-13
        assert exp.grad.abs().max() > 0
    full = torch.zeros_like(x)
    full[:written] = xr.grad
    assert rel_err(xg.cpu().transpose(-1, -2), full) < 1e-13


def _ref_config(steps, batch, prec, jump):
    import json
    with open("tests/golden/sim_quad.json") as fh:
        raw = json.load(fh)
    raw["process"].update(max_steps=steps, run_period=20, turn_period=60, jump_period=jump)
    raw["nn"]["batch_size"] = batch
    raw["precision"] = prec
    raw["train"]["task"] = ["velocity", "height", "pose", "actuation", "rotation", "crawl"]
    return dl.ConfigSim(raw)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_trainer_every_loss_matches_oracle(prec):
    """simulate(train=True) with loss_enable = every reference loss (the
    per-step kernels, tape replay), targets chosen so the height (target +1
    at a jump-period end), actuation (target_h below 0.1) and pose losses
    all contribute, against the oracle rollout with the same losses."""
    _say(f"trainer every loss {prec}")
    T, B, J = 130, 3, 40
    config = _ref_config(T, B, prec, J)
    trainer = dl.DiffPhyTrainer(config)
    env = trainer.taichi_env
    env.setup_robot()
    trainer.nn.weights_init(seed=33)
    W = [w.data.double().cpu().clone() for w in (trainer.nn.weights1, trainer.nn.bias1,
                                                  trainer.nn.weights2, trainer.nn.bias2)]
    # segments of 60 steps: robot k jumps (h=+1) in segment k; the others run
    # (vx / vz) or crouch (h=-1)
    seg = [[(0.3, 0.5), (0.9, 0.8), (0.1, 0.5)], [(0.6, 0.2), (0.4, 0.5), (0.8, 0.7)],
           [(0.1, 0.5), (0.7, 0.3), (0.3, 0.5)]]
    pool = np.zeros(B * 3 * 3)
    for s_ in range(3):
        for k in range(B):
            q = (s_ * B + k) * 3
            pool[q], pool[q + 1] = seg[s_][k]
    trainer.simulate(T, iter=0, max_speed=0.1, max_height=0.1, pool=pool)
    torch.cuda.synchronize()
    parts = {n: env.loss_dict[f"loss_{n}"][None] for n in ("velocity", "height", "pose",
                                                             "actuation", "rotation")}
    _say("  " + " ".join(f"{n}={v:.3e}" for n, v in parts.items()))
    assert all(v > 0 for v in parts.values())
    loss = env.loss[None]
    rob = oracle_reference_robot()
    cfg = om.RefConfig(steps=T, run_period=20, turn_period=60, jump_period=J,
                       layout=om.SensorLayout(run_period=20))
    tv, th, tc = om.train_targets(pool, T, B, 60, 0.1)
    Wr = [w.clone().requires_grad_() for w in W]
    x0, v0 = om.initial_state(rob, B)
    lo, _, _ = om.rollout_ref(rob, cfg, Wr, tv, th, tc, x0, v0,
                              losses=("velocity", "height", "pose", "actuation", "rotation",
                                      "crawl"))
    lo.backward()
    ltol, gtol = (1e-10, 1e-8) if prec == "f64" else (1e-4, 1e-3)
    assert abs(loss - lo.item()) / abs(lo.item()) < ltol
    G = [trainer.nn.weights1.grad, trainer.nn.bias1.grad, trainer.nn.weights2.grad,
         trainer.nn.bias2.grad]
    for a, b in zip(G, Wr):
        assert rel_err(a.data.double().cpu(), b.grad) < gtol
