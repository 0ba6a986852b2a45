"""Trainer paths beyond one plain rollout: the zero-order hold of the
actuation between network evaluations (diff_copy, control_length > 1,
diffphy_trainer.py:215-218, 249-254) and the restarts from the previous
rollout's final state after change_iter iterations (rounded_train,
diffphy_trainer.py:270-281), against the oracle (FP64)."""

import json
import sys

import numpy as np
import pytest
import torch

import diffloco as dl
from oracle import mass_spring as om

from helpers import oracle_reference_robot, rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _config(T, B, control_length=1):
    with open("tests/golden/sim_quad.json") as fh:
        raw = json.load(fh)
    raw["process"].update(max_steps=T, run_period=20, turn_period=60, jump_period=60)
    raw["nn"]["batch_size"] = B
    raw["robot"]["control_length"] = control_length
    raw["precision"] = "f64"
    return dl.ConfigSim(raw)


def _weights(nn):
    return [w.data.double().cpu().clone() for w in (nn.weights1, nn.bias1, nn.weights2, nn.bias2)]


def _grads(nn):
    return [w.grad.data.double().cpu() for w in (nn.weights1, nn.bias1, nn.weights2, nn.bias2)]


def _cfg(T):
    return om.RefConfig(steps=T, run_period=20, turn_period=60, jump_period=60,
                        layout=om.SensorLayout(run_period=20))


@pytest.mark.parametrize("cl", [3, 7])
def test_control_length_hold_matches_oracle(cl):
    _say(f"control_length {cl}")
    T, B = 130, 3
    tr = dl.DiffPhyTrainer(_config(T, B, cl))
    assert tr.control_length == cl
    tr.taichi_env.setup_robot()
    tr.nn.weights_init(seed=41)
    W = _weights(tr.nn)
    pool = torch.rand(B * 3 * 3, dtype=torch.float64, generator=torch.Generator().manual_seed(42))
    tr.simulate(T, iter=0, max_speed=0.1, max_height=0.1, pool=pool)
    torch.cuda.synchronize()
    # the held steps really repeat the last evaluated output
    a = tr.taichi_env.actuation.data.double().cpu()
    assert torch.equal(a[cl + 1], a[cl]) and not torch.equal(a[cl], a[cl - 1])
    rob = oracle_reference_robot()
    tv, th, tc = om.train_targets(pool.numpy(), T, B, 60, 0.1)
    Wr = [w.clone().requires_grad_() for w in W]
    x0, v0 = om.initial_state(rob, B)
    lo, _, _ = om.rollout_ref(rob, _cfg(T), Wr, tv, th, tc, x0, v0, control_length=cl)
    lo.backward()
    assert abs(tr.taichi_env.loss[None] - lo.item()) <= 1e-10 * abs(lo.item())
    for g, w in zip(_grads(tr.nn), Wr):
        assert rel_err(g, w.grad) < 1e-8


def test_rounded_train_restarts_match_oracle():
    """After change_iter iterations run_step restarts every robot from the
    previous rollout's final state except robots start, start + step, ...,
    whose positions return to the initial pose (velocities stay copied)."""
    _say("rounded_train")
    T, B = 130, 5
    tr = dl.DiffPhyTrainer(_config(T, B))
    env = tr.taichi_env
    env.setup_robot()
    tr.nn.weights_init(seed=43)
    W = _weights(tr.nn)
    g = torch.Generator().manual_seed(44)
    pools = [torch.rand(B * 3 * 3, dtype=torch.float64, generator=g) for _ in range(2)]
    tr.simulate(T, iter=0, max_speed=0.1, max_height=0.1, pool=pools[0])
    tr.iter = tr.change_iter + 1                     # 5001: stride 2, start = 1
    tr.run_step(pool=pools[1])
    torch.cuda.synchronize()
    loss = tr.losses_list[-1]
    grads = _grads(tr.nn)
    # oracle: the first rollout's final state, robots 1 and 3 reset to the pose
    rob = oracle_reference_robot()
    cfg = _cfg(T)
    x0, v0 = om.initial_state(rob, B)
    tv, th, tc = om.train_targets(pools[0].numpy(), T, B, 60, 0.1)
    _, _, traj = om.rollout_ref(rob, cfg, W, tv, th, tc, x0, v0, record=True)
    xr, vr = traj["x"][T - 1].clone(), traj["v"][T - 1].clone()
    xr[[1, 3]] = x0[[1, 3]]
    tv, th, tc = om.train_targets(pools[1].numpy(), T, B, 60, 0.1)
    Wr = [w.clone().requires_grad_() for w in W]
    lo, _, _ = om.rollout_ref(rob, cfg, Wr, tv, th, tc, xr, vr)
    lo.backward()
    assert abs(loss - lo.item()) <= 1e-10 * abs(lo.item())
    for a, w in zip(grads, Wr):
        assert rel_err(a, w.grad) < 1e-8
    # the reset stride grows by one every 500 iterations up to max_reset_step
    tr.iter = 5500
    tr.run_step(pool=pools[0])
    assert tr.reset_step == 3
