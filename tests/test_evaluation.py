"""Evaluation path (diffphy_trainer.py:304-435, taichi_env.py:54-72,
333-383): validation matrix, scripted / validation / interactive targets,
OBJ export (CPU); the evaluation rollout's per-robot losses against the
oracle (GPU)."""

import json

import numpy as np
import pytest
import torch

import diffloco as dl
from oracle import mass_spring as om

# (t, tv_x, tv_z, th) of initialize_script(-s, 0, 0, s, s, 0, 0, -s), s = 0.08,
# read off the branches of taichi_env.py:340-376
S = 0.08
SCRIPT = [(0, -S, 0.0, 0.0), (999, -S, 0.0, 0.0), (1000, 0.0, S, 0.0), (1999, 0.0, S, 0.0),
          (2000, S, 0.0, 0.0), (2999, S, 0.0, 0.0), (3000, 0.0, -S, 0.0), (3999, 0.0, -S, 0.0),
          (4000, 0.0, 0.0, -1.0), (4999, 0.0, 0.0, -1.0), (5000, 0.0, 0.0, 1.0), (5999, 0.0, 0.0, 1.0)]


def _config(steps=1000, batch=2, validation=None, task=("velocity", "rotation"), prec="f64"):
    with open("tests/golden/sim_quad.json") as fh:
        raw = json.load(fh)
    raw["process"].update(max_steps=steps, run_period=20, turn_period=60, jump_period=60)
    raw["nn"]["batch_size"] = batch
    raw["train"]["task"] = list(task)
    if validation is not None:
        raw["validation"] = validation
    raw["precision"] = prec
    return dl.ConfigSim(raw)


def test_validation_matrix_is_the_product_of_trained_targets():
    val = {"velocity": [0.02, 0.05], "height": [0.1, 0.2, 0.3], "crawl": [1.0]}
    env = dl.TaichiEnv(_config(batch=6, validation=val, task=("velocity", "height")),
                                  device="cpu")
    vv = env.validate_targets_values
    assert set(vv) == {"velocity", "height"}          # crawl is not trained: dropped
    assert list(zip(vv["velocity"], vv["height"])) == [
        (0.02, 0.1), (0.02, 0.2), (0.02, 0.3), (0.05, 0.1), (0.05, 0.2), (0.05, 0.3)]


def test_script_validate_and_interactive_targets():
    env = dl.TaichiEnv(_config(steps=6000, batch=2), device="cpu")
    env.initialize_script(6000, -S, 0, 0, S, S, 0, 0, -S)
    tvo, tho, _ = om.script_targets(6000, 2, -S, 0, 0, S, S, 0, 0, -S)
    for t, vx, vz, h in SCRIPT:
        for tv, th in ((env.target_v.data, env.target_h.data), (tvo, tho)):
            assert tv[t, 1, 0].item() == pytest.approx(vx) and tv[t, 1, 2].item() == pytest.approx(vz)
            assert th[t, 1].item() == h
    env.initialize_validate(10, [0.01, 0.02], [0.1, 0.2], [0.0, 1.0])
    assert env.target_v.data[9, 1, 0].item() == pytest.approx(0.02)
    assert env.target_h.data[0, 0].item() == pytest.approx(0.1) and env.target_c.data[3, 1].item() == 1.0
    env.initialize_interactive(10, 0.03, 0.15, 0.0)
    assert env.target_v.data[5, 0, 0].item() == pytest.approx(0.03)
    assert env.target_h.data[5, 1].item() == pytest.approx(0.15)


def test_obj_export(tmp_path):
    tr = dl.DiffPhyTrainer(_config(steps=40, batch=1), device="cpu")
    env = tr.taichi_env
    N = env.n_objects_()
    g = torch.Generator().manual_seed(0)
    env.x.view[:] = torch.rand(env.x.view.shape, generator=g, dtype=env.x.view.dtype)
    paths = tr.export_obj(str(tmp_path), steps=30, every=10)
    assert len(paths) == 4                             # states 0, 10, 20, 30
    lines = open(paths[2]).read().splitlines()
    verts = [l for l in lines if l.startswith("v ")]
    faces = [l for l in lines if l.startswith("f ")]
    assert len(verts) == N and len(faces) == len(tr.config.get_config()["robot"]["faces"]) > 0
    x = np.array([[float(c) for c in l.split()[1:]] for l in verts])
    assert np.allclose(x, env.x.view[20, 0].numpy(), atol=1e-6)
    assert min(int(c) for l in faces for c in l.split()[1:]) == 1
    assert max(int(c) for l in faces for c in l.split()[1:]) <= N


@pytest.mark.gpu
def test_evaluation_rollout_matches_oracle():
    """simulate(train=False) on the reference preset (3D: scripted targets,
    steps+1 advances, velocity + rotation losses per robot) against the
    oracle's evaluation rollout, FP64."""
    print("evaluation rollout", flush=True)
    from helpers import oracle_reference_robot
    steps, B = 125, 1
    tr = dl.DiffPhyTrainer(_config(steps=steps + 1, batch=B,
                                           validation={"velocity": [0.05]}))
    tr.nn.weights_init(seed=21)
    W = [w.data.double().cpu().clone() for w in (tr.nn.weights1, tr.nn.bias1, tr.nn.weights2,
                                                 tr.nn.bias2)]
    res = tr.evaluate_weights(steps=steps)
    torch.cuda.synchronize()
    rob = oracle_reference_robot()
    cfg = om.RefConfig(steps=steps, run_period=20, turn_period=60, jump_period=60,
                       layout=om.SensorLayout(run_period=20))
    ms = tr.max_speed
    tv, th, tc = om.script_targets(steps + 1, B, -ms, 0, 0, ms, ms, 0, 0, -ms)
    x0, v0 = om.initial_state(rob, B)
    lo, lob = om.rollout_eval(rob, cfg, W, tv, th, tc, x0, v0)
    got = res["task_loss_velocity_0.05"]
    assert abs(got - lob[0].item()) <= 1e-10 * abs(lob[0].item())
    assert abs(tr.taichi_env.loss[None] - lo.item()) <= 1e-10 * abs(lo.item())
