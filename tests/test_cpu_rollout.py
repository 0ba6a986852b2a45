"""The C/OpenMP restatement of the training rollout (oracle/cpu_rollout.c, the
CPU baseline) against the torch oracle, whose autograd gradients are pinned
by the closed-form and finite-difference tests of tests/test_oracle_*.py.

The FP64 build must reproduce the oracle's loss and every weight gradient to
rounding (its adjoint is hand-written, the oracle's is autograd), with
checkpoints and partial final segments; the FP32 build (the timed baseline)
to FP32 accuracy; Adam to the oracle's update; results must not depend on the
thread count."""

import numpy as np
import pytest
import torch

from oracle import build_cpu
from oracle import builders as ob
from oracle import mass_spring as om
from oracle.cpu_rollout import CpuRollout

if not (build_cpu.lib_path("f32").exists() and build_cpu.lib_path("f64").exists()):
    build_cpu.build()


def _case(B, T, W, seed=1):
    objects, springs, groups = ob.build_quadruped()
    rob = om.make_robot(objects, springs, groups)
    cfg = om.TaskConfig(steps=T, window=W)
    I = cfg.layout.n_input(rob.n_vert)
    w64 = om.init_weights(I, 64, 4, seed=0)
    g = torch.Generator().manual_seed(seed)
    w64[1] = 0.1 * torch.randn(64, dtype=torch.float64, generator=g)
    w64[3] = 0.1 * torch.randn(4, dtype=torch.float64, generator=g)
    rng = np.random.default_rng(seed)
    tv = np.stack([rng.uniform(-0.08, 0.08, B), np.zeros(B), rng.uniform(-0.08, 0.08, B)], 1)
    th = rng.uniform(0.1, 0.25, B)
    x0, v0 = om.initial_state(rob, B)
    x0 = x0 + 0.01 * torch.randn(x0.shape, dtype=torch.float64, generator=g)
    return rob, cfg, w64, tv, th, x0, v0


def _oracle(rob, cfg, w64, tv, th, x0, v0):
    wr = [w.clone().requires_grad_() for w in w64]
    lo, lb, _ = om.rollout_task(rob, cfg, wr, torch.tensor(tv), torch.tensor(th), x0, v0)
    lo.backward()
    return lo.item(), lb.detach().numpy(), [w.grad.numpy() for w in wr]


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("T,W,K", [(150, 60, 16), (300, 150, 50)])
def test_fp64_port_matches_oracle(T, W, K):
    rob, cfg, w64, tv, th, x0, v0 = _case(3, T, W)
    lo, lbo, go = _oracle(rob, cfg, w64, tv, th, x0, v0)
    cr = CpuRollout(rob, cfg, K=K, dtype="f64")
    cr.set_weights(*[w.numpy() for w in w64])
    loss, lb, g = cr.run(x0.numpy(), tv, th, threads=2)
    assert abs(loss - lo) / abs(lo) < 1e-12
    assert np.abs(lb - lbo).max() / np.abs(lbo).max() < 1e-12
    for a, b in zip(cr.split(g), go):
        assert np.linalg.norm(b) > 0
        assert _rel(a, b) < 1e-10


def test_fp32_port_close_to_oracle():
    rob, cfg, w64, tv, th, x0, v0 = _case(3, 150, 60)
    lo, _, go = _oracle(rob, cfg, w64, tv, th, x0, v0)
    cr = CpuRollout(rob, cfg, K=50, dtype="f32")
    cr.set_weights(*[w.numpy() for w in w64])
    loss, _, g = cr.run(x0.numpy(), tv, th)
    assert abs(loss - lo) / abs(lo) < 1e-5
    for a, b in zip(cr.split(g), go):
        assert _rel(a, b) < 1e-4


def test_threads_do_not_change_results():
    rob, cfg, w64, tv, th, x0, _ = _case(5, 120, 60)
    runs = []
    for threads in (1, 3):
        cr = CpuRollout(rob, cfg, K=50, dtype="f32")
        cr.set_weights(*[w.numpy() for w in w64])
        runs.append(cr.run(x0.numpy(), tv, th, threads=threads))
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])


def test_checkpoint_interval_does_not_change_results():
    """Segment recompute reproduces the forward bit for bit, so K changes
    nothing (one segment of the whole horizon against segments of 7)."""
    rob, cfg, w64, tv, th, x0, _ = _case(2, 100, 40)
    runs = []
    for K in (100, 7):
        cr = CpuRollout(rob, cfg, K=K, dtype="f64")
        cr.set_weights(*[w.numpy() for w in w64])
        runs.append(cr.run(x0.numpy(), tv, th))
    assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][2], runs[1][2])


def test_adam_matches_oracle_update():
    rob, cfg, w64, tv, th, x0, _ = _case(2, 60, 30)
    cr = CpuRollout(rob, cfg, K=50, dtype="f64")
    cr.set_weights(*[w.numpy() for w in w64])
    ref = [w.numpy().copy() for w in w64]
    m = [np.zeros_like(w) for w in ref]
    v = [np.zeros_like(w) for w in ref]
    rng = np.random.default_rng(3)
    for it in range(3):
        gs = [rng.standard_normal(w.shape) * 1e-2 for w in ref]
        cr.adam(np.concatenate([g.ravel() for g in gs]))
        for w, g, mm, vv in zip(ref, gs, m, v):
            om.adam_update(w, g, mm, vv, it)
    for a, b in zip(cr.W, ref):
        assert np.abs(a - b).max() < 1e-15
