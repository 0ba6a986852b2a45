"""Oracle functions pinned by closed forms and invariants (no GPU):
``pre_advance`` height / rotation (solver_mass_spring.py:158-192), the
quadruped task loss window against free fall, and the MPM sensor vector
against hand-built particle sets."""

import math

import numpy as np
import pytest
import torch

from oracle import mass_spring as om
from oracle import mpm as omp

from helpers import oracle_quadruped, oracle_reference_robot

F64 = torch.float64


def _ry(theta):
    """Rotation about +y: (x, z) -> (x cos + z sin, -x sin + z cos)."""
    c, s = math.cos(theta), math.sin(theta)
    return torch.tensor([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]], dtype=F64)


def _wrap(a):
    return (a + math.pi) % (2 * math.pi) - math.pi


def test_height_window_and_clamps():
    """height[t] = running max of min_i y over the jump window, reset at
    t % jump_period == 0; min starts from 10, upper from -10 (:158-173)."""
    B, N, J = 2, 5, 4
    mins = [0.30, 0.10, 0.20, 0.05, 0.40, 0.15, 0.35, 0.60]      # min_i y per step, robot 0
    height = None
    for t, m in enumerate(mins):
        y = torch.full((B, N), 0.9, dtype=F64)
        y[0, t % N] = m
        y[1] = torch.linspace(12.0, 13.0, N)          # all above 10: min clamps to 10
        x = torch.zeros(B, N, 3, dtype=F64)
        x[..., 1] = y
        _, height, upper, _ = om.pre_advance(t, x, x, x.mean(1), height, J)
        w0 = t // J * J
        assert height[0].item() == max(mins[w0:t + 1])
        assert height[1].item() == 10.0
        assert upper[0].item() == 0.9
        assert upper[1].item() == 13.0
    x = torch.full((1, N, 3), -20.0, dtype=F64)        # all below -10: upper clamps to -10
    _, h, up, _ = om.pre_advance(0, x, x, x.mean(1), None, J)
    assert up.item() == -10.0 and h.item() == -20.0


def test_rotation_two_point_closed_form():
    """Head/tail branch: a head vertex left of the initial centre and a tail
    vertex right of it; rotation = atan2(z, x) of (tail - head)."""
    x0 = torch.tensor([[[0.0, 0.1, 0.0], [1.0, 0.1, 0.0]]], dtype=F64)
    c0 = x0.mean(1)
    for phi in (0.0, 0.3, -1.2, 2.9, math.pi / 2):
        xt = torch.tensor([[[0.5, 0.2, 0.5], [0.5 + math.cos(phi), 0.2, 0.5 + math.sin(phi)]]],
                          dtype=F64)
        rot = om.pre_advance(7, xt, x0, c0, None, 500, robot_id=0)[3]
        assert abs(_wrap(rot.item() - phi)) < 1e-14


@pytest.mark.parametrize("robot", ["ref100", "quad"])
def test_rotation_turns_with_the_body(robot):
    """Turning the whole body by theta about +y (and translating it) moves the
    rotation by -theta: robot 100's vertex ids 45, 46, 34, 36 and the
    head/tail split alike."""
    rob = oracle_reference_robot(100) if robot == "ref100" else oracle_quadruped()
    rid = 100 if robot == "ref100" else 0
    x0, _ = om.initial_state(rob, 1)
    c0 = x0.mean(1)
    g = torch.Generator().manual_seed(3)
    base = x0 + 0.002 * torch.randn(x0.shape, dtype=F64, generator=g)
    r0 = om.pre_advance(1, base, x0, c0, None, 500, robot_id=rid)[3].item()
    for theta in (0.4, -0.9, 2.0):
        c = base.mean(1, keepdim=True)
        moved = (base - c) @ _ry(theta).T + c + torch.tensor([0.3, 0.1, -0.2], dtype=F64)
        r = om.pre_advance(1, moved, x0, c0, None, 500, robot_id=rid)[3].item()
        assert abs(_wrap(r - (r0 - theta))) < 1e-12


def _free_fall_com(T, dt, g, c0y):
    """Semi-implicit Euler COM: v_t = t dt g, c_t = c_0 + dt^2 g t (t + 1) / 2."""
    t = np.arange(T + 1, dtype=np.float64)
    return c0y + dt * dt * g * t * (t + 1) / 2, t * dt * g


def task_loss_free_fall(T, W, dt, g, c0y, tv, th):
    """Closed-form task loss of a body in free fall from rest: the COM moves
    like a point mass (internal impulses cancel), v_xz = 0."""
    cy, _ = _free_fall_com(T, dt, g, c0y)
    ts = np.arange(T - W + 1, T + 1)
    return np.array([np.mean(tv[k, 0] ** 2 + tv[k, 2] ** 2 + (cy[ts] - th[k]) ** 2)
                     for k in range(len(th))])


def test_task_loss_window_against_free_fall():
    """The quadruped workload's loss (window (T-W, T], |v_xz - tv_xz|^2 +
    (c_y - th)^2, mean over the window) against the closed form of a
    free-falling body (released 1 m up, no contact within T)."""
    rob = oracle_quadruped()
    B, T, W = 3, 60, 25
    cfg = om.TaskConfig(steps=T, window=W)
    x0, v0 = om.initial_state(rob, B, shift=(0.4, 1.0, 0.0))
    tv = torch.tensor([[0.05, 0.0, -0.02], [-0.07, 0.01, 0.03], [0.0, 0.0, 0.08]], dtype=F64)
    th = torch.tensor([0.12, 0.2, 0.25], dtype=F64)
    w = om.init_weights(cfg.layout.n_input(rob.n_vert), 64, 4, seed=5)
    loss, lb, _ = om.rollout_task(rob, cfg, w, tv, th, x0, v0)
    c0y = x0[0, :, 1].mean().item()
    want = task_loss_free_fall(T, W, cfg.physics.dt, cfg.physics.gravity, c0y, tv.numpy(),
                               th.numpy())
    assert np.allclose(lb.numpy(), want, rtol=1e-11, atol=0)
    assert abs(loss.item() - want.mean()) < 1e-11 * want.mean()
    # one state more or less in the window is visible
    off = task_loss_free_fall(T, W + 1, cfg.physics.dt, cfg.physics.gravity, c0y, tv.numpy(),
                              th.numpy())
    assert not np.allclose(lb.numpy(), off, rtol=1e-6)


def test_mpm_sensors_hand_built_groups():
    """MPM sensor vector on a hand-built particle set whose group centroids,
    COM and COM velocity are known: [sin(2 pi t / period + 2 pi j / n_sin) |
    v_com | (centroid_g - c) / scale per group | tv / max_speed | th]."""
    P, G = 12, 3
    centroids = torch.tensor([[0.2, 0.3, 0.4], [0.6, 0.3, 0.4], [0.4, 0.5, 0.2]], dtype=F64)
    offs = torch.tensor([[0.01, 0.0, 0.0], [-0.01, 0.0, 0.0], [0.0, 0.02, 0.0], [0.0, -0.02, 0.0]],
                        dtype=F64)
    x = torch.cat([centroids[g] + offs for g in range(G)])[None]           # [1, 12, 3]
    group = torch.arange(P) // 4
    v = torch.zeros(1, P, 3, dtype=F64)
    v[0, :, 0] = torch.arange(P, dtype=F64) * 0.01                          # mean 0.055
    v[0, :, 2] = -0.03
    tv = torch.tensor([[0.04, 0.0, -0.08]], dtype=F64)
    th = torch.tensor([0.2], dtype=F64)
    t, n_sin, period, scale, vmax = 37, 10, 100, 0.2, 0.08
    inp, c, vc = omp.sensors(t, x, v, group, G, n_sin, period, scale, vmax, tv, th)
    com = centroids.mean(0)
    assert torch.allclose(c[0], com, atol=1e-15)
    assert torch.allclose(vc[0], torch.tensor([0.055, 0.0, -0.03], dtype=F64), atol=1e-15)
    want = [math.sin(2 * math.pi * t / period + 2 * math.pi * j / n_sin) for j in range(n_sin)]
    want += [0.055, 0.0, -0.03]
    for g in range(G):
        want += ((centroids[g] - com) / scale).tolist()
    want += [0.5, 0.0, -1.0, 0.2]
    assert inp.shape == (1, n_sin + 3 + 3 * G + 4)
    assert torch.allclose(inp[0], torch.tensor(want, dtype=F64), atol=1e-13)


def test_mpm_task_loss_against_free_fall():
    """The MPM rollout's task loss against the closed form of a body in free
    fall (momentum is conserved by P2G / G2P, gravity adds dt g per step; the
    grid's 1e-10 mass guard is negligible at p_mass = 1)."""
    p = omp.MPMParams(dim=3, n_grid=16, dt=1e-3, gravity=-1.8, p_vol=1.0, p_mass=1.0, mu=20.0,
                      la=20.0, act_strength=2.0, act_axis=0, bound=3)
    ax = torch.arange(3, dtype=F64)
    pos = (torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) * 0.5
           + torch.tensor([7.0, 9.0, 7.0])) / 16
    ids = torch.tensor([(i % 3) - 1 for i in range(len(pos))])
    B, T, W = 2, 12, 5
    tv = torch.tensor([[0.03, 0.0, -0.05], [0.0, 0.0, 0.07]], dtype=F64)
    th = torch.tensor([0.1, 0.25], dtype=F64)
    I = 10 + 3 + 3 * 3 + 4
    w = om.init_weights(I, 64, 2, seed=4)
    loss = omp.rollout_task(pos[None].expand(B, -1, -1).clone(), ids[None].expand(B, -1), 2, p,
                            w, tv, th, T, W)
    want = task_loss_free_fall(T, W, p.dt, p.gravity, pos[:, 1].mean().item(), tv.numpy(),
                               th.numpy())
    assert abs(loss.item() - want.mean()) < 1e-9 * want.mean()


def test_height_loss_closed_form():
    """compute_loss_height: only t % jump == jump - 1 with target_h above
    0.1 counts; (h - th)^2 / B / (steps // jump) * 100; rows past the
    written ones read zero."""
    T, B, J, steps = 13, 2, 5, 12
    h = torch.zeros(T, B, dtype=F64)
    h[4] = torch.tensor([0.30, 0.12], dtype=F64)
    h[9] = torch.tensor([0.25, 0.50], dtype=F64)
    th = torch.full((T, B), 0.2, dtype=F64)
    th[9, 1] = 0.05                               # below 0.1: excluded
    loss, lb, hmax = om.loss_height(h[:11], th, steps, J)
    per = [(0.30 - 0.2) ** 2 + (0.25 - 0.2) ** 2, (0.12 - 0.2) ** 2]
    want = [p / B / (steps // J) * 100 for p in per]
    assert torch.allclose(lb, torch.tensor(want, dtype=F64) * B, rtol=1e-14)
    assert abs(loss.item() - sum(want)) < 1e-15
    assert hmax == 0.30


def test_pose_loss_closed_form():
    """compute_loss_pose of a body shifted by delta from its rest pose at
    every step: N |delta|^2 per jump-period step, / B / (steps // jump)."""
    N, B, T, J, steps = 5, 3, 11, 4, 10
    obj = torch.rand(N, 3, dtype=F64, generator=torch.Generator().manual_seed(1))
    delta = torch.tensor([0.01, -0.02, 0.005], dtype=F64)
    x = (obj + delta).expand(T, B, N, 3).clone()
    loss, lb = om.loss_pose(x, obj, steps, J)
    n_t = len([t for t in range(1, steps + 1) if t % J == 0])           # t = 4, 8
    want = N * float((delta ** 2).sum()) * n_t / (steps // J)
    assert abs(loss.item() - want) < 1e-15 and torch.allclose(lb, torch.full((B,), want, dtype=F64))
    assert om.loss_pose(obj.expand(T, B, N, 3).clone(), obj, steps, J)[0].item() == 0.0


def test_actuation_loss_closed_form():
    """compute_loss_actuation: a constant actuation a above the allowance
    sqrt(|tv_x| / 0.08) costs 10 (|a| - allowance), summed over the batch's
    mean; steps with target_h above 0.1 are free."""
    T, B, O, steps = 8, 2, 3, 8
    a, v = 0.7, 0.02
    act = torch.full((T, B, O), -a, dtype=F64)
    tv = torch.zeros(T, B, 3, dtype=F64)
    tv[..., 0] = v
    th = torch.zeros(T, B, dtype=F64)
    loss, _ = om.loss_actuation(act, tv, th, steps)
    assert abs(loss.item() - 10 * (a - math.sqrt(v / 0.08))) < 1e-14
    th[:4] = 0.2                                   # half the steps exempt
    assert abs(om.loss_actuation(act, tv, th, steps)[0].item()
               - 5 * (a - math.sqrt(v / 0.08))) < 1e-14


def test_crawl_loss_closed_form():
    T, B, steps = 9, 2, 8
    up = torch.full((T, B), 0.3, dtype=F64)
    tc = torch.ones(T, B, dtype=F64)
    tc[:, 1] = 0.0
    loss, lb = om.loss_crawl(up, tc, steps)
    assert abs(loss.item() - (0.3 - 0.1) ** 2 / B * 5) < 1e-15
    assert lb[1].item() == 0.0
    up[:] = 0.05                                   # below 0.1: free
    assert om.loss_crawl(up, tc, steps)[0].item() == 0.0
