"""Pins the FP64 oracle's mass-spring rollout against what the mathematics
fixes: closed forms, invariants, hand-worked contact cases, library routines
(torch.nn.functional.linear, torch.optim.Adam) and finite differences."""

import math

import numpy as np
import pytest
import torch

from oracle import mass_spring as om

from helpers import oracle_quadruped, oracle_reference_robot

F = torch.float64


def _two_mass(L0=0.05, k=3e4, coeff=0.0):
    objects = [(0.0, 1.0, 0.0), (L0, 1.0, 0.0)]
    return om.make_robot(objects, [[0, 1, L0, k, coeff]], [0])


def test_free_fall_com_closed_form():
    """Springs are internal: the COM follows semi-implicit Euler under gravity,
    y_t = y0 + dt * (t v0 + dt g t (t+1) / 2)."""
    rob = oracle_quadruped()
    phys = om.TaskConfig().physics
    B = 2
    x, v = om.initial_state(rob, B, shift=(0.4, 1.0, 0.0))
    g = torch.Generator().manual_seed(0)
    x = x + 0.003 * torch.randn(x.shape, generator=g, dtype=F)
    v = 0.05 * torch.randn(v.shape, generator=g, dtype=F)
    act = torch.rand(B, 4, generator=g, dtype=F) * 2 - 1
    c0, v0 = x.mean(1), v.mean(1)
    T = 120
    for _ in range(T):
        x, v = om.step(x, v, act, rob, phys)
    cy = c0[:, 1] + phys.dt * (T * v0[:, 1] + phys.dt * phys.gravity * T * (T + 1) / 2)
    assert torch.allclose(x.mean(1)[:, 1], cy, rtol=0, atol=1e-12)
    cx = c0[:, 0] + phys.dt * T * v0[:, 0]
    assert torch.allclose(x.mean(1)[:, 0], cx, rtol=0, atol=1e-12)


def test_spring_impulses_conserve_momentum():
    rob = oracle_reference_robot()
    g = torch.Generator().manual_seed(1)
    x, v = om.initial_state(rob, 3)
    x = x + 0.01 * torch.randn(x.shape, generator=g, dtype=F)
    v = torch.randn(v.shape, generator=g, dtype=F)
    act = torch.rand(3, rob.n_spring, generator=g, dtype=F) * 2 - 1
    vi = om.spring_impulses(x, v, act, rob, om.Physics())
    assert vi.abs().max() > 1e-3
    assert vi.sum(1).abs().max() < 1e-14


def test_rest_pose_impulse_value():
    """At rest length with zero actuation only the norm offsets act:
    L = sqrt(L0^2 + 1e-8) + 1e-4, J = dt k (L - L0) / L * (xa - xb)."""
    L0, k, dt = 0.05, 3e4, 0.004
    rob = _two_mass(L0, k)
    phys = om.Physics(dt=dt, stiffness_scale=0.1, dashpot=0.1)
    x = torch.tensor([[[0.0, 1.0, 0.0], [L0, 1.0, 0.0]]], dtype=F)
    vi = om.spring_impulses(x, torch.zeros_like(x), torch.zeros(1, 1, dtype=F), rob, phys)
    L = math.sqrt(L0 * L0 + 1e-8) + 1e-4
    J = dt * (k * 0.1) * (L - L0) / L * (-L0)          # d = xa - xb = (-L0, 0, 0)
    assert abs(vi[0, 0, 0].item() - (-J)) < 1e-18
    assert abs(vi[0, 1, 0].item() - J) < 1e-18
    assert vi[0, 0, 0] > 0                                # pair is pulled together


def test_actuated_rest_length():
    """L* = L0 (1 + coeff a): a spring at L0 (1 + coeff a) - offsets carries no impulse."""
    L0, coeff, a = 0.05, 0.25, 0.6
    target = L0 * (1 + coeff * a)
    sep = math.sqrt((target - 1e-4) ** 2 - 1e-8)          # makes L == L*
    rob = _two_mass(L0, 3e4, coeff)
    x = torch.tensor([[[0.0, 1.0, 0.0], [sep, 1.0, 0.0]]], dtype=F)
    vi = om.spring_impulses(x, torch.zeros_like(x), torch.tensor([[a]], dtype=F), rob, om.Physics())
    assert vi.abs().max() < 1e-15


def test_dashpot_impulse():
    """c ((va - vb) . d^) d^ with no dt; perpendicular relative velocity gives none."""
    rob = _two_mass(0.05, 0.0)
    phys = om.Physics(dashpot=0.1)
    x = torch.tensor([[[0.0, 1.0, 0.0], [0.05, 1.0, 0.0]]], dtype=F)
    v = torch.tensor([[[0.0, 0.0, 0.0], [0.3, 0.7, 0.0]]], dtype=F)
    vi = om.spring_impulses(x, v, torch.zeros(1, 1, dtype=F), rob, phys)
    # d^ = (-1,0,0); vrel = (va - vb).d^ = 0.3 ; J = 0.1*0.3*(-1,0,0); v_inc[a] = -J
    assert torch.allclose(vi[0, 0], torch.tensor([0.03, 0.0, 0.0], dtype=F), atol=1e-17)
    assert torch.allclose(vi[0, 1], torch.tensor([-0.03, 0.0, 0.0], dtype=F), atol=1e-17)


def test_two_mass_oscillator_period():
    """Semi-implicit Euler on the relative coordinate has the discrete
    frequency 2/dt asin(w dt / 2), w^2 = 2 k_eff / m (k_eff = k * 0.1)."""
    L0, k, dt = 0.05, 3e4, 0.004
    rob = _two_mass(L0, k)
    phys = om.Physics(dt=dt, gravity=0.0, ground=-10.0, dashpot=0.0, stiffness_scale=0.1)
    x = torch.tensor([[[0.0, 1.0, 0.0], [L0 + 1e-3, 1.0, 0.0]]], dtype=F)
    v = torch.zeros_like(x)
    act = torch.zeros(1, 1, dtype=F)
    sep = []
    for _ in range(400):
        x, v = om.step(x, v, act, rob, phys)
        sep.append((x[0, 1, 0] - x[0, 0, 0]).item())
    s = np.array(sep) - np.mean(sep)
    crossings = np.nonzero(np.diff(np.sign(s)) > 0)[0]
    period_steps = np.mean(np.diff(crossings))
    w = math.sqrt(2 * k * 0.1)
    wd = 2 / dt * math.asin(w * dt / 2)
    assert abs(period_steps * dt - 2 * math.pi / wd) / (2 * math.pi / wd) < 0.02


def _one_vertex(x, v, ground=0.1, friction=0.4):
    rob = om.make_robot([(0.0, 0.0, 0.0)], [], [])
    phys = om.Physics(ground=ground, friction=friction)
    xs = torch.tensor([[x]], dtype=F)
    vs = torch.tensor([[v]], dtype=F)
    vi = torch.zeros_like(xs)
    return om.advance_toi(xs, vs, vi, phys)


@pytest.mark.parametrize("vx,vz", [(0.3, -0.05), (-0.3, 0.05), (0.1, -0.6), (-0.05, 0.0)])
def test_contact_toi_and_friction(vx, vz):
    """Hand-worked ground hit: toi, per-axis Coulomb clamp, y velocity zeroed."""
    dt, g, mu, G = 0.004, -1.8, 0.4, 0.1
    x = (0.2, 0.1005, 0.3)
    ov = (vx, -0.5 + dt * g, vz)
    toi = -(x[1] - G) / ov[1]
    fr = mu * (-ov[1])

    def clamp(c):
        return min(0.0, c + fr) if c < 0 else max(0.0, c - fr)
    nv = (clamp(ov[0]), 0.0, clamp(ov[2]))
    ex = [x[d] + toi * ov[d] + (dt - toi) * nv[d] for d in range(3)]
    xn, vn = _one_vertex(x, (vx, -0.5, vz))
    assert np.allclose(vn[0, 0].numpy(), nv, rtol=0, atol=1e-15)
    assert np.allclose(xn[0, 0].numpy(), ex, rtol=0, atol=1e-15)
    assert abs(xn[0, 0, 1].item() - (x[1] + toi * ov[1])) < 1e-15   # lands on the ground


def test_no_contact_branches():
    # below the ground but moving up slower than the threshold: no contact
    xn, vn = _one_vertex((0.0, 0.099, 0.0), (0.0, 0.0072 - 0.5e-4, 0.0))
    assert abs(vn[0, 0, 1].item() - (-0.5e-4)) < 1e-15
    # above the ground after the step: plain Euler
    xn, vn = _one_vertex((0.0, 0.5, 0.0), (0.1, -0.2, 0.0))
    assert abs(xn[0, 0, 1].item() - (0.5 + 0.004 * (-0.2 - 0.0072))) < 1e-15


def test_mlp_matches_linear():
    g = torch.Generator().manual_seed(2)
    inp = torch.randn(5, 17, generator=g, dtype=F)
    W1, b1 = torch.randn(8, 17, generator=g, dtype=F), torch.randn(8, generator=g, dtype=F)
    W2, b2 = torch.randn(3, 8, generator=g, dtype=F), torch.randn(3, generator=g, dtype=F)
    for act, f in (("sin", torch.sin), ("tanh", torch.tanh)):
        h, o = om.mlp_forward(inp, W1, b1, W2, b2, act)
        eh = f(torch.nn.functional.linear(inp, W1, b1))
        assert torch.allclose(h, eh, atol=1e-15)
        assert torch.allclose(o, f(torch.nn.functional.linear(eh, W2, b2)), atol=1e-15)


def test_adam_matches_torch():
    g = torch.Generator().manual_seed(3)
    w0 = torch.randn(20, generator=g, dtype=F)
    w = w0.clone().numpy()
    m, v = np.zeros(20), np.zeros(20)
    p = torch.nn.Parameter(w0.clone())
    opt = torch.optim.Adam([p], lr=3e-4, betas=(0.9, 0.9), eps=1e-8)
    for it in range(5):
        grad = torch.randn(20, generator=g, dtype=F)
        om.adam_update(w, grad.numpy(), m, v, it)
        p.grad = grad.clone()
        opt.step()
    assert np.allclose(w, p.detach().numpy(), atol=1e-15)


def test_sensor_width_and_layout():
    ref = om.SensorLayout()
    assert ref.n_input(48) == 838            # config_sim.py:138-140 for robot 100
    assert om.TaskConfig().layout.n_input(54) == 179
    rob = oracle_reference_robot()
    B = 2
    x, v = om.initial_state(rob, B)
    x = x + 0.01 * torch.randn(x.shape, dtype=F, generator=torch.Generator().manual_seed(4))
    c = x.mean(1)
    tv = torch.tensor([[0.05, 0.0, -0.02], [0.0, 0.0, 0.1]], dtype=F)
    th = torch.tensor([1.0, -1.0], dtype=F)
    inp = om.nn_input(7, x, v, c, tv, th, torch.zeros(B, dtype=F), ref)
    # brute force from taichi_env.py:150-183 index arithmetic, plain floats
    k, N = 1, 48
    for j in (0, 37, 99):
        exp = math.sin(2 * math.pi / 0.004 / 100 * 7 * 0.004 + 2 * math.pi / 100 * j)
        assert abs(inp[k, j].item() - exp) < 1e-12
    for i in (0, 20, 47):
        for d in range(3):
            assert abs(inp[k, i * 6 + 100 + d].item() - (x[k, i, d] - c[k, d]).item() / 0.2) < 1e-15
            assert inp[k, i * 6 + 100 + 3 + d].item() == 0.0
    base = N * 6 + 100
    for j in (0, 149):
        assert inp[k, base + 2 * j].item() == tv[k, 0].item() / 0.1
        assert inp[k, base + 2 * j + 1].item() == tv[k, 2].item() / 0.1
    assert inp[k, base + 300].item() == th[k].item() and inp[k, base + 449].item() == th[k].item()


def test_train_targets_branches():
    B, turn = 2, 10
    pool = np.array([0.1, 0.9, 0.0, 0.3, 0.2, 0.0,         # segment 0: k=0 -> h=-1, k=1 -> h=+1
                     0.6, 0.75, 0.0, 0.8, 0.25, 0.0,       # segment 1: k=0 -> vz, k=1 -> vx
                     0.0, 0.0, 0.0, 0.0, 0.0, 0.0])
    tv, th, _ = om.train_targets(pool, 20, B, turn, 0.1)
    assert th[3, 0] == -1 and th[3, 1] == 1 and tv[3].abs().max() == 0
    assert abs(tv[12, 0, 2].item() - (0.75 * 2 - 1) * 0.1) < 1e-17 and tv[12, 0, 0] == 0
    assert abs(tv[12, 1, 0].item() - (0.25 * 2 - 1) * 0.1) < 1e-17 and tv[12, 1, 2] == 0
    assert th[12, 0] == 0 and th[12, 1] == 0


def test_losses_match_brute_force():
    """Vectorised losses against the reference's loops run over plain floats."""
    g = torch.Generator().manual_seed(5)
    T, B, N, rp, tp = 14, 2, 3, 3, 7
    center = torch.randn(T, B, 3, generator=g, dtype=F)
    x = torch.randn(T, B, N, 3, generator=g, dtype=F)
    tv = torch.randn(T, B, 3, generator=g, dtype=F)
    th = torch.randn(T, B, generator=g, dtype=F)
    steps = T - 1
    lv, _ = om.loss_velocity(center, tv, steps, rp, tp)
    lr, _ = om.loss_rotation(x, center, th, steps, rp, tp)
    c, X, V, H = center.tolist(), x.tolist(), tv.tolist(), th.tolist()
    bv = br = 0.0
    for t in range(rp, steps + 1):
        for k in range(B):
            if t % tp > rp:
                bv += (c[t][k][0] - c[t - rp][k][0] - V[t - rp][k][0]) ** 2 / B / steps * 100
                bv += (c[t][k][2] - c[t - rp][k][2] - V[t - rp][k][2]) ** 2 / B / steps * 100
    for t in range(1, steps + 1):
        for k in range(B):
            for i in range(N):
                if t % tp > rp:
                    t0 = [X[t][k][i][d] - c[t][k][d] for d in range(3)]
                    t1 = [X[t - rp][k][i][d] - c[t - rp][k][d] for d in range(3)]
                    a = H[t - rp][k] * 72.0 / 180.0 * 3.1415926535
                    t2 = [math.cos(a) * t1[0] - math.sin(a) * t1[2], t1[1],
                          math.sin(a) * t1[0] + math.cos(a) * t1[2]]
                    br += sum((t0[d] - t2[d]) ** 2 for d in range(3)) / steps / 40
    assert abs(lv.item() - bv) < 1e-12 * max(1, abs(bv))
    assert abs(lr.item() - br) < 1e-12 * max(1, abs(br))


def test_free_fall_gradient_is_zero():
    """Without ground contact internal forces cannot move the COM, so the COM
    loss does not depend on the controller weights at all."""
    rob = oracle_quadruped()
    cfg = om.TaskConfig(steps=40, window=20)
    I = cfg.layout.n_input(rob.n_vert)
    W = [w.requires_grad_() for w in om.init_weights(I, 64, 4, seed=6)]
    x0, v0 = om.initial_state(rob, 2, shift=(0.4, 0.5, 0.0))
    loss, _, _ = om.rollout_task(rob, cfg, W, torch.zeros(2, 3, dtype=F),
                                 torch.full((2,), 0.2, dtype=F), x0, v0)
    loss.backward()
    assert loss.item() > 0
    assert all(w.grad.abs().max().item() < 1e-16 for w in W)


def test_autograd_gradient_matches_central_differences():
    """dL/dW from autograd of the restated rollout against FP64 central
    differences (eps 1e-4) on 16 random weights, with ground contact."""
    rob = oracle_quadruped()
    cfg = om.TaskConfig(steps=140, window=60)
    I = cfg.layout.n_input(rob.n_vert)
    W = om.init_weights(I, 64, 4, seed=7)
    B = 2
    tv = torch.tensor([[0.05, 0.0, -0.03], [-0.02, 0.0, 0.06]], dtype=F)
    th = torch.tensor([0.15, 0.2], dtype=F)
    x0, v0 = om.initial_state(rob, B)
    Wr = [w.clone().requires_grad_() for w in W]
    loss, _, _ = om.rollout_task(rob, cfg, Wr, tv, th, x0, v0)
    loss.backward()
    g = torch.Generator().manual_seed(8)
    rows = torch.randint(0, 64, (12,), generator=g)
    cols = torch.randint(0, I, (12,), generator=g)
    picks = [(0, (int(r), int(c))) for r, c in zip(rows, cols)]
    picks += [(2, (int(o), int(j))) for o, j in zip(torch.randint(0, 4, (4,), generator=g),
                                                    torch.randint(0, 64, (4,), generator=g))]
    eps = 1e-4
    for wi, idx in picks:
        vals = []
        for sgn in (1, -1):
            Wp = [w.clone() for w in W]
            Wp[wi][idx] += sgn * eps
            lp, _, _ = om.rollout_task(rob, cfg, Wp, tv, th, x0, v0)
            vals.append(lp.item())
        fd = (vals[0] - vals[1]) / (2 * eps)
        ad = Wr[wi].grad[idx].item()
        assert abs(fd - ad) <= 1e-3 * abs(ad) + 1e-10, (wi, idx, fd, ad)
