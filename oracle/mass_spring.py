"""Restated mass-spring rollout in FP64 (TEST INFRASTRUCTURE ONLY -- see
oracle/__init__.py).

Everything is written with torch FP64 CPU tensors, vectorised over the batch
axis ``k``; ``torch.autograd`` over this forward is the stand-in for Taichi's
generated adjoints (``ti.ad.Tape``, ``multitask/diffphy_trainer.py:246``).

Reference functions restated here (file:line in the reference repo):

* ``spring_impulses``  -- ``SolverMassSpring.apply_spring_force``
  (``multitask/solver_mass_spring.py:86-105``)
* ``advance_toi``      -- ``SolverMassSpring.advance_toi`` (``:117-148``)
* ``pre_advance``      -- ``compute_center/height/rotation`` (``:150-198``)
* ``nn_input``         -- ``TaichiEnv.nn_input`` (``multitask/taichi_env.py:149-183``),
  generalised by a ``SensorLayout`` so the same index arithmetic also
  describes the quadruped workload's sensor vector
* ``mlp_forward``      -- ``Model.nn1/nn2/forward`` (``multitask/nn.py:130-150``)
* ``loss_velocity``    -- ``compute_loss_velocity`` (``taichi_env.py:185-211``)
* ``loss_rotation``    -- ``compute_loss_rotation`` (``taichi_env.py:235-252``)
* ``loss_height`` / ``loss_pose`` / ``loss_actuation`` / ``loss_crawl`` --
  ``compute_loss_height/pose/actuation/crawl`` (``taichi_env.py:213-233, 276-293``)
* ``train_targets``    -- ``initialize_train`` (``taichi_env.py:385-447``) with an
  injected random pool instead of ``ti.random``
* ``rollout_ref``      -- ``DiffPhyTrainer.simulate(train=True)``
  (``diffphy_trainer.py:219-256``): ``T-1`` advances, ``center[T-1]`` never
  written (it stays zero and the losses read it)
* ``rollout_task``     -- the quadruped workload: ``T`` advances with per-task
  targets and the windowed COM-velocity/height loss
* ``adam_update``      -- ``Model.adam_update`` (``nn.py:184-191``)

Pinning (tests/test_oracle_mass_spring.py): free fall against the closed form
of semi-implicit Euler; momentum conservation of the spring impulses; the
rest-pose impulse ``dt*k*(L-L0)/L*(xa-xb)`` with ``L-L0 = sqrt(L0^2+1e-8)+1e-4-L0``;
a two-mass oscillator's period; hand-worked ground-contact cases for each
friction branch; the MLP against ``torch.nn.functional.linear``; Adam against
``torch.optim.Adam``; losses against brute-force loops on tiny inputs; every
gradient against FP64 central finite differences.
"""

import math
from dataclasses import dataclass, field

import numpy as np
import torch

F64 = torch.float64


@dataclass
class Physics:
    dt: float = 0.004
    gravity: float = -1.8
    ground: float = 0.1
    drag: float = 0.0
    dashpot: float = 0.1          # impulse coefficient, no dt (solver_mass_spring.py:103)
    friction: float = 0.4         # hard-coded 0.4 in the reference (:136)
    stiffness_scale: float = 0.1  # springs[i][3] / 10 (:64)


@dataclass
class SensorLayout:
    n_sin: int = 100
    run_period: int = 100
    vcom: bool = False            # 3 slots of COM velocity after the sinusoids
    vert_pad: int = 3             # zero slots after each vertex offset (ref: 3)
    offset_scale: float = 0.2     # (x - c) / 0.2
    tv_comps: tuple = (0, 2)      # components of target_v fed per copy
    dup_v: int = 150
    max_speed: float = 0.1
    dup_h: int = 150
    dup_c: int = 0

    def n_input(self, n_vert):
        return (self.n_sin + 3 * int(self.vcom) + n_vert * (3 + self.vert_pad)
                + self.dup_v * len(self.tv_comps) + self.dup_h + self.dup_c)


@dataclass
class Robot:
    objects: torch.Tensor        # [N,3] rest positions (builder output, f64)
    a: torch.Tensor              # [S] int64
    b: torch.Tensor
    L0: torch.Tensor             # [S]
    k: torch.Tensor              # [S] stiffness as stored by the builder (unscaled)
    act_coeff: torch.Tensor      # [S]
    act_index: torch.Tensor      # [S] int64 -- MLP output driving spring s

    @property
    def n_vert(self):
        return self.objects.shape[0]

    @property
    def n_spring(self):
        return self.a.shape[0]

    def to(self, dtype):
        """The same robot with real-valued fields in ``dtype`` (an FP32 run
        of the restatement measures what FP32 arithmetic alone changes)."""
        return Robot(objects=self.objects.to(dtype), a=self.a, b=self.b, L0=self.L0.to(dtype),
                     k=self.k.to(dtype), act_coeff=self.act_coeff.to(dtype),
                     act_index=self.act_index)


def make_robot(objects, springs, act_index=None):
    S = len(springs)
    if act_index is None:                    # reference: one MLP output per spring
        act_index = list(range(S))
    return Robot(objects=torch.tensor(objects, dtype=F64),
                 a=torch.tensor([s[0] for s in springs], dtype=torch.int64),
                 b=torch.tensor([s[1] for s in springs], dtype=torch.int64),
                 L0=torch.tensor([s[2] for s in springs], dtype=F64),
                 k=torch.tensor([s[3] for s in springs], dtype=F64),
                 act_coeff=torch.tensor([s[4] for s in springs], dtype=F64),
                 act_index=torch.tensor(act_index, dtype=torch.int64))


# ----------------------------------------------------------------------------
# physics
# ----------------------------------------------------------------------------

def spring_impulses(x, v, act, robot, phys):
    """Spring + dashpot impulses accumulated into ``v_inc`` (solver_mass_spring.py:86-105).

    x, v: [B,N,3]; act: [B,O] MLP outputs.  Returns v_inc [B,N,3]."""
    dist = x[:, robot.a] - x[:, robot.b]
    nsq = (dist * dist).sum(-1)
    length = torch.sqrt(nsq + 1e-8) + 1e-4                      # dist.norm(1e-8) + 1e-4
    target = robot.L0 * (1.0 + robot.act_coeff * act[:, robot.act_index])
    kk = robot.k * phys.stiffness_scale
    imp = (phys.dt * (length - target) * kk / length)[..., None] * dist
    d = dist * (1.0 / torch.sqrt(nsq))[..., None]               # normalized()
    vrel = ((v[:, robot.a] - v[:, robot.b]) * d).sum(-1)
    imp = imp + phys.dashpot * vrel[..., None] * d
    v_inc = torch.zeros_like(x)
    v_inc = v_inc.index_add(1, robot.a, -imp)
    v_inc = v_inc.index_add(1, robot.b, imp)
    return v_inc


def advance_toi(x, v, v_inc, phys):
    """Semi-implicit Euler with time-of-impact ground contact and per-axis
    friction clamp (solver_mass_spring.py:117-148).  Returns (x', v')."""
    s = math.exp(-phys.dt * phys.drag)
    g = torch.zeros(3, dtype=x.dtype)
    g[1] = phys.dt * phys.gravity
    old_v = s * v + g + v_inc
    new_x = x + phys.dt * old_v
    vy = old_v[..., 1]
    hit = (new_x[..., 1] < phys.ground) & (vy < -1e-4)
    safe_vy = torch.where(hit, vy, torch.ones_like(vy))
    toi = torch.where(hit, -(x[..., 1] - phys.ground) / safe_vy, torch.zeros_like(vy))
    fr = phys.friction * (-vy)
    zero = torch.zeros_like(vy)

    def clamp(c):
        return torch.where(c < 0, torch.minimum(zero, c + fr), torch.maximum(zero, c - fr))

    hv = torch.stack([clamp(old_v[..., 0]), zero, clamp(old_v[..., 2])], -1)
    new_v = torch.where(hit[..., None], hv, old_v)
    new_x = x + toi[..., None] * old_v + (phys.dt - toi)[..., None] * new_v
    return new_x, new_v


def step(x, v, act, robot, phys):
    """``advance(t)`` = apply_spring_force(t) then advance_toi(t+1) (:200-202)."""
    return advance_toi(x, v, spring_impulses(x, v, act, robot, phys), phys)


def pre_advance(t, x_t, x_0, center_0, height_prev, jump_period, robot_id=None):
    """compute_center / compute_height / compute_rotation (:150-192).

    Returns (center [B,3], height [B], upper_height [B], rotation [B])."""
    N = x_t.shape[1]
    center = (x_t / N).sum(1)
    h = torch.clamp(x_t[..., 1].min(1).values, max=10.0)
    if t % jump_period == 0 or height_prev is None:
        height = h
    else:
        height = torch.maximum(height_prev, h)
    upper = torch.clamp(x_t[..., 1].max(1).values, min=-10.0)
    if robot_id == 100:
        d = x_t[:, 45] + x_t[:, 46] - x_t[:, 34] - x_t[:, 36]
    else:
        head = (x_0[..., 0] < center_0[:, None, 0]).to(x_t.dtype)      # [B,N]
        tail = 1.0 - head
        hc = (head[..., None] * x_t).sum(1)
        tc = (tail[..., None] * x_t).sum(1)
        d = -hc * tail.sum(1, keepdim=True) + tc * head.sum(1, keepdim=True)
    rotation = torch.atan2(d[:, 2], d[:, 0])
    return center, height, upper, rotation


# ----------------------------------------------------------------------------
# controller
# ----------------------------------------------------------------------------

def nn_input(t, x_t, v_t, center, tv, th, tc, layout, offset=0):
    """Sensor vector (taichi_env.py:149-183) for one step.  tv: [B,3];
    th, tc: [B].  Returns [B, I]."""
    B, N, _ = x_t.shape
    parts = []
    if layout.n_sin > 0:
        omega = 2 * math.pi / layout.run_period        # spring_omega * dt (config_sim.py:114)
        j = torch.arange(layout.n_sin, dtype=x_t.dtype)
        s = torch.sin(omega * (t + offset) + 2 * math.pi / layout.n_sin * j)
        parts.append(s.expand(B, -1))
    if layout.vcom:
        parts.append((v_t / N).sum(1))
    off = (x_t - center[:, None, :]) / layout.offset_scale
    if layout.vert_pad:
        off = torch.cat([off, torch.zeros(B, N, layout.vert_pad, dtype=x_t.dtype)], -1)
    parts.append(off.reshape(B, -1))
    if layout.dup_v:
        comp = tv[:, list(layout.tv_comps)] / layout.max_speed
        parts.append(comp.repeat(1, layout.dup_v))
    if layout.dup_h:
        parts.append(th[:, None].expand(B, layout.dup_h))
    if layout.dup_c:
        parts.append(tc[:, None].expand(B, layout.dup_c))
    return torch.cat(parts, -1)


def activation(name):
    return {"sin": torch.sin, "tanh": torch.tanh}[name]


def mlp_forward(inp, W1, b1, W2, b2, act="sin"):
    """nn1 + nn2 (nn.py:130-146).  W1 [H,I], W2 [O,H].  Returns (hidden_act, out_act)."""
    f = activation(act)
    h = f(inp @ W1.T + b1)
    return h, f(h @ W2.T + b2)


def adam_update(w, grad, m, v, it, lr=3e-4, b1=0.9, b2=0.9, eps=1e-8):
    """nn.py:184-191, in place on numpy/torch arrays."""
    m[...] = b1 * m + (1 - b1) * grad
    v[...] = b2 * v + (1 - b2) * grad * grad
    m_cap = m / (1 - b1 ** (it + 1))
    v_cap = v / (1 - b2 ** (it + 1))
    w[...] = w - (lr * m_cap) / (v_cap ** 0.5 + eps)


# ----------------------------------------------------------------------------
# targets and losses
# ----------------------------------------------------------------------------

def train_targets(pool, steps, batch, turn_period, max_velocity):
    """initialize_train, 3D branch (taichi_env.py:385-447).  pool: flat array of
    uniforms of length >= batch*(steps//turn_period+1)*3.  Returns
    tv [T,B,3], th [T,B], tc [T,B]."""
    pool = np.asarray(pool, dtype=np.float64)
    tv = np.zeros((steps, batch, 3))
    th = np.zeros((steps, batch))
    tc = np.zeros((steps, batch))
    for t in range(steps):
        for k in range(batch):
            q = (t // turn_period * batch + k) * 3
            if pool[q] < 0.25:
                th[t, k] = -1.0
            elif pool[q] < 0.5:
                th[t, k] = 1.0
            elif pool[q] < 0.75:
                tv[t, k, 2] = (pool[q + 1] * 2.0 - 1.0) * max_velocity
            else:
                tv[t, k, 0] = (pool[q + 1] * 2.0 - 1.0) * max_velocity
    return torch.tensor(tv), torch.tensor(th), torch.tensor(tc)


def _at(arr, t):
    """Field read with never-written entries (t >= len) reading zero."""
    return arr[t] if t < arr.shape[0] else torch.zeros_like(arr[0])


def loss_velocity(center, tv, steps, run_period, turn_period):
    """compute_loss_velocity, 3D branch (taichi_env.py:185-211).
    center [T',B,3] (unwritten steps are zero), tv [T,B,3].
    Returns (loss, loss_batch [B])."""
    B = center.shape[1]
    total = torch.zeros(B, dtype=F64)
    for t in range(run_period, steps + 1):
        if t % turn_period > run_period:
            c1, c0, tgt = _at(center, t), _at(center, t - run_period), tv[t - run_period]
            lx = (c1[:, 0] - c0[:, 0] - tgt[:, 0]) ** 2 / B / steps * 100
            lz = (c1[:, 2] - c0[:, 2] - tgt[:, 2]) ** 2 / B / steps * 100
            total = total + lx + lz
    return total.sum(), total * B


def loss_rotation(x, center, th, steps, run_period, turn_period):
    """compute_loss_rotation, active loop (taichi_env.py:244-252).
    x [T',B,N,3], center [T',B,3], th [T,B]."""
    B = center.shape[1]
    total = torch.zeros(B, dtype=F64)
    for t in range(1, steps + 1):
        if t % turn_period > run_period:
            t0 = _at(x, t) - _at(center, t)[:, None, :]
            t1 = _at(x, t - run_period) - _at(center, t - run_period)[:, None, :]
            ang = th[t - run_period] * 72.0 / 180.0 * 3.1415926535
            ca, sa = torch.cos(ang)[:, None], torch.sin(ang)[:, None]
            t2 = torch.stack([ca * t1[..., 0] - sa * t1[..., 2], t1[..., 1],
                              sa * t1[..., 0] + ca * t1[..., 2]], -1)
            total = total + ((t0 - t2) ** 2).sum(-1).sum(-1) / steps / 40
    return total.sum(), total * B


def loss_height(height, th, steps, jump_period):
    """compute_loss_height (taichi_env.py:213-222): at t in [1, steps] with
    t % jump == jump - 1 and target_h > 0.1 + 1e-6, (height - target_h)^2 /
    B / (steps // jump) * 100.  height [T',B] (unwritten rows read zero),
    th [T,B].  Returns (loss, loss_batch [B], max contributing height)."""
    B = th.shape[1]
    total = torch.zeros(B, dtype=th.dtype)
    hmax = None
    # target rows past the given ones (t = steps in training) read zero, as
    # the fields' never-written rows do
    for t in range(1, steps + 1):
        if t % jump_period != jump_period - 1:
            continue
        h, tg = _at(height, t), _at(th, t)
        on = tg > 0.1 + 1e-6
        term = (h - tg) ** 2 / B / (steps // jump_period) * 100
        total = total + torch.where(on, term, torch.zeros_like(term))
        if on.any():
            m = h[on].max().item()
            hmax = m if hmax is None else max(hmax, m)
    return total.sum(), total * B, hmax


def loss_pose(x, objects, steps, jump_period):
    """compute_loss_pose (taichi_env.py:224-233): at t in [1, steps] with
    t % jump == 0, sum_i |x_i - initial_object_i|^2 / B / (steps // jump).
    x [T',B,N,3] (unwritten rows read zero), objects [N,3]."""
    B = x.shape[1]
    total = torch.zeros(B, dtype=x.dtype)
    for t in range(1, steps + 1):
        if t % jump_period == 0:
            d = _at(x, t) - objects[None]
            total = total + (d * d).sum(-1).sum(-1) / B / (steps // jump_period)
    return total.sum(), total * B


def loss_actuation(act, tv, th, steps):
    """compute_loss_actuation (taichi_env.py:276-285): at t in [0, steps)
    with target_h < 0.1 + 1e-4, per actuation output max(|a| -
    sqrt(|tv_x| / 0.08), 0) / n_out / B / steps * 10.  act [T',B,O]."""
    B, O = act.shape[1], act.shape[2]
    total = torch.zeros(B, dtype=act.dtype)
    for t in range(steps):
        a = _at(act, t)
        lim = torch.sqrt(tv[t, :, 0].abs() / 0.08)
        term = torch.clamp(a.abs() - lim[:, None], min=0.0).sum(-1) / O / B / steps * 10
        total = total + torch.where(th[t] < 0.1 + 1e-4, term, torch.zeros_like(term))
    return total.sum(), total * B


def loss_crawl(upper, tc, steps):
    """compute_loss_crawl (taichi_env.py:287-293): at t in [1, steps] with
    target_c > 1 - 1e-4, max(upper_height - 0.1, 0)^2 / B / steps * 5."""
    B = tc.shape[1]
    total = torch.zeros(B, dtype=tc.dtype)
    for t in range(1, steps + 1):
        u = torch.clamp(_at(upper, t) - 0.1, min=0.0)
        term = u * u / B / steps * 5
        total = total + torch.where(_at(tc, t) > 1 - 1e-4, term, torch.zeros_like(term))
    return total.sum(), total * B


def loss_task_terms(center, vcom, tv, th):
    """Per-step term of the quadruped workload's loss:
    |v_com,xz - v_tgt,xz|^2 + (c_y - h_tgt)^2, per robot."""
    return ((vcom[:, 0] - tv[:, 0]) ** 2 + (vcom[:, 2] - tv[:, 2]) ** 2
            + (center[:, 1] - th) ** 2)


# ----------------------------------------------------------------------------
# rollouts
# ----------------------------------------------------------------------------

@dataclass
class RefConfig:
    """cfg3d/sim_quad.json as derived by config_sim.py:67-147."""
    steps: int = 1000
    run_period: int = 100
    turn_period: int = 500
    jump_period: int = 500
    max_speed: float = 0.1
    physics: Physics = field(default_factory=Physics)
    layout: SensorLayout = field(default_factory=SensorLayout)
    act: str = "sin"


def initial_state(robot, batch, shift=(0.4, 0.0, 0.0)):
    """setup_robot / reset_robot (taichi_env.py:487-504): x[0] = objects + shift,
    v[0] = 0."""
    x0 = robot.objects + torch.tensor(shift, dtype=F64)
    return x0.expand(batch, -1, -1).clone(), torch.zeros(batch, robot.n_vert, 3, dtype=F64)


def rollout_ref(robot, cfg, weights, tv, th, tc, x0, v0, record=False,
                losses=("velocity", "rotation"), control_length=1, objects=None):
    """simulate(train=True) (diffphy_trainer.py:231-256); the default
    ``losses`` are sim_quad's loss_enable = {velocity, rotation}
    (sim_quad.json:7), any of height, pose, actuation and crawl may be added
    (taichi_env.py:304-331).  pre_advance fills center / height / upper rows
    0..T-2 only; rows T-1.. stay zero and the losses read them.  With
    control_length > 1 the network runs at t % control_length == 0 and its
    output is held in between (diff_copy, diffphy_trainer.py:215-218,
    249-254).  ``objects``: the pose loss's rest positions (default x0[0]).
    Returns (loss, loss_batch, traj)."""
    T = cfg.steps
    W1, b1, W2, b2 = weights
    B, N, _ = x0.shape
    xs, vs, cs, acts, hs, us = [x0], [v0], [], [], [], []
    x, v = x0, v0
    height = None
    for t in range(T - 1):
        c, height, upper, _ = pre_advance(t, x, x0, (x0 / N).sum(1), height, cfg.jump_period)
        inp = nn_input(t, x, v, c, tv[t], th[t], tc[t], cfg.layout)
        if t % control_length == 0:
            _, a = mlp_forward(inp, W1, b1, W2, b2, cfg.act)
        x, v = step(x, v, a, robot, cfg.physics)
        xs.append(x); vs.append(v); cs.append(c); acts.append(a); hs.append(height); us.append(upper)
    # center[T-1] is never written (the loop stops at T-2) and reads as zero
    center = torch.stack(cs + [torch.zeros_like(cs[0])])         # [T,B,3]
    X = torch.stack(xs)                                           # [T,B,N,3]
    loss = torch.zeros((), dtype=x0.dtype)
    lb = torch.zeros(B, dtype=x0.dtype)
    parts = []
    if "velocity" in losses:
        parts.append(loss_velocity(center, tv, T, cfg.run_period, cfg.turn_period))
    if "height" in losses:
        parts.append(loss_height(torch.stack(hs), th, T, cfg.jump_period)[:2])
    if "pose" in losses:
        parts.append(loss_pose(X, x0[0] if objects is None else objects, T, cfg.jump_period))
    if "actuation" in losses:
        parts.append(loss_actuation(torch.stack(acts), tv, th, T))
    if "rotation" in losses:
        parts.append(loss_rotation(X, center, th, T, cfg.run_period, cfg.turn_period))
    if "crawl" in losses:
        parts.append(loss_crawl(torch.stack(us), tc, T))
    for l, b in parts:
        loss = loss + l
        lb = lb + b
    traj = dict(x=X, v=torch.stack(vs), center=center, act=torch.stack(acts),
                height=torch.stack(hs), upper=torch.stack(us)) if record else None
    return loss, lb, traj


def script_targets(steps, B, x0, y0, x1, y1, x2, y2, x3, y3):
    """initialize_script (taichi_env.py:340-376): (x, z) velocity targets from
    the four pairs for steps [0,1000), [1000,2000), [2000,3000), [3000,4000);
    then zero velocity with height target -1 below 5000 and +1 after."""
    tv = torch.zeros(steps, B, 3, dtype=F64)
    th = torch.zeros(steps, B, dtype=F64)
    pairs = [(x0, y0), (x1, y1), (x2, y2), (x3, y3)]
    for t in range(steps):
        if t < 4000:
            tv[t, :, 0], tv[t, :, 2] = pairs[t // 1000]
        else:
            th[t] = -1.0 if t < 5000 else 1.0
    return tv, th, torch.zeros(steps, B, dtype=F64)


def rollout_eval(robot, cfg, weights, tv, th, tc, x0, v0):
    """simulate(train=False) (diffphy_trainer.py:257-268): steps+1 advances
    without a tape, every center written; get_loss(steps) with the
    velocity and rotation losses.  Returns (loss, loss_batch)."""
    T = cfg.steps
    W1, b1, W2, b2 = weights
    B, N, _ = x0.shape
    xs, cs = [x0], []
    x, v = x0, v0
    for t in range(T + 1):
        c = (x / N).sum(1)
        inp = nn_input(t, x, v, c, tv[t], th[t], tc[t], cfg.layout)
        _, a = mlp_forward(inp, W1, b1, W2, b2, cfg.act)
        x, v = step(x, v, a, robot, cfg.physics)
        xs.append(x)
        cs.append(c)
    center = torch.stack(cs)                                      # rows 0..T
    X = torch.stack(xs[: T + 1])
    lv, lvb = loss_velocity(center, tv, T, cfg.run_period, cfg.turn_period)
    lr, lrb = loss_rotation(X, center, th, T, cfg.run_period, cfg.turn_period)
    return lv + lr, lvb + lrb


@dataclass
class TaskConfig:
    """The quadruped workload: T advances, per-task constant targets, loss =
    mean over the last ``window`` states of |v_com,xz - v_tgt|^2 + (c_y - h)^2."""
    steps: int = 1000
    window: int = 500
    physics: Physics = field(default_factory=lambda: Physics(
        dt=0.004, gravity=-1.8, ground=0.0, drag=0.0, dashpot=0.12, friction=0.5,
        stiffness_scale=0.1))
    layout: SensorLayout = field(default_factory=lambda: SensorLayout(
        n_sin=10, run_period=100, vcom=True, vert_pad=0, offset_scale=0.2,
        tv_comps=(0, 1, 2), dup_v=1, max_speed=0.08, dup_h=1, dup_c=0))
    act: str = "tanh"


def rollout_task(robot, cfg, weights, tv, th, x0, v0, record=False):
    """T advances; loss over states t in (T-window, T].  tv [B,3], th [B]."""
    T, Wn = cfg.steps, cfg.window
    W1, b1, W2, b2 = weights
    B, N, _ = x0.shape
    x, v = x0, v0
    zero_c = torch.zeros(B, dtype=x0.dtype)
    total = torch.zeros(B, dtype=x0.dtype)
    xs, vs = [x0], [v0]
    for t in range(T + 1):
        c = (x / N).sum(1)
        vc = (v / N).sum(1)
        if t > T - Wn:
            total = total + loss_task_terms(c, vc, tv, th)
        if t == T:
            break
        inp = nn_input(t, x, v, c, tv, th, zero_c, cfg.layout)
        _, a = mlp_forward(inp, W1, b1, W2, b2, cfg.act)
        x, v = step(x, v, a, robot, cfg.physics)
        if record:
            xs.append(x); vs.append(v)
    loss_b = total / Wn
    traj = dict(x=torch.stack(xs), v=torch.stack(vs)) if record else None
    return loss_b.sum() / B, loss_b, traj


def init_weights(n_in, n_hidden, n_out, seed):
    """N(0, 1/fan_in) weights, zero biases, from a numpy seed."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((n_hidden, n_in)) / math.sqrt(n_in)
    W2 = rng.standard_normal((n_out, n_hidden)) / math.sqrt(n_hidden)
    return [torch.tensor(W1), torch.zeros(n_hidden, dtype=F64),
            torch.tensor(W2), torch.zeros(n_out, dtype=F64)]
