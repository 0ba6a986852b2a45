"""Restated MLS-MPM step in FP64 (TEST INFRASTRUCTURE ONLY -- see
oracle/__init__.py).

Follows ``multitask/solver_mpm.py`` (2D) and generalises it to 3D with the
same rules:

* ``p2g``      -- ``SolverMPM.p2g`` (``solver_mpm.py:83-124``): F update,
  polar decomposition, fixed-corotated stress plus active stress
  ``F A F^T`` with ``A = act_strength * a * e e^T`` (reference: e = y in 2D),
  APIC scatter over the 3^D quadratic B-spline stencil.
* ``grid_op``  -- ``SolverMPM.grid_op`` (``:126-144``): normalise, gravity on
  y, every component zeroed within ``bound`` cells of any wall.
* ``g2p``      -- ``SolverMPM.g2p`` (``:146-165``).
* ``center``   -- ``SolverMPM.compute_center`` (``:167-173``).

The polar factor comes from an unscaled Newton iteration differentiated by
autograd (finite at F = I, where the SVD's backward is not), pinned against
``torch.linalg.svd``; the kernels use a closed form (2D) and a scaled Newton
iteration with a closed-form adjoint (3D).  Gradients come from
``torch.autograd`` over this forward.

Pinned by tests/test_oracle_mpm.py: B-spline partition of unity, mass
conservation, exact transfer of an affine velocity field, zero stress at
F = I without actuation, sticky boundary zeroing, the 2D builder's particle
count, and central finite differences.
"""

import itertools
from dataclasses import dataclass

import torch

F64 = torch.float64


@dataclass
class MPMParams:
    dim: int = 2
    n_grid: int = 32
    dt: float = 1e-3
    gravity: float = -10.0
    p_vol: float = 1.0
    p_mass: float = 1.0
    mu: float = 40.0
    la: float = 40.0
    act_strength: float = 4.0
    act_axis: int = 1
    bound: int = 3

    @property
    def dx(self):
        return 1.0 / self.n_grid

    @property
    def inv_dx(self):
        return float(self.n_grid)


def offsets(dim):
    return torch.tensor(list(itertools.product(range(3), repeat=dim)), dtype=torch.int64)


def _stencil(x, p):
    """base [B,P,D] int, fx [B,P,D], weights W [B,P,3^D], dpos [B,P,3^D,D] (in cells)."""
    xi = x * p.inv_dx
    base = (xi - 0.5).to(torch.int64)                       # truncation, like ti.cast(..., ti.i32)
    fx = xi - base.to(x.dtype)
    w = torch.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2], -2)
    off = offsets(p.dim)                                     # [K,D]
    # W_o = prod_d w[o_d][d]
    W = torch.ones(x.shape[:2] + (off.shape[0],), dtype=x.dtype)
    for d in range(p.dim):
        W = W * w[..., off[:, d], d]
    dpos = off.to(x.dtype)[None, None] - fx[:, :, None, :]
    return base, fx, W, dpos, off


def _node_index(base, off, n, dim):
    idx = base[:, :, None, :] + off[None, None]              # [B,P,K,D]
    flat = torch.zeros(idx.shape[:3], dtype=torch.int64)
    for d in range(dim):
        flat = flat * n + idx[..., d]
    return flat


def polar_svd(Fm):
    """Rotation factor from the SVD, R = U V^T with det R = +1."""
    U, S, Vh = torch.linalg.svd(Fm)
    det = torch.det(U @ Vh)
    fix = torch.ones_like(S)
    fix[..., -1] = torch.sign(det)
    return (U * fix[..., None, :]) @ Vh


def polar(Fm, iters=25):
    """Rotation factor by the Newton iteration X <- (X + X^{-T}) / 2, which
    autograd differentiates through; unlike the SVD's backward it stays
    finite at repeated singular values (F = I at the start of a rollout).
    Its value is pinned against ``polar_svd``."""
    X = Fm
    for _ in range(iters):
        X = 0.5 * (X + torch.linalg.inv(X).transpose(-1, -2))
    return X


def p2g(x, v, C, Fm, act, act_id, p):
    """Returns (grid_v_in [B,n^D,D], grid_m_in [B,n^D], F_new [B,P,D,D])."""
    B, P, D = x.shape
    n = p.n_grid
    eye = torch.eye(D, dtype=x.dtype)
    Fn = (eye + p.dt * C) @ Fm
    J = torch.det(Fn)
    R = polar(Fn)
    a = torch.where(act_id >= 0, act.gather(1, act_id.clamp_min(0)), torch.zeros_like(x[..., 0]))
    e = torch.zeros(D, dtype=x.dtype)
    e[p.act_axis] = 1.0
    A = (p.act_strength * a)[..., None, None] * torch.outer(e, e)
    cauchy = 2 * p.mu * (Fn - R) @ Fn.transpose(-1, -2) + (p.la * (J - 1) * J)[..., None, None] * eye
    cauchy = cauchy + Fn @ A @ Fn.transpose(-1, -2)
    stress = -(p.dt * p.p_vol * 4 * p.inv_dx * p.inv_dx) * cauchy
    affine = stress + p.p_mass * C
    base, fx, W, dpos, off = _stencil(x, p)
    contrib = W[..., None] * (p.p_mass * v[:, :, None, :] +
                              (affine[:, :, None] @ (dpos * p.dx)[..., None])[..., 0])
    flat = _node_index(base, off, n, D)
    gv = torch.zeros(B, n ** D, D, dtype=x.dtype)
    gm = torch.zeros(B, n ** D, dtype=x.dtype)
    for k in range(B):
        gv[k] = gv[k].index_add(0, flat[k].reshape(-1), contrib[k].reshape(-1, D))
        gm[k] = gm[k].index_add(0, flat[k].reshape(-1), (W[k] * p.p_mass).reshape(-1))
    return gv, gm, Fn


def boundary_mask(p):
    n, D = p.n_grid, p.dim
    idx = torch.stack(torch.meshgrid(*[torch.arange(n)] * D, indexing="ij"), -1).reshape(-1, D)
    outside = ((idx < p.bound) | (idx > n - p.bound)).any(-1)
    return outside


def grid_op(gv, gm, p):
    vo = gv / (gm[..., None] + 1e-10)
    g = torch.zeros(p.dim, dtype=gv.dtype)
    g[1] = p.dt * p.gravity
    vo = vo + g
    return torch.where(boundary_mask(p)[None, :, None], torch.zeros_like(vo), vo)


def g2p(x, gvo, p):
    B, P, D = x.shape
    base, fx, W, dpos, off = _stencil(x, p)
    flat = _node_index(base, off, p.n_grid, D)
    gvp = torch.stack([gvo[k][flat[k]] for k in range(B)])         # [B,P,K,D]
    new_v = (W[..., None] * gvp).sum(2)
    new_C = (4 * p.inv_dx * W[..., None, None] * gvp[..., :, None] * dpos[..., None, :]).sum(2)
    return x + p.dt * new_v, new_v, new_C


def step(x, v, C, Fm, act, act_id, p):
    """advance_core (solver_mpm.py:175-180): clear grid, p2g, grid_op, g2p."""
    gv, gm, Fn = p2g(x, v, C, Fm, act, act_id, p)
    gvo = grid_op(gv, gm, p)
    xn, vn, Cn = g2p(x, gvo, p)
    return xn, vn, Cn, Fn


def center(x):
    return x.mean(1)


def sensors(t, x, v, group, n_groups, n_sin, run_period, offset_scale, max_speed, tv, th):
    """Sensor vector of the MPM quadruped: [sinusoids | COM velocity | per
    group mean offset from the COM / scale | tv / max_speed | th]."""
    import math
    B, P, D = x.shape
    c = x.mean(1)
    vc = v.mean(1)
    j = torch.arange(n_sin, dtype=F64)
    s = torch.sin(2 * math.pi / run_period * t + 2 * math.pi / n_sin * j).expand(B, -1)
    offs = []
    for g in range(n_groups):
        m = group == g
        gm = x[:, m].mean(1) if m.any() else torch.zeros(B, D, dtype=F64)
        offs.append((gm - c) / offset_scale)
    return torch.cat([s, vc, torch.cat(offs, -1), tv / max_speed, th[:, None]], -1), c, vc


def rollout_task(x0, act_id, n_act, p, weights, tv, th, T, window, n_sin=10, run_period=100,
                 offset_scale=0.2, max_speed=0.08):
    """T MPM steps under the MLP controller; loss = mean over robots and over
    states t in (T - window, T] of |v_com,xz - tv_xz|^2 + (c_y - th)^2."""
    W1, b1, W2, b2 = weights
    B, P, D = x0.shape
    G = n_act + 1
    group = torch.where(act_id[0] >= 0, act_id[0], torch.tensor(n_act))
    aid = act_id
    x, v = x0, torch.zeros_like(x0)
    C = torch.zeros(B, P, D, D, dtype=F64)
    Fm = torch.eye(D, dtype=F64).expand(B, P, D, D).clone()
    total = torch.zeros(B, dtype=F64)
    for t in range(T + 1):
        inp, c, vc = sensors(t, x, v, group, G, n_sin, run_period, offset_scale, max_speed, tv, th)
        if t > T - window:
            total = total + (vc[:, 0] - tv[:, 0]) ** 2 + (vc[:, 2] - tv[:, 2]) ** 2 + (c[:, 1] - th) ** 2
        if t == T:
            break
        h = torch.tanh(inp @ W1.T + b1)
        act = torch.tanh(h @ W2.T + b2)
        x, v, C, Fm = step(x, v, C, Fm, act, aid, p)
    return (total / window).mean()
