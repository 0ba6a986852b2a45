"""Shared test helpers: oracle-side robots and seeded inputs."""

import numpy as np
import torch

from oracle import builders as ob
from oracle import mass_spring as om


def oracle_reference_robot(robot_id=100):
    b = ob.MassSpring3D(spring_stiffness=1e4, spring_actuation=0.25)
    objects, springs, _ = b.build(robot_id)
    return om.make_robot(objects, springs)


def oracle_quadruped():
    objects, springs, groups = ob.build_quadruped()
    return om.make_robot(objects, springs, groups)


def perturbed_state(robot, B, seed, ground=None, shift=(0.4, 0.0, 0.0), xs=0.004, vs=0.2):
    """Seeded state near the rest pose; with ``ground`` set, a few vertices are
    pushed below it with downward velocity so every contact branch runs."""
    g = torch.Generator().manual_seed(seed)
    x0, v0 = om.initial_state(robot, B, shift)
    x = x0 + xs * torch.randn(x0.shape, generator=g, dtype=torch.float64)
    v = vs * torch.randn(v0.shape, generator=g, dtype=torch.float64)
    if ground is not None:
        N = x.shape[1]
        idx = torch.randperm(N, generator=g)[: max(4, N // 4)]
        x[:, idx, 1] = ground + 0.002 * torch.rand(B, len(idx), generator=g, dtype=torch.float64) - 0.0005
        v[:, idx, 1] = -0.3 - 0.5 * torch.rand(B, len(idx), generator=g, dtype=torch.float64)
    return x, v


def to_soa(x, dtype=torch.float64, device="cuda"):
    """[B,N,3] -> device [B,3,N]."""
    return x.transpose(1, 2).contiguous().to(device=device, dtype=dtype)


def from_soa(x):
    return x.double().cpu().transpose(-1, -2)


def rel_err(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return ((a - b).norm() / b.norm().clamp_min(1e-300)).item()


def max_rel(a, b, floor):
    """max |a-b| / max(|b|, floor) elementwise."""
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return ((a - b).abs() / b.abs().clamp_min(floor)).max().item()


def task_targets(B, seed):
    rng = np.random.default_rng(seed)
    tv = np.stack([rng.uniform(-0.08, 0.08, B), np.zeros(B), rng.uniform(-0.08, 0.08, B)], 1)
    th = rng.uniform(0.1, 0.25, B)
    return tv, th
