"""Pins the FP64 MLS-MPM oracle: B-spline partition of unity, mass
conservation of P2G, exact transfer of an affine velocity field through
P2G + G2P (APIC), zero elastic stress at F = I, sticky boundary, the 2D
builder's particle count, the polar factor, and central differences."""

import numpy as np
import torch

from oracle import builders as ob
from oracle import mpm as omp

F64 = torch.float64


def _block(dim, n, B=1, seed=0):
    side = torch.arange(5, dtype=F64)
    grid = torch.stack(torch.meshgrid(*[side] * dim, indexing="ij"), -1).reshape(-1, dim)
    pos = (grid * 0.5 + n / 2 - 1.0) / n
    g = torch.Generator().manual_seed(seed)
    x = pos[None].expand(B, -1, -1) + 0.003 * torch.randn(B, pos.shape[0], dim, generator=g, dtype=F64)
    return x


def test_partition_of_unity():
    p = omp.MPMParams(dim=3, n_grid=16)
    x = _block(3, 16)
    _, fx, W, dpos, _ = omp._stencil(x, p)
    assert torch.allclose(W.sum(-1), torch.ones_like(W[..., 0]), atol=1e-14)
    # first moment: sum_o W_o (o - fx) = 0 for the quadratic B-spline
    assert (W[..., None] * dpos).sum(2).abs().max() < 1e-14


def test_p2g_conserves_mass_and_momentum():
    p = omp.MPMParams(dim=2, n_grid=32, p_mass=0.7)
    x = _block(2, 32, B=2)
    B, P, D = x.shape
    g = torch.Generator().manual_seed(1)
    v = torch.randn(B, P, D, generator=g, dtype=F64)
    C = torch.randn(B, P, D, D, generator=g, dtype=F64)
    Fm = torch.eye(D, dtype=F64).expand(B, P, D, D).clone()
    act = torch.zeros(B, 4, dtype=F64)
    aid = torch.full((B, P), -1, dtype=torch.int64)
    gv, gm, _ = omp.p2g(x, v, C, Fm, act, aid, p)
    assert torch.allclose(gm.sum(1), torch.full((B,), P * 0.7, dtype=F64), atol=1e-12)
    # affine and stress terms carry no net momentum (first moment vanishes)
    assert torch.allclose(gv.sum(1), 0.7 * v.sum(1), atol=1e-12)


def test_affine_field_transfers_exactly():
    """v(x) = b + A x sampled at the particles with C = A: P2G then normalise
    then G2P reproduces v and C exactly (no gravity, no stress at F = I)."""
    p = omp.MPMParams(dim=3, n_grid=16, gravity=0.0, mu=0.0, la=0.0, dt=1e-4)
    x = _block(3, 16)
    A = torch.tensor([[0.3, -0.2, 0.1], [0.05, 0.4, -0.3], [0.2, 0.1, -0.1]], dtype=F64)
    b = torch.tensor([0.5, -0.25, 0.1], dtype=F64)
    v = b + x @ A.T
    B, P, D = x.shape
    C = A.expand(B, P, D, D).clone()
    Fm = torch.eye(D, dtype=F64).expand(B, P, D, D).clone()
    gv, gm, _ = omp.p2g(x, v, C, Fm, torch.zeros(B, 1, dtype=F64),
                        torch.full((B, P), -1, dtype=torch.int64), p)
    gvo = gv / (gm[..., None] + 1e-10)
    _, vn, Cn = omp.g2p(x, gvo, p)
    assert torch.allclose(vn, v, atol=1e-9)
    assert torch.allclose(Cn, C, atol=1e-7)


def test_zero_stress_at_identity():
    p = omp.MPMParams(dim=3, n_grid=16)
    x = _block(3, 16)
    B, P, D = x.shape
    Fm = torch.eye(D, dtype=F64).expand(B, P, D, D).clone()
    zero = torch.zeros(B, P, D, D, dtype=F64)
    gv, gm, Fn = omp.p2g(x, torch.zeros(B, P, D, dtype=F64), zero, Fm,
                         torch.zeros(B, 1, dtype=F64), torch.full((B, P), -1, dtype=torch.int64), p)
    assert gv.abs().max() < 1e-13 and torch.allclose(Fn, Fm)


def test_polar_is_rotation_times_spd():
    g = torch.Generator().manual_seed(2)
    Fm = torch.eye(3, dtype=F64) + 0.3 * torch.randn(50, 3, 3, generator=g, dtype=F64)
    Fm = Fm[torch.det(Fm) > 0]
    R = omp.polar(Fm)
    assert torch.allclose(R, omp.polar_svd(Fm), atol=1e-12)
    assert torch.allclose(R @ R.transpose(-1, -2), torch.eye(3, dtype=F64).expand_as(R), atol=1e-12)
    assert torch.allclose(torch.det(R), torch.ones(R.shape[0], dtype=F64), atol=1e-12)
    S = R.transpose(-1, -2) @ Fm
    assert torch.allclose(S, S.transpose(-1, -2), atol=1e-12)
    assert (torch.linalg.eigvalsh(S) > 0).all()


def test_sticky_boundary():
    p = omp.MPMParams(dim=2, n_grid=32)
    n = 32
    gv = torch.ones(1, n * n, 2, dtype=F64)
    gm = torch.ones(1, n * n, dtype=F64)
    vo = omp.grid_op(gv, gm, p).reshape(n, n, 2)
    assert vo[:3].abs().max() == 0 and vo[:, :3].abs().max() == 0
    assert vo[n - 2:].abs().max() == 0 and vo[:, n - 2:].abs().max() == 0   # i > n - bound
    assert torch.allclose(vo[3:n - 2, 3:n - 2, 0], torch.ones(n - 5, n - 5, dtype=F64), atol=1e-9)
    assert torch.allclose(vo[3:n - 2, 3:n - 2, 1], torch.full((n - 5, n - 5), 1 - 1e-2, dtype=F64), atol=1e-9)


def test_2d_design_builder():
    pos, act, n_act = ob.MPM2D(ob.load_json("tests/golden/robot_mpm_1.json")).build()
    assert len(pos) == 320 and n_act == 4
    from diffloco import mpm as dm
    p2, a2, n2 = dm.particles_2d(ob.load_json("tests/golden/robot_mpm_1.json"))
    assert np.array_equal(np.asarray(pos), p2) and list(a2) == list(act) and n2 == 4


def test_step_gradient_matches_central_differences():
    p = omp.MPMParams(dim=3, n_grid=16, act_axis=0, dt=5e-4, mu=200.0, la=150.0,
                      act_strength=5.0, p_vol=1e-3, p_mass=1e-3)
    x = _block(3, 16)
    B, P, D = x.shape
    g = torch.Generator().manual_seed(3)
    v = 0.2 * torch.randn(B, P, D, generator=g, dtype=F64)
    C = torch.randn(B, P, D, D, generator=g, dtype=F64)
    Fm = torch.eye(D, dtype=F64) + 0.05 * torch.randn(B, P, D, D, generator=g, dtype=F64)
    act = torch.rand(B, 2, generator=g, dtype=F64) - 0.5
    aid = torch.tensor([i % 3 - 1 for i in range(P)])[None]
    wx = torch.randn(B, P, D, generator=g, dtype=F64)

    def f(xx, vv, CC, FF, aa):
        out = omp.step(xx, vv, CC, FF, aa, aid, p)
        return (out[0] * wx).sum() + out[1].pow(2).sum() + out[3].sum()

    args = [t.clone().requires_grad_() for t in (x, v, C, Fm, act)]
    f(*args).backward()
    eps = 1e-6
    for ai, idx in [(0, (0, 7, 1)), (1, (0, 11, 2)), (2, (0, 3, 0, 2)), (3, (0, 20, 1, 1)), (4, (0, 1))]:
        hi = [t.clone() for t in (x, v, C, Fm, act)]
        lo = [t.clone() for t in (x, v, C, Fm, act)]
        hi[ai][idx] += eps
        lo[ai][idx] -= eps
        fd = (f(*hi) - f(*lo)).item() / (2 * eps)
        ad = args[ai].grad[idx].item()
        assert abs(fd - ad) <= 1e-5 * max(1.0, abs(ad)), (ai, idx, fd, ad)


def test_polar_gradient_finite_differences_at_identity():
    """d R / d F at F = I (repeated singular values) from the Newton oracle
    against central differences: R(I + eps E) for skew E rotates, for
    symmetric E stays I to first order."""
    g = torch.Generator().manual_seed(4)
    Fm = torch.eye(3, dtype=F64)[None].clone().requires_grad_()
    Rb = torch.randn(1, 3, 3, generator=g, dtype=F64)
    (omp.polar(Fm) * Rb).sum().backward()
    eps = 1e-6
    for i in range(3):
        for j in range(3):
            E = torch.zeros(1, 3, 3, dtype=F64)
            E[0, i, j] = eps
            fd = ((omp.polar_svd(torch.eye(3, dtype=F64)[None] + E) * Rb).sum() -
                  (omp.polar_svd(torch.eye(3, dtype=F64)[None] - E) * Rb).sum()) / (2 * eps)
            assert abs(fd.item() - Fm.grad[0, i, j].item()) < 1e-8
