"""HIP MLS-MPM step (P2G, grid, G2P) and its adjoint against the FP64 oracle,
2D (the reference's robot_mpm_1 design) and 3D.  FP64 agrees to rounding;
FP32 within 1e-5 relative on positions (atomics reorder the P2G sums, the
only run-to-run variation of this path)."""

import json
import sys

import numpy as np
import pytest
import torch

from diffloco import mpm as dm
from oracle import builders as ob
from oracle import mpm as omp

from helpers import rel_err

pytestmark = pytest.mark.gpu
DT = {"f64": torch.float64, "f32": torch.float32}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(dim, B, seed):
    g = torch.Generator().manual_seed(seed)
    if dim == 2:
        pos, ids, n_act = ob.MPM2D(ob.load_json("tests/golden/robot_mpm_1.json")).build()
        pos = torch.tensor(pos, dtype=torch.float64) + torch.tensor([0.2, 0.15])
        params = dm.reference_params()
        op = omp.MPMParams()
    else:
        n = 16
        side = torch.arange(6, dtype=torch.float64)
        grid = torch.stack(torch.meshgrid(side, side, side, indexing="ij"), -1).reshape(-1, 3)
        pos = (grid * 0.5 + 5.0) / n
        ids = [(i % 5) - 1 for i in range(len(pos))]
        n_act = 4
        params = dm.quadruped3d_params(n_grid=n)
        params.dt = 1e-3
        op = omp.MPMParams(dim=3, n_grid=n, dt=params.dt, gravity=params.gravity,
                           p_vol=params.p_vol, p_mass=params.p_mass, mu=params.mu, la=params.la,
                           act_strength=params.act_strength, act_axis=0, bound=3)
    P, D = pos.shape
    x = pos[None].expand(B, -1, -1) + 0.002 * torch.randn(B, P, D, generator=g, dtype=torch.float64)
    v = 0.3 * torch.randn(B, P, D, generator=g, dtype=torch.float64)
    C = 2.0 * torch.randn(B, P, D, D, generator=g, dtype=torch.float64)
    F = torch.eye(D, dtype=torch.float64) + 0.05 * torch.randn(B, P, D, D, generator=g, dtype=torch.float64)
    act = torch.rand(B, n_act, generator=g, dtype=torch.float64) * 2 - 1
    act_id = torch.tensor(ids, dtype=torch.int64)[None].expand(B, -1)
    return params, op, x, v, C, F, act, act_id, n_act


def _vsoa(a, dt):
    return a.transpose(1, 2).contiguous().to("cuda", dt)


def _msoa(a, dt):
    B, P, D, _ = a.shape
    return a.reshape(B, P, D * D).transpose(1, 2).contiguous().to("cuda", dt)


def _vback(a):
    return a.double().cpu().transpose(1, 2)


def _mback(a, D):
    B, _, P = a.shape
    return a.double().cpu().transpose(1, 2).reshape(B, P, D, D)


@pytest.mark.parametrize("dim,prec", [(2, "f64"), (2, "f32"), (3, "f64"), (3, "f32")])
def test_mpm_step_matches_oracle(dim, prec):
    _say(f"mpm step {dim}d {prec}")
    B = 2
    params, op, x, v, C, F, act, act_id, n_act = _case(dim, B, seed=20 + dim)
    dt = DT[prec]
    st = dm.MPMStep(params, B, x.shape[1], n_act, act_id[0].numpy(), dt)
    ins = [_vsoa(x, dt), _vsoa(v, dt), _msoa(C, dt), _msoa(F, dt), act.to("cuda", dt)]
    outs = [torch.empty_like(ins[0]), torch.empty_like(ins[1]), torch.empty_like(ins[2]),
            torch.empty_like(ins[3])]
    st.forward(*ins, *outs)
    torch.cuda.synchronize()
    st.check()
    xo, vo, Co, Fo = omp.step(_vback(ins[0]), _vback(ins[1]), _mback(ins[2], dim),
                              _mback(ins[3], dim), ins[4].double().cpu(), act_id, op)
    tol = 1e-10 if prec == "f64" else 2e-5
    assert rel_err(_vback(outs[0]), xo) < tol
    assert rel_err(_vback(outs[1]), vo) < tol * 10
    assert rel_err(_mback(outs[2], dim), Co) < tol * 10
    assert rel_err(_mback(outs[3], dim), Fo) < tol


@pytest.mark.parametrize("dim", [2, 3])
def test_mpm_adjoint_matches_autograd(dim):
    _say(f"mpm adjoint {dim}d")
    B = 2
    params, op, x, v, C, F, act, act_id, n_act = _case(dim, B, seed=30 + dim)
    dt = torch.float64
    P = x.shape[1]
    g = torch.Generator().manual_seed(40 + dim)
    xb1, vb1 = torch.randn(B, P, dim, generator=g, dtype=dt), torch.randn(B, P, dim, generator=g, dtype=dt)
    Cb1, Fb1 = torch.randn(B, P, dim, dim, generator=g, dtype=dt), torch.randn(B, P, dim, dim, generator=g, dtype=dt)
    st = dm.MPMStep(params, B, P, n_act, act_id[0].numpy(), dt)
    ins = [_vsoa(x, dt), _vsoa(v, dt), _msoa(C, dt), _msoa(F, dt), act.to("cuda", dt)]
    ups = [_vsoa(xb1, dt), _vsoa(vb1, dt), _msoa(Cb1, dt), _msoa(Fb1, dt)]
    xb, vb = torch.zeros_like(ins[0]), torch.zeros_like(ins[1])
    Cb, Fb = torch.zeros_like(ins[2]), torch.zeros_like(ins[3])
    ab = torch.zeros_like(ins[4])
    st.backward(*ins, *ups, xb, vb, Cb, Fb, ab)
    torch.cuda.synchronize()
    st.check()
    xr, vr, Cr, Fr, ar = [t.clone().requires_grad_() for t in (x, v, C, F, act)]
    xo, vo, Co, Fo = omp.step(xr, vr, Cr, Fr, ar, act_id, op)
    ((xo * xb1).sum() + (vo * vb1).sum() + (Co * Cb1).sum() + (Fo * Fb1).sum()).backward()
    assert rel_err(_vback(xb), xr.grad) < 1e-9
    assert rel_err(_vback(vb), vr.grad) < 1e-9
    assert rel_err(_mback(Cb, dim), Cr.grad) < 1e-9
    assert rel_err(_mback(Fb, dim), Fr.grad) < 1e-8
    assert rel_err(ab.cpu(), ar.grad) < 1e-8


def _batched_case(offsets=((5.0, 3.2, 5.0),), n=16, side=5, perm_seed=None):
    """Blobs of side^3 particles, half a cell apart, at the given cell
    offsets; the 3D quadruped material at dt = 2e-4 (stable for these
    blobs; the state stays well inside the grid)."""
    params = dm.quadruped3d_params(n_grid=n)
    params.dt = 2e-4
    ax = torch.arange(side, dtype=torch.float64)
    grid = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    pos = torch.cat([(grid * 0.5 + torch.tensor(o)) / n for o in offsets])
    ids = torch.tensor([(i % 5) - 1 for i in range(len(pos))])
    if perm_seed is not None:
        perm = torch.randperm(len(pos), generator=torch.Generator().manual_seed(perm_seed))
        pos, ids = pos[perm], ids[perm]
    op = omp.MPMParams(dim=3, n_grid=n, dt=params.dt, gravity=params.gravity, p_vol=params.p_vol,
                       p_mass=params.p_mass, mu=params.mu, la=params.la,
                       act_strength=params.act_strength, act_axis=0, bound=3)
    return params, op, pos, ids


def _weights(I, seed=6):
    g = torch.Generator().manual_seed(seed)
    return [torch.randn(64, I, generator=g, dtype=torch.float64) / I ** 0.5,
            0.1 * torch.randn(64, generator=g, dtype=torch.float64),
            torch.randn(4, 64, generator=g, dtype=torch.float64) / 8.0,
            0.1 * torch.randn(4, generator=g, dtype=torch.float64)]


def _run_batched(params, op, pos, ids, B, T, K, W, dtype, sort=True, **kw):
    batch = dm.MPMBatch(params, pos.numpy(), ids.numpy(), 4, B, T, K=K, window=W, dtype=dtype,
                        sort_particles=sort, **kw)
    rng = np.random.default_rng(5)
    tv = np.stack([rng.uniform(-0.08, 0.08, B), np.zeros(B), rng.uniform(-0.08, 0.08, B)], 1)
    th = rng.uniform(0.1, 0.25, B)
    batch.set_targets(tv, th)
    Wt = _weights(batch.I)
    batch.set_weights(*[w.cuda() for w in Wt])
    loss = batch.forward()
    grads = batch.backward()
    torch.cuda.synchronize()
    batch.check()
    Wr = [w.clone().requires_grad_() for w in Wt]
    x0 = pos[None].expand(B, -1, -1).clone()
    lo = omp.rollout_task(x0, ids[None].expand(B, -1), 4, op, Wr, torch.tensor(tv),
                          torch.tensor(th), T, W)
    lo.backward()
    return batch, loss.item(), grads, lo.item(), [w.grad for w in Wr]


@pytest.mark.parametrize("prec,K,scatter,tile,save,unroll", [
    ("f64", 7, dm.SCATTER_ATOMIC, None, True, 3), ("f64", 24, dm.SCATTER_STORE, None, True, 3),
    ("f32", 7, dm.SCATTER_ATOMIC, None, True, 3), ("f32", 7, dm.SCATTER_STORE, None, True, 3),
    ("f64", 7, dm.SCATTER_STORE, 64, True, 3), ("f64", 7, dm.SCATTER_ATOMIC, None, False, 3),
    ("f64", 7, dm.SCATTER_ATOMIC, None, True, 1), ("f32", 7, dm.SCATTER_ATOMIC, None, True, 27)])
def test_mpm_batched_rollout_matches_oracle(prec, K, scatter, tile, save, unroll):
    """Checkpointed MPM training rollout (sensors, MLP, P2G/grid/G2P, windowed
    COM loss) and its adjoint through every segment, against autograd of the
    FP64 oracle.  Particles are block-sorted internally (order-independent
    results); K = 7 exercises partial segments and odd segment lengths; both
    tile commit modes (global atomics, store + gather); a 64-node tile forces
    the slab path; with and without the saved per-step grids."""
    _say(f"mpm batched rollout {prec} K={K} scatter={scatter} tile={tile} save={save} "
         f"unroll={unroll}")
    params, op, pos, ids = _batched_case()
    B, T, W = 2, 24, 12
    batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, B, T, K, W, DT[prec],
                                                scatter=scatter, tile_nodes=tile, save_grid=save,
                                                unroll=unroll)
    tl, tg = (1e-10, 1e-7) if prec == "f64" else (1e-4, 1e-3)
    assert abs(loss - lo) / abs(lo) < tl
    for a, b in zip(grads, gref):
        assert rel_err(a.double().cpu(), b) < tg


@pytest.mark.parametrize("hblocks", [None, 1])
def test_mpm_batched_replay_matches_full_recompute(hblocks):
    """Segments recomputed from the horizon grid save (saved grid velocities
    restored, F updated, G2P) give the loss and gradients of the full
    recompute (P2G, grid update, G2P) and of the oracle.  With a one-block
    capacity no step fits the save and every segment falls back to the full
    recompute."""
    _say(f"mpm batched replay hblocks={hblocks}")
    params, op, pos, ids = _batched_case()
    runs = []
    for save in (False, True):
        batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, 2, 24, 7, 12,
                                                    torch.float64, save_horizon=save,
                                                    horizon_blocks=hblocks if save else None)
        assert abs(loss - lo) / abs(lo) < 1e-10
        for a, b in zip(grads, gref):
            assert rel_err(a.double().cpu(), b) < 1e-7
        if save:
            assert batch.replay == [hblocks is None] * batch.nck
        runs.append((loss, grads))
    assert abs(runs[0][0] - runs[1][0]) / abs(runs[0][0]) < 1e-12
    for a, b in zip(runs[0][1], runs[1][1]):
        assert rel_err(a.double().cpu(), b.double().cpu()) < 1e-10


@pytest.mark.parametrize("shuffled", [False, True])
def test_mpm_batched_overlapped_backward_matches_sequential(shuffled):
    """The backward that replays segment s - 1 on a side stream (its own
    segment buffer and velocity grid) while the adjoint of segment s runs
    gives the gradients of the one-stream segment loop and of the oracle,
    also with a segment forced to the full recompute between replayed ones
    (the full recompute runs on the main stream and main grid)."""
    _say(f"mpm batched overlapped backward shuffled={shuffled}")
    if shuffled:
        params, op, pos, ids = _batched_case(offsets=((5.0, 3.2, 5.0), (8.0, 3.2, 5.0)),
                                             perm_seed=5)
        T, K, W, sort = 13, 4, 6, False
    else:
        params, op, pos, ids = _batched_case()
        T, K, W, sort = 24, 7, 12, True
    batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, 2, T, K, W, torch.float64,
                                                sort=sort)
    assert batch.overlap_replay and batch.nck >= 3
    assert batch.replay == [True] * batch.nck
    for a, b in zip(grads, gref):
        assert rel_err(a.double().cpu(), b) < 1e-7
    runs = []
    for overlapped in (True, False):
        batch.forward()
        batch.start_backward()
        batch.replay[1] = False
        if overlapped:
            batch._backward_overlapped()
        else:
            for s in reversed(range(batch.nck)):
                batch.backward_segment(s)
        g = [x.clone() for x in batch.finish_backward()]
        torch.cuda.synchronize()
        batch.check()
        runs.append(g)
    for a, b, c in zip(runs[0], runs[1], gref):
        assert rel_err(a.cpu(), b.cpu()) < 1e-12
        assert rel_err(a.cpu(), c) < 1e-7


def _blob(side, origin, n=16):
    ax = torch.arange(side, dtype=torch.float64)
    g = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    pos = (g * 0.5 + torch.tensor(origin)) / n
    ids = torch.tensor([(i % 5) - 1 for i in range(len(pos))], dtype=torch.int32)
    return pos.numpy(), ids.numpy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mpm_uneven_batch_matches_robots_alone(prec):
    """One batch of three different bodies (125, 343 and 64 particles:
    padded to 343), each with its own controller and horizon (24, 13 and 20
    steps of a 24-step batch; checkpoints every 7 steps, so two robots stop
    inside a segment), against each robot run alone in an even one-robot
    batch: per-robot loss and every weight gradient."""
    _say(f"mpm uneven batch {prec}")
    dt = DT[prec]
    params = dm.quadruped3d_params(n_grid=16)
    params.dt = 2e-4
    bodies = [_blob(5, (5.0, 3.2, 5.0)), _blob(7, (4.0, 3.2, 4.0)), _blob(4, (6.0, 3.2, 6.0))]
    hz, T, K, W = [24, 13, 20], 24, 7, 12
    rng = np.random.default_rng(11)
    tv = np.stack([rng.uniform(-0.08, 0.08, 3), np.zeros(3), rng.uniform(-0.08, 0.08, 3)], 1)
    th = rng.uniform(0.1, 0.25, 3)
    ub = dm.MPMBatch(params, [b[0] for b in bodies], [b[1] for b in bodies], 4, 3, T, K=K,
                     window=W, dtype=dt, save_horizon=False, horizons=hz,
                     per_robot_weights=True)
    Ws = [_weights(ub.I, seed=20 + k) for k in range(3)]
    ub.set_targets(tv, th)
    for k in range(3):
        ub.set_robot_weights(k, *[w.cuda() for w in Ws[k]])
    ub.forward()
    grads = ub.backward(loss_grad=torch.ones(3, dtype=torch.float64))
    torch.cuda.synchronize()
    ub.check()
    # FP32 gradients: two identical runs of the 343-particle robot differ by
    # up to 1.5e-4 relative (P2G atomic order over 13 steps), so the bound is
    # the 1e-3 gradient tolerance; FP64 agrees to rounding.
    tl, tg = (1e-11, 1e-9) if prec == "f64" else (1e-5, 1e-3)
    for k in range(3):
        one = dm.MPMBatch(params, bodies[k][0], bodies[k][1], 4, 1, hz[k], K=K,
                          window=min(W, hz[k]), dtype=dt, save_horizon=False)
        one.set_targets(tv[k:k + 1], th[k:k + 1])
        one.set_weights(*[w.cuda() for w in Ws[k]])
        l1 = one.forward().item()
        g1 = one.backward()
        torch.cuda.synchronize()
        one.check()
        lk = ub.loss_b[k].item()
        assert abs(lk - l1) / abs(l1) < tl, (k, lk, l1)
        for a, b in zip(grads, g1):
            assert b.abs().max() > 0
            assert rel_err(a[k].double().cpu(), b.double().cpu()) < tg, k


def test_mpm_scatter_tuner_picks_and_caches(tmp_path):
    """The scatter autotuner times every candidate, returns the fastest, and
    answers the same key from its cache (after one verifying run)."""
    from diffloco.autotune import Autotuner
    _say("mpm scatter tuner")
    params, _, pos, ids = _batched_case()
    cands = dm.scatter_candidates(torch.float32)[:4]
    tuner = Autotuner(path=tmp_path / "tune.json")
    pick = dm.tune_scatter(params, pos.numpy(), ids.numpy(), 4, 2, dtype=torch.float32, K=5,
                           tuner=tuner, candidates=cands)
    assert pick in cands
    (entry,) = tuner.cache.values()
    ms = {json.dumps(t["candidate"], sort_keys=True): t["ms"] for t in entry["timings"]}
    assert len(ms) == len(cands) and ms[json.dumps(pick, sort_keys=True)] == min(ms.values())
    again = dm.tune_scatter(params, pos.numpy(), ids.numpy(), 4, 2, dtype=torch.float32, K=5,
                            tuner=Autotuner(path=tmp_path / "tune.json"), candidates=cands)
    assert again == pick


@pytest.mark.parametrize("hblocks,shuffled", [(None, False), (1, False), (None, True)])
def test_mpm_batched_stored_states_match_replay(hblocks, shuffled):
    """Every step's particle state kept by the forward, the adjoint run on
    those states with no recompute, gives the loss and gradients of the
    replayed segments and of the oracle.  With a one-block horizon save every
    segment still falls back to the full recompute from its first stored
    state; the shuffled case re-sorts particles at every checkpoint."""
    _say(f"mpm batched stored states hblocks={hblocks} shuffled={shuffled}")
    if shuffled:
        params, op, pos, ids = _batched_case(offsets=((5.0, 3.2, 5.0), (8.0, 3.2, 5.0)),
                                             perm_seed=5)
        T, K, W, sort = 13, 4, 6, False
    else:
        params, op, pos, ids = _batched_case()
        T, K, W, sort = 24, 7, 12, True
    runs = []
    for stored in (False, True):
        batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, 2, T, K, W,
                                                    torch.float64, sort=sort,
                                                    horizon_blocks=hblocks, store_states=stored)
        assert abs(loss - lo) / abs(lo) < 1e-10
        for a, b in zip(grads, gref):
            assert rel_err(a.double().cpu(), b) < 1e-7
        assert batch.replay == [hblocks is None] * batch.nck
        if stored:
            # the last slot of each segment is the next segment's start state
            # (before that segment's re-sort)
            for s in range(batch.nck - 1):
                nxt = batch.views[s + 1].x[0]
                if batch.resort:
                    pi = batch.perms[s + 1]
                    cur = torch.gather(batch.views[s].x[-1], 2, pi[:, None, :].expand_as(nxt))
                else:
                    cur = batch.views[s].x[-1]
                assert torch.equal(cur, nxt)
        runs.append((loss, grads))
    assert abs(runs[0][0] - runs[1][0]) / abs(runs[0][0]) < 1e-12
    for a, b in zip(runs[0][1], runs[1][1]):
        assert rel_err(a.double().cpu(), b.double().cpu()) < 1e-10


def test_mpm_batched_resort_matches_fixed_order():
    """Shuffled particles re-sorted by block at every checkpoint (state,
    actuator ids and groups permuted; the adjoint permuted back) give the
    loss and gradients of the fixed order and of the oracle."""
    _say("mpm batched resort vs fixed order")
    params, op, pos, ids = _batched_case(offsets=((5.0, 3.2, 5.0), (8.0, 3.2, 5.0)), perm_seed=5)
    out = []
    for resort in (False, True):
        batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, 2, 13, 4, 6,
                                                    torch.float64, sort=False, resort=resort)
        assert abs(loss - lo) / abs(lo) < 1e-10
        for a, b in zip(grads, gref):
            assert rel_err(a.double().cpu(), b) < 1e-7
        out.append((loss, grads))
        if resort:
            # the orders are permutations, and the re-sort really moved particles
            o = batch.orders.cpu()
            assert torch.equal(o.sort(-1).values, torch.arange(o.shape[-1]).expand_as(o))
            assert not torch.equal(o[1], o[0])
    assert abs(out[0][0] - out[1][0]) / abs(out[0][0]) < 1e-12
    for a, b in zip(out[0][1], out[1][1]):
        assert rel_err(a.double().cpu(), b.double().cpu()) < 1e-10


def test_mpm_batched_plain_chunks_match_windows():
    """Several workgroups per robot (3 blobs, 375 particles): without a
    window table (plain 256-particle chunks) and with box-limited windows
    the rollout gives the same loss and gradients, and no error flag."""
    _say("mpm batched plain chunks vs windows")
    params, op, pos, ids = _batched_case(offsets=((5.0, 3.2, 5.0), (8.0, 3.2, 5.0), (5.0, 3.2, 8.0)))
    out = []
    for limit in (None, 200):
        batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, 2, 10, 5, 5,
                                                    torch.float64, box_limit=limit)
        assert (batch.windows is None) == (limit is None)
        assert abs(loss - lo) / abs(lo) < 1e-10
        out.append((loss, grads))
    assert abs(out[0][0] - out[1][0]) / abs(out[0][0]) < 1e-12
    for a, b in zip(out[0][1], out[1][1]):
        assert rel_err(a.double().cpu(), b.double().cpu()) < 1e-10


@pytest.mark.parametrize("case,scatter", [("slabs", dm.SCATTER_ATOMIC), ("direct", dm.SCATTER_ATOMIC),
                                          ("slabs", dm.SCATTER_STORE), ("staged", dm.SCATTER_ATOMIC),
                                          ("windows", dm.SCATTER_ATOMIC),
                                          ("windows", dm.SCATTER_STORE)])
def test_mpm_batched_scatter_paths_and_invariants(case, scatter):
    """Two blobs far apart in one workgroup (shuffled, unsorted), so the
    workgroup's node box exceeds the LDS tile: "slabs" (box 21 x 5 x 21)
    scatters in several x-slab passes through the tile, "direct" (a y-z plane
    of 45 x 45 nodes) with direct global atomics.  "staged": the slabs case in
    FP32 with the largest tile, where the per-run stage fits: nearly every
    lane is its own run and each cell's runs (about 8) are chained in the
    per-cell lists.  "windows": the slabs layout cut into box-limited scatter
    windows (most of them a few particles wide, the shuffle alternating
    between the blobs).  Results match the oracle, and the grid,
    active-block and actuation-adjoint buffers are zero again afterwards."""
    _say(f"mpm batched scatter path {case} scatter={scatter}")
    dtype, tile, (tl, tg), limit, resort = torch.float64, 1024, (1e-10, 1e-7), None, False
    if case == "staged":
        dtype, tile, (tl, tg) = torch.float32, dm.MPM_TILE_MAX, (1e-4, 1e-3)
    if case == "windows":
        limit = 512
    if case in ("slabs", "staged", "windows"):
        params, op, pos, ids = _batched_case(offsets=((5.0, 3.2, 5.0), (20.0, 3.2, 20.0)), n=32,
                                             perm_seed=3)
    else:
        params, op, pos, ids = _batched_case(offsets=((5.0, 4.0, 5.0), (5.0, 44.0, 44.0)), n=64,
                                             perm_seed=3)
    B, T, K, W = 2, 12, 5, 6
    batch, loss, grads, lo, gref = _run_batched(params, op, pos, ids, B, T, K, W, dtype,
                                                sort=False, scatter=scatter, tile_nodes=tile,
                                                box_limit=limit, resort=resort)
    if case == "windows":
        assert len(batch.windows) > 20
    else:
        assert batch.windows is None
    assert abs(loss - lo) / abs(lo) < tl
    for a, b in zip(grads, gref):
        assert rel_err(a.double().cpu(), b) < tg
    for t in (batch.grid, batch.grid_vo_grad, batch.act_grad):
        assert t.abs().max().item() == 0.0
    assert batch.blk_flag.abs().max().item() == 0 and batch.blk_count.abs().max().item() == 0
    assert batch.blk_tile_count.abs().max().item() == 0


def _mpm_config(B, T):
    import json
    import diffloco as dl
    with open("tests/golden/sim_quad.json") as fh:
        raw = json.load(fh)
    raw["robot"].update(simulator="mpm", dim=2, design_file="robot_mpm_1.json")
    raw["process"]["max_steps"] = T
    raw["nn"]["batch_size"] = B
    raw["precision"] = "f64"
    return dl.ConfigSim(raw, base_dir="tests/golden")


def test_solver_mpm_config_dropin_under_tape():
    """SolverMPM(config) driven like the reference (pre_advance / advance per
    step inside a Tape; advance_core records advance_core_grad): loss on the
    final state and on every step's centre, gradients with respect to the
    actuation of every step and the initial state against the oracle's
    autograd (2D, robot_mpm_1, FP64)."""
    _say("SolverMPM(config) drop-in")
    from diffloco.fields import Field, Tape, record
    B, T = 2, 6
    config = _mpm_config(B, T)
    sv = dm.SolverMPM(config)
    assert (sv.n_particles, sv.n_springs, sv.n_grid) == (320, 320, 32)
    sv.initialize_robot()
    sv.clear_states(T + 1)
    g = torch.Generator().manual_seed(61)
    pos = torch.tensor(np.asarray(config.get_config()["robot"]["objects"])) + torch.tensor([0.2, 0.15])
    x0 = pos[None].expand(B, -1, -1) + 0.002 * torch.randn(B, 320, 2, generator=g, dtype=torch.float64)
    v0 = 0.2 * torch.randn(B, 320, 2, generator=g, dtype=torch.float64)
    act = torch.rand(T, B, 4, generator=g, dtype=torch.float64) * 2 - 1
    sv.x.view[0] = x0
    sv.v.view[0] = v0
    sv.actuation.data[:T, :, :4] = act.cuda()
    Wx = torch.randn(B, 320, 2, generator=g, dtype=torch.float64)
    Wc = torch.randn(T, B, 2, generator=g, dtype=torch.float64)
    loss = Field((), torch.float64, "cuda", name="loss")
    Wxd, Wcd = Wx.transpose(1, 2).contiguous().cuda(), Wc.cuda()
    with Tape(loss):
        for t in range(T):
            sv.pre_advance(t)
            sv.advance(t)
        loss.data.copy_((sv.x.at(T) * Wxd).sum() + (sv.center.data[:T] * Wcd).sum())

        def adj():
            sv.x.grad.at(T).add_(Wxd)
            sv.center.grad.data[:T].add_(Wcd)
        record(adj)
    torch.cuda.synchronize()
    sv.check()
    op = omp.MPMParams()
    act_id = torch.tensor(config.get_config()["robot"]["springs"])[None].expand(B, -1)
    xr, vr, ar = x0.clone().requires_grad_(), v0.clone().requires_grad_(), act.clone().requires_grad_()
    x, v = xr, vr
    C = torch.zeros(B, 320, 2, 2, dtype=torch.float64)
    F = torch.eye(2, dtype=torch.float64).expand(B, 320, 2, 2).clone()
    lo = torch.zeros((), dtype=torch.float64)
    for t in range(T):
        lo = lo + (x.mean(1) * Wc[t]).sum()
        x, v, C, F = omp.step(x, v, C, F, ar[t], act_id, op)
    lo = lo + (x * Wx).sum()
    lo.backward()
    assert abs(loss[None] - lo.item()) < 1e-10 * abs(lo.item())
    assert rel_err(sv.actuation.grad.data[:T, :, :4].cpu(), ar.grad) < 1e-9
    assert rel_err(sv.x.grad.view[0].cpu(), xr.grad) < 1e-9
    assert rel_err(sv.v.grad.view[0].cpu(), vr.grad) < 1e-9
    assert sv.actuation.grad.data[:, :, 4:].abs().max() == 0


def test_solver_mpm_kernel_calls_record_their_adjoints():
    """Direct clear_grid / p2g / grid_op / g2p calls inside a Tape record
    each kernel's adjoint (Taichi's kernel autodiff): one step's gradients
    with respect to x, v, C, F and the actuation against the oracle."""
    _say("SolverMPM kernel-level adjoints")
    from diffloco.fields import Field, Tape, record
    B = 2
    config = _mpm_config(B, 2)
    sv = dm.SolverMPM(config)
    sv.initialize_robot()
    sv.clear_states(3)
    params, op, x, v, C, F, act, act_id, n_act = _case(2, B, seed=62)
    sv.x.view[0] = x
    sv.v.view[0] = v
    sv.C.data[0] = _msoa(C, torch.float64)
    sv.F.data[0] = _msoa(F, torch.float64)
    sv.actuation.data[0, :, :4] = act.cuda()
    ids = torch.as_tensor(act_id.numpy(), dtype=torch.int32).cuda()
    sv.actuator_id.copy_(ids)
    g = torch.Generator().manual_seed(63)
    ws = [torch.randn(B, 320, 2, generator=g, dtype=torch.float64),
          torch.randn(B, 320, 2, generator=g, dtype=torch.float64),
          torch.randn(B, 320, 2, 2, generator=g, dtype=torch.float64),
          torch.randn(B, 320, 2, 2, generator=g, dtype=torch.float64)]
    wd = [_vsoa(ws[0], torch.float64), _vsoa(ws[1], torch.float64), _msoa(ws[2], torch.float64),
          _msoa(ws[3], torch.float64)]
    loss = Field((), torch.float64, "cuda", name="loss")
    with Tape(loss):
        sv.clear_grid()
        sv.p2g(0)
        sv.grid_op()
        sv.g2p(0)
        flds = (sv.x, sv.v, sv.C, sv.F)
        loss.data.copy_(sum((f_.at(1) * w).sum() for f_, w in zip(flds, wd)))

        def adj():
            for f_, w in zip(flds, wd):
                f_.grad.at(1).add_(w)
        record(adj)
    torch.cuda.synchronize()
    xr, vr, Cr, Fr, ar = [t.clone().requires_grad_() for t in (x, v, C, F, act)]
    outs = omp.step(xr, vr, Cr, Fr, ar, act_id, op)
    sum((o * w).sum() for o, w in zip(outs, ws)).backward()
    assert rel_err(_vback(sv.x.grad.at(0)), xr.grad) < 1e-9
    assert rel_err(_vback(sv.v.grad.at(0)), vr.grad) < 1e-9
    assert rel_err(_mback(sv.C.grad.at(0), 2), Cr.grad) < 1e-9
    assert rel_err(_mback(sv.F.grad.at(0), 2), Fr.grad) < 1e-9
    assert rel_err(sv.actuation.grad.data[0, :, :4].cpu(), ar.grad) < 1e-9


def test_mpm_rollout_autograd_function():
    """mpm_rollout_loss: the batched MPM rollout behind torch.autograd
    (weights in, mean task loss out) gives MPMBatch.backward's gradients and
    matches the oracle (3D, FP64)."""
    _say("MPM autograd.Function")
    from oracle import mass_spring as om
    n, B, T, K, W = 16, 2, 12, 5, 6
    params = dm.quadruped3d_params(n_grid=n)
    params.dt = 2e-4
    ax = torch.arange(5, dtype=torch.float64)
    pos = (torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) * 0.5
           + torch.tensor([5.0, 3.2, 5.0])) / n
    ids = torch.tensor([(i % 5) - 1 for i in range(len(pos))])
    mb = dm.MPMBatch(params, pos.numpy(), ids.numpy(), 4, B, T, K=K, window=W, dtype=torch.float64)
    tv = np.array([[0.05, 0.0, -0.02], [-0.03, 0.0, 0.06]])
    th = np.array([0.12, 0.2])
    mb.set_targets(tv, th)
    w64 = om.init_weights(mb.I, 64, 4, seed=64)
    wd = [w.cuda().requires_grad_() for w in w64]
    loss = dm.mpm_rollout_loss(mb, *wd)
    loss.sum().backward()
    torch.cuda.synchronize()
    mb.check()
    op = omp.MPMParams(dim=3, n_grid=n, dt=params.dt, gravity=params.gravity, p_vol=params.p_vol,
                       p_mass=params.p_mass, mu=params.mu, la=params.la,
                       act_strength=params.act_strength, act_axis=0, bound=3)
    wr = [w.clone().requires_grad_() for w in w64]
    lo = omp.rollout_task(pos[None].expand(B, -1, -1).clone(), ids[None].expand(B, -1), 4, op, wr,
                          torch.tensor(tv), torch.tensor(th), T, W)
    lo.backward()
    assert abs(loss.item() - lo.item()) < 1e-10 * abs(lo.item())
    for a, b in zip(wd, wr):
        assert rel_err(a.grad.cpu(), b.grad) < 1e-7


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mpm_sensor_entry_points_match_oracle(prec):
    """dl_mpm_sensor / dl_mpm_sensor_adj (an MPM robot's controller input,
    COM and COM velocity, and the input adjoint) against the oracle's sensor
    vector and torch autograd of it; one actuator group is left empty (its
    offset is -COM / scale)."""
    from diffloco import _native as nat
    _say(f"mpm sensor entry points {prec}")
    dt, f64 = DT[prec], torch.float64
    g = torch.Generator().manual_seed(12)
    B, P, G, n_sin, period, t, scale, vmax = 3, 700, 5, 10, 100, 37, 0.2, 0.08
    x = (0.2 + 0.5 * torch.rand(B, P, 3, generator=g, dtype=f64)).requires_grad_()
    v = (0.1 * torch.randn(B, P, 3, generator=g, dtype=f64)).requires_grad_()
    group = torch.randint(0, G - 1, (P,), generator=g)          # group G - 1 empty
    tv = 0.08 * torch.rand(B, 3, generator=g, dtype=f64)
    th = 0.1 + 0.1 * torch.rand(B, generator=g, dtype=f64)
    inp, c, vc = omp.sensors(t, x, v, group, G, n_sin, period, scale, vmax, tv, th)
    I = inp.shape[1]
    gin = torch.randn(B, I, generator=g, dtype=f64)
    xg, vg = torch.autograd.grad((inp * gin).sum(), (x, v))
    cnt = torch.bincount(group, minlength=G).to(f64)
    inv = torch.where(cnt > 0, 1.0 / cnt.clamp_min(1), torch.zeros_like(cnt))

    def dev(a):
        return a.detach().to("cuda", dt).contiguous()

    xd, vd = dev(x.transpose(1, 2)), dev(v.transpose(1, 2))         # [B][3][P]
    gd = group.to(torch.int32)[None].expand(B, -1).contiguous().cuda()
    tgd = dev(torch.cat([tv, th[:, None], torch.zeros(B, 1, dtype=f64)], 1))
    invd = dev(inv[None].expand(B, -1))
    ind = torch.empty(B, I, device="cuda", dtype=dt)
    cd = torch.empty(B, 3, device="cuda", dtype=dt)
    vcd = torch.empty(B, 3, device="cuda", dtype=dt)
    dc = nat.dtype_code(dt)
    nat.call("dl_mpm_sensor", dc, B, P, G, n_sin, period, t, scale, vmax,
             *map(nat.ptr, (xd, vd, gd, tgd, invd, ind, cd, vcd)), nat.stream())
    xgd, vgd = torch.zeros_like(xd), torch.zeros_like(vd)
    nat.call("dl_mpm_sensor_adj", dc, B, P, G, n_sin, scale, nat.ptr(gd), nat.ptr(invd),
             nat.ptr(dev(gin)), nat.ptr(xgd), nat.ptr(vgd), nat.stream())
    torch.cuda.synchronize()
    tol = 1e-12 if prec == "f64" else 1e-5
    assert rel_err(ind, inp.detach()) < tol
    assert rel_err(cd, c.detach()) < tol and rel_err(vcd, vc.detach()) < tol
    assert rel_err(xgd.transpose(1, 2), xg) < tol
    assert rel_err(vgd.transpose(1, 2), vg) < tol
