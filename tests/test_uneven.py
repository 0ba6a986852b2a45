"""Uneven batches: size classes, autotuner choice/caching (CPU), and on the
GPU the bucketed multi-stream iteration and its HIP-graph replay against a
single-launch reference."""

import numpy as np
import pytest
import torch

import diffloco as dl
from diffloco import robot_design as rd
from diffloco.autotune import Autotuner
from diffloco.solver_mass_spring import _physics
from diffloco.taichi_env import sensor_layout

from helpers import rel_err


def _setup(n=10, seed=7, cube_range=(2, 30)):
    designs, horizons = rd.random_robot_batch(n, seed, horizon_range=(40, 90), cube_range=cube_range,
                                              group_range=(4, 12))
    qc = dl.quadruped_config(batch_size=n, steps=int(max(horizons)), window=30).get_config()
    rng = np.random.default_rng(seed + 1)
    tg = np.zeros((n, 5))
    tg[:, 0] = rng.uniform(-0.08, 0.08, n)
    tg[:, 2] = rng.uniform(-0.08, 0.08, n)
    tg[:, 3] = rng.uniform(0.1, 0.25, n)
    return designs, horizons, _physics(qc), sensor_layout(qc), tg


def test_size_classes_partition_every_robot_once():
    from diffloco.uneven import size_classes, n_inputs, WAVE_MAX_IN
    designs, horizons, phys, lay, tg = _setup(n=40, cube_range=(2, 200))
    cls = size_classes(designs, lay, torch.float32)
    ids = np.concatenate([np.asarray(i) for _, i in cls])
    assert sorted(ids.tolist()) == list(range(40))
    for name, rs in cls:
        for r in rs:
            d = designs[r]
            if name == "wave":
                assert n_inputs(lay, d.n_vert) <= WAVE_MAX_IN and d.n_act <= 8 and d.n_vert <= 64
            elif name.startswith("n<="):
                assert d.n_vert <= int(name[3:])
    assert size_classes(designs, lay, torch.float64)[0][0] != "wave"


def test_autotuner_picks_fastest_and_caches(tmp_path):
    t = Autotuner(tmp_path / "tune.json")
    cands = [{"block": 64}, {"block": 128}, {"block": 256}]
    cost = {64: 3.0, 128: 1.0, 256: 2.0}
    calls = []

    def measure(c):
        calls.append(c["block"])
        if c["block"] == 256:
            raise RuntimeError("shape not supported")
        return cost[c["block"]]

    assert t.pick("k", cands, measure) == {"block": 128}
    n = len(calls)
    t2 = Autotuner(tmp_path / "tune.json")
    # cached: one verification run of the cached choice, no re-tune
    assert t2.pick("k", cands, measure) == {"block": 128} and calls[n:] == [128]
    assert "error" in str(t2.cache["k"]["timings"])
    assert t2.pick("k", cands, measure, verify=False) == {"block": 128} and len(calls) == n + 1


def test_autotuner_drops_a_stale_entry(tmp_path):
    """A cached choice that no longer runs (the workload or library changed
    under the same key) is dropped and the key re-tuned."""
    t = Autotuner(tmp_path / "tune.json")
    cands = [{"block": 64}, {"block": 128}]
    t.pick("k", cands, lambda c: {64: 2.0, 128: 1.0}[c["block"]])
    assert t.cache["k"]["choice"] == {"block": 128}

    def measure(c):
        if c["block"] == 128:
            raise RuntimeError("LDS limit exceeded")
        return 5.0

    assert Autotuner(tmp_path / "tune.json").pick("k", cands, measure) == {"block": 64}
    assert Autotuner(tmp_path / "tune.json").cache["k"]["choice"] == {"block": 64}


def test_tuning_key_names_sizes_and_library():
    from diffloco.autotune import library_tag
    tag = library_tag()
    assert len(tag) == 12 and tag != "none"
    k1 = Autotuner.key(cls="n<=400", max_n=400, lib=tag)
    assert k1 != Autotuner.key(cls="n<=400", max_n=399, lib=tag)


@pytest.mark.gpu
@pytest.mark.parametrize("stored", [True, False])
def test_uneven_iteration_and_graph_replay_match_single_launch(tmp_path, stored):
    """Size classes on streams (rows kept by the forwards, or checkpoint
    segments recomputed), tuned, then captured and replayed, against one
    block-kernel launch over all robots."""
    from diffloco.rollout import MassSpringBatch, init_weights
    from diffloco.uneven import UnevenBatch, n_inputs
    print(f"uneven batch: tune, streams, graph (stored={stored})", flush=True)
    designs, horizons, phys, lay, tg = _setup()
    R = len(designs)
    ub = UnevenBatch(designs, horizons, phys, lay, tg, ckpt_every=16, loss_window=30, stored=stored)
    assert len(ub.buckets) >= 2 and ub.buckets[0]["name"] == "wave"
    W = [init_weights(n_inputs(lay, d.n_vert), 64, d.n_act, seed=100 + r) for r, d in enumerate(designs)]
    for r in range(R):
        ub.set_weights(r, *W[r])
    ub.tune(Autotuner(tmp_path / "tune.json"))
    ub.prepare()
    ub.run_iteration()
    torch.cuda.synchronize()
    loss_e = ub.losses().clone()
    grads_e = [[g.clone() for g in ub.weight_grads(r)] for r in range(R)]

    ref = MassSpringBatch(designs, list(range(R)), horizons, phys, lay, tg,
                          [d.objects + np.array([0.4, 0.0, 0.0]) for d in designs], ckpt_every=16,
                          dtype=torch.float32, loss_window=30, variant=1)
    for r in range(R):
        ref.set_weights(r, *W[r])
    ref.prepare_iteration(np.ones(R))
    ref.run_iteration()
    torch.cuda.synchronize()
    assert rel_err(loss_e.double().cpu(), ref.loss.double().cpu()) < 1e-4
    for r in range(R):
        for a, b in zip(grads_e[r], ref.weight_grads(r)):
            assert rel_err(a.double().cpu(), b.double().cpu()) < 1e-3, r

    ub.capture()
    for rep in range(3):
        ref.run_iteration()                 # unrelated eager work between replays
        torch.cuda.synchronize()
        ub.replay()
        torch.cuda.synchronize()
        assert torch.equal(ub.losses(), loss_e), rep
        for r in range(R):
            for a, b in zip(ub.weight_grads(r), grads_e[r]):
                assert torch.equal(a, b), (rep, r)


@pytest.mark.gpu
def test_schedule_search_picks_bounds_and_results_do_not_depend_on_them(tmp_path):
    """The robot-to-launch schedule search times every class-boundary
    candidate and returns the fastest; the losses and gradients do not
    depend on which boundaries (so which kernels and launch shapes) run."""
    from diffloco.rollout import init_weights
    from diffloco.uneven import UnevenBatch, n_inputs, tune_schedule
    print("uneven schedule search", flush=True)
    designs, horizons, phys, lay, tg = _setup(n=12, cube_range=(2, 60))
    W = [init_weights(n_inputs(lay, d.n_vert), 64, d.n_act, seed=200 + r)
         for r, d in enumerate(designs)]

    def load(ub):
        for r in range(len(designs)):
            ub.set_weights(r, *W[r])

    cands = ((128, 400), (60,))
    kw = dict(ckpt_every=16, loss_window=30, stored=False)
    bounds, timings = tune_schedule(designs, horizons, phys, lay, tg, set_weights=load,
                                    tuner=Autotuner(tmp_path / "t.json"), schedules=cands, **kw)
    assert bounds in cands and len([t for t in timings if "ms" in t]) == 2
    runs = []
    for b in cands:
        ub = UnevenBatch(designs, horizons, phys, lay, tg, bounds=b, **kw)
        load(ub)
        ub.prepare()
        ub.run_iteration()
        torch.cuda.synchronize()
        runs.append((ub.losses().double().cpu(),
                     [[g.double().cpu() for g in ub.weight_grads(r)] for r in range(len(designs))]))
    assert rel_err(runs[0][0], runs[1][0]) < 1e-4
    for r in range(len(designs)):
        for a, b in zip(runs[0][1][r], runs[1][1][r]):
            assert rel_err(a, b) < 1e-3, r
