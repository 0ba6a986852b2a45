"""Request loop: admission order and HBM budget with a fake executor and
clock (CPU); on the GPU, each request's result is the same alone and in a
mix, and the slot pool agrees with an independent batched rollout."""

import math

import numpy as np
import pytest
import torch

from diffloco import serving as sv


class FakeExec:
    """Takes requests of one kind; a request needs ceil(T/chunk) ticks
    (twice that with gradients); bytes = horizon."""

    def __init__(self, kind, slots, log):
        self.kind, self.slots, self.log = kind, slots, log
        self.active = {}

    def fits(self, r):
        return r.kind == self.kind

    def bytes_for(self, r):
        return r.horizon

    def has_slot(self):
        return len(self.active) < self.slots

    def admit(self, r):
        self.active[r.rid] = [r, 0]
        self.log.append(("admit", r.rid))

    def busy(self):
        return len(self.active) > 0

    def tick(self, chunk):
        fin = []
        for k, (r, n) in list(self.active.items()):
            n += 1
            self.active[k][1] = n
            if n == math.ceil(r.horizon / chunk) * (2 if r.grad else 1):
                fin.append(k)
        return fin

    def collect(self, k):
        r, _ = self.active.pop(k)
        r.loss = float(r.rid)
        return r


class Clock:
    """Fake time: a tick of device work takes 1 s; waiting advances time."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-3)

    # events: a tick launched now completes one second later
    def record(self, ex):
        return self.t + 1.0

    def done(self, ev):
        return self.t >= ev - 1e-9


def _reqs(spec):
    return [sv.Request(i, kind, T, grad, np.zeros(5), None, arrival=a)
            for i, (kind, T, grad, a) in enumerate(spec)]


def test_fifo_admission_budget_and_latency():
    log = []
    ck = Clock()
    ex = [FakeExec("ms", 2, log), FakeExec("mpm", 1, log)]
    reqs = _reqs([("ms", 100, True, 0.0), ("ms", 50, False, 0.1), ("mpm", 200, False, 0.2),
                  ("ms", 300, True, 0.3), ("ms", 40, False, 5.0), ("mpm", 50, True, 5.5)])
    srv = sv.Server(ex, budget_bytes=400, chunk=50, clock=ck, sleep=ck.sleep, events=ck,
                    poll_s=0.05)
    done = srv.run(reqs)
    assert sorted(r.rid for r in done) == list(range(6))
    # admission strictly in arrival order
    order = [rid for what, rid in log if what == "admit"]
    assert order == sorted(order)
    for r in done:
        need = math.ceil(r.horizon / 50) * (2 if r.grad else 1)
        assert r.admitted >= r.arrival - 1e-12 and r.latency >= need - 1e-12
    # the 300-step request (needs 300 B) waits until the budget frees up
    r3 = next(r for r in done if r.rid == 3)
    assert r3.admitted > r3.arrival
    st = sv.latency_stats(done, ck())
    assert st["requests"] == 6 and st["p99_s"] >= st["p50_s"] > 0


def test_request_over_the_whole_budget_is_rejected():
    ck = Clock()
    srv = sv.Server([FakeExec("ms", 4, [])], budget_bytes=100, chunk=50, clock=ck, sleep=ck.sleep,
                    events=ck)
    with pytest.raises(ValueError):
        srv.run(_reqs([("ms", 500, False, 0.0)]))


def test_unfit_request_is_rejected_and_the_rest_served():
    """A request no executor takes is turned away with a reason; the loop
    goes on serving the others."""
    ck = Clock()
    log = []
    srv = sv.Server([FakeExec("ms", 4, log)], budget_bytes=10000, chunk=50, clock=ck,
                    sleep=ck.sleep, events=ck)
    done = srv.run(_reqs([("ms", 100, False, 0.0), ("mpm", 100, False, 0.1),
                          ("ms", 100, True, 0.2)]))
    assert sorted(r.rid for r in done) == [0, 2]
    assert [r.rid for r in srv.rejected] == [1] and "no executor" in srv.rejected[0].rejected


def test_every_generated_robot_fits_a_pool():
    """The mass-spring robots of the request generator (8-400 cubes
    log-uniform, 4-64 groups, horizons 100-1000) all fit one of the two
    pools' caps: 2000 draws from the generator's distribution."""
    import diffloco as dl
    from diffloco.robot_design import build_voxels, random_voxel_cubes
    from diffloco.taichi_env import sensor_layout
    lay = sensor_layout(dl.quadruped_config(batch_size=1).get_config())
    caps = sv.pool_caps(lay)
    rng = np.random.default_rng(4)
    biggest = 0
    for _ in range(2000):
        nc = int(round(math.exp(rng.uniform(math.log(8), math.log(400)))))
        cubes, g = random_voxel_cubes(rng, nc, int(rng.integers(4, 65)))
        d = build_voxels(cubes, n_act=g)
        T = int(rng.integers(100, 1001))
        assert sv.fits_caps(d, T, lay, caps["wave"]) or sv.fits_caps(d, T, lay, caps["block"]), \
            (nc, d.n_vert, d.n_spring, d.n_act)
        biggest = max(biggest, d.n_vert)
    assert biggest > 800          # the old 800-vertex cap would have refused some


def test_every_generated_mpm_body_fits_a_pool():
    """The request generator's MPM bodies (8k-60k particles, 30k at most on
    32^3 grids) never exceed their grid size's pool capacity, and stay off
    the side walls while resting on the floor."""
    import diffloco as dl
    from diffloco.taichi_env import sensor_layout
    lay = sensor_layout(dl.quadruped_config(batch_size=1).get_config())
    reqs = sv.make_requests(50, None, seed=4, layout=lay, count=300, device="cpu")
    mpm = [r for r in reqs if r.kind == "mpm"]
    assert len(mpm) > 50
    for r in mpm:
        n = r.mpm["params"].n_grid
        pos = r.mpm["positions"]
        assert len(pos) <= sv.MPM_POOL_PARTICLES[n]
        cell = pos * n
        assert cell[:, 1].min() < 3.5 and cell.min() >= 3.0 and cell.max() <= n - 3.0


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix():
    import diffloco as dl
    from diffloco.mpm import quadruped3d_params
    from diffloco.robot_design import random_voxel_cubes, build_voxels
    from diffloco.rollout import MassSpringBatch, init_weights
    from diffloco.solver_mass_spring import _physics
    from diffloco.taichi_env import sensor_layout
    from diffloco.uneven import n_inputs
    print("serving: alone vs mix", flush=True)
    qc = dl.quadruped_config(batch_size=1, steps=200, window=100).get_config()
    lay, phys = sensor_layout(qc), _physics(qc)
    phys.dashpot = 0.03
    rng = np.random.default_rng(11)

    def ms_req(rid, n_cubes, n_groups, T, grad):
        cubes, g = random_voxel_cubes(rng, n_cubes, n_groups)
        d = build_voxels(cubes, n_act=g)
        tgt = np.array([0.05, 0.0, -0.03, 0.2, 0.0])
        return sv.Request(rid, "ms", T, grad, tgt,
                          init_weights(n_inputs(lay, d.n_vert), 64, d.n_act, seed=rid), design=d)

    def mpm_req(rid, T):
        pos, ids, na = sv.mpm_particle_robot(rng, 32, 8000)
        return sv.Request(rid, "mpm", T, True, np.array([0.03, 0.0, 0.0, 0.15, 0.0]),
                          init_weights(10 + 3 + 3 * (na + 1) + 4, 64, na, seed=rid),
                          mpm={"params": quadruped3d_params(32), "positions": pos, "act_id": ids,
                               "n_act": na})

    A = ms_req(0, 6, 4, 130, True)          # small robot (wave pool)
    B = ms_req(1, 60, 10, 170, True)        # large robot (block pool)
    C = mpm_req(2, 150)
    D = ms_req(3, 5, 4, 120, False)
    E = mpm_req(4, 100)                     # shares C's MPM batch in the mix

    def serve(reqs):
        ex = sv.default_executors(phys, lay, slots_small=8, slots_large=4, mpm_active=2)
        for r in reqs:
            r.done = r.loss = r.grads = None
        out = sv.Server(ex, budget_bytes=8 << 30, chunk=50).run(reqs)
        return {r.rid: (r.loss, [g.clone() for g in r.grads] if r.grads else None) for r in out}

    alone = {**serve([A]), **serve([B]), **serve([C])}
    mix = serve([A, B, C, D, E])
    for rid in (0, 1):
        la, ga = alone[rid]
        lm, gm = mix[rid]
        assert la == lm, rid
        for a, b in zip(ga, gm):
            assert torch.equal(a, b), rid
    la, ga = alone[2]
    lm, gm = mix[2]
    # MPM: alone, C is a one-robot batch; in the mix it shares a batch with E
    # (its body padded to E's particle count).  FP32 atomic summation order in
    # the scatter kernels differs run to run; the body touches the sticky
    # floor from step 0, so its gradients are far from vanishing
    print("mpm dW1 norm alone", ga[0].norm().item(), flush=True)
    assert ga[0].norm() > 1e-6
    assert abs(la - lm) <= 1e-4 * abs(la)
    for a, b in zip(ga, gm):
        assert (a - b).norm() <= 1e-3 * a.norm()
    assert mix[3][0] is not None and mix[3][1] is None

    # the slot pool against an independent single-robot batched rollout
    for req in (A, B):
        d = req.design
        ref = MassSpringBatch([d], [0], req.horizon, phys, lay, req.target[None],
                              [d.objects + np.array([0.4, 0.0, 0.0])], ckpt_every=50,
                              dtype=torch.float32, loss_window=100, variant=1)
        ref.set_weights(0, *req.weights)
        ref.prepare_iteration(np.ones(1))
        ref.run_iteration()
        torch.cuda.synchronize()
        lr = float(ref.loss[0].item())
        assert abs(mix[req.rid][0] - lr) <= 1e-4 * abs(lr), req.rid
        for a, b in zip(mix[req.rid][1], ref.weight_grads(0)):
            assert (a - b).norm() <= 1e-3 * b.norm(), req.rid
