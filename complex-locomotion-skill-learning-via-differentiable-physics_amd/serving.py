"""Request loop: rollout and gradient requests for different robot designs,
served with continuous batching on one GPU.

A request is one robot (a mass-spring design or an MPM particle body), its
controller weights, its target, a horizon, and whether it wants weight
gradients.  Requests arrive over time; the loop admits them in arrival order
while the HBM budget and a slot allow, keeps their state and weights
resident on the device, and advances every admitted request by one chunk
per tick:

* mass-spring requests live in slots of a ``MSPool`` -- preallocated
  per-slot regions of the shared topology / weight / state / checkpoint /
  scratch arrays addressed by the rollout kernels' offset table.  One tick
  is one persistent forward launch (every slot in its forward phase advances
  ``chunk`` steps) and one adjoint launch (every slot in its adjoint phase
  goes back ``chunk`` steps).  Requests join and leave between ticks; a
  request's slot is fixed for its lifetime, so its result does not depend
  on what else is in the batch.  Robots that fit the wave-per-robot kernel
  and the rest use two pools.
* MPM requests are staged per grid size and start together as one uneven
  ``MPMBatch`` (every robot its own grid, body, controller and horizon),
  which advances one checkpoint segment per tick on its own stream,
  overlapping the mass-spring launches.

The reference has no serving path; the closest is the batch of tasks that
``DiffPhyTrainer.simulate`` advances together (``diffphy_trainer.py:219-268``).
"""

import collections
import ctypes
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from .uneven import WAVE_MAX_IN, WAVE_MAX_OUT, WAVE_MAX_VERT, n_inputs


@dataclass
class Request:
    rid: int
    kind: str                  # "ms" (mass-spring) or "mpm"
    horizon: int
    grad: bool                 # forward + adjoint (True) or forward only
    target: np.ndarray         # [5]: tv (3), th, unused
    weights: list              # W1 [H][I], b1 [H], W2 [O][H], b2 [O] (device tensors)
    design: object = None      # robot_design.Design (ms)
    mpm: dict = None           # {"params", "positions", "act_id", "n_act"} (mpm)
    arrival: float = 0.0       # seconds after the start of the run
    admitted: float = None
    done: float = None
    loss: float = None
    grads: list = None
    nbytes: int = 0
    rejected: str = None       # why the request was turned away (no executor fits it)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def _rows(n):
    return np.zeros(n, dtype=nat.ROBOT_DTYPE)


class MSPool:
    """Fixed-capacity slots for mass-spring requests (one persistent launch
    pair per tick).  caps = (max vertices, springs, inputs, outputs,
    horizon); every slot owns a region sized for the caps in each shared
    array."""

    def __init__(self, name, slots, caps, physics, layout, ckpt_every=50, loss_window=100,
                 n_hidden=64, act="tanh", dtype=torch.float32, variant=1, block=256,
                 lds_policy=0, device="cuda"):
        self.name, self.C = name, slots
        self.Nc, self.Sc, self.Ic, self.Oc, self.Tc = caps
        if variant in (4, 5):
            # stored rows: the forward keeps every step's adjoint rows, so the
            # scratch (and the checkpoint interval) spans the horizon cap
            ckpt_every = max(ckpt_every, self.Tc)
        self.phys, self.lay = physics, layout
        self.K, self.W, self.H = ckpt_every, loss_window, n_hidden
        self.act = nat.ACT_SIN if act == "sin" else nat.ACT_TANH
        self.dtype, self.device = dtype, torch.device(device)
        self.variant, self.block, self.lds_policy = variant, block, lds_policy
        L = nat.lib()
        H, K = n_hidden, ckpt_every
        self.sz = dict(
            spring=self.Sc, vptr=self.Nc + 1, vidx=2 * self.Sc, aptr=self.Oc + 1, aidx=self.Sc,
            weights=L.dl_weights_size(self.Ic, H, self.Oc), wgrad=L.dl_wgrad_size(self.Ic, H, self.Oc),
            state=6 * self.Nc, ckpt=(self.Tc + K - 1) // K * 6 * self.Nc,
            scratch=L.dl_scratch_size(self.Nc, self.Sc, self.Ic, H, self.Oc, K), targets=5)
        C1 = slots + 1                    # region `slots` holds the idle dummy robot
        i32 = lambda n: torch.zeros(C1 * n, dtype=torch.int32, device=self.device)
        rl = lambda n: torch.zeros(C1 * n, dtype=dtype, device=self.device)
        sz = self.sz
        self.sa, self.sb, self.sact = i32(sz["spring"]), i32(sz["spring"]), i32(sz["spring"])
        self.L0, self.kk, self.co = rl(sz["spring"]), rl(sz["spring"]), rl(sz["spring"])
        self.vptr, self.vidx = i32(sz["vptr"]), i32(sz["vidx"])
        self.aptr, self.aidx = i32(sz["aptr"]), i32(sz["aidx"])
        self.weights, self.wgrad = rl(sz["weights"]), rl(sz["wgrad"])
        self.state, self.adjoint = rl(sz["state"]), rl(sz["state"])
        self.ckpt, self.scratch = rl(sz["ckpt"]), rl(sz["scratch"])
        self.targets = rl(sz["targets"])
        self.loss, self.loss_grad = rl(1), rl(1)
        self.table = _rows(slots)
        self.t_table = torch.empty(slots * nat.ROBOT_DTYPE.itemsize, dtype=torch.uint8,
                                   device=self.device)
        self.req = [None] * slots
        self.phase = [None] * slots         # "fwd" / "bwd" / None
        self.free = collections.deque(range(slots))
        self._dummy = None
        self.stream = torch.cuda.Stream(device=self.device)
        self.pool_bytes = sum(t.numel() * t.element_size() for t in (
            self.sa, self.sb, self.sact, self.L0, self.kk, self.co, self.vptr, self.vidx,
            self.aptr, self.aidx, self.weights, self.wgrad, self.state, self.adjoint, self.ckpt,
            self.scratch, self.targets))

    # ------------------------------------------------------------ admission
    def fits(self, req):
        return req.kind == "ms" and fits_caps(req.design, req.horizon, self.lay,
                                              (self.Nc, self.Sc, self.Ic, self.Oc, self.Tc))

    def bytes_for(self, req):
        """Device bytes this request actually uses (admission accounting)."""
        d, L, H, K = req.design, nat.lib(), self.H, self.K
        I = n_inputs(self.lay, d.n_vert)
        es = torch.finfo(self.dtype).bits // 8
        reals = (3 * d.n_spring + L.dl_weights_size(I, H, d.n_act) + L.dl_wgrad_size(I, H, d.n_act)
                 + 12 * d.n_vert + (req.horizon + K - 1) // K * 6 * d.n_vert
                 + L.dl_scratch_size(d.n_vert, d.n_spring, I, H, d.n_act, K) + 5)
        ints = 3 * d.n_spring + d.n_vert + 1 + 2 * d.n_spring + d.n_act + 1 + d.n_spring
        return reals * es + ints * 4

    def has_slot(self):
        return len(self.free) > 0

    def _write_design(self, k, d, stiffness_scale=0.1):
        """Topology of design d into region k."""
        sz, S = self.sz, d.n_spring
        o = k * sz["spring"]
        put = lambda dst, off, a, dt: dst[off:off + len(a)].copy_(torch.as_tensor(np.asarray(a), dtype=dt))
        put(self.sa, o, d.spring_a, torch.int32)
        put(self.sb, o, d.spring_b, torch.int32)
        put(self.sact, o, d.act_index, torch.int32)
        put(self.L0, o, d.L0, self.dtype)
        put(self.kk, o, np.asarray(d.stiffness) * stiffness_scale, self.dtype)
        put(self.co, o, d.coeff, self.dtype)
        vp, vx = d.vertex_csr()
        ap, ax = d.actuator_csr()
        put(self.vptr, k * sz["vptr"], np.asarray(vp) + k * sz["vidx"], torch.int32)
        put(self.vidx, k * sz["vidx"], vx, torch.int32)
        put(self.aptr, k * sz["aptr"], np.asarray(ap) + k * sz["aidx"], torch.int32)
        if len(ax):
            put(self.aidx, k * sz["aidx"], ax, torch.int32)

    def _row(self, k, d, T):
        sz, e = self.sz, _rows(1)[0]
        e["n_vert"], e["n_spring"], e["n_act"] = d.n_vert, d.n_spring, d.n_act
        e["n_in"] = n_inputs(self.lay, d.n_vert)
        e["horizon"], e["t_fwd"], e["t_bwd"] = T, 0, 0
        e["spring0"], e["vptr0"], e["aptr0"] = k * sz["spring"], k * sz["vptr"], k * sz["aptr"]
        for f in ("weights", "wgrad", "state", "ckpt", "scratch", "targets"):
            e[f] = k * sz[f]
        e["target_stride"] = 0
        e["traj"] = -1
        return e

    def _idle_row(self):
        """A tiny valid robot with nothing to do, for free slots."""
        if self._dummy is None:
            from .robot_design import build_voxels
            d = build_voxels([((0, 0, 0), True, 0), ((1, 0, 0), True, 0)], n_act=1)
            self._write_design(self.C, d)
            self._dummy = self._row(self.C, d, 0)
        return self._dummy

    def start(self):
        for k in range(self.C):
            self.table[k] = self._idle_row()

    def admit(self, req):
        k = self.free.popleft()
        d = req.design
        self._write_design(k, d)
        sz, I, H, O = self.sz, n_inputs(self.lay, d.n_vert), self.H, d.n_act
        W1, b1, W2, b2 = [torch.as_tensor(w, device=self.device, dtype=self.dtype) for w in req.weights]
        blk = torch.cat([W1.reshape(-1), W1.t().contiguous().reshape(-1), b1.reshape(-1),
                         W2.reshape(-1), W2.t().contiguous().reshape(-1), b2.reshape(-1)])
        self.weights[k * sz["weights"]:k * sz["weights"] + blk.numel()].copy_(blk)
        x0 = np.asarray(d.objects, np.float64) + np.array([0.4, 0.0, 0.0])
        st = np.zeros(6 * d.n_vert)
        st[:3 * d.n_vert] = x0.T.reshape(-1)                       # SoA [3][N], v = 0
        so = k * sz["state"]
        self.state[so:so + 6 * d.n_vert].copy_(torch.as_tensor(st, dtype=self.dtype))
        self.adjoint[so:so + 6 * d.n_vert].zero_()
        self.wgrad[k * sz["wgrad"]:(k + 1) * sz["wgrad"]].zero_()
        self.targets[k * 5:(k + 1) * 5].copy_(torch.as_tensor(np.asarray(req.target, np.float64), dtype=self.dtype))
        self.loss[k].zero_()
        self.loss_grad[k].fill_(1.0)
        self.table[k] = self._row(k, d, req.horizon)
        self.req[k], self.phase[k] = req, "fwd"

    def active(self):
        return [k for k in range(self.C) if self.req[k] is not None]

    def busy(self):
        return any(p in ("fwd", "bwd") for p in self.phase)

    # ------------------------------------------------------------ ticks
    def _desc(self, chunk):
        D = nat.Rollout()
        D.dtype = nat.dtype_code(self.dtype)
        D.act = self.act
        D.n_robots = self.C
        D.n_hidden = self.H
        D.ckpt_every = self.K
        D.max_steps = chunk
        D.block = self.block
        D.loss_kind = 1
        D.loss_window = self.W
        D.store_traj = 0
        D.max_n, D.max_s, D.max_i, D.max_o = self.Nc, self.Sc, self.Ic, self.Oc
        D.lds_policy = self.lds_policy
        D.variant = self.variant
        D.phys, D.lay = self.phys, self.lay
        P = nat.ptr
        D.robots = P(self.t_table)
        D.spring_a, D.spring_b, D.spring_act = P(self.sa), P(self.sb), P(self.sact)
        D.spring_L0, D.spring_k, D.spring_coeff = P(self.L0), P(self.kk), P(self.co)
        D.vert_ptr, D.vert_idx = P(self.vptr), P(self.vidx)
        D.act_ptr, D.act_idx = P(self.aptr), P(self.aidx)
        D.weights, D.wgrad = P(self.weights), P(self.wgrad)
        D.state, D.adjoint, D.ckpt, D.scratch = P(self.state), P(self.adjoint), P(self.ckpt), P(self.scratch)
        D.targets = P(self.targets)
        D.loss, D.loss_grad = P(self.loss), P(self.loss_grad)
        D.traj, D.traj_grad = None, None
        return D

    def _upload(self):
        host = torch.from_numpy(self.table.view(np.uint8).copy()).pin_memory()
        self.t_table.copy_(host, non_blocking=True)

    def tick(self, chunk):
        """Launch one tick on the pool's stream; returns the requests whose
        work completes with it (host clocks are exact, no device read)."""
        if not any(r is not None for r in self.req):
            return []
        finished = []
        self.stream.wait_stream(torch.cuda.current_stream())     # admissions
        with torch.cuda.stream(self.stream):
            self._upload()
            D = self._desc(chunk)
            nat.call("dl_ms_rollout_forward", ctypes.byref(D), nat.stream())
            for k in self.active():
                if self.phase[k] != "fwd":
                    continue
                e = self.table[k]
                e["t_fwd"] = min(int(e["horizon"]), int(e["t_fwd"]) + chunk)
                if e["t_fwd"] == e["horizon"]:
                    if self.req[k].grad:
                        self.phase[k] = "bwd"
                        e["t_bwd"] = e["horizon"]
                    else:
                        self.phase[k] = "done"
                        finished.append(k)
            if any(p == "bwd" for p in self.phase):
                self._upload()
                nat.call("dl_ms_rollout_backward", ctypes.byref(D), nat.stream())
                for k in self.active():
                    if self.phase[k] != "bwd":
                        continue
                    e = self.table[k]
                    e["t_bwd"] = max(0, int(e["t_bwd"]) - chunk)
                    if e["t_bwd"] == 0:
                        self.phase[k] = "done"
                        finished.append(k)
        return finished

    def collect(self, k):
        """Results of slot k (after the tick's work has completed); frees it."""
        req = self.req[k]
        req.loss = float(self.loss[k].item())
        if req.grad:
            d, H = req.design, self.H
            I, O = n_inputs(self.lay, d.n_vert), d.n_act
            o = k * self.sz["wgrad"]
            g = self.wgrad[o:o + H * I + H + O * H + O].clone()
            req.grads = [g[:H * I].view(H, I), g[H * I:H * I + H],
                         g[H * I + H:H * I + H + O * H].view(O, H), g[H * I + H + O * H:]]
        self.req[k], self.phase[k] = None, None
        self.table[k] = self._idle_row()
        self.free.append(k)
        return req


class MPMPool:
    """MPM requests on one grid size, packed into shared uneven MPMBatch
    launches.  Admitted requests are staged; at a tick, when fewer than
    ``max_cohorts`` batches are running, the staged requests (up to
    ``max_robots``, in admission order) start together as one batch: every
    robot on its own grid with its own body (padded to the largest),
    controller, target and horizon.  A tick advances every running batch by
    one checkpoint segment (forward, then the adjoint when some request of
    the batch wants gradients), each batch on its own stream; forward-only
    requests complete when their batch's forward does."""

    def __init__(self, n_grid, max_robots=8, max_cohorts=2, ckpt_every=25, window=100,
                 max_particles=60000, dtype=torch.float32, device="cuda"):
        self.n_grid, self.max_robots, self.max_cohorts = n_grid, max_robots, max_cohorts
        self.K, self.window, self.max_particles = ckpt_every, window, max_particles
        self.dtype, self.device = dtype, device
        self.staged, self.cohorts = [], []
        self.stream = torch.cuda.Stream(device=device)          # joins the batch streams
        self.free_streams = [torch.cuda.Stream(device=device) for _ in range(max_cohorts)]

    def fits(self, req):
        return (req.kind == "mpm" and req.mpm["params"].n_grid == self.n_grid
                and len(req.mpm["positions"]) <= self.max_particles)

    def has_slot(self):
        return (len(self.staged) + sum(len(c["reqs"]) for c in self.cohorts)
                < self.max_robots * self.max_cohorts)

    def busy(self):
        return bool(self.staged) or any(c["phase"] in ("fwd", "bwd") for c in self.cohorts)

    @staticmethod
    def estimate_bytes(P, n, horizon, ckpt_every=25, es=4):
        """Device bytes of one robot of an MPMBatch with P particle slots on an
        n^3 grid: state slots (ring, checkpoints, segment, adjoint) of 24 reals
        per particle, four 16-byte grid records per node, the segment's saved
        grids, block tables, per-checkpoint particle orders."""
        nck = (horizon + ckpt_every - 1) // ckpt_every
        slots = 2 + nck + ckpt_every + 1 + 2
        nblk = (n // 4) ** 3
        saved = ckpt_every * nblk * 64 * 4 * 2
        reorder = 2 * 8 * nck * P
        return (es * (slots * 24 * P + 4 * 4 * n ** 3 + saved) + 4 * (3 + ckpt_every) * nblk
                + reorder)

    def bytes_for(self, req):
        # a robot's slots are padded to the largest body of its batch
        return self.estimate_bytes(self.max_particles, self.n_grid, req.horizon, self.K,
                                   torch.finfo(self.dtype).bits // 8)

    def admit(self, req):
        self.staged.append(req)

    def active(self):
        return self.staged + [r for c in self.cohorts for r in c["reqs"]]

    def _start(self):
        from .mpm import MPMBatch
        reqs, self.staged = self.staged[:self.max_robots], self.staged[self.max_robots:]
        m0 = reqs[0].mpm
        B, T = len(reqs), max(r.horizon for r in reqs)
        st = self.free_streams.pop()
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            b = MPMBatch(m0["params"], [r.mpm["positions"] for r in reqs],
                         [r.mpm["act_id"] for r in reqs], m0["n_act"], B, T, K=self.K,
                         window=self.window, dtype=self.dtype, device=self.device,
                         save_horizon=False, horizons=[r.horizon for r in reqs],
                         per_robot_weights=True)
            tg = np.stack([np.asarray(r.target, np.float64) for r in reqs])
            b.set_targets(tg[:, :3], tg[:, 3])
            for k, r in enumerate(reqs):
                b.set_robot_weights(k, *r.weights)
            b.start_forward()
        self.cohorts.append({"reqs": list(reqs), "all": list(reqs), "b": b, "phase": "fwd",
                             "seg": 0, "stream": st})

    def tick(self, chunk=None):
        finished = []
        cur = torch.cuda.current_stream()
        while self.staged and self.free_streams:
            self._start()
        for c in self.cohorts:
            if c["phase"] not in ("fwd", "bwd"):
                continue
            st, b = c["stream"], c["b"]
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                if c["phase"] == "fwd":
                    b.forward_segment(c["seg"])
                    c["seg"] += 1
                    if c["seg"] == b.nck:
                        b.finish_forward()
                        finished += [(c, r) for r in c["all"] if not r.grad]
                        want = [float(r.grad) for r in c["all"]]
                        if any(want):
                            b.start_backward(loss_grad=torch.tensor(want, dtype=torch.float64))
                            c["phase"], c["seg"] = "bwd", b.nck - 1
                        else:
                            c["phase"] = "done"
                else:
                    b.backward_segment(c["seg"])
                    c["seg"] -= 1
                    if c["seg"] < 0:
                        b.finish_backward()
                        c["phase"] = "done"
                        finished += [(c, r) for r in c["all"] if r.grad]
            self.stream.wait_stream(st)
        return finished

    def collect(self, item):
        c, req = item
        b = c["b"]
        b.check()
        k = c["all"].index(req)
        req.loss = float(b.loss_b[k].item())
        if req.grad:
            req.grads = [g[k].clone() for g in b.grads]
        c["reqs"].remove(req)
        if not c["reqs"]:
            self.cohorts.remove(c)
            self.free_streams.append(c["stream"])
        return req


class CudaEvents:
    """Completion of an executor's launched tick: an event on its stream."""

    def record(self, ex):
        ev = torch.cuda.Event()
        ev.record(ex.stream)
        return ev

    def done(self, ev):
        return ev.query()


class Server:
    """Continuous-batching loop over executors (MSPool, MPMPool) with an
    HBM admission budget.  Each executor ticks at its own pace: a tick is
    launched as soon as the executor's previous one has completed on the
    device, so mass-spring slots are not held back by MPM segments.
    ``run(requests)`` plays the arrivals in real time and returns the
    requests with admission / completion times and results."""

    def __init__(self, executors, budget_bytes, chunk=50, clock=time.perf_counter,
                 sleep=time.sleep, events=None, max_active=None, poll_s=2e-5):
        self.ex = executors
        self.budget = budget_bytes
        self.chunk = chunk
        self.clock, self.sleep = clock, sleep
        self.events = events or CudaEvents()
        self.max_active = max_active
        self.poll_s = poll_s
        self.ticks = 0
        self.rejected = []

    def _executor(self, req):
        for e in self.ex:
            if e.fits(req):
                return e
        return None

    def run(self, requests):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        used, n_active = 0, 0
        inflight = {}
        for e in self.ex:
            if hasattr(e, "start"):
                e.start()
        t0 = self.clock()
        now = lambda: self.clock() - t0
        done = []
        while pending or queue or n_active:
            t = now()
            while pending and pending[0].arrival <= t:
                queue.append(pending.popleft())
            # admit in arrival order while the head fits (no overtaking); a
            # request no executor can take is turned away, the others go on
            while queue:
                r = queue[0]
                e = self._executor(r)
                if e is None:
                    queue.popleft()
                    r.rejected = "no executor takes this robot (sizes beyond every pool's caps)"
                    r.done = now()
                    self.rejected.append(r)
                    continue
                need = e.bytes_for(r)
                if need > self.budget:
                    raise ValueError(f"request {r.rid} needs {need} B, over the whole budget")
                if (used + need > self.budget or not e.has_slot()
                        or (self.max_active is not None and n_active >= self.max_active)):
                    break
                queue.popleft()
                r.nbytes, r.admitted = need, now()
                e.admit(r)
                used += need
                n_active += 1
            # launch the next tick of every idle executor with work
            for i, e in enumerate(self.ex):
                if i not in inflight and e.busy():
                    fin = e.tick(self.chunk)
                    inflight[i] = (self.events.record(e), fin)
                    self.ticks += 1
            progressed = False
            for i in list(inflight):
                ev, fin = inflight[i]
                if not self.events.done(ev):
                    continue
                del inflight[i]
                progressed = True
                t = now()
                for k in fin:
                    r = self.ex[i].collect(k)
                    r.done = t
                    used -= r.nbytes
                    n_active -= 1
                    done.append(r)
            if progressed:
                continue
            if inflight:
                self.sleep(self.poll_s)
            elif pending and n_active == 0:
                self.sleep(max(0.0, pending[0].arrival - now()))
        return done


def latency_stats(reqs, wall):
    lat = np.array([r.latency for r in reqs])
    return {"requests": len(reqs), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "max_s": float(lat.max()), "throughput_req_s": len(reqs) / wall,
            "robot_steps": int(sum(r.horizon for r in reqs))}


# ------------------------------------------------------------------ workload
def mpm_particle_robot(rng, n_grid, n_particles, n_act=4, ppc_edge=2):
    """A compact random voxel body for the MPM solver: cubes of c^3 cells
    (c = 3 on 32^3 grids, 4 on 64^3), ppc_edge^3 particles per cell, grown by
    connected accretion inside a box that keeps the body off the walls;
    actuator groups = nearest of n_act seed cubes."""
    c = 3 if n_grid <= 32 else 4
    per_cube = c ** 3 * ppc_edge ** 3
    n_cubes = max(1, n_particles // per_cube)        # at most n_particles
    box = (n_grid - 8) // c                      # cubes per axis, 4 cells margin
    cells = [(0, 0, 0)]
    have = {(0, 0, 0)}
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    while len(cells) < min(n_cubes, box ** 3):
        b = cells[int(rng.integers(len(cells)))]
        d = dirs[int(rng.integers(6))]
        q = (b[0] + d[0], b[1] + d[1], b[2] + d[2])
        if q in have or q[1] < 0:
            continue
        span = [max(max(p[i] for p in cells), q[i]) - min(min(p[i] for p in cells), q[i]) + 1
                for i in range(3)]
        if max(span) > box:
            continue
        have.add(q)
        cells.append(q)
    cells = np.array(cells)
    cells -= cells.min(0)
    g = min(n_act, len(cells))
    seeds = cells[rng.choice(len(cells), size=g, replace=False)]
    grp = ((cells[:, None, :] - seeds[None]) ** 2).sum(-1).argmin(1)
    dx = 1.0 / n_grid
    sub = (np.arange(ppc_edge) + 0.5) / ppc_edge
    local = np.stack(np.meshgrid(*[np.arange(c)] * 3, indexing="ij"), -1).reshape(-1, 3)
    subs = np.stack(np.meshgrid(sub, sub, sub, indexing="ij"), -1).reshape(-1, 3)
    # 4 cells from the side walls; the bottom cubes start on the floor (the
    # sticky boundary's 3 cells), so the body is in contact from step 0
    lift = np.array([4.0, 3.0, 4.0])
    pos = ((cells[:, None, None, :] * c + local[None, :, None, :] + subs[None, None] + lift) * dx)
    pos = pos.reshape(-1, 3)
    ids = np.repeat(grp, c ** 3 * ppc_edge ** 3).astype(np.int32)
    return pos, ids, n_act


def make_requests(rate, duration, seed, layout, frac_mpm=0.3, grad_frac=0.5, n_hidden=64,
                  ms_horizon=(100, 1000), mpm_horizon=(200, 2000), mpm_particles=(8000, 60000),
                  device="cuda", count=None):
    """Poisson arrivals at `rate` per second over `duration` seconds (or
    exactly `count` arrivals): a fraction frac_mpm of MPM bodies (the soft
    quadruped's physics, 32^3 or 64^3 grids), the rest random voxel
    mass-spring robots (the 512-robot generator's distribution), half of
    each wanting gradients.  The robots drawn do not depend on `rate` (its
    inter-arrival draws scale one unit-rate sequence)."""
    from .mpm import quadruped3d_params
    from .robot_design import random_voxel_cubes, build_voxels
    from .rollout import init_weights
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0) / rate
        if (count is None and t > duration) or (count is not None and len(reqs) >= count):
            break
        rid = len(reqs)
        grad = bool(rng.random() < grad_frac)
        tgt = np.array([rng.uniform(-0.08, 0.08), 0.0, rng.uniform(-0.08, 0.08),
                        rng.uniform(0.1, 0.25), 0.0])
        if rng.random() < frac_mpm:
            n_grid = int(rng.choice([32, 64]))
            P = int(rng.integers(mpm_particles[0], mpm_particles[1] + 1))
            if n_grid == 32:
                P = min(P, 30000)
            pos, ids, n_act = mpm_particle_robot(rng, n_grid, P)
            T = int(rng.integers(mpm_horizon[0], mpm_horizon[1] + 1))
            I = 10 + 3 + 3 * (n_act + 1) + 4
            reqs.append(Request(rid, "mpm", T, grad, tgt, init_weights(I, n_hidden, n_act, seed=seed * 7919 + rid, device=device),
                                mpm={"params": quadruped3d_params(n_grid), "positions": pos,
                                     "act_id": ids, "n_act": n_act}, arrival=t))
        else:
            nc = int(round(math.exp(rng.uniform(math.log(8), math.log(400)))))
            ng = int(rng.integers(4, 65))
            cubes, g = random_voxel_cubes(rng, nc, ng)
            d = build_voxels(cubes, n_act=g)
            T = int(rng.integers(ms_horizon[0], ms_horizon[1] + 1))
            reqs.append(Request(rid, "ms", T, grad, tgt,
                                init_weights(n_inputs(layout, d.n_vert), n_hidden, d.n_act,
                                             seed=seed * 7919 + rid, device=device),
                                design=d, arrival=t))
    return reqs


def pool_caps(layout):
    """Caps (vertices, springs, inputs, outputs, horizon) of the two
    mass-spring pools.  The wave pool's are the wave kernel's limits.  The
    request generator draws 8-400 connected cubes, 4-64 actuator groups and
    horizons up to 1000; its 400-cube robots reach ~820 vertices and ~6700
    springs, so the block pool takes 1024 vertices and 8192 springs."""
    return {"wave": (WAVE_MAX_VERT, 512, WAVE_MAX_IN, WAVE_MAX_OUT, 1000),
            "block": (1024, 8192, n_inputs(layout, 1024), 64, 1000)}


# particle capacity of the MPM pools per grid size (the request generator's
# ranges: up to 30k particles on 32^3 grids, 60k on 64^3)
MPM_POOL_PARTICLES = {32: 30000, 64: 60000}


def fits_caps(design, horizon, layout, caps):
    N, S, I, O, T = caps
    return (design.n_vert <= N and design.n_spring <= S and n_inputs(layout, design.n_vert) <= I
            and design.n_act <= O and horizon <= T)


def default_executors(physics, layout, slots_small=256, slots_large=128, mpm_active=8,
                      ckpt_every=50, loss_window=100, dtype=torch.float32, device="cuda",
                      stored=True):
    """Two mass-spring pools (wave-per-robot kernel for small robots, block
    kernel up to 1024 vertices / 8192 springs / 64 actuators / 1000 steps)
    and two MPM pools (32^3 and 64^3 grids, batches of up to mpm_active
    robots).  With `stored` the mass-spring pools' forwards keep every
    step's adjoint rows (variants 4 and 5), else checkpoints every
    ckpt_every steps are recomputed (variants 3 and 1)."""
    caps = pool_caps(layout)
    small = MSPool("wave", slots_small, caps["wave"],
                   physics, layout, ckpt_every, loss_window, dtype=dtype,
                   variant=4 if stored else 3, block=64, device=device)
    large = MSPool("block", slots_large, caps["block"], physics,
                   layout, ckpt_every, loss_window, dtype=dtype, variant=5 if stored else 1,
                   block=256, device=device)
    return [small, large] + [MPMPool(n, mpm_active, max_particles=cap, window=loss_window,
                                     dtype=dtype, device=device)
                             for n, cap in MPM_POOL_PARTICLES.items()]
