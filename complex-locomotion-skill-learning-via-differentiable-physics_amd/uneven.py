"""Uneven mass-spring batches: robots of different designs, sizes, actuator
counts and horizons, each with its own MLP, trained together.

The reference trains one design at a time, the batch axis holding tasks of
that design (``diffphy_trainer.py:219-268``).  Here every robot may be a
different design.  Robots are grouped into size classes; each class runs as
one persistent forward launch and one adjoint launch (``MassSpringBatch``,
driven by the per-robot offset table) with a launch shape chosen by the
autotuner, and the classes run concurrently on separate streams.  A whole
iteration can be captured in HIP graphs and replayed.

Size classes: robots that fit the wave-per-robot kernel (FP32, hidden 64,
network input <= 192, outputs <= 8, <= 64 vertices) form one class; the rest
are split by vertex count (the block-per-robot kernel, whose best block size
grows with the robot).
"""

import numpy as np
import torch

from .autotune import Autotuner, library_tag
from .rollout import GraphedIteration, MassSpringBatch

WAVE_MAX_IN, WAVE_MAX_OUT, WAVE_MAX_VERT = 192, 8, 64


_CLASS_STREAMS = {}


def _class_streams(n, device):
    """The first n of a process-wide list of streams for size classes.  The
    HIP runtime binds a stream to one of its hardware queues (four by
    default) when the stream is first used, in round-robin order; streams
    created per batch, and so handed out by torch's pool in whatever order,
    can land two classes on one queue, where their launches run one after
    the other.  Reused streams, first used together (classes are issued
    largest first), keep the mapping the same for every batch."""
    key = str(torch.device(device))
    lst = _CLASS_STREAMS.setdefault(key, [])
    while len(lst) < n:
        lst.append(torch.cuda.Stream(device=device))
    return lst[:n]


def n_inputs(layout, n_vert):
    """Network input width of a robot with n_vert vertices (taichi_env.py:149-183 order)."""
    lay = layout
    return (lay.n_sin + 3 * lay.vcom + n_vert * (3 + lay.vert_pad) + lay.dup_v * lay.n_tv
            + lay.dup_h + lay.dup_c)


def size_classes(designs, layout, dtype, n_hidden=64, bounds=(128, 400)):
    """Ordered list of (name, robot indices)."""
    out = {}
    for r, d in enumerate(designs):
        if (dtype == torch.float32 and n_hidden == 64 and n_inputs(layout, d.n_vert) <= WAVE_MAX_IN
                and d.n_act <= WAVE_MAX_OUT and d.n_vert <= WAVE_MAX_VERT):
            name = "wave"
        else:
            name = next((f"n<={b}" for b in bounds if d.n_vert <= b), f"n>{bounds[-1]}")
        out.setdefault(name, []).append(r)
    order = ["wave"] + [f"n<={b}" for b in bounds] + [f"n>{bounds[-1]}"]
    return [(k, out[k]) for k in order if k in out]


class UnevenBatch:
    """R robots, robot r = design designs[r] with horizon horizons[r].

    targets: [R, 5] (tv (3), th, unused); x0: list of [N_r, 3] (default: the
    design's rest positions shifted +0.4 in x, taichi_env.py:500-501).
    stored: the forwards keep every step's adjoint rows and the adjoints read
    them (variants 4 / 5, scratch for each class's longest horizon) instead
    of recomputing checkpoint segments of ckpt_every steps."""

    def __init__(self, designs, horizons, physics, layout, targets, x0=None, ckpt_every=50,
                 loss_window=100, dtype=torch.float32, act="tanh", n_hidden=64, streams=True,
                 device="cuda", bounds=(128, 400), stored=True):
        self.R = len(designs)
        self.designs, self.horizons = designs, np.asarray(horizons, np.int64)
        if x0 is None:
            x0 = [d.objects + np.array([0.4, 0.0, 0.0]) for d in designs]
        targets = np.asarray(targets, np.float64)
        self.dtype, self.streams_on, self.stored = dtype, streams, stored
        self.buckets = []
        self.slot = {}
        for name, ids in size_classes(designs, layout, dtype, n_hidden, bounds):
            b = MassSpringBatch([designs[r] for r in ids], list(range(len(ids))),
                                self.horizons[ids], physics, layout, targets[ids],
                                [x0[r] for r in ids], n_hidden=n_hidden, act=act,
                                ckpt_every=ckpt_every, dtype=dtype, loss_window=loss_window,
                                device=device, variant=self._variant(name),
                                block=64 if name == "wave" else 256)
            for j, r in enumerate(ids):
                self.slot[r] = (len(self.buckets), j)
            self.buckets.append({"name": name, "ids": np.asarray(ids), "batch": b, "config": None})
        self.graphs = None
        n = len(self.buckets)
        # the largest class on the first shared stream, the next on the second ...
        self._streams = _class_streams(n, device)[::-1] if streams else None

    # ------------------------------------------------------------ weights
    def set_weights(self, r, W1, b1, W2, b2):
        bi, j = self.slot[r]
        self.buckets[bi]["batch"].set_weights(j, W1, b1, W2, b2)

    def weight_grads(self, r):
        bi, j = self.slot[r]
        return self.buckets[bi]["batch"].weight_grads(j)

    def losses(self):
        out = torch.empty(self.R, dtype=self.dtype, device=self.buckets[0]["batch"].loss.device)
        for bk in self.buckets:
            out[torch.as_tensor(bk["ids"], device=out.device)] = bk["batch"].loss
        return out

    @property
    def robot_steps(self):
        return int(self.horizons.sum())

    # ------------------------------------------------------------ launch shapes
    def _variant(self, name):
        if name == "wave":
            return 4 if self.stored else 3
        return 5 if self.stored else 1

    def candidates(self, name, quick=False):
        """Launch shapes of a size class; ``quick``: block sizes 256 and up
        with the LDS policy 4 (what wins on the random-robot batches), for the
        schedule search."""
        if name == "wave":
            vs = (4,) if self.stored else (3, 2)
            return [{"variant": v, "block": 64, "lds_policy": 0} for v in vs][:1 if quick else 2]
        # 1024-thread blocks (FP32) win for the largest robots alone but lose
        # in the concurrent iteration: a workgroup that needs a whole CU waits
        # for the other classes' workgroups to leave one (measured: 246 ->
        # 360 ms for the 512 random robots), so the tuner does not offer them
        blocks = (256, 512)
        if quick:
            return [{"variant": self._variant(name), "block": b, "lds_policy": 4}
                    for b in blocks]
        return [{"variant": self._variant(name), "block": b, "lds_policy": p}
                for b in (64, 128) + blocks for p in (0, 1, 4)]

    def _apply(self, bk, cfg):
        b = bk["batch"]
        b.variant, b.block, b.lds_policy = cfg["variant"], cfg["block"], cfg["lds_policy"]
        bk["config"] = cfg

    def tune(self, tuner=None, quick=False):
        """Pick each size class's launch shape by timing a full iteration of
        that class with every candidate (cached by the autotuner); without
        ``quick``, keep these picks only if the concurrent iteration is faster
        with them than with the quick candidates' picks."""
        tuner = tuner or Autotuner()
        for bk in self.buckets:
            b = bk["batch"]
            nv = np.array([b.designs[int(d)].n_vert for d in b.robot_design])
            key = Autotuner.key(kernel="ms_rollout", cls=bk["name"], dtype=str(self.dtype),
                                robots=len(bk["ids"]), mean_n=int(nv.mean()),
                                mean_T=int(b.horizons.mean()), stored=int(self.stored),
                                max_n=b.max_n, max_s=b.max_s, max_i=b.max_i, max_o=b.max_o,
                                lib=library_tag())

            def measure(cfg, bk=bk):
                self._apply(bk, cfg)
                bk["batch"].prepare_iteration(np.ones(len(bk["ids"])))
                bk["batch"].run_iteration()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                bk["batch"].run_iteration()
                e1.record()
                torch.cuda.synchronize()
                return e0.elapsed_time(e1)

            self._apply(bk, tuner.pick(key + f"|quick={int(quick)}",
                                       self.candidates(bk["name"], quick), measure))
        if not quick:
            # each class was timed alone; concurrently a shape that wins alone
            # can lose (a wider workgroup waiting for a CU the other classes
            # hold), so the whole iteration is timed with these shapes and with
            # the quick candidates' picks, and the faster set is kept
            full = [bk["config"] for bk in self.buckets]
            quick_set = self.tune(tuner, quick=True)
            if full != quick_set:
                t_quick = self._iteration_ms()
                for bk, cfg in zip(self.buckets, full):
                    self._apply(bk, cfg)
                # the classes' placement on the CUs varies from run to run, so
                # the per-class picks must win clearly to be kept
                if self._iteration_ms() > 0.97 * t_quick:
                    for bk, cfg in zip(self.buckets, quick_set):
                        self._apply(bk, cfg)
        return [bk["config"] for bk in self.buckets]

    def _iteration_ms(self, reps=3):
        """Median of ``reps`` concurrent iterations (after one warm-up), ms."""
        self.prepare()
        self.run_iteration()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ts = []
        for _ in range(reps):
            e0.record()
            self.run_iteration()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    def class_times(self):
        """Milliseconds of one forward + adjoint iteration of each size class
        run alone (after prepare), by class name."""
        out = {}
        for bk in self.buckets:
            bk["batch"].run_iteration()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            bk["batch"].run_iteration()
            e1.record()
            torch.cuda.synchronize()
            out[bk["name"]] = e0.elapsed_time(e1)
        return out

    # ------------------------------------------------------------ iterations
    def prepare(self, loss_grad=None):
        """Device-side setup for full iterations; loss_grad: [R] weights of
        the per-robot losses (default 1 each)."""
        lg = np.ones(self.R) if loss_grad is None else np.asarray(loss_grad, np.float64)
        for bk in self.buckets:
            if bk["config"] is None:
                self._apply(bk, self.candidates(bk["name"])[0])
            bk["batch"].prepare_iteration(lg[bk["ids"]])

    def _on_streams(self, fns):
        """Run one function per size class, each on its own stream, issued
        largest class first: its robots have the longest per-step chains and
        set the iteration's length, so they take CUs before the small ones."""
        order = list(range(len(fns)))[::-1]
        if not self.streams_on:
            for i in order:
                fns[i]()
            return
        cur = torch.cuda.current_stream()
        for i in order:
            self._streams[i].wait_stream(cur)
            with torch.cuda.stream(self._streams[i]):
                fns[i]()
        for s in self._streams:
            cur.wait_stream(s)

    def run_iteration(self):
        """One forward + adjoint of every robot (size classes concurrently)."""
        self._on_streams([bk["batch"].run_iteration for bk in self.buckets])

    def capture(self):
        """Record each size class's iteration in a HIP graph."""
        self.graphs = [None] * len(self.buckets)
        for i in range(len(self.buckets))[::-1]:
            self.graphs[i] = GraphedIteration(self.buckets[i]["batch"].run_iteration)

    def replay(self):
        self._on_streams([g.replay for g in self.graphs])


# Size-class boundaries (vertex counts) the schedule search tries: which
# robots share a launch, and so which kernel and launch shape each robot
# runs with.  The first is the default.
SCHEDULES = ((128, 400), (96, 256, 512), (160, 480), (200,), (128, 256, 512))


def tune_schedule(designs, horizons, physics, layout, targets, set_weights=None, tuner=None,
                  schedules=SCHEDULES, **kw):
    """Robot-to-launch schedule of an uneven batch: for every class-boundary
    candidate, build the batch, pick each class's launch shape among the
    quick candidates, and time a whole iteration (classes concurrent);
    returns the fastest boundaries (cached by the autotuner per robot set
    and library build).  ``set_weights(batch)`` loads the controllers."""
    tuner = tuner or Autotuner()
    nv = np.array([d.n_vert for d in designs])
    key = Autotuner.key(kernel="ms_uneven_schedule", robots=len(designs), sum_n=int(nv.sum()),
                        max_n=int(nv.max()), sum_T=int(np.sum(horizons)),
                        stored=int(kw.get("stored", True)), lib=library_tag())

    def measure(c):
        ub = UnevenBatch(designs, horizons, physics, layout, targets, bounds=tuple(c["bounds"]),
                         **kw)
        if set_weights is not None:
            set_weights(ub)
        ub.tune(tuner, quick=True)
        ub.prepare()
        ub.run_iteration()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(2):          # the better of two iterations (concurrent classes vary)
            e0.record()
            ub.run_iteration()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        del ub
        torch.cuda.empty_cache()
        return best

    pick = tuner.pick(key, [{"bounds": list(b)} for b in schedules], measure, repeat=1)
    return tuple(pick["bounds"]), tuner.cache.get(key, {}).get("timings")
