"""Batched persistent rollout: many robots, one launch.

``MassSpringBatch`` packs robots of any mix of designs, sizes and horizons
into one offset table (``dl_robot``) and device buffers, and drives the
persistent kernels: ``forward`` advances every robot through its horizon
(one launch, checkpoints every K steps, in-kernel windowed COM loss or a
stored trajectory for external losses), ``backward`` runs the hand-written
adjoint segment by segment from the checkpoints (one launch) and reduces the
per-robot weight-gradient partials per design in a fixed order.

``rollout_loss`` wraps forward/backward in a ``torch.autograd.Function`` over
the weights of each design.

This replaces the reference's per-step Python loop with Tape replay
(``multitask/diffphy_trainer.py:231-256``) for training rollouts.
"""

import ctypes
import math

import numpy as np
import torch

from . import _native as nat


class DesignWeights:
    """Weight tensors of one design: W1 [H,I], b1 [H], W2 [O,H], b2 [O]."""

    def __init__(self, W1, b1, W2, b2):
        self.W1, self.b1, self.W2, self.b2 = W1, b1, W2, b2


def init_weights(n_in, n_hidden, n_out, seed, dtype=torch.float32, device="cuda"):
    """N(0, 1/fan_in) weights, zero biases."""
    g = torch.Generator(device="cpu")
    g.manual_seed(int(seed))
    W1 = torch.randn(n_hidden, n_in, generator=g, dtype=torch.float64) / math.sqrt(n_in)
    W2 = torch.randn(n_out, n_hidden, generator=g, dtype=torch.float64) / math.sqrt(n_hidden)
    mk = lambda t: t.to(dtype=dtype, device=device)
    return [mk(W1), torch.zeros(n_hidden, dtype=dtype, device=device), mk(W2),
            torch.zeros(n_out, dtype=dtype, device=device)]


class MassSpringBatch:
    """Persistent batched mass-spring rollout.

    designs:      list of ``robot_design.Design``
    robot_design: design index of every robot
    horizons:     advances per robot (int or list)
    physics:      ``_native.MSPhysics``; stiffness_scale multiplies the
                  builder stiffness (reference: 0.1)
    layout:       ``_native.SensorLayoutC``
    targets:      [R,5] (tv xyz, th, tc) constant per robot, or a list of
                  [T_r,5] per-step arrays
    x0:           list of [N_r,3] initial positions (v0 = 0)
    loss_window:  >0 enables the in-kernel windowed COM loss
    variant:      0 auto, 1 workgroup per robot, 2 wavefront per robot,
                  3 as 2 with the adjoint's segment recompute on a second
                  wavefront, 4 as 3 without recompute (the forward keeps
                  every step's adjoint rows; ckpt_every becomes the longest
                  horizon), 5 as 1 with the forward's rows kept likewise
    """

    def __init__(self, designs, robot_design, horizons, physics, layout, targets, x0,
                 n_hidden=64, act="tanh", ckpt_every=50, dtype=torch.float32, block=256,
                 loss_window=500, store_traj=False, stiffness_scale=0.1, device="cuda",
                 lds_policy=0, variant=0):
        self.designs = designs
        self.robot_design = np.asarray(robot_design, np.int64)
        R = self.R = len(self.robot_design)
        self.horizons = np.broadcast_to(np.asarray(horizons, np.int64), (R,)).copy()
        if variant in (4, 5):
            # the forward keeps every step's adjoint rows: one scratch buffer
            # (and one checkpoint interval) spans the longest horizon
            ckpt_every = max(int(ckpt_every), int(self.horizons.max()))
        self.dtype, self.device = dtype, torch.device(device)
        self.H, self.K, self.block = n_hidden, ckpt_every, block
        self.lds_policy = lds_policy
        self.variant = variant
        self.act = nat.ACT_SIN if act == "sin" else nat.ACT_TANH
        self.phys, self.layout = physics, layout
        self.loss_window = loss_window
        self.store_traj = store_traj
        lay = layout
        n_tv = lay.n_tv

        def n_in(d):
            return (lay.n_sin + 3 * lay.vcom + d.n_vert * (3 + lay.vert_pad) + lay.dup_v * n_tv
                    + lay.dup_h + lay.dup_c)

        # ---- topology of every design, concatenated
        sa, sb, sact, L0, kk, co, vptr, vidx, aptr, aidx = ([] for _ in range(10))
        self.spring0, self.vptr0, self.aptr0 = [], [], []
        s_off = v_off = a_off = vi_off = ai_off = 0
        for d in designs:
            self.spring0.append(s_off)
            self.vptr0.append(v_off)
            self.aptr0.append(a_off)
            vp, vx = d.vertex_csr()
            ap, ax = d.actuator_csr()
            sa.append(d.spring_a); sb.append(d.spring_b); sact.append(d.act_index)
            L0.append(d.L0); kk.append(d.stiffness * stiffness_scale); co.append(d.coeff)
            vptr.append(vp + vi_off); vidx.append(vx)
            aptr.append(ap + ai_off); aidx.append(ax)
            s_off += d.n_spring; v_off += d.n_vert + 1; a_off += d.n_act + 1
            vi_off += len(vx); ai_off += len(ax)
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(np.concatenate(a), np.int32),
                                        device=self.device)
        rl = lambda a: torch.as_tensor(np.concatenate(a), dtype=dtype, device=self.device)
        self.t_sa, self.t_sb, self.t_sact = i32(sa), i32(sb), i32(sact)
        self.t_L0, self.t_k, self.t_co = rl(L0), rl(kk), rl(co)
        self.t_vptr, self.t_vidx = i32(vptr), i32(vidx)
        self.t_aptr = i32(aptr)
        self.t_aidx = i32(aidx) if ai_off > 0 else torch.zeros(1, dtype=torch.int32, device=self.device)

        # ---- weights (per design) and gradient partials (per robot)
        self.n_in = [n_in(d) for d in designs]
        self.w_off, off = [], 0
        for d, I in zip(designs, self.n_in):
            self.w_off.append(off)
            off += nat.lib().dl_weights_size(I, n_hidden, d.n_act)
        self.weights = torch.zeros(off, dtype=dtype, device=self.device)
        tab = np.zeros(R, dtype=nat.ROBOT_DTYPE)
        g_off = s_off_ = c_off = sc_off = tr_off = 0
        tg_rows = []
        const_targets = not isinstance(targets, (list, tuple))
        if const_targets:
            targets = np.asarray(targets, np.float64).reshape(R, 5)
        self.g_size = []
        for r in range(R):
            di = int(self.robot_design[r])
            d, I = designs[di], self.n_in[di]
            N, Th = d.n_vert, int(self.horizons[r])
            e = tab[r]
            e["n_vert"], e["n_spring"], e["n_act"], e["n_in"] = N, d.n_spring, d.n_act, I
            e["horizon"], e["t_fwd"], e["t_bwd"] = Th, 0, Th
            e["spring0"], e["vptr0"], e["aptr0"] = self.spring0[di], self.vptr0[di], self.aptr0[di]
            e["weights"] = self.w_off[di]
            gs = nat.lib().dl_wgrad_size(I, n_hidden, d.n_act)
            self.g_size.append(gs)
            e["wgrad"] = g_off; g_off += gs
            e["state"] = s_off_; s_off_ += 6 * N
            e["ckpt"] = c_off; c_off += (Th + ckpt_every - 1) // ckpt_every * 6 * N
            e["scratch"] = sc_off
            sc_off += nat.lib().dl_scratch_size(N, d.n_spring, I, n_hidden, d.n_act, ckpt_every)
            e["targets"] = sum(len(t) for t in tg_rows) * 5 if not const_targets else r * 5
            if not const_targets:
                tg_rows.append(np.asarray(targets[r], np.float64).reshape(-1, 5))
            e["target_stride"] = 0 if const_targets else 5
            if store_traj:
                e["traj"] = tr_off
                tr_off += (Th + 1) * (3 * N + 3)
            else:
                e["traj"] = -1
        self.table = tab
        tg = targets.reshape(-1) if const_targets else np.concatenate(tg_rows).reshape(-1)
        z = lambda n: torch.zeros(max(int(n), 1), dtype=dtype, device=self.device)
        self.targets = torch.as_tensor(tg, dtype=dtype, device=self.device)
        self.wgrad = z(g_off)
        self.state = z(s_off_)
        self.adjoint = z(s_off_)
        self.ckpt = z(c_off)
        self.scratch = z(sc_off)
        self.traj = z(tr_off) if store_traj else None
        self.traj_grad = None
        self.loss = z(R)
        self.loss_grad = z(R)
        self.err = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.x0 = x0
        self.max_n = max(d.n_vert for d in designs)
        self.max_s = max(d.n_spring for d in designs)
        self.max_i = max(self.n_in)
        self.max_o = max(d.n_act for d in designs)
        self.t_table = torch.empty(R * nat.ROBOT_DTYPE.itemsize, dtype=torch.uint8, device=self.device)
        # design -> robots lists for the gradient reduction
        order = np.argsort(self.robot_design, kind="stable")
        counts = np.bincount(self.robot_design, minlength=len(designs))
        lptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.t_lptr = torch.as_tensor(lptr, device=self.device)
        self.t_list = torch.as_tensor(order.astype(np.int32), device=self.device)
        self.t_roff = torch.as_tensor(tab["wgrad"].astype(np.int64), device=self.device)
        dsz = [nat.lib().dl_wgrad_size(I, n_hidden, d.n_act) for d, I in zip(designs, self.n_in)]
        self.d_goff = np.concatenate([[0], np.cumsum(dsz)[:-1]]).astype(np.int64)
        self.t_doff = torch.as_tensor(self.d_goff, device=self.device)
        self.t_dsize = torch.as_tensor(np.asarray(dsz, np.int64), device=self.device)
        self.grad_out = z(sum(dsz))
        self._state0 = None
        self.reset()

    # ------------------------------------------------------------ state
    def reset(self):
        """Initial state into every slot, clocks to 0, losses and grads zeroed."""
        if self._state0 is None:
            st = np.zeros(self.state.numel(), np.float64)
            for r in range(self.R):
                N = self.designs[int(self.robot_design[r])].n_vert
                o = int(self.table[r]["state"])
                x0 = np.asarray(self.x0[r], np.float64).reshape(N, 3)
                st[o:o + 3 * N] = x0.T.reshape(-1)           # SoA [3][N], v = 0
            self._state0 = torch.as_tensor(st, dtype=self.dtype, device=self.device)
        self.state.copy_(self._state0)
        self.table["t_fwd"] = 0
        self.table["t_bwd"] = self.horizons
        self._upload_table()
        self.loss.zero_()
        self.adjoint.zero_()
        self.wgrad.zero_()
        self.err.zero_()
        if self.traj is not None:
            self.traj.zero_()

    def _upload_table(self):
        self.t_table.copy_(torch.from_numpy(self.table.view(np.uint8).copy()))

    def set_weights(self, di, W1, b1, W2, b2):
        """Write design di's weights and their transposed copies."""
        d, I, H = self.designs[di], self.n_in[di], self.H
        O = d.n_act
        blk = torch.cat([W1.reshape(-1), W1.t().contiguous().reshape(-1), b1.reshape(-1),
                         W2.reshape(-1), W2.t().contiguous().reshape(-1), b2.reshape(-1)])
        o = self.w_off[di]
        self.weights[o:o + blk.numel()] = blk.to(self.dtype)

    def weight_grads(self, di):
        d, I, H = self.designs[di], self.n_in[di], self.H
        O = d.n_act
        g = self.grad_out[int(self.d_goff[di]):]
        W1 = g[: H * I].view(H, I)
        b1 = g[H * I: H * I + H]
        W2 = g[H * I + H: H * I + H + O * H].view(O, H)
        b2 = g[H * I + H + O * H: H * I + H + O * H + O]
        return W1, b1, W2, b2

    # ------------------------------------------------------------ launches
    def _desc(self, max_steps=0):
        D = nat.Rollout()
        D.dtype = nat.dtype_code(self.dtype)
        D.act = self.act
        D.n_robots = self.R
        D.n_hidden = self.H
        D.ckpt_every = self.K
        D.max_steps = max_steps
        D.block = self.block
        D.loss_kind = 1 if self.loss_window > 0 else 0
        D.loss_window = max(self.loss_window, 0)
        D.store_traj = int(self.store_traj)
        D.max_n, D.max_s, D.max_i, D.max_o = self.max_n, self.max_s, self.max_i, self.max_o
        D.lds_policy = self.lds_policy
        D.variant = self.variant
        if self.variant in (2, 3, 4) or (self.variant == 0 and self.wave_eligible()):
            # one wavefront per robot; 4 per workgroup when each aligned group of
            # 4 robots shares one design (topology and W1 are staged per workgroup)
            D.block = 256 if self.packed4() else 64
        D.phys = self.phys
        D.lay = self.layout
        P = nat.ptr
        D.robots = P(self.t_table)
        D.spring_a, D.spring_b, D.spring_act = P(self.t_sa), P(self.t_sb), P(self.t_sact)
        D.spring_L0, D.spring_k, D.spring_coeff = P(self.t_L0), P(self.t_k), P(self.t_co)
        D.vert_ptr, D.vert_idx = P(self.t_vptr), P(self.t_vidx)
        D.act_ptr, D.act_idx = P(self.t_aptr), P(self.t_aidx)
        D.weights, D.wgrad = P(self.weights), P(self.wgrad)
        D.state, D.adjoint, D.ckpt, D.scratch = P(self.state), P(self.adjoint), P(self.ckpt), P(self.scratch)
        D.targets = P(self.targets)
        D.loss, D.loss_grad = P(self.loss), P(self.loss_grad)
        D.traj = P(self.traj) if self.traj is not None else None
        D.traj_grad = P(self.traj_grad) if self.traj_grad is not None else None
        D.err = P(self.err)
        return D

    def _check_contract(self):
        """Host-side checks of the launch contract (the kernels would mark the
        offending robots with NaN and an error bit instead)."""
        if self.variant in (4, 5) and int(self.horizons.max()) > self.K:
            raise ValueError(f"stored rows (variant {self.variant}) need ckpt_every >= every "
                             f"horizon ({self.K} < {int(self.horizons.max())})")

    def check(self):
        """Raise if a launch reported a contract violation (reads the device
        error word: synchronises)."""
        e = int(self.err.item())
        if e:
            why = []
            if e & nat.ROLLOUT_ERR_HORIZON:
                why.append("a robot's horizon exceeds the stored-rows scratch (ckpt_every)")
            if e & nat.ROLLOUT_ERR_PACKING:
                why.append("robots packed into one workgroup do not share a design")
            raise nat.NativeError("rollout: " + "; ".join(why or [f"error bits {e:#x}"]))

    def wave_eligible(self):
        """The wave-per-robot kernel covers FP32, hidden 64, inputs <= 192, outputs <= 8."""
        return self.dtype == torch.float32 and self.H == 64 and self.max_i <= 192 and self.max_o <= 8

    def packed4(self):
        rd = self.robot_design
        n = len(rd) - len(rd) % 4
        groups_ok = np.all(rd[:n].reshape(-1, 4) == rd[:n].reshape(-1, 4)[:, :1])
        tail_ok = len(rd) % 4 == 0 or np.all(rd[n:] == rd[n])
        return bool(groups_ok and tail_ok)

    def forward(self, max_steps=0):
        """Advance every robot from its clock by up to max_steps (0: to the end)."""
        self._check_contract()
        D = self._desc(max_steps)
        nat.call("dl_ms_rollout_forward", ctypes.byref(D), nat.stream())
        step = self.horizons - self.table["t_fwd"] if max_steps <= 0 else \
            np.minimum(max_steps, self.horizons - self.table["t_fwd"])
        self.table["t_fwd"] += step
        self._upload_table()
        return self.loss

    def backward(self, loss_grad=None, max_steps=0, reduce=True):
        """Adjoint from each robot's backward clock by up to max_steps."""
        if np.any(self.table["t_fwd"] != self.horizons):
            raise RuntimeError("backward needs every robot's forward rollout to be complete")
        if loss_grad is not None:
            self.loss_grad.copy_(torch.as_tensor(loss_grad, dtype=self.dtype).reshape(-1))
        self._check_contract()
        D = self._desc(max_steps)
        nat.call("dl_ms_rollout_backward", ctypes.byref(D), nat.stream())
        step = self.table["t_bwd"] if max_steps <= 0 else np.minimum(max_steps, self.table["t_bwd"])
        self.table["t_bwd"] -= step
        self._upload_table()
        if reduce:
            self.reduce_grads()

    # ------------------------------------------------ whole training iterations
    def prepare_iteration(self, loss_grad=None):
        """Set up full-horizon training iterations (forward to every robot's
        horizon, adjoint back to 0): the initial state, the reset clock table
        (the forward reads only t_fwd, the adjoint only t_bwd) and the loss
        adjoint stay on the device, so ``run_iteration`` issues device work
        only and can be captured in a HIP graph."""
        self._check_contract()
        self.reset()
        tab = self.table.copy()
        tab["t_fwd"] = 0
        tab["t_bwd"] = self.horizons
        self.t_table_iter = torch.from_numpy(tab.view(np.uint8).copy()).to(self.device)
        if loss_grad is not None:
            self.loss_grad.copy_(torch.as_tensor(loss_grad, dtype=self.dtype).reshape(-1))

    def run_iteration(self):
        """Reset state, forward rollout, adjoint and per-design gradient
        reduction -- launches and device copies only."""
        self.state.copy_(self._state0)
        self.loss.zero_()
        self.adjoint.zero_()
        self.wgrad.zero_()
        if self.traj is not None:
            self.traj.zero_()
        D = self._desc()
        D.robots = nat.ptr(self.t_table_iter)
        nat.call("dl_ms_rollout_forward", ctypes.byref(D), nat.stream())
        nat.call("dl_ms_rollout_backward", ctypes.byref(D), nat.stream())
        self.reduce_grads()

    def reduce_grads(self):
        nat.call("dl_wgrad_reduce", nat.dtype_code(self.dtype), self.R, nat.ptr(self.t_lptr),
                 nat.ptr(self.t_list), len(self.designs), nat.ptr(self.t_roff),
                 nat.ptr(self.t_doff), nat.ptr(self.t_dsize), nat.ptr(self.wgrad),
                 nat.ptr(self.grad_out), nat.stream())

    def trajectory(self, r):
        """(x [Th+1, N, 3], center [Th+1, 3]) of robot r (store_traj=True)."""
        N = self.designs[int(self.robot_design[r])].n_vert
        Th = int(self.horizons[r])
        o = int(self.table[r]["traj"])
        rows = self.traj[o:o + (Th + 1) * (3 * N + 3)].view(Th + 1, 3 * N + 3)
        return rows[:, :3 * N].view(Th + 1, 3, N).transpose(1, 2), rows[:, 3 * N:]


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *flat):
        for di in range(len(batch.designs)):
            batch.set_weights(di, *flat[4 * di: 4 * di + 4])
        batch.reset()
        batch.forward()
        ctx.batch = batch
        return batch.loss.clone()

    @staticmethod
    def backward(ctx, g):
        b = ctx.batch
        b.backward(loss_grad=g.contiguous())
        out = [None]
        for di in range(len(b.designs)):
            out += [t.clone() for t in b.weight_grads(di)]
        return tuple(out)


def rollout_loss(batch, weights):
    """Per-robot losses [R] as a differentiable function of the design weights
    (``weights``: list over designs of [W1, b1, W2, b2])."""
    flat = [t for w in weights for t in w]
    return _RolloutFn.apply(batch, *flat)


class GraphedIteration:
    """A device-only callable (e.g. ``MassSpringBatch.run_iteration``)
    captured once in a HIP graph (torch's graph API over the HIP runtime) and
    replayed: one launch of the whole recorded sequence instead of one host
    call per kernel.  The callable is run ``warmup`` times eagerly first on a
    side stream, so that one-time setup (LDS limits, allocator growth)
    happens outside the capture."""

    def __init__(self, fn, warmup=1):
        self.fn = fn
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()

    def replay(self):
        self.graph.replay()
