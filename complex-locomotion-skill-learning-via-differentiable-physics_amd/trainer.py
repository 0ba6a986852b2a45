"""Drop-in rollout driver (``DiffPhyTrainer.simulate``,
``multitask/diffphy_trainer.py:215-302``), training path only.

``simulate`` runs the reference's per-step loop -- ``pre_advance(t)``,
``nn_input(t)``, ``nn.forward(t)`` (or the zero-order hold ``diff_copy``),
``advance(t)`` for ``t < steps - 1`` -- inside a ``Tape`` so the adjoint of
every launched kernel replays in reverse on exit.  ``simulate_task`` runs the
quadruped workload's loop (``steps`` advances, windowed COM loss).
"""

import numpy as np
import torch

from . import _native as nat
from .fields import Tape, record
from .nn import Model
from .taichi_env import TaichiEnv


class DiffPhyTrainer:
    def __init__(self, config, device="cuda"):
        self.config = config
        cfg = config.get_config()
        self.taichi_env = env = TaichiEnv(config, device=device)
        self.nn = Model(config, env.max_steps, env.batch_size, env.n_input_states,
                        env.n_actuators, env.input_state, env.actuation, env.n_hidden,
                        method=cfg["nn"].get("optimizer", "adam"))
        self.max_height = cfg["process"]["max_height"]
        self.max_speed = cfg["process"]["max_speed"]
        self.control_length = cfg["robot"].get("control_length", 1)
        self.loss_enable = set(env.task)
        # restarts from the previous rollout's final state after change_iter
        # iterations (diffphy_trainer.py:186-193, 270-281)
        self.change_iter = 5000
        self.reset_step = 2
        self.max_reset_step = cfg["nn"].get("max_reset_step", 4)
        self.iter = 0
        self.losses_list = []
        self.total_norm_sqr = 0.0

    def diff_copy(self, t):
        """actuation[t] = actuation[t-1] (diffphy_trainer.py:215-218)."""
        a = self.taichi_env.actuation
        a.data[t] = a.data[t - 1]

        def adj():
            a.grad.data[t - 1] += a.grad.data[t]
            a.grad.data[t] = 0
        record(adj)

    def _clear_losses(self):
        env = self.taichi_env
        env.loss.fill(0)
        for l in env.losses:
            l.fill(0)
        for l in env.loss_dict_batch.values():
            l.fill(0)

    def simulate(self, steps, output_v=None, output_h=None, output_c=None, train=True, iter=0,
                 max_speed=0.08, max_height=0.1, loss_enable=None, pool=None):
        """diffphy_trainer.py:219-268.  Training: T-1 advances under the tape.
        Evaluation (train=False): validation targets (2D) or the scripted
        target schedule (3D), steps+1 advances without a tape, losses of
        every row."""
        env, nn = self.taichi_env, self.nn
        loss_enable = self.loss_enable if loss_enable is None else set(loss_enable)
        if not train:
            if steps + 1 > env.max_steps:
                raise ValueError(f"evaluation writes state {steps + 1}; the fields hold "
                                 f"{env.max_steps + 1} rows")
            env.solver.clear_states(steps)
            nn.clear()
            if env.dim == 2:
                if output_c is None:
                    output_c = np.zeros(env.batch_size)
                env.initialize_validate(steps, output_v, output_h, output_c)
            else:
                env.initialize_script(steps + 1, -max_speed, 0, 0, max_speed, max_speed, 0, 0,
                                      -max_speed)
            self._clear_losses()
            with torch.no_grad():
                for t in range(steps + 1):
                    env.solver.pre_advance(t)
                    env.nn_input(t, 0, max_speed, max_height)
                    if t % self.control_length == 0:
                        nn.forward(t)
                    else:
                        self.diff_copy(t)
                    env.solver.advance(t)
                env.get_loss(steps, loss_enable, written=steps + 1)
            return
        env.solver.clear_states(steps)
        nn.clear()
        env.initialize_train(iter, steps, max_speed, max_height, pool=pool)
        self._clear_losses()
        with Tape(env.loss):
            for t in range(steps - 1):
                env.solver.pre_advance(t)
                env.nn_input(t, 0, max_speed, max_height)
                if t % self.control_length == 0:
                    nn.forward(t)
                else:
                    self.diff_copy(t)
                env.solver.advance(t)
            # center[steps-1] is never computed (the loop stops at steps-2)
            env.solver.center.data[steps - 1:].zero_()
            env.get_loss(steps, loss_enable, written=steps - 1)

    def simulate_task(self, steps, tv, th):
        """Quadruped workload: ``steps`` advances, loss over the last window."""
        env, nn = self.taichi_env, self.nn
        env.solver.clear_states(steps + 1)
        nn.clear()
        env.initialize_task(steps, tv, th)
        self._clear_losses()
        with Tape(env.loss):
            for t in range(steps):
                env.solver.compute_center(t)
                env.nn_input(t, 0, self.max_speed, self.max_height)
                nn.forward(t)
                env.solver.advance(t)
            env.solver.compute_center(steps)
            env.get_loss(steps, {"task"})

    # ------------------------------------------------------------ evaluation
    def evaluate_weights(self, weights=None, loss_enable=None, steps=None):
        """One model of ``_evaluate`` (diffphy_trainer.py:319-435): set up the
        robot, load the weights (a ``[W1, W2, b1, b2]`` pickle path or
        tensors), run the evaluation rollout and return the per-robot losses,
        keyed like the reference's metrics: ``task_loss{suffix}`` and
        ``{name}_loss{suffix}`` with suffix ``_{target}_{value}...`` of the
        validation matrix (or ``_{k}`` without one)."""
        env = self.taichi_env
        loss_enable = set(env.task) if loss_enable is None else set(loss_enable)
        steps = env.max_steps - 1 if steps is None else steps
        env.setup_robot()
        if isinstance(weights, str):
            self.nn.load_weights(weights)
        elif weights is not None:
            self.nn.set_weights(*weights)
        vv = env.validate_targets_values
        keys = list(vv)
        B = env.batch_size
        if keys and len(vv[keys[0]]) != B:
            raise ValueError(f"validation matrix has {len(vv[keys[0]])} cases, batch is {B}")
        suffix = ["".join(f"_{k}_{vv[k][i]}" for k in keys) if keys else f"_{i}" for i in range(B)]
        out_v = np.asarray(vv.get("velocity", np.zeros(B)), np.float64)
        out_h = np.asarray(vv.get("height", np.zeros(B)), np.float64)
        out_c = np.asarray(vv.get("crawl", np.zeros(B)), np.float64)
        self.simulate(steps, output_v=out_v, output_h=out_h, output_c=out_c, train=False,
                      max_speed=self.max_speed, max_height=self.max_height,
                      loss_enable=loss_enable)
        lb = env.loss_batch.data.double().cpu().numpy()
        res = {f"task_loss{suffix[k]}": float(lb[k]) for k in range(B)}
        for name in keys:
            v = env.loss_dict_batch[f"loss_{name}"].data.double().cpu().numpy()
            res.update({f"{name}_loss{suffix[k]}": float(v[k]) for k in range(B)})
        return res

    def export_obj(self, folder, steps, every=10, robot=0):
        """Mesh frames of one robot every ``every`` states as OBJ files
        ``{t:06d}.obj`` (vertices and the design's triangles, 1-based), the
        3D branch of the evaluation visualizer (diffphy_trainer.py:336-347)."""
        import os
        os.makedirs(folder, exist_ok=True)
        x = self.taichi_env.x.view[: steps + 1: every, robot].double().cpu().numpy()
        faces = self.config.get_config()["robot"].get("faces", [])
        paths = []
        for i, frame in enumerate(x):
            fn = os.path.join(folder, f"{i:06d}.obj")
            with open(fn, "w") as f:
                f.writelines(f"v {a:.6f} {b:.6f} {c:.6f}\n" for a, b, c in frame)
                f.writelines(f"f {p0 + 1} {p1 + 1} {p2 + 1}\n" for p0, p1, p2 in faces)
            paths.append(fn)
        return paths

    def rounded_train(self, steps, iter, reset_step):
        """diffphy_trainer.py:270-275: every robot starts the next rollout from
        the previous rollout's final state, except robots start, start + step,
        ... (start = iter % reset_step) whose positions go back to the
        initial pose (their velocities stay copied, as in the reference).
        The final state of a training rollout is row ``steps - 1``: the
        reference's copy_robot(steps) reads row ``steps``, one past its
        fields and never written by training."""
        env = self.taichi_env
        env.copy_robot(steps - 1)
        start = iter % reset_step
        times = (env.batch_size + reset_step - start - 1) // reset_step
        env.reset_robot(start, reset_step, times)

    def run_step(self, pool=None):
        """diffphy_trainer.py:277-302 without logging: after change_iter
        iterations the restarts of rounded_train (the reset stride growing by
        one every 500 iterations up to max_reset_step), then one rollout, TNS,
        Adam.  ``pool`` optionally fixes the target draw."""
        if self.iter > self.change_iter:
            if self.iter % 500 == 0 and self.reset_step < self.max_reset_step:
                self.reset_step += 1
            self.rounded_train(self.taichi_env.max_steps, self.iter, reset_step=self.reset_step)
        self.simulate(self.taichi_env.max_steps, iter=self.iter, max_height=self.max_height,
                      max_speed=self.max_speed, loss_enable=self.loss_enable, pool=pool)
        self.total_norm_sqr = self.nn.get_TNS()
        self.nn.gradient_update(self.iter)
        self.losses_list.append(self.taichi_env.loss[None])
        self.iter += 1
