"""Drop-in rollout driver (``DiffPhyTrainer.simulate``,
``multitask/diffphy_trainer.py:215-302``), training path only.

``simulate`` runs the reference's per-step loop -- ``pre_advance(t)``,
``nn_input(t)``, ``nn.forward(t)`` (or the zero-order hold ``diff_copy``),
``advance(t)`` for ``t < steps - 1`` -- inside a ``Tape`` so the adjoint of
every launched kernel replays in reverse on exit.  ``simulate_task`` runs the
quadruped workload's loop (``steps`` advances, windowed COM loss).
"""

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

    def simulate(self, steps, train=True, iter=0, max_speed=0.08, max_height=0.1,
                 loss_enable=None, pool=None):
        """diffphy_trainer.py:219-256 (training branch)."""
        env, nn = self.taichi_env, self.nn
        if not train:
            raise NotImplementedError("evaluation rollouts are not part of this path")
        loss_enable = self.loss_enable if loss_enable is None else set(loss_enable)
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

    def run_step(self):
        """diffphy_trainer.py:277-302 without logging: one rollout, TNS, Adam."""
        self.simulate(self.taichi_env.max_steps, iter=self.iter, max_height=self.max_height,
                      max_speed=self.max_speed, loss_enable=self.loss_enable)
        self.total_norm_sqr = self.nn.get_TNS()
        self.nn.gradient_update(self.iter)
        self.losses_list.append(self.taichi_env.loss[None])
        self.iter += 1
