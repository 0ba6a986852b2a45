"""Drop-in ``TaichiEnv`` (``multitask/taichi_env.py:20-516``), training path.

Holds the loss fields, the sensor vector ``input_state``, the targets and the
solver; ``nn_input``, the losses and their adjoints run as HIP kernels.
Random targets come from a seeded device generator instead of ``ti.random``.
The GUI/visualiser and evaluation-only target scripts are not part of this
path.
"""

import collections
import ctypes
import itertools

import numpy as np
import torch

from . import _native as nat
from .fields import Field, record
from .solver_mass_spring import SolverMassSpring, dtype_of


def sensor_layout(cfg):
    nn = cfg["nn"]
    comps = list(nn.get("tv_comps", [0, 2])) + [0, 0, 0]
    return nat.SensorLayoutC(
        n_sin=nn["n_sin_waves"], run_period=cfg["process"]["run_period"], vcom=nn.get("vcom", 0),
        vert_pad=nn.get("vert_pad", 3), n_tv=len(nn.get("tv_comps", [0, 2])),
        tv_comp=(ctypes.c_int32 * 3)(*comps[:3]), dup_v=nn["duplicate_v"],
        dup_h=nn["duplicate_h"], dup_c=nn["duplicate_c"],
        offset_scale=nn.get("offset_scale", 0.2), max_speed=cfg["process"]["max_speed"])


class TaichiEnv:
    LOSS_NAMES = ("loss_velocity", "loss_height", "loss_pose", "loss_rotation", "loss_weight",
                  "loss_actuation", "loss_crawl", "loss_task")

    def __init__(self, config, device="cuda"):
        self.config_ = config
        self.config = cfg = config.get_config()
        self.dim = cfg["robot"]["dim"]
        self.max_steps = cfg["process"]["max_steps"]
        self.batch_size = cfg["nn"]["batch_size"]
        self.n_hidden = cfg["nn"]["n_hidden"]
        self.robot_id = cfg["robot"]["robot_id"]
        self.n_input_states = cfg["nn"]["n_input_states"]
        self.n_springs = cfg["robot"]["n_springs"]
        self.n_actuators = cfg["robot"]["n_actuators"]
        self.objects = cfg["robot"]["objects"]
        self.turn_period = cfg["process"]["turn_period"]
        self.run_period = cfg["process"]["run_period"]
        self.jump_period = cfg["process"]["jump_period"]
        self.n_sin_waves = cfg["nn"]["n_sin_waves"]
        self.dt = cfg["process"]["dt"]
        self.task = list(cfg["train"]["task"])
        self.window = cfg["train"].get("window", 500)
        self.dtype = dtype_of(cfg)
        self.device = torch.device(device)
        self.layout = sensor_layout(cfg)

        T, B = self.max_steps + 1, self.batch_size
        mk = lambda shape, name, soa=False: Field(shape, self.dtype, self.device, soa=soa, name=name)
        self.loss = mk((), "loss")
        self.loss_dict = {n: mk((), n) for n in self.LOSS_NAMES}
        self.losses = list(self.loss_dict.values())
        self.loss_batch = mk((B,), "loss_batch")
        self.loss_dict_batch = {n: mk((B,), n + "_batch") for n in self.LOSS_NAMES}
        self.initial_objects = mk((self.n_objects_(), 3), "initial_objects")
        self.initial_center = mk((3,), "initial_center")
        self.input_state = mk((T, B, self.n_input_states), "input_state")
        self.target_v = mk((T, B, 3), "target_v")
        self.target_h = mk((T, B), "target_h")
        self.target_c = mk((T, B), "target_c")
        self.current_max_height = mk((), "current_max_height")

        # validation matrix: every combination of the validation targets of
        # the trained loss types (taichi_env.py:54-72)
        self.validate_targets = {k: list(v) for k, v in cfg.get("validation", {}).items()
                                 if k in self.task}
        self.validate_targets_values = collections.defaultdict(list)
        keys = list(self.validate_targets)
        for element in itertools.product(*[self.validate_targets[k] for k in keys]):
            for i, name in enumerate(keys):
                self.validate_targets_values[name].append(element[i])

        self.solver = SolverMassSpring(config, device=device)
        s = self.solver
        self.x, self.v, self.center, self.height = s.x, s.v, s.center, s.height
        self.upper_height, self.rotation, self.actuation = s.upper_height, s.rotation, s.actuation
        self._work = torch.zeros(B, dtype=self.dtype, device=self.device)
        self._work2 = torch.zeros(2 * B, dtype=self.dtype, device=self.device)
        self._gen = torch.Generator(device=self.device)

    def n_objects_(self):
        return self.config["robot"]["n_objects"]

    @property
    def _dc(self):
        return nat.dtype_code(self.dtype)

    # ------------------------------------------------------------ sensors
    def nn_input(self, t, offset, max_speed, max_height):
        """taichi_env.py:149-183 (max_speed is taken from the layout)."""
        B, N = self.batch_size, self.n_objects_()
        s = self.solver
        nat.call("dl_nn_input", self._dc, ctypes.byref(self.layout), B, N, t, offset,
                 nat.ptr(s.x.at(t)), nat.ptr(s.v.at(t)), nat.ptr(s.center.at(t)),
                 nat.ptr(self.target_v.at(t)), nat.ptr(self.target_h.at(t)),
                 nat.ptr(self.target_c.at(t)), nat.ptr(self.input_state.at(t)), nat.stream())

        def adj():
            nat.call("dl_nn_input_grad", self._dc, ctypes.byref(self.layout), B, N,
                     nat.ptr(self.input_state.grad.at(t)), nat.ptr(s.x.grad.at(t)),
                     nat.ptr(s.v.grad.at(t)), nat.ptr(s.center.grad.at(t)), nat.stream())
        record(adj)

    # ------------------------------------------------------------ targets
    def seed(self, seed):
        self._gen.manual_seed(int(seed))

    def initialize_train(self, iter, steps, max_velocity, max_height, pool=None):
        """taichi_env.py:385-447 (3D branch); ``pool`` (uniforms) may be given
        to reproduce a specific draw."""
        B = self.batch_size
        times = steps // self.turn_period + 1
        if pool is None:
            pool = torch.rand(B * times * 3, generator=self._gen, device=self.device,
                              dtype=torch.float64)
        pool = torch.as_tensor(pool, device=self.device, dtype=torch.float64)
        t = torch.arange(steps, device=self.device)
        k = torch.arange(B, device=self.device)
        q = ((t[:, None] // self.turn_period) * B + k[None, :]) * 3
        p0, r = pool[q], pool[q + 1] * 2.0 - 1.0
        tv = torch.zeros(steps, B, 3, dtype=torch.float64, device=self.device)
        th = torch.zeros(steps, B, dtype=torch.float64, device=self.device)
        th = torch.where(p0 < 0.25, -1.0, torch.where(p0 < 0.5, 1.0, 0.0)).to(torch.float64)
        in_z = (p0 >= 0.5) & (p0 < 0.75)
        in_x = p0 >= 0.75
        tv[..., 2] = torch.where(in_z, r * max_velocity, 0.0)
        tv[..., 0] = torch.where(in_x, r * max_velocity, 0.0)
        self.target_v.data[:steps] = tv.to(self.dtype)
        self.target_h.data[:steps] = th.to(self.dtype)
        self.target_c.data[:steps] = 0

    def initialize_task(self, steps, tv, th):
        """Per-task constant targets of the quadruped workload: tv [B,3], th [B]."""
        tv = torch.as_tensor(tv, dtype=self.dtype, device=self.device)
        th = torch.as_tensor(th, dtype=self.dtype, device=self.device)
        self.target_v.data[: steps + 1] = tv[None]
        self.target_h.data[: steps + 1] = th[None]
        self.target_c.data[: steps + 1] = 0

    def initialize_validate(self, steps, output_v, output_h, output_c):
        """taichi_env.py:378-383: per-robot constant targets (x velocity,
        height, crawl) for every step."""
        f = lambda a: torch.as_tensor(np.asarray(a, np.float64), dtype=self.dtype, device=self.device)
        self.target_v.data[:steps, :, 0] = f(output_v)[None]
        self.target_h.data[:steps] = f(output_h)[None]
        self.target_c.data[:steps] = f(output_c)[None]

    def initialize_interactive(self, steps, output_v, output_h, output_c):
        """taichi_env.py:333-338: one x velocity, height and crawl target for all."""
        self.target_v.data[:steps, :, 0] = output_v
        self.target_h.data[:steps] = output_h
        self.target_c.data[:steps] = output_c

    def initialize_script(self, steps, x0, y0, x1, y1, x2, y2, x3, y3):
        """taichi_env.py:340-376: velocity targets (x, z) switching every 1000
        steps through the four given pairs, then height -1 (crouch) for steps
        4000-4999 and +1 (jump) after."""
        t = torch.arange(steps, device=self.device)
        seg = torch.clamp(t // 1000, max=5)
        vx = torch.tensor([x0, x1, x2, x3, 0.0, 0.0], dtype=self.dtype, device=self.device)[seg]
        vz = torch.tensor([y0, y1, y2, y3, 0.0, 0.0], dtype=self.dtype, device=self.device)[seg]
        h = torch.tensor([0.0, 0.0, 0.0, 0.0, -1.0, 1.0], dtype=self.dtype, device=self.device)[seg]
        self.target_v.data[:steps, :, 0] = vx[:, None]
        self.target_v.data[:steps, :, 2] = vz[:, None]
        self.target_h.data[:steps] = h[:, None]

    # ------------------------------------------------------------ losses
    def compute_loss_velocity(self, steps, written):
        B, s = self.batch_size, self.solver
        Tn = s.center.data.shape[0]
        nat.call("dl_loss_velocity", self._dc, B, Tn, written, steps, self.run_period,
                 self.turn_period, nat.ptr(s.center.data), nat.ptr(self.target_v.data),
                 nat.ptr(self.loss_dict["loss_velocity"].data),
                 nat.ptr(self.loss_dict_batch["loss_velocity"].data), nat.stream())

        def adj():
            g = self.loss_dict["loss_velocity"].grad[None]
            nat.call("dl_loss_velocity_grad", self._dc, B, Tn, written, steps, self.run_period,
                     self.turn_period, nat.ptr(s.center.data), nat.ptr(self.target_v.data),
                     float(g), nat.ptr(s.center.grad.data), nat.stream())
        record(adj)

    def compute_loss_rotation(self, steps, written):
        B, N, s = self.batch_size, self.n_objects_(), self.solver
        Tn = s.center.data.shape[0]
        nat.call("dl_loss_rotation", self._dc, B, N, Tn, written, steps, self.run_period,
                 self.turn_period, nat.ptr(s.x.data), nat.ptr(s.center.data),
                 nat.ptr(self.target_h.data), nat.ptr(self.loss_dict["loss_rotation"].data),
                 nat.ptr(self.loss_dict_batch["loss_rotation"].data), nat.ptr(self._work),
                 nat.stream())

        def adj():
            g = self.loss_dict["loss_rotation"].grad[None]
            nat.call("dl_loss_rotation_grad", self._dc, B, N, Tn, written, steps,
                     self.run_period, self.turn_period, nat.ptr(s.x.data),
                     nat.ptr(s.center.data), nat.ptr(self.target_h.data), float(g),
                     nat.ptr(s.x.grad.data), nat.ptr(s.center.grad.data), nat.stream())
        record(adj)

    def compute_loss_task(self, steps):
        B, s = self.batch_size, self.solver
        nat.call("dl_loss_task", self._dc, B, steps, self.window, nat.ptr(s.center.data),
                 nat.ptr(s.vcom.data), nat.ptr(self.target_v.at(0)), nat.ptr(self.target_h.at(0)),
                 nat.ptr(self.loss_dict["loss_task"].data),
                 nat.ptr(self.loss_dict_batch["loss_task"].data), nat.stream())

        def adj():
            g = self.loss_dict["loss_task"].grad[None]
            nat.call("dl_loss_task_grad", self._dc, B, steps, self.window,
                     nat.ptr(s.center.data), nat.ptr(s.vcom.data), nat.ptr(self.target_v.at(0)),
                     nat.ptr(self.target_h.at(0)), float(g), nat.ptr(s.center.grad.data),
                     nat.ptr(s.vcom.grad.data), nat.stream())
        record(adj)

    def compute_loss_height(self, steps, written):
        """taichi_env.py:213-222; also raises current_max_height."""
        B, s = self.batch_size, self.solver
        Tn = s.height.data.shape[0]
        nat.call("dl_loss_height", self._dc, B, Tn, written, steps, self.jump_period,
                 nat.ptr(s.height.data), nat.ptr(self.target_h.data),
                 nat.ptr(self.loss_dict["loss_height"].data),
                 nat.ptr(self.loss_dict_batch["loss_height"].data),
                 nat.ptr(self.current_max_height.data), nat.ptr(self._work2), nat.stream())

        def adj():
            g = self.loss_dict["loss_height"].grad[None]
            nat.call("dl_loss_height_grad", self._dc, B, Tn, written, steps, self.jump_period,
                     nat.ptr(s.height.data), nat.ptr(self.target_h.data), float(g),
                     nat.ptr(s.height.grad.data), nat.stream())
        record(adj)

    def compute_loss_pose(self, steps, written):
        """taichi_env.py:224-233 (x rows up to ``written`` hold states)."""
        B, N, s = self.batch_size, self.n_objects_(), self.solver
        Tn = s.x.data.shape[0]
        Tx = min(written + 1, Tn)
        nat.call("dl_loss_pose", self._dc, B, N, Tn, Tx, steps, self.jump_period,
                 nat.ptr(s.x.data), nat.ptr(self.initial_objects.data),
                 nat.ptr(self.loss_dict["loss_pose"].data),
                 nat.ptr(self.loss_dict_batch["loss_pose"].data), nat.ptr(self._work), nat.stream())

        def adj():
            g = self.loss_dict["loss_pose"].grad[None]
            nat.call("dl_loss_pose_grad", self._dc, B, N, Tn, Tx, steps, self.jump_period,
                     nat.ptr(s.x.data), nat.ptr(self.initial_objects.data), float(g),
                     nat.ptr(s.x.grad.data), nat.stream())
        record(adj)

    def compute_loss_actuation(self, steps, written):
        """taichi_env.py:276-285 over the network outputs (the reference's
        robots have one output per spring)."""
        B, s = self.batch_size, self.solver
        O = s.actuation.data.shape[-1]
        Tn = s.actuation.data.shape[0]
        nat.call("dl_loss_actuation", self._dc, B, O, Tn, written, steps,
                 nat.ptr(s.actuation.data), nat.ptr(self.target_v.data),
                 nat.ptr(self.target_h.data), nat.ptr(self.loss_dict["loss_actuation"].data),
                 nat.ptr(self.loss_dict_batch["loss_actuation"].data), nat.ptr(self._work),
                 nat.stream())

        def adj():
            g = self.loss_dict["loss_actuation"].grad[None]
            nat.call("dl_loss_actuation_grad", self._dc, B, O, Tn, written, steps,
                     nat.ptr(s.actuation.data), nat.ptr(self.target_v.data),
                     nat.ptr(self.target_h.data), float(g), nat.ptr(s.actuation.grad.data),
                     nat.stream())
        record(adj)

    def compute_loss_crawl(self, steps, written):
        """taichi_env.py:287-293."""
        B, s = self.batch_size, self.solver
        Tn = s.upper_height.data.shape[0]
        nat.call("dl_loss_crawl", self._dc, B, Tn, written, steps, nat.ptr(s.upper_height.data),
                 nat.ptr(self.target_c.data), nat.ptr(self.loss_dict["loss_crawl"].data),
                 nat.ptr(self.loss_dict_batch["loss_crawl"].data), nat.ptr(self._work),
                 nat.stream())

        def adj():
            g = self.loss_dict["loss_crawl"].grad[None]
            nat.call("dl_loss_crawl_grad", self._dc, B, Tn, written, steps,
                     nat.ptr(s.upper_height.data), nat.ptr(self.target_c.data), float(g),
                     nat.ptr(s.upper_height.grad.data), nat.stream())
        record(adj)

    def compute_loss_final(self, l):
        """loss += l (taichi_env.py:295-297)."""
        self.loss.data += l.data

        def adj():
            l.grad.data += self.loss.grad.data
        record(adj)

    def get_loss(self, steps, loss_enable, written=None):
        """taichi_env.py:304-331: the velocity, height, pose, actuation,
        rotation and crawl losses (in the reference's order) and the
        quadruped task loss.
        ``written``: number of leading center/x rows the rollout wrote
        (rows at and after it read as zero, as the unwritten field entries do)."""
        if written is None:
            written = steps
        if "velocity" in loss_enable:
            self.compute_loss_velocity(steps, written)
        if "height" in loss_enable:
            self.compute_loss_height(steps, written)
        if "pose" in loss_enable:
            self.compute_loss_pose(steps, written)
        if "actuation" in loss_enable:
            self.compute_loss_actuation(steps, written)
        if "rotation" in loss_enable:
            self.compute_loss_rotation(steps, written)
        if "crawl" in loss_enable:
            self.compute_loss_crawl(steps, written)
        if "task" in loss_enable:
            self.compute_loss_task(steps)
        for l in self.losses:
            self.compute_loss_final(l)
        self.loss_batch.data.zero_()
        for l in self.loss_dict_batch.values():
            self.loss_batch.data += l.data

    # ------------------------------------------------------------ robot
    def get_center(self):
        self.initial_center.data.copy_(self.initial_objects.data.mean(0))

    def reset_robot(self, start, step, times):
        """taichi_env.py:487-490: x[0, k*step+start] = initial objects."""
        ks = [k * step + start for k in range(times)]
        self.solver.x.view[0, ks] = self.initial_objects.data[None].expand(len(ks), -1, -1)

    def setup_robot(self, shift=(0.4, 0.0, 0.0)):
        """taichi_env.py:497-504: objects shifted +0.4 in x."""
        obj = torch.tensor(self.objects, dtype=torch.float64)
        obj = obj + torch.tensor(shift, dtype=torch.float64)
        self.initial_objects.data.copy_(obj)
        self.get_center()
        self.reset_robot(0, 1, self.batch_size)
        self.solver.initialize_robot()

    def copy_robot(self, steps):
        """taichi_env.py:506-510."""
        s = self.solver
        s.x.data[0] = s.x.data[steps]
        s.v.data[0] = s.v.data[steps]
