"""Drop-in ``SolverMassSpring`` (``multitask/solver_mass_spring.py:6-202``).

Same constructor, methods and fields as the reference class; every method
launches HIP kernels of ``libdiffloco.so`` on the current torch stream and,
inside a ``Tape``, records the launch of its hand-written adjoint.

Fields (user layout ``[t, k, i, d]``, storage structure-of-arrays):
``x, v`` [T, B, N, 3]; ``center`` (and ``vcom``) [T, B, 3]; ``actuation``
[T, B, O]; ``height, upper_height, rotation`` [T, B].
"""

import ctypes

import numpy as np
import torch

from . import _native as nat
from .fields import Field, record


def _physics(cfg):
    s = cfg["simulator"]
    return nat.MSPhysics(dt=cfg["process"]["dt"], gravity=s["gravity"], ground=s["ground_height"],
                         drag=s["drag_damping"], dashpot=s["dashpot_damping"],
                         friction=s["friction"])


def dtype_of(cfg):
    return torch.float64 if cfg.get("precision", "f64") == "f64" else torch.float32


class DeviceTopology:
    """Spring table and CSR maps of one design on the device (``dl_ms_topology``)."""

    def __init__(self, design, dtype, device, stiffness_scale=0.1):
        d = design
        vptr, vidx = d.vertex_csr()
        aptr, aidx = d.actuator_csr()
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=device)
        self.spring_a = i32(d.spring_a)
        self.spring_b = i32(d.spring_b)
        self.spring_act = i32(d.act_index)
        self.L0 = torch.as_tensor(d.L0, dtype=dtype, device=device)
        self.k = torch.as_tensor(d.stiffness * stiffness_scale, dtype=dtype, device=device)
        self.coeff = torch.as_tensor(d.coeff, dtype=dtype, device=device)
        self.vptr, self.vidx = i32(vptr), i32(vidx)
        self.aptr = i32(aptr)
        self.aidx = i32(aidx) if len(aidx) else torch.zeros(1, dtype=torch.int32, device=device)
        self.c = nat.MSTopology(
            n_vert=d.n_vert, n_spring=d.n_spring, n_act=d.n_act,
            spring_a=nat.ptr(self.spring_a), spring_b=nat.ptr(self.spring_b),
            spring_act=nat.ptr(self.spring_act), spring_L0=nat.ptr(self.L0),
            spring_k=nat.ptr(self.k), spring_coeff=nat.ptr(self.coeff),
            vert_ptr=nat.ptr(self.vptr), vert_idx=nat.ptr(self.vidx),
            act_ptr=nat.ptr(self.aptr), act_idx=nat.ptr(self.aidx))


class SolverMassSpring:
    def __init__(self, config, device="cuda"):
        cfg = config.get_config()
        self.config = config
        self.max_steps = cfg["process"]["max_steps"]
        self.dt = cfg["process"]["dt"]
        self.ground_height = cfg["simulator"]["ground_height"]
        self.gravity = cfg["simulator"]["gravity"]
        self.drag_damping = cfg["simulator"]["drag_damping"]
        self.dashpot_damping = cfg["simulator"]["dashpot_damping"]
        self.batch_size = cfg["nn"]["batch_size"]
        self.n_objects = cfg["robot"]["n_objects"]
        self.n_springs = cfg["robot"]["n_springs"]
        self.n_actuators = cfg["robot"]["n_actuators"]
        self.robot_id = cfg["robot"]["robot_id"]
        self.springs = cfg["robot"]["springs"]
        self.design = cfg["robot"]["design_obj"]
        self.dim = cfg["robot"]["dim"]
        self.jump_period = cfg["process"]["jump_period"]
        self.stiffness_scale = cfg["simulator"]["stiffness_scale"]
        self.dtype = dtype_of(cfg)
        self.device = torch.device(device)
        self.phys = _physics(cfg)
        self.act_list = []

        T, B, N, O = self.max_steps + 1, self.batch_size, self.n_objects, self.n_actuators
        mk = lambda shape, soa=False, name="": Field(shape, self.dtype, self.device, soa=soa, name=name)
        self.x = mk((T, B, N, 3), soa=True, name="x")
        self.v = mk((T, B, N, 3), soa=True, name="v")
        self.center = mk((T, B, 3), name="center")
        self.vcom = mk((T, B, 3), name="vcom")
        self.actuation = mk((T, B, O), name="actuation")
        self.height = mk((T, B), name="height")
        self.upper_height = mk((T, B), name="upper_height")
        self.rotation = mk((T, B), name="rotation")
        self._topo = None

    # ------------------------------------------------------------ topology
    def initialize_robot(self):
        """Upload the spring table (solver_mass_spring.py:58-67): stiffness is
        stored scaled by the solver's stiffness scale (reference: /10)."""
        self.topology = DeviceTopology(self.design, self.dtype, self.device, self.stiffness_scale)
        self._topo = self.topology.c
        self.act_list = [i for i in range(self.design.n_spring) if self.design.coeff[i] > 0]

    def clear_states(self, steps):
        """solver_mass_spring.py:69-84 (impulses are never stored, nothing
        accumulates across steps; rotation is reset)."""
        self.rotation.data[:steps].zero_()

    # ------------------------------------------------------------ kernels
    @property
    def _dc(self):
        return nat.dtype_code(self.dtype)

    def compute_center(self, t):
        B, N = self.batch_size, self.n_objects
        nat.call("dl_ms_center", self._dc, B, N, nat.ptr(self.x.at(t)), nat.ptr(self.v.at(t)),
                 nat.ptr(self.center.at(t)), nat.ptr(self.vcom.at(t)), nat.stream())

        def adj():
            nat.call("dl_ms_center_grad", self._dc, B, N, nat.ptr(self.center.grad.at(t)),
                     nat.ptr(self.vcom.grad.at(t)), nat.ptr(self.x.grad.at(t)),
                     nat.ptr(self.v.grad.at(t)), nat.stream())
        record(adj)

    def compute_height(self, t):
        B, N = self.batch_size, self.n_objects
        prev = self.height.at(t - 1) if t > 0 else None
        nat.call("dl_ms_height", self._dc, B, N, t, self.jump_period, nat.ptr(self.x.at(t)),
                 nat.ptr(prev), nat.ptr(self.height.at(t)), nat.ptr(self.upper_height.at(t)),
                 nat.stream())

        def adj():
            pg = self.height.grad.at(t - 1) if t > 0 else None
            nat.call("dl_ms_height_grad", self._dc, B, N, t, self.jump_period,
                     nat.ptr(self.x.at(t)), nat.ptr(prev), nat.ptr(self.height.grad.at(t)),
                     nat.ptr(self.upper_height.grad.at(t)), nat.ptr(self.x.grad.at(t)),
                     nat.ptr(pg), nat.stream())
        record(adj)

    def compute_rotation(self, t):
        B, N = self.batch_size, self.n_objects
        nat.call("dl_ms_rotation", self._dc, B, N, self.robot_id, nat.ptr(self.x.at(t)),
                 nat.ptr(self.x.at(0)), nat.ptr(self.center.at(0)),
                 nat.ptr(self.rotation.at(t)), nat.stream())

        def adj():
            nat.call("dl_ms_rotation_grad", self._dc, B, N, self.robot_id, nat.ptr(self.x.at(t)),
                     nat.ptr(self.x.at(0)), nat.ptr(self.center.at(0)),
                     nat.ptr(self.rotation.grad.at(t)), nat.ptr(self.x.grad.at(t)), nat.stream())
        record(adj)

    def pre_advance(self, t):
        """solver_mass_spring.py:194-198."""
        self.compute_center(t)
        self.compute_height(t)
        if self.dim == 3:
            self.compute_rotation(t)

    def advance(self, t):
        """apply_spring_force(t) + advance_toi(t+1) (solver_mass_spring.py:200-202)."""
        if self._topo is None:
            raise RuntimeError("initialize_robot() must run before advance()")
        B = self.batch_size
        nat.call("dl_ms_advance", self._dc, ctypes.byref(self._topo), ctypes.byref(self.phys), B,
                 nat.ptr(self.x.at(t)), nat.ptr(self.v.at(t)), nat.ptr(self.actuation.at(t)),
                 nat.ptr(self.x.at(t + 1)), nat.ptr(self.v.at(t + 1)), nat.stream())

        def adj():
            nat.call("dl_ms_advance_grad", self._dc, ctypes.byref(self._topo),
                     ctypes.byref(self.phys), B, nat.ptr(self.x.at(t)), nat.ptr(self.v.at(t)),
                     nat.ptr(self.actuation.at(t)), nat.ptr(self.x.grad.at(t + 1)),
                     nat.ptr(self.v.grad.at(t + 1)), nat.ptr(self.x.grad.at(t)),
                     nat.ptr(self.v.grad.at(t)), nat.ptr(self.actuation.grad.at(t)),
                     nat.stream())
        record(adj)
