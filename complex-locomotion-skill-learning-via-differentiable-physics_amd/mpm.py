"""MLS-MPM path: parameters, particle robots, the drop-in per-step
``SolverMPM`` and the batched checkpointed rollout ``MPMBatch``.

Kernels (``csrc/mpm.hip``, 2D and 3D, FP32/FP64): ``dl_mpm_p2g``,
``dl_mpm_grid``, ``dl_mpm_g2p`` and their adjoints, replacing
``SolverMPM.p2g/grid_op/g2p`` (``multitask/solver_mpm.py:83-165``) and the
Taichi-generated gradients that ``advance_core_grad`` (``:182-190``) calls
after re-running the grid transfer -- the same recompute scheme is used here.
"""

import ctypes
import itertools
import math
from ctypes import Structure, c_double, c_int32

import numpy as np
import torch

from . import _native as nat
from .fields import record


class MPMParamsC(Structure):
    """dl_mpm_params."""
    _fields_ = [("dim", c_int32), ("n_grid", c_int32), ("bound", c_int32), ("act_axis", c_int32),
                ("dt", c_double), ("gravity", c_double), ("p_vol", c_double), ("p_mass", c_double),
                ("mu", c_double), ("la", c_double), ("act_strength", c_double)]


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def reference_params():
    """cfg3d/sim_quad.json + config_sim.py for the MPM solver (2D)."""
    return MPMParamsC(dim=2, n_grid=32, bound=3, act_axis=1, dt=1e-3, gravity=-10.0, p_vol=1.0,
                      p_mass=1.0, mu=40.0, la=40.0, act_strength=4.0)


def quadruped3d_params(n_grid=64):
    """3D soft quadruped: E = 5e3, nu = 0.2, dt = 2e-4, gravity -1.8, active
    stress along x scaled by 5, particle volume (dx/2)^3 at density 1."""
    mu, la = lame(5e3, 0.2)
    vol = (0.5 / n_grid) ** 3
    return MPMParamsC(dim=3, n_grid=n_grid, bound=3, act_axis=0, dt=2e-4, gravity=-1.8,
                      p_vol=vol, p_mass=vol, mu=mu, la=la, act_strength=5.0)


def particles_2d(design):
    """``RobotDesignMPM.build`` (robot_design.py:65-103): n_num^2 particles at
    the sub-square centres of every square of every rectangle."""
    p = design["design_parameter"]
    size, n_num = p["square_size"], p["particle_in_square_edge"]
    off = p.get("offset", {"x": 0, "y": 0})
    dx = size / n_num
    pos, ids = [], []
    d = design["design"]
    for anchor, shape, act in zip(d["anchor"], d["shape"], d["actuator"]):
        for i, j in itertools.product(range(shape[0]), range(shape[1])):
            sx, sy = anchor[0] + i, anchor[1] + j
            for a, b in itertools.product(range(n_num), range(n_num)):
                pos.append((off["x"] + (sx * size + (a + 0.5) * dx),
                            off["y"] + (sy * size + (b + 0.5) * dx)))
                ids.append(act)
    return np.array(pos), np.array(ids, np.int32), p["actuator_num"]


def quadruped3d_particles(cells_per_cube=6, ppc_edge=2, n_grid=64, origin_cells=(6, 4, 6)):
    """Particles filling the quadruped's voxel body (4x1x2 body on four 1x2x1
    legs, leg g = actuator group g, body passive), each unit cube spanning
    ``cells_per_cube`` grid cells, ``ppc_edge^3`` particles per cell."""
    from .robot_design import quadruped_cubes
    dx = 1.0 / n_grid
    s = cells_per_cube
    sub = (np.arange(ppc_edge) + 0.5) / ppc_edge
    pos, ids = [], []
    for (ci, cj, ck), act, g in quadruped_cubes():
        for a, b, c in itertools.product(range(s), repeat=3):
            cell = np.array([origin_cells[0] + ci * s + a, origin_cells[1] + cj * s + b,
                             origin_cells[2] + ck * s + c], np.float64)
            for u, v, w in itertools.product(sub, sub, sub):
                pos.append((cell + np.array([u, v, w])) * dx)
                ids.append(g if act else -1)
    return np.array(pos), np.array(ids, np.int32), 4


def _soa(a, B, dtype, device):
    """[P,D] host -> [B,D,P] device copies."""
    t = torch.as_tensor(np.ascontiguousarray(np.asarray(a).T), dtype=dtype)
    return t[None].expand(B, -1, -1).contiguous().to(device)


class MPMStep:
    """Launch helpers for one batch geometry (B robots, P particles each)."""

    def __init__(self, params, B, P, n_act, act_id, dtype, device="cuda"):
        self.p, self.B, self.P, self.n_act = params, B, P, n_act
        self.D = params.dim
        self.G = params.n_grid ** params.dim
        self.dtype, self.device = dtype, torch.device(device)
        self.dc = nat.dtype_code(dtype)
        ids = torch.as_tensor(np.asarray(act_id, np.int32))
        self.act_id = (ids[None].expand(B, -1) if ids.dim() == 1 else ids).contiguous().to(self.device)
        z = lambda *s: torch.zeros(*s, dtype=dtype, device=self.device)
        self.gv, self.gm, self.gvo = z(B, self.D, self.G), z(B, self.G), z(B, self.D, self.G)
        self.gvo_b, self.gv_b, self.gm_b = z(B, self.D, self.G), z(B, self.D, self.G), z(B, self.G)
        self.err = torch.zeros(1, dtype=torch.int32, device=self.device)

    def fill(self, t, value=0.0):
        nat.call("dl_mpm_fill", self.dc, nat.ptr(t), t.numel(), float(value), nat.stream())

    def forward(self, x, v, C, F, act, x1, v1, C1, F1):
        """advance_core: clear grid, P2G (writes F1), grid op, G2P."""
        self.fill(self.gv)
        self.fill(self.gm)
        pp = ctypes.byref(self.p)
        nat.call("dl_mpm_p2g", self.dc, pp, self.B, self.P, *map(nat.ptr, (x, v, C, F, act)),
                 nat.ptr(self.act_id), self.n_act, nat.ptr(F1), nat.ptr(self.gv), nat.ptr(self.gm),
                 nat.ptr(self.err), nat.stream())
        nat.call("dl_mpm_grid", self.dc, pp, self.B, nat.ptr(self.gv), nat.ptr(self.gm),
                 nat.ptr(self.gvo), nat.stream())
        nat.call("dl_mpm_g2p", self.dc, pp, self.B, self.P, nat.ptr(x), nat.ptr(self.gvo),
                 nat.ptr(x1), nat.ptr(v1), nat.ptr(C1), nat.stream())

    def backward(self, x, v, C, F, act, xb1, vb1, Cb1, Fb1, xb, vb, Cb, Fb, actb):
        """Recompute the grid of step t, then G2P, grid and P2G adjoints.
        Writes xb, vb, Cb, Fb (adjoint of state t); accumulates actb."""
        scratchF = self.scratch_F(F)
        self.fill(self.gv)
        self.fill(self.gm)
        pp = ctypes.byref(self.p)
        nat.call("dl_mpm_p2g", self.dc, pp, self.B, self.P, *map(nat.ptr, (x, v, C, F, act)),
                 nat.ptr(self.act_id), self.n_act, nat.ptr(scratchF), nat.ptr(self.gv),
                 nat.ptr(self.gm), nat.ptr(self.err), nat.stream())
        nat.call("dl_mpm_grid", self.dc, pp, self.B, nat.ptr(self.gv), nat.ptr(self.gm),
                 nat.ptr(self.gvo), nat.stream())
        self.fill(self.gvo_b)
        nat.call("dl_mpm_g2p_adj", self.dc, pp, self.B, self.P, nat.ptr(x), nat.ptr(self.gvo),
                 nat.ptr(xb1), nat.ptr(vb1), nat.ptr(Cb1), nat.ptr(xb), nat.ptr(self.gvo_b),
                 nat.stream())
        nat.call("dl_mpm_grid_adj", self.dc, pp, self.B, nat.ptr(self.gv), nat.ptr(self.gm),
                 nat.ptr(self.gvo_b), nat.ptr(self.gv_b), nat.ptr(self.gm_b), nat.stream())
        nat.call("dl_mpm_p2g_adj", self.dc, pp, self.B, self.P, *map(nat.ptr, (x, v, C, F, act)),
                 nat.ptr(self.act_id), self.n_act, nat.ptr(self.gv_b), nat.ptr(self.gm_b),
                 nat.ptr(Fb1), nat.ptr(xb), nat.ptr(vb), nat.ptr(Cb), nat.ptr(Fb), nat.ptr(actb),
                 nat.stream())

    def scratch_F(self, F):
        if getattr(self, "_sF", None) is None or self._sF.shape != F.shape:
            self._sF = torch.empty_like(F)
        return self._sF

    def check(self):
        if int(self.err.item()) != 0:
            raise nat.NativeError("MPM: a particle left the grid (stencil out of bounds)")


class SolverMPM:
    """Drop-in ``SolverMPM`` (``multitask/solver_mpm.py:5-196``), built from a
    configuration like the reference's (``ConfigSim`` with ``robot.simulator
    = "mpm"``; ``config_sim.py:80-129`` derives the keys read here).

    Fields (user layout as the reference's, storage structure-of-arrays):
    ``x, v`` [T+1, B, P, D] (``Field``, storage [T+1][B][D][P]); ``C, F``
    storage [T+1][B][D*D][P] (``C_view`` / ``F_view`` give [T+1, B, P, D, D]);
    ``center`` [T+1, B, D]; ``actuation`` [T+1, B, n_springs] (the reference
    sizes it by n_springs, which its config sets to the particle count);
    ``actuator_id``, ``particle_type`` [B, P]; per-robot grids
    ``grid_v_in`` [B, D, n^D], ``grid_m_in`` [B, n^D], ``grid_v_out``.

    ``p2g(f)``, ``grid_op()``, ``g2p(f)`` launch the HIP kernels and, inside a
    ``Tape``, record their adjoints like Taichi's kernel autodiff (reading the
    grid as it is when the adjoint replays).  ``advance_core(t)`` is the
    reference's ``ti.ad.grad_replaced`` step: clear grid, P2G, grid update,
    G2P, with ``advance_core_grad(t)`` recorded in their place -- it rebuilds
    the step's grid (clear, P2G, grid update) and runs the G2P, grid and P2G
    adjoints, accumulating into the gradients of state t and actuation t.
    Only solid particles (``particle_type`` 1, the reference's fill) are
    supported."""

    def __init__(self, config, device="cuda"):
        from .fields import Field
        cfg = config.get_config()
        self.config = config
        self.simulator = "mpm"
        proc, sim, rob = cfg["process"], cfg["simulator"], cfg["robot"]
        self.max_steps = proc["max_steps"]
        self.dt = proc["dt"]
        self.ground_height = sim["ground_height"]
        self.gravity = sim["gravity"]
        self.drag_damping = sim["drag_damping"]
        self.dashpot_damping = sim["dashpot_damping"]
        self.batch_size = cfg["nn"]["batch_size"]
        self.n_objects = rob["n_objects"]
        self.n_springs = rob["n_springs"]
        self.springs = rob["springs"]
        self.dim = rob["dim"]
        self.n_particles = sim["n_particles"]
        self.n_grid = sim["mpm"]["n_grid"]
        self.dx = sim["mpm"]["dx"]
        self.inv_dx = sim["mpm"]["inv_dx"]
        self.act_strength = sim["act_strength"]
        self.p_vol = sim["p_vol"]
        self.E = sim["E"]
        self.mu = sim["mu"]
        self.la = sim["la"]
        self.bound = sim["bound"]
        self.height = None
        self.rotation = None
        self.dtype = torch.float64 if cfg.get("precision", "f64") == "f64" else torch.float32
        self.device = torch.device(device)
        # solid particles have mass 1 (solver_mpm.py:104-107); the active
        # stress acts along y in 2D (:105), along x for the 3D bodies
        self.params = MPMParamsC(dim=self.dim, n_grid=self.n_grid, bound=self.bound,
                                 act_axis=sim.get("act_axis", 1 if self.dim == 2 else 0),
                                 dt=self.dt, gravity=self.gravity, p_vol=self.p_vol, p_mass=1.0,
                                 mu=self.mu, la=self.la, act_strength=self.act_strength)
        D, P, B, T = self.dim, self.n_particles, self.batch_size, self.max_steps + 1
        f = lambda shape, soa=False, name="": Field(shape, self.dtype, self.device, soa=soa,
                                                    name=name)
        self.x = f((T, B, P, D), soa=True, name="x")
        self.v = f((T, B, P, D), soa=True, name="v")
        self.C = f((T, B, D * D, P), name="C")
        self.F = f((T, B, D * D, P), name="F")
        self.center = f((T, B, D), name="center")
        self.actuation = f((T, B, self.n_springs), name="actuation")
        self.actuator_id = torch.zeros(B, P, dtype=torch.int32, device=self.device)
        self.particle_type = torch.zeros(B, P, dtype=torch.int32, device=self.device)
        G = self.n_grid ** D
        z = lambda *sh: torch.zeros(*sh, dtype=self.dtype, device=self.device)
        self.grid_v_in, self.grid_m_in, self.grid_v_out = z(B, D, G), z(B, G), z(B, D, G)
        self._gvo_b, self._gv_b, self._gm_b = z(B, D, G), z(B, D, G), z(B, G)
        self._err = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._dc = nat.dtype_code(self.dtype)

    @property
    def C_view(self):
        T, B, _, P = self.C.data.shape
        return self.C.data.view(T, B, self.dim, self.dim, P).permute(0, 1, 4, 2, 3)

    @property
    def F_view(self):
        T, B, _, P = self.F.data.shape
        return self.F.data.view(T, B, self.dim, self.dim, P).permute(0, 1, 4, 2, 3)

    # ------------------------------------------------------------ set-up
    def initialize_robot(self):
        """solver_mpm.py:54-58: actuator ids from the design (the config's
        ``springs``), every particle solid."""
        ids = torch.as_tensor(np.asarray(self.springs, np.int32)[: self.n_particles])
        self.actuator_id.copy_(ids[None].expand(self.batch_size, -1))
        self.particle_type.fill_(1)
        if int(self.actuator_id.max()) >= self.n_springs:
            raise ValueError("actuator id beyond the actuation field's width")

    def clear_states(self, steps):
        """solver_mpm.py:60-71: C = 0, F = I and the x, v, C, F gradients of the
        first ``steps`` rows zeroed."""
        D = self.dim
        self.C.data[:steps].zero_()
        self.F.data[:steps].zero_()
        for i in range(D):
            self.F.data[:steps, :, i * D + i] = 1.0
        for fld in (self.x, self.v, self.C, self.F):
            fld.grad.data[:steps].zero_()

    def clear_grid(self):
        """solver_mpm.py:73-81 (graph-capture-safe fill kernels)."""
        for t in (self.grid_v_in, self.grid_m_in, self.grid_v_out, self._gvo_b, self._gv_b,
                  self._gm_b):
            nat.call("dl_mpm_fill", self._dc, nat.ptr(t), t.numel(), 0.0, nat.stream())

    def _check_types(self):
        if int(self.particle_type.min()) != 1:
            raise NotImplementedError("fluid particles (particle_type 0) are not supported")

    # ------------------------------------------------------------ kernels
    def _p2g(self, f, F_out):
        P = nat.ptr
        nat.call("dl_mpm_p2g", self._dc, ctypes.byref(self.params), self.batch_size,
                 self.n_particles, P(self.x.at(f)), P(self.v.at(f)), P(self.C.at(f)),
                 P(self.F.at(f)), P(self.actuation.at(f)), P(self.actuator_id), self.n_springs,
                 P(F_out), P(self.grid_v_in), P(self.grid_m_in), P(self._err), nat.stream())

    def _grid_op(self):
        nat.call("dl_mpm_grid", self._dc, ctypes.byref(self.params), self.batch_size,
                 nat.ptr(self.grid_v_in), nat.ptr(self.grid_m_in), nat.ptr(self.grid_v_out),
                 nat.stream())

    def _g2p(self, f):
        P = nat.ptr
        nat.call("dl_mpm_g2p", self._dc, ctypes.byref(self.params), self.batch_size,
                 self.n_particles, P(self.x.at(f)), P(self.grid_v_out), P(self.x.at(f + 1)),
                 P(self.v.at(f + 1)), P(self.C.at(f + 1)), nat.stream())

    def _g2p_adj(self, f):
        """x.grad[f] += G2P adjoint; grid_v_out adjoint accumulated."""
        P = nat.ptr
        xb = torch.zeros_like(self.x.at(f))
        nat.call("dl_mpm_g2p_adj", self._dc, ctypes.byref(self.params), self.batch_size,
                 self.n_particles, P(self.x.at(f)), P(self.grid_v_out), P(self.x.grad.at(f + 1)),
                 P(self.v.grad.at(f + 1)), P(self.C.grad.at(f + 1)), P(xb), P(self._gvo_b),
                 nat.stream())
        self.x.grad.at(f).add_(xb)

    def _grid_adj(self):
        nat.call("dl_mpm_grid_adj", self._dc, ctypes.byref(self.params), self.batch_size,
                 nat.ptr(self.grid_v_in), nat.ptr(self.grid_m_in), nat.ptr(self._gvo_b),
                 nat.ptr(self._gv_b), nat.ptr(self._gm_b), nat.stream())

    def _p2g_adj(self, f):
        """x/v/C/F.grad[f] += P2G adjoint (from the grid adjoints and
        F.grad[f+1]); actuation.grad[f] accumulated."""
        P = nat.ptr
        xb = torch.zeros_like(self.x.at(f))
        vb, Cb, Fb = torch.empty_like(xb), torch.empty_like(self.C.at(f)), torch.empty_like(self.F.at(f))
        nat.call("dl_mpm_p2g_adj", self._dc, ctypes.byref(self.params), self.batch_size,
                 self.n_particles, P(self.x.at(f)), P(self.v.at(f)), P(self.C.at(f)),
                 P(self.F.at(f)), P(self.actuation.at(f)), P(self.actuator_id), self.n_springs,
                 P(self._gv_b), P(self._gm_b), P(self.F.grad.at(f + 1)), P(xb), P(vb), P(Cb),
                 P(Fb), P(self.actuation.grad.at(f)), nat.stream())
        self.x.grad.at(f).add_(xb)
        self.v.grad.at(f).add_(vb)
        self.C.grad.at(f).add_(Cb)
        self.F.grad.at(f).add_(Fb)

    def p2g(self, f):
        """solver_mpm.py:83-124: writes F[f+1] and the (cleared) grid."""
        self._check_types()
        self._p2g(f, self.F.at(f + 1))
        record(lambda: self._p2g_adj(f))

    def grid_op(self):
        """solver_mpm.py:126-144."""
        self._grid_op()
        record(self._grid_adj)

    def g2p(self, f):
        """solver_mpm.py:146-165: writes x, v, C of row f+1."""
        self._g2p(f)
        record(lambda: self._g2p_adj(f))

    def compute_center(self, t):
        """solver_mpm.py:167-173: mean particle position per robot."""
        with torch.no_grad():
            self.center.data[t].copy_(self.x.at(t).mean(-1))

        def adj():
            self.x.grad.at(t).add_(self.center.grad.at(t)[..., None] / self.n_particles)
        record(adj)

    def advance_core(self, t):
        """solver_mpm.py:175-180 (grad_replaced): the adjoint recorded for
        the step is advance_core_grad(t)."""
        self._check_types()
        self.clear_grid()
        self._p2g(t, self.F.at(t + 1))
        self._grid_op()
        self._g2p(t)
        record(lambda: self.advance_core_grad(t))

    def advance_core_grad(self, t):
        """solver_mpm.py:182-190: rebuild step t's grid, then the G2P, grid
        and P2G adjoints."""
        self.clear_grid()
        scratch = torch.empty_like(self.F.at(t + 1))
        self._p2g(t, scratch)                   # F[t+1] is already in place
        self._grid_op()
        self._g2p_adj(t)
        self._grid_adj()
        self._p2g_adj(t)

    def pre_advance(self, s):
        self.compute_center(s)

    def advance(self, s):
        self.advance_core(s)

    def check(self):
        if int(self._err.item()) != 0:
            raise nat.NativeError("MPM: a particle left the grid (stencil out of bounds)")


class MPMBatchC(Structure):
    """dl_mpm_batch."""
    _fields_ = ([(n, c_int32) for n in ("dtype", "B", "P", "n_act", "n_groups", "n_sin",
                                        "run_period", "n_hidden", "act", "horizon",
                                        "tile_nodes", "scatter")] +
                [("offset_scale", c_double), ("max_speed", c_double), ("params", MPMParamsC)] +
                [(n, ctypes.c_void_p) for n in (
                    "act_id", "group", "inv_count", "targets", "W1", "b1", "W2", "b2",
                    "grid", "grid_vo", "grid_vo_grad", "grid_grad", "blk_flag", "blk_list",
                    "blk_count", "err", "sens", "center", "vcom", "inp", "hid", "act_out",
                    "act_grad", "in_grad", "center_grad", "vcom_grad", "wgrad", "tiles",
                    "tile_hdr", "blk_tile_count", "blk_tiles", "grid_save", "vo_save",
                    "list_save", "count_save")] + [("save_steps", c_int32), ("scatter_unroll", c_int32)] +
                [(n, ctypes.c_void_p) for n in ("hid_grad", "out_grad", "windows")] +
                [("n_windows", c_int32), ("pad_", c_int32)] +
                [(n, ctypes.c_void_p) for n in ("hgrid", "hvo", "hlist", "hcount", "hover")] +
                [("hcap", c_int32), ("restore_from", c_int32)] +
                [(n, ctypes.c_void_p) for n in ("n_part", "horizons")] +
                [("per_robot_weights", c_int32), ("pad2_", c_int32)])


class MPMSlotsC(Structure):
    """dl_mpm_slots."""
    _fields_ = [("x", ctypes.c_void_p), ("v", ctypes.c_void_p), ("C", ctypes.c_void_p),
                ("F", ctypes.c_void_p), ("n_slots", c_int32), ("pad_", c_int32)]


MPM_CONTROL, MPM_SENSE_FIRST, MPM_SAVE_GRID, MPM_SAVE_HORIZON, MPM_REPLAY = 1, 2, 4, 8, 16
MPM_PB, MPM_TILE_MAX, MPM_TILE_CAP = 256, 8192, 32
SCATTER_ATOMIC, SCATTER_STORE = 0, 1


def block_order(positions, n_grid, dim=3):
    """Permutation sorting particles by grid block (4^3 nodes in 3D, 8^2 in
    2D) of their stencil base cell, then by cell: consecutive particles then
    share stencil nodes, which keeps the LDS tiles of P2G and of the G2P
    adjoint small."""
    E = 4 if dim == 3 else 8
    base = np.floor(np.asarray(positions) * n_grid - 0.5).astype(np.int64)
    nb = n_grid // E
    blk = np.zeros(len(base), np.int64)
    cell = np.zeros(len(base), np.int64)
    for d in range(dim):
        blk = blk * nb + base[:, d] // E
        cell = cell * E + base[:, d] % E
    return np.argsort(blk * E ** dim + cell, kind="stable")


def scatter_windows(positions, n_grid, box_limit, dim=3, pb=MPM_PB):
    """Cut the particle order into windows (first, end) of at most pb
    particles whose stencil box (base cells' bounding box plus the 2-node
    stencil reach) holds at most box_limit nodes: each window grows until
    the next particle would exceed either bound (a single particle always
    fits).  Returns an int32 array [n_windows, 2] partitioning [0, P)."""
    base = np.floor(np.asarray(positions, np.float64) * n_grid - 0.5).astype(np.int64)[:, :dim]
    P, out, i = len(base), [], 0
    while i < P:
        seg = base[i:i + pb]
        lo, hi = np.minimum.accumulate(seg, 0), np.maximum.accumulate(seg, 0)
        over = np.nonzero(np.prod(hi - lo + 3, 1) > box_limit)[0]
        n = len(seg) if len(over) == 0 else max(int(over[0]), 1)
        out.append((i, i + n))
        i += n
    return np.asarray(out, np.int32).reshape(-1, 2)


class _Slots:
    """S state slots: x, v [S,B,D,P]; C, F [S,B,D*D,P].  With ``parent`` the
    slots are a view of the parent's slots [base, base + S)."""

    def __init__(self, S, B, D, P, dtype, device, parent=None, base=0):
        if parent is None:
            z = lambda *s: torch.zeros(*s, dtype=dtype, device=device)
            self.x, self.v = z(S, B, D, P), z(S, B, D, P)
            self.C, self.F = z(S, B, D * D, P), z(S, B, D * D, P)
        else:
            self.x, self.v, self.C, self.F = (f[base:base + S] for f in parent.fields())
        self.c = MPMSlotsC(nat.ptr(self.x), nat.ptr(self.v), nat.ptr(self.C), nat.ptr(self.F), S, 0)

    def fields(self):
        return (self.x, self.v, self.C, self.F)

    def copy_slot(self, dst, src, other=None):
        for a, b in zip(self.fields(), (other or self).fields()):
            a[dst].copy_(b[src])


class MPMBatch:
    """Batched, checkpointed MLS-MPM training rollout (3D).

    B robots of one particle design, each on its own grid, driven by the
    sensor -> MLP -> actuation loop of the quadruped workload (sinusoids,
    COM velocity, per-group mean offsets from the COM, targets) and trained
    on the windowed COM loss.  ``forward`` advances all robots ``T`` steps,
    checkpointing particle state every ``K`` steps, one native call per
    segment; ``backward`` recomputes each segment from its checkpoint and runs
    the hand-written adjoints in reverse.  Weight gradients land in ``grads``
    (W1, b1, W2, b2).  Particles are reordered by grid block (see
    ``block_order``; ``order`` maps internal to given indices) -- the sensors,
    COM and loss do not depend on particle order.  The scatter kernels run one
    workgroup per window of at most 256 consecutive particles whose stencil
    box holds at most ``box_limit`` nodes (``scatter_windows``; None: plain
    256-particle chunks); ``tile_nodes`` sizes their LDS buffer (default
    2048 nodes).  With ``resort`` each robot's particles are re-sorted by
    block at every checkpoint (the state, actuator ids and sensor groups
    permuted together), so the scatter keeps compact cells as the robots
    move; the adjoint is permuted back at each checkpoint.  With
    ``save_horizon`` the forward also saves every step's active grid blocks
    (at most ``horizon_blocks`` per robot and step; default three times the
    rest shape's blocks), so the backward recomputes a segment from its
    checkpoint by restoring the saved grid velocities and running only the
    update of F and G2P; a segment where some step did not fit is recomputed
    in full.  With ``store_states`` (and ``save_horizon``) the forward keeps
    every step's particle state (T + T/K slots of x, v, C, F: one view of
    K + 1 slots per segment), and the backward of a segment whose every
    step's grid was saved runs the adjoint on those states with no
    recompute at all.

    Uneven batches: ``positions`` and ``act_id`` given as lists of B arrays
    (one body per robot, any particle counts) pad every robot to the largest
    count (padding is never read); ``horizons`` (B values, at most T) stop
    each robot at its own step; ``per_robot_weights`` gives every robot its
    own controller (``set_weights`` then takes [B, ...] arrays, ``grads`` are
    [B, ...]).  Each robot's loss is the windowed COM loss over its own last
    ``min(window, horizon)`` states (``loss_b``; ``loss`` is their mean).
    Uneven batches run with plain 256-particle windows and without the
    horizon grid save."""

    def __init__(self, params, positions, act_id, n_act, B, T, K=100, window=None,
                 dtype=torch.float32, n_sin=10, run_period=100, offset_scale=0.2,
                 max_speed=0.08, activation="tanh", n_hidden=64, device="cuda",
                 sort_particles=True, tile_nodes=None, scatter=SCATTER_ATOMIC, save_grid=True,
                 unroll=3, box_limit=512, resort=True, save_horizon=True, horizon_blocks=None,
                 store_states=False, overlap_replay=True, horizons=None, per_robot_weights=False):
        if params.dim != 3:
            raise ValueError("MPMBatch drives 3D robots (the controller's sensors are 3D)")
        if params.n_grid % 4:
            raise ValueError("MPMBatch: n_grid must be a multiple of 4")
        self.p, self.B, self.T, self.K = params, B, T, K
        self.D = D = 3
        self.uneven = isinstance(positions, (list, tuple))
        self.per_robot_weights = bool(per_robot_weights)
        self.horizons = None if horizons is None else np.asarray(horizons, np.int64)
        if self.uneven or self.horizons is not None or self.per_robot_weights:
            if store_states or save_horizon:
                raise ValueError("MPMBatch: uneven batches run without the horizon grid save "
                                 "and stored states (save_horizon=False, store_states=False)")
            box_limit = None
        if self.horizons is not None:
            if self.horizons.shape != (B,) or self.horizons.min() < 1 or self.horizons.max() > T:
                raise ValueError(f"MPMBatch: horizons must be {B} values in [1, {T}]")
        if self.uneven:
            if len(positions) != B or len(act_id) != B:
                raise ValueError(f"MPMBatch: {B} bodies expected")
            bodies = []
            for pk, ik in zip(positions, act_id):
                pk, ik = np.asarray(pk, np.float64), np.asarray(ik, np.int32)
                ok = block_order(pk, params.n_grid) if sort_particles else np.arange(len(pk))
                bodies.append((pk[ok], ik[ok], ok))
            self.n_parts = np.array([len(b[0]) for b in bodies], np.int64)
            P = int(self.n_parts.max())
            # padding: the robot's first particle's position, no actuator
            pos_b = np.stack([np.concatenate([b[0], np.repeat(b[0][:1], P - len(b[0]), 0)])
                              for b in bodies])
            ids_b = np.stack([np.concatenate([b[1], np.full(P - len(b[1]), -1, np.int32)])
                              for b in bodies])
            self.order = [b[2] for b in bodies]
            pos, ids = bodies[0][0], bodies[0][1]
        else:
            pos = np.asarray(positions, np.float64)
            ids = np.asarray(act_id, np.int32)
            self.order = block_order(pos, params.n_grid) if sort_particles else np.arange(len(pos))
            pos, ids = pos[self.order], ids[self.order]
            P = len(pos)
            self.n_parts = None
        self.positions = pos
        self.P = P
        self.n_act = n_act
        self.window = window if window is not None else T // 2
        self.dtype, self.device = dtype, torch.device(device)
        self.dc = nat.dtype_code(dtype)
        self.n_sin, self.run_period = n_sin, run_period
        self.offset_scale, self.max_speed = offset_scale, max_speed
        self.act = nat.ACT_SIN if activation == "sin" else nat.ACT_TANH
        self.G = n_act + 1                                  # actuator groups + passive body
        self.I = n_sin + 3 + 3 * self.G + 4
        self.H, self.O = n_hidden, n_act
        dev = self.device
        z = lambda *s: torch.zeros(*s, dtype=dtype, device=dev)
        zi = lambda *s: torch.zeros(*s, dtype=torch.int32, device=dev)
        bt = lambda a, dt_: torch.as_tensor(a, dtype=dt_)[None].expand(B, *np.shape(a)).contiguous().to(dev)
        inv = lambda c: np.where(c > 0, 1.0 / np.maximum(c, 1), 0.0)
        if self.uneven:
            grp_b = np.where(ids_b >= 0, ids_b, n_act).astype(np.int32)
            counts = np.stack([np.bincount(g[:n], minlength=self.G)
                               for g, n in zip(grp_b, self.n_parts)]).astype(np.float64)
            self.act_id = torch.as_tensor(ids_b, dtype=torch.int32).to(dev)
            self.group = torch.as_tensor(grp_b, dtype=torch.int32).to(dev)
            self.inv_count = torch.as_tensor(inv(counts), dtype=dtype).to(dev)
        else:
            grp = np.where(ids >= 0, ids, n_act).astype(np.int32)
            counts = np.bincount(grp, minlength=self.G).astype(np.float64)
            self.act_id = bt(ids, torch.int32)
            self.group = bt(grp, torch.int32)
            self.inv_count = bt(inv(counts), dtype)
        self.resort = resort
        ng = params.n_grid ** 3
        nblk = (params.n_grid // 4) ** 3
        self.grid, self.grid_vo, self.grid_vo_grad, self.grid_grad = (z(B, ng, 4) for _ in range(4))
        self.blk_flag, self.blk_list, self.blk_count = zi(B, nblk), zi(B, nblk), zi(B)
        self.err = zi(1)
        # scatter windows: box-limited cuts of the (sorted) particle order
        win = (scatter_windows(pos, params.n_grid, box_limit) if box_limit is not None
               else None)
        self.windows = torch.as_tensor(win).to(dev) if win is not None else None
        wpr = len(win) if win is not None else (P + MPM_PB - 1) // MPM_PB
        if tile_nodes is None:      # 32 KiB of LDS in FP32, 64 KiB in FP64
            tile_nodes = 2048
        self.tile_nodes, self.scatter = tile_nodes, scatter
        self.tiles = z(B, wpr, tile_nodes if scatter == SCATTER_STORE else 1, 4)
        self.tile_hdr = zi(B, wpr, 8)
        self.blk_tile_count, self.blk_tiles = zi(B, nblk), zi(B, nblk, MPM_TILE_CAP)
        # per-step grids of a segment, saved by the segment recompute so the
        # adjoint does not re-run P2G and the grid update (K x B x nblk
        # blocks of 64 records, written only for active blocks)
        self.save_grid = save_grid
        emp = lambda *sh, dt=dtype: torch.empty(*sh, dtype=dt, device=dev)
        self.grid_save = emp(K, B * nblk * 64 * 4) if save_grid else None
        self.vo_save = emp(K, B * nblk * 64 * 4) if save_grid else None
        self.list_save = emp(K, B * nblk, dt=torch.int32) if save_grid else None
        self.count_save = emp(K, B, dt=torch.int32) if save_grid else None
        NS = 3 * self.G + 6
        self.sens = z(T + 1, B, NS)
        self.center, self.vcom = z(T + 1, B, 3), z(T + 1, B, 3)
        self.inp, self.hid, self.act_out = z(T, B, self.I), z(T, B, self.H), z(T, B, self.O)
        self.act_grad, self.in_grad = z(B, self.O), z(T + 1, B, self.I)
        self.cg, self.vg = z(T + 1, B, 3), z(T + 1, B, 3)
        H, I, O = self.H, self.I, self.O
        self.nw = H * I + H + O * H + O
        self.wgrad = z(B, self.nw)
        self.hid_grad, self.out_grad = emp(T, B * H), emp(T, B * O)
        self.W = ([z(B, H, I), z(B, H), z(B, O, H), z(B, O)] if self.per_robot_weights
                  else [z(H, I), z(H), z(O, H), z(O)])
        self.targets, self.tv, self.th = z(B, 5), z(B, 3), z(B)
        self.loss, self.loss_b = z(1), z(B)
        if self.uneven:                                                             # [B,3,P]
            self.x0 = torch.as_tensor(np.ascontiguousarray(pos_b.transpose(0, 2, 1)),
                                      dtype=dtype).to(dev)
        else:                                                                       # [3,P]
            self.x0 = torch.as_tensor(np.ascontiguousarray(pos.T), dtype=dtype).to(dev)
        self.n_part_t = (torch.as_tensor(self.n_parts, dtype=torch.int32).to(dev)
                         if self.uneven else None)
        self.horizons_t = (torch.as_tensor(self.horizons, dtype=torch.int32).to(dev)
                           if self.horizons is not None else None)
        self.nck = (T + K - 1) // K
        if resort:
            # segment s runs in order orders[s] (indices into the initial
            # order); perms[s]: state_s[i] = state_{s-1}[perms[s][i]]
            self.act_id0, self.group0 = self.act_id.clone(), self.group.clone()
            self.orders = torch.empty(self.nck, B, P, dtype=torch.int64, device=dev)
            self.perms = torch.empty(self.nck, B, P, dtype=torch.int64, device=dev)
            self.orders[0].copy_(torch.arange(P, device=dev)[None].expand(B, -1))
        self.ring = _Slots(2, B, D, P, dtype, dev)
        self.adj = _Slots(2, B, D, P, dtype, dev)
        self.store_states = bool(store_states)
        if self.store_states:
            # segment s = steps [t0, t0 + n) owns slots [t0 + s, t0 + s + n + 1)
            self.states = _Slots(T + self.nck, B, D, P, dtype, dev)
            self.views = [_Slots(n + 1, B, D, P, dtype, dev, parent=self.states, base=t0 + s)
                          for s, (t0, n) in enumerate(self._segments())]
            self.ck = self.seg = None
        else:
            self.ck = _Slots(self.nck, B, D, P, dtype, dev)
            self.seg = _Slots(K + 1, B, D, P, dtype, dev)
        # Replayed recomputes (saved grid velocities + G2P) of segment s - 1
        # run on a side stream while the adjoint of segment s runs: a second
        # segment buffer, and the replay's own velocity grid and active-block
        # count (the only grid state the replay writes), so nothing it
        # touches is shared with the adjoint in flight.
        self.overlap_replay = bool(overlap_replay and save_horizon and not self.store_states)
        if self.overlap_replay:
            self.seg2 = _Slots(K + 1, B, D, P, dtype, dev)
            self.grid_vo2 = z(B, ng, 4)
            self.blk_count2 = zi(B)
            self.replay_stream = torch.cuda.Stream(device=dev)
        P_ = lambda t: nat.ptr(t)
        self.desc = MPMBatchC(
            dtype=self.dc, B=B, P=P, n_act=n_act, n_groups=self.G, n_sin=n_sin,
            run_period=run_period, n_hidden=H, act=self.act, horizon=T,
            tile_nodes=tile_nodes, scatter=scatter,
            offset_scale=offset_scale, max_speed=max_speed, params=params,
            act_id=P_(self.act_id), group=P_(self.group), inv_count=P_(self.inv_count),
            targets=P_(self.targets), W1=P_(self.W[0]), b1=P_(self.W[1]), W2=P_(self.W[2]),
            b2=P_(self.W[3]), grid=P_(self.grid), grid_vo=P_(self.grid_vo),
            grid_vo_grad=P_(self.grid_vo_grad), grid_grad=P_(self.grid_grad),
            blk_flag=P_(self.blk_flag), blk_list=P_(self.blk_list), blk_count=P_(self.blk_count),
            err=P_(self.err), sens=P_(self.sens), center=P_(self.center), vcom=P_(self.vcom),
            inp=P_(self.inp), hid=P_(self.hid), act_out=P_(self.act_out),
            act_grad=P_(self.act_grad), in_grad=P_(self.in_grad), center_grad=P_(self.cg),
            vcom_grad=P_(self.vg), wgrad=P_(self.wgrad), tiles=P_(self.tiles),
            tile_hdr=P_(self.tile_hdr), blk_tile_count=P_(self.blk_tile_count),
            blk_tiles=P_(self.blk_tiles),
            grid_save=P_(self.grid_save) if save_grid else None,
            vo_save=P_(self.vo_save) if save_grid else None,
            list_save=P_(self.list_save) if save_grid else None,
            count_save=P_(self.count_save) if save_grid else None,
            save_steps=K if save_grid else 0, scatter_unroll=unroll,
            hid_grad=P_(self.hid_grad), out_grad=P_(self.out_grad),
            windows=P_(self.windows) if win is not None else None, n_windows=wpr,
            n_part=P_(self.n_part_t) if self.uneven else None,
            horizons=P_(self.horizons_t) if self.horizons is not None else None,
            per_robot_weights=int(self.per_robot_weights))
        self.save_horizon = save_horizon
        if save_horizon:
            if horizon_blocks is None:
                horizon_blocks = max(64, 3 * self.rest_blocks())
            hcap = int(min(nblk, max(1, horizon_blocks)))
            self.hcap = hcap
            self.hgrid, self.hvo = emp(T, B * hcap * 64 * 4), emp(T, B * hcap * 64 * 4)
            self.hlist = emp(T, B * hcap, dt=torch.int32)
            self.hcount = emp(T, B, dt=torch.int32)
            self.hover = zi(T)
            self.desc.hgrid, self.desc.hvo = P_(self.hgrid), P_(self.hvo)
            self.desc.hlist, self.desc.hcount = P_(self.hlist), P_(self.hcount)
            self.desc.hover, self.desc.hcap = P_(self.hover), hcap
        self.replay = [False] * self.nck
        self.grads = None

    def rest_blocks(self):
        """Grid blocks (4^3 nodes) under some particle's stencil at rest."""
        base = np.floor(self.positions * self.p.n_grid - 0.5).astype(np.int64)
        offs = np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 3)
        blk = ((base[:, None, :] + offs[None]) // 4).reshape(-1, 3)
        nb = self.p.n_grid // 4
        return len(np.unique((blk[:, 0] * nb + blk[:, 1]) * nb + blk[:, 2]))

    def set_targets(self, tv, th):
        t = np.zeros((self.B, 5))
        t[:, :3] = np.asarray(tv)
        t[:, 3] = np.asarray(th)
        self.targets.copy_(torch.as_tensor(t, dtype=self.dtype))
        self.tv.copy_(self.targets[:, 0:3])
        self.th.copy_(self.targets[:, 3])

    def set_weights(self, W1, b1, W2, b2):
        """Controller weights W1 [H, I], b1 [H], W2 [O, H], b2 [O]; with
        per-robot weights each with a leading robot axis (or without it: the
        same weights for every robot)."""
        for d, s in zip(self.W, (W1, b1, W2, b2)):
            d.copy_(torch.as_tensor(s, dtype=self.dtype, device=d.device).expand_as(d))

    def set_robot_weights(self, k, W1, b1, W2, b2):
        """Robot k's controller (per-robot weights)."""
        if not self.per_robot_weights:
            raise ValueError("MPMBatch: set_robot_weights needs per_robot_weights=True")
        for d, s in zip(self.W, (W1, b1, W2, b2)):
            d[k].copy_(torch.as_tensor(s, dtype=self.dtype, device=d.device))

    def active_nodes(self):
        """Grid nodes inside some particle's stencil in the initial state (per robot)."""
        base = np.floor(self.positions * self.p.n_grid - 0.5).astype(np.int64)
        offs = np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 3)
        nodes = (base[:, None, :] + offs[None]).reshape(-1, 3)
        n = self.p.n_grid
        return len(np.unique((nodes[:, 0] * n + nodes[:, 1]) * n + nodes[:, 2]))

    def _segments(self):
        return [(s * self.K, min(self.K, self.T - s * self.K)) for s in range(self.nck)]

    # A rollout as resumable pieces (a serving loop advances one segment per
    # tick); forward() / backward() run them all.
    def _block_keys(self, x):
        """Sort keys (grid block, then cell; as ``block_order``) of [B, 3, P]
        positions."""
        n = self.p.n_grid
        base = torch.floor(x * n - 0.5).to(torch.int64).clamp_(0, n - 1)
        nb = n // 4
        blk = ((base[:, 0] // 4) * nb + base[:, 1] // 4) * nb + base[:, 2] // 4
        cell = ((base[:, 0] % 4) * 4 + base[:, 1] % 4) * 4 + base[:, 2] % 4
        key = blk * 64 + cell
        if self.uneven:       # padding slots stay behind each robot's particles
            pad = torch.arange(self.P, device=x.device)[None] >= self.n_part_t[:, None]
            key = torch.where(pad, torch.full_like(key, 1 << 62), key)
        return key

    def _set_order(self, s):
        self.act_id.copy_(torch.gather(self.act_id0, 1, self.orders[s]))
        self.group.copy_(torch.gather(self.group0, 1, self.orders[s]))

    def start_forward(self):
        r = self.ring
        if self.resort:
            self._set_order(0)
        r.x[0].copy_(self.x0 if self.uneven else self.x0[None].expand(self.B, -1, -1))
        r.v[0].zero_()
        r.C[0].zero_()
        r.F[0].zero_()
        for i in range(3):
            r.F[0][:, i * 3 + i] = 1.0
        self.sens.zero_()
        if self.save_horizon:
            self.hover.zero_()

    def forward_segment(self, s):
        r = self.ring
        t0, n = self._segments()[s]
        if self.resort and s > 0:
            pi = torch.argsort(self._block_keys(r.x[0]), dim=1)
            for f in r.fields():
                f[0].copy_(torch.gather(f[0], 2, pi[:, None, :].expand_as(f[0])))
            self.perms[s].copy_(pi)
            torch.gather(self.orders[s - 1], 1, pi, out=self.orders[s])
            self._set_order(s)
        flags = MPM_CONTROL | (MPM_SENSE_FIRST if s == 0 else 0) | \
            (MPM_SAVE_HORIZON if self.save_horizon else 0)
        if self.store_states:
            v = self.views[s]
            v.copy_slot(0, 0, r)
            nat.call("dl_mpm_forward", ctypes.byref(self.desc), ctypes.byref(v.c), t0, n, flags,
                     nat.stream())
            r.copy_slot(0, n, v)
            return
        self.ck.copy_slot(s, 0, r)
        nat.call("dl_mpm_forward", ctypes.byref(self.desc), ctypes.byref(r.c), t0, n, flags,
                 nat.stream())
        if n % 2:
            r.copy_slot(0, 1)

    def _loss_mask(self):
        """[T+1, B] mask of each robot's loss states (t in (T_k - W_k, T_k],
        W_k = min(window, T_k)) and the windows W_k."""
        hz = self.horizons_t.to(torch.int64)
        wk = torch.clamp(hz, max=self.window)
        t = torch.arange(self.T + 1, device=hz.device)[:, None]
        return ((t > hz - wk) & (t <= hz)).to(self.dtype), wk.to(self.dtype)

    def finish_forward(self):
        if self.horizons is not None:
            # per-robot horizons: each robot's windowed loss over its own
            # last states (the terms of dl_loss_task; torch on the device)
            m, wk = self._loss_mask()
            c, vc = self.center, self.vcom
            term = ((vc[..., 0] - self.tv[:, 0]) ** 2 + (vc[..., 2] - self.tv[:, 2]) ** 2
                    + (c[..., 1] - self.th) ** 2)
            self.loss_b.copy_((term * m).sum(0) / wk)
            self.loss.copy_(self.loss_b.mean().reshape(1))
            return self.loss
        nat.call("dl_loss_task", self.dc, self.B, self.T, self.window, nat.ptr(self.center),
                 nat.ptr(self.vcom), nat.ptr(self.tv), nat.ptr(self.th), nat.ptr(self.loss),
                 nat.ptr(self.loss_b), nat.stream())
        return self.loss

    def forward(self):
        self.start_forward()
        for s in range(self.nck):
            self.forward_segment(s)
        return self.finish_forward()

    def start_backward(self, loss_grad=1.0):
        """Seeds the adjoint with d(loss_grad * loss)/d(COM terms); with
        per-robot horizons ``loss_grad`` may be B values, robot k's loss_b[k]
        then weighted by loss_grad[k] (not divided by B)."""
        if self.save_horizon:
            # segments whose every step's grid was saved are replayed
            over = self.hover.cpu().numpy()
            self.replay = [not over[t0:t0 + n].any() for t0, n in self._segments()]
        self.cg.zero_()
        self.vg.zero_()
        self.in_grad.zero_()
        self.wgrad.zero_()
        for f in self.adj.fields():
            f[0].zero_()
            if self.horizons is not None:
                f[1].zero_()     # robots past their horizon keep a zero adjoint
        if self.horizons is not None:
            m, wk = self._loss_mask()
            g = torch.as_tensor(loss_grad, dtype=self.dtype, device=m.device)
            g = g / self.B if g.dim() == 0 else g
            s = m * (2.0 * g / wk)
            self.cg[..., 1].copy_(s * (self.center[..., 1] - self.th))
            self.vg[..., 0].copy_(s * (self.vcom[..., 0] - self.tv[:, 0]))
            self.vg[..., 2].copy_(s * (self.vcom[..., 2] - self.tv[:, 2]))
            return
        nat.call("dl_loss_task_grad", self.dc, self.B, self.T, self.window, nat.ptr(self.center),
                 nat.ptr(self.vcom), nat.ptr(self.tv), nat.ptr(self.th), float(loss_grad),
                 nat.ptr(self.cg), nat.ptr(self.vg), nat.stream())

    def backward_segment(self, s):
        """Adjoint of segment s (segments run last to first)."""
        t0, n = self._segments()[s]
        if self.resort:
            self._set_order(s)
        if self.store_states:
            seg = self.views[s]
            if self.replay[s]:
                self.desc.restore_from = 1      # states stored, grids saved: adjoint only
            else:
                nat.call("dl_mpm_forward", ctypes.byref(self.desc), ctypes.byref(seg.c), t0, n,
                         MPM_SAVE_GRID if self.save_grid else 0, nat.stream())
        else:
            seg = self.seg
            seg.copy_slot(0, s, self.ck)
            if self.replay[s]:
                self.desc.restore_from = 1
                nat.call("dl_mpm_forward", ctypes.byref(self.desc), ctypes.byref(seg.c), t0, n,
                         MPM_REPLAY, nat.stream())
            else:
                nat.call("dl_mpm_forward", ctypes.byref(self.desc), ctypes.byref(seg.c), t0, n,
                         MPM_SAVE_GRID if self.save_grid else 0, nat.stream())
        nat.call("dl_mpm_backward", ctypes.byref(self.desc), ctypes.byref(seg.c),
                 ctypes.byref(self.adj.c), t0, n, nat.stream())
        self.desc.restore_from = 0
        if self.resort and s > 0:
            # adjoint of state t0 back to the previous segment's order
            pi = self.perms[s]
            for f in self.adj.fields():
                f[0].copy_(torch.empty_like(f[0]).scatter_(2, pi[:, None, :].expand_as(f[0]), f[0]))

    def finish_backward(self):
        H, I, O, B = self.H, self.I, self.O, self.B
        if self.per_robot_weights:
            parts = torch.split(self.wgrad, [H * I, H, O * H, O], dim=1)
            self.grads = [parts[0].reshape(B, H, I), parts[1], parts[2].reshape(B, O, H),
                          parts[3]]
            return self.grads
        g = self.wgrad.sum(0)
        parts = torch.split(g, [H * I, H, O * H, O])
        self.grads = [parts[0].view(H, I), parts[1], parts[2].view(O, H), parts[3]]
        return self.grads

    def backward(self, loss_grad=1.0):
        self.start_backward(loss_grad)
        if self.overlap_replay and any(self.replay):
            self._backward_overlapped()
        else:
            for s in reversed(range(self.nck)):
                self.backward_segment(s)
        return self.finish_backward()

    def _adjoint(self, s, seg):
        """Adjoint of segment s over the recomputed states in ``seg``."""
        t0, n = self._segments()[s]
        self.desc.restore_from = 1 if self.replay[s] else 0
        nat.call("dl_mpm_backward", ctypes.byref(self.desc), ctypes.byref(seg.c),
                 ctypes.byref(self.adj.c), t0, n, nat.stream())
        self.desc.restore_from = 0
        if self.resort and s > 0:
            pi = self.perms[s]
            for f in self.adj.fields():
                f[0].copy_(torch.empty_like(f[0]).scatter_(2, pi[:, None, :].expand_as(f[0]), f[0]))

    def _backward_overlapped(self):
        """Segments last to first; the replayed recompute of segment s - 1 is
        issued on the side stream before the adjoint of segment s, into the
        other segment buffer once the adjoint that last read that buffer
        (segment s + 1) has finished.  Segments that could not be replayed
        are recomputed in full on the main stream (they use the main grid)."""
        main, rs = torch.cuda.current_stream(), self.replay_stream
        bufs = [self.seg, self.seg2]
        d2 = MPMBatchC.from_buffer_copy(self.desc)
        d2.grid_vo, d2.blk_count, d2.restore_from = nat.ptr(self.grid_vo2), nat.ptr(self.blk_count2), 0
        done_adj, done_rec = {}, {}
        rs.wait_stream(main)

        def replay(s):
            t0, n = self._segments()[s]
            with torch.cuda.stream(rs):
                if s + 2 in done_adj:
                    rs.wait_event(done_adj[s + 2])
                buf = bufs[s % 2]
                buf.copy_slot(0, s, self.ck)
                nat.call("dl_mpm_forward", ctypes.byref(d2), ctypes.byref(buf.c), t0, n,
                         MPM_REPLAY, nat.stream())
                ev = torch.cuda.Event()
                ev.record(rs)
                done_rec[s] = ev

        last = self.nck - 1
        if self.replay[last]:
            replay(last)
        for s in range(last, -1, -1):
            if s > 0 and self.replay[s - 1]:
                replay(s - 1)
            if self.resort:
                self._set_order(s)
            buf = bufs[s % 2]
            if self.replay[s]:
                main.wait_event(done_rec[s])
            else:
                t0, n = self._segments()[s]
                if s + 2 in done_adj:
                    main.wait_event(done_adj[s + 2])
                buf.copy_slot(0, s, self.ck)
                nat.call("dl_mpm_forward", ctypes.byref(self.desc), ctypes.byref(buf.c), t0, n,
                         MPM_SAVE_GRID if self.save_grid else 0, nat.stream())
            self._adjoint(s, buf)
            ev = torch.cuda.Event()
            ev.record(main)
            done_adj[s] = ev
        main.wait_stream(rs)

    def nbytes(self):
        """Device bytes held by this rollout's buffers."""
        own = sum(t.numel() * t.element_size() for t in vars(self).values()
                  if isinstance(t, torch.Tensor))
        own += sum(t.numel() * t.element_size() for t in self.W)
        return own + sum(f.numel() * f.element_size()
                         for sl in (self.ring, self.ck, self.seg, self.adj,
                                    getattr(self, "seg2", None),
                                    getattr(self, "states", None))
                         if sl is not None for f in sl.fields())

    def check(self):
        e = int(self.err.item())
        if e & 1:
            raise nat.NativeError("MPM: a particle left the grid (stencil out of bounds)")
        if e:
            raise nat.NativeError(f"MPM: error flags {e:#x} (2: a scatter window wider than "
                                  f"{MPM_PB} particles)")


class _MPMRolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, W1, b1, W2, b2):
        batch.set_weights(W1, b1, W2, b2)
        loss = batch.forward()
        ctx.batch = batch
        return loss.clone()

    @staticmethod
    def backward(ctx, g):
        grads = ctx.batch.backward(loss_grad=float(g.reshape(-1)[0].item()))
        return (None,) + tuple(x.clone() for x in grads)


def mpm_rollout_loss(batch, W1, b1, W2, b2):
    """The batched MPM rollout's task loss (mean over robots, shape [1]) as a
    differentiable function of the controller weights: the forward runs
    ``MPMBatch.forward`` (checkpoints every K steps), the backward
    ``MPMBatch.backward`` (segment recompute + hand-written adjoints).  The
    per-robot losses are in ``batch.loss_b``."""
    return _MPMRolloutFn.apply(batch, W1, b1, W2, b2)


# ------------------------------------------------------------------ autotuning
def scatter_candidates(dtype=torch.float32):
    """Scatter configurations of the batched MPM kernels: commit mode (global
    atomics or tile store + gather), the box limit of the scatter windows
    (stencil-box nodes per workgroup), the LDS tile capacity (nodes; the
    largest that fits 128 KiB of LDS in FP64 is 4096) and the run-sum unroll.
    The first is the batch's default."""
    tiles = (2048, 4096) if dtype == torch.float64 else (2048, 4096, 8192)
    out = [{"scatter": SCATTER_ATOMIC, "box_limit": 512, "tile_nodes": 2048, "unroll": 3}]
    for sc, box, tn, un in itertools.product((SCATTER_ATOMIC, SCATTER_STORE), (256, 512, 1024),
                                             tiles, (1, 3)):
        c = {"scatter": sc, "box_limit": box, "tile_nodes": tn, "unroll": un}
        if c not in out and tn >= box:
            out.append(c)
    return out


def tune_scatter(params, positions, act_id, n_act, B, dtype=torch.float32, K=100, steps=None,
                 tuner=None, candidates=None, device="cuda"):
    """Pick the scatter configuration of a B-robot batch of this body by
    timing the forward and adjoint of ``steps`` (default 2 K) steps for every
    candidate of ``scatter_candidates`` (best of two runs after a warm-up);
    cached by the autotuner per body size, grid, batch, dtype and library
    build.  Returns the chosen keyword arguments of ``MPMBatch``."""
    from .autotune import Autotuner, library_tag
    tuner = tuner or Autotuner()
    cands = candidates or scatter_candidates(dtype)
    T = steps or 2 * K
    key = Autotuner.key(kernel="mpm_batch_scatter", P=len(positions), n_grid=params.n_grid,
                        B=B, dtype=str(dtype), K=K, T=T, lib=library_tag())
    rng = np.random.default_rng(0)
    tv = np.stack([rng.uniform(-0.08, 0.08, B), np.zeros(B), rng.uniform(-0.08, 0.08, B)], 1)
    th = rng.uniform(0.1, 0.25, B)

    def measure(c):
        mb = MPMBatch(params, positions, act_id, n_act, B, T, K=K, dtype=dtype, device=device,
                      save_horizon=False, **c)
        mb.set_targets(tv, th)
        g = torch.Generator().manual_seed(1)
        mb.set_weights(*[0.1 * torch.randn(w.shape, generator=g, dtype=torch.float64)
                         for w in mb.W])
        best = None
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for it in range(3):
            e0.record()
            mb.forward()
            mb.backward()
            e1.record()
            torch.cuda.synchronize()
            if it:
                best = min(best or 1e30, e0.elapsed_time(e1))
        mb.check()
        del mb
        torch.cuda.empty_cache()
        return best

    return tuner.pick(key, cands, measure, repeat=1)
