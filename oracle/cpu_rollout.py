"""ctypes binding of the C restatement of the training rollout
(oracle/cpu_rollout.c, built by oracle/build_cpu.py).

TEST INFRASTRUCTURE AND CPU BASELINE ONLY (see oracle/__init__.py): used by
tests/ against the torch oracle and by the cpu_baseline leg of bench.py.
"""

import ctypes
from ctypes import POINTER, Structure, c_double, c_float, c_int32, c_void_p

import numpy as np

from . import build_cpu

_LIBS = {}


def _cfg_type(real):
    class Cfg(Structure):
        _fields_ = ([(n, c_int32) for n in ("N", "S", "O", "H", "I")]
                    + [(n, c_void_p) for n in ("a", "b", "act_index", "L0", "k", "act_coeff")]
                    + [(n, real) for n in ("dt", "gravity", "ground", "drag", "dashpot",
                                           "friction", "stiffness_scale")]
                    + [("n_sin", c_int32), ("run_period", c_int32),
                       ("offset_scale", real), ("max_speed", real),
                       ("act_tanh", c_int32), ("T", c_int32), ("W", c_int32), ("K", c_int32)])
    return Cfg


def lib(dtype="f32"):
    if dtype not in _LIBS:
        path = build_cpu.lib_path(dtype)
        if not path.exists():
            raise RuntimeError(f"{path} is not built (python oracle/build_cpu.py)")
        L = ctypes.CDLL(str(path))
        real = c_float if dtype == "f32" else c_double
        Cfg = _cfg_type(real)
        P = c_void_p
        L.cr_rollout.argtypes = [POINTER(Cfg), c_int32, P, P, P, P, P, P, P, P, P, c_int32]
        L.cr_rollout.restype = c_int32
        L.cr_adam.argtypes = [c_int32, P, P, P, P, c_int32, real, real, real, real]
        L.cr_adam.restype = None
        L.cr_param_count.argtypes = [c_int32, c_int32, c_int32]
        L.cr_param_count.restype = c_int32
        _LIBS[dtype] = (L, Cfg, np.float32 if dtype == "f32" else np.float64)
    return _LIBS[dtype]


class CpuRollout:
    """One design (an oracle ``Robot``) on the quadruped task configuration
    (an oracle ``TaskConfig``), checkpoints every K steps.  ``run`` returns
    (mean loss, per-task losses, flat gradient in the order W1, b1, W2, b2);
    ``adam`` applies one Adam step to the held weights."""

    def __init__(self, robot, cfg, K, dtype="f32"):
        self.L, Cfg, self.np_t = lib(dtype)
        f = lambda t: np.ascontiguousarray(np.asarray(t, dtype=self.np_t))
        i32 = lambda t: np.ascontiguousarray(np.asarray(t, dtype=np.int32))
        self._keep = [i32(robot.a), i32(robot.b), i32(robot.act_index), f(robot.L0),
                      f(robot.k), f(robot.act_coeff)]
        lay, ph = cfg.layout, cfg.physics
        assert (lay.vcom and lay.vert_pad == 0 and tuple(lay.tv_comps) == (0, 1, 2)
                and lay.dup_v == 1 and lay.dup_h == 1 and lay.dup_c == 0), \
            "the C restatement covers the quadruped task's sensor layout"
        self.N, self.O = robot.n_vert, int(max(robot.act_index)) + 1
        self.I = lay.n_input(self.N)
        self.cfg = Cfg(self.N, robot.n_spring, self.O, 0, self.I,
                       *[a.ctypes.data for a in self._keep],
                       ph.dt, ph.gravity, ph.ground, ph.drag, ph.dashpot, ph.friction,
                       ph.stiffness_scale, lay.n_sin, lay.run_period, lay.offset_scale,
                       lay.max_speed, 1 if cfg.act == "tanh" else 0, cfg.steps, cfg.window, K)
        self.W = None

    def set_weights(self, W1, b1, W2, b2):
        f = lambda t: np.ascontiguousarray(np.array(t, dtype=self.np_t))
        self.W = [f(W1), f(b1), f(W2), f(b2)]
        self.cfg.H = self.W[0].shape[0]
        assert self.W[0].shape[1] == self.I and self.W[2].shape == (self.O, self.cfg.H)
        n = sum(w.size for w in self.W)
        self.m, self.v = np.zeros(n, self.np_t), np.zeros(n, self.np_t)
        self.it = 0

    def run(self, x0, tv, th, threads=0):
        """x0 [B, N, 3] initial positions (velocities zero), tv [B, 3], th [B]."""
        f = lambda t: np.ascontiguousarray(np.asarray(t, dtype=self.np_t))
        x0, tv, th = f(x0), f(tv), f(th)
        B = x0.shape[0]
        assert x0.shape == (B, self.N, 3) and tv.shape == (B, 3) and th.shape == (B,)
        loss_b = np.zeros(B, self.np_t)
        P = self.L.cr_param_count(self.I, self.cfg.H, self.O)
        grad = np.zeros(P, self.np_t)
        rc = self.L.cr_rollout(ctypes.byref(self.cfg), B, x0.ctypes.data, tv.ctypes.data,
                               th.ctypes.data, *[w.ctypes.data for w in self.W],
                               loss_b.ctypes.data, grad.ctypes.data, threads)
        if rc != 0:
            raise MemoryError("cr_rollout: out of memory")
        return float(loss_b.astype(np.float64).sum() / B), loss_b, grad

    def split(self, grad):
        out, o = [], 0
        for w in self.W:
            out.append(grad[o:o + w.size].reshape(w.shape))
            o += w.size
        return out

    def adam(self, grad, lr=3e-4, b1=0.9, b2=0.9, eps=1e-8):
        flat = np.concatenate([w.ravel() for w in self.W])
        g = np.ascontiguousarray(grad, dtype=self.np_t)
        self.L.cr_adam(flat.size, flat.ctypes.data, g.ctypes.data, self.m.ctypes.data,
                       self.v.ctypes.data, self.it, lr, b1, b2, eps)
        self.it += 1
        o = 0
        for w in self.W:
            w[...] = flat[o:o + w.size].reshape(w.shape)
            o += w.size
