"""ctypes binding of ``libdiffloco.so`` (C ABI declared in ``include/diffloco.h``).

The library is the only compute path: every call raises ``NativeError`` when
the library is missing or an entry point reports a failure.  There is no
Python or CPU fallback.
"""

import ctypes
from ctypes import POINTER, Structure, c_double, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

import numpy as np
import torch

ABI_VERSION = 1
LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libdiffloco.so"

F32, F64 = 0, 1
ACT_SIN, ACT_TANH = 0, 1


class NativeError(RuntimeError):
    pass


class MSPhysics(Structure):
    _fields_ = [("dt", c_double), ("gravity", c_double), ("ground", c_double),
                ("drag", c_double), ("dashpot", c_double), ("friction", c_double)]


class SensorLayoutC(Structure):
    _fields_ = [("n_sin", c_int32), ("run_period", c_int32), ("vcom", c_int32),
                ("vert_pad", c_int32), ("n_tv", c_int32), ("tv_comp", c_int32 * 3),
                ("dup_v", c_int32), ("dup_h", c_int32), ("dup_c", c_int32), ("pad_", c_int32),
                ("offset_scale", c_double), ("max_speed", c_double)]


class MSTopology(Structure):
    _fields_ = [("n_vert", c_int32), ("n_spring", c_int32), ("n_act", c_int32),
                ("pad_", c_int32),
                ("spring_a", c_void_p), ("spring_b", c_void_p), ("spring_act", c_void_p),
                ("spring_L0", c_void_p), ("spring_k", c_void_p), ("spring_coeff", c_void_p),
                ("vert_ptr", c_void_p), ("vert_idx", c_void_p),
                ("act_ptr", c_void_p), ("act_idx", c_void_p)]


# dl_robot as a numpy structured dtype (the table is built on the host and
# copied to the device as raw bytes); layout checked against ctypes below.
ROBOT_DTYPE = np.dtype([
    ("n_vert", "<i4"), ("n_spring", "<i4"), ("n_act", "<i4"), ("n_in", "<i4"),
    ("horizon", "<i4"), ("t_fwd", "<i4"), ("t_bwd", "<i4"), ("spring0", "<i4"),
    ("vptr0", "<i4"), ("aptr0", "<i4"),
    ("weights", "<i8"), ("wgrad", "<i8"), ("state", "<i8"), ("ckpt", "<i8"),
    ("scratch", "<i8"), ("targets", "<i8"), ("traj", "<i8"),
    ("target_stride", "<i4"), ("pad_", "<i4"),
])


class RobotC(Structure):
    _fields_ = [(n, c_int32) for n in ("n_vert", "n_spring", "n_act", "n_in", "horizon", "t_fwd",
                                       "t_bwd", "spring0", "vptr0", "aptr0")] + \
               [(n, c_int64) for n in ("weights", "wgrad", "state", "ckpt", "scratch",
                                       "targets", "traj")] + \
               [("target_stride", c_int32), ("pad_", c_int32)]


assert ctypes.sizeof(RobotC) == ROBOT_DTYPE.itemsize, "dl_robot layout mismatch"


class Rollout(Structure):
    _fields_ = [(n, c_int32) for n in ("dtype", "act", "n_robots", "n_hidden", "ckpt_every",
                                       "max_steps", "block", "loss_kind", "loss_window",
                                       "store_traj", "max_n", "max_s", "max_i", "max_o",
                                       "lds_policy", "variant")] + \
               [("phys", MSPhysics), ("lay", SensorLayoutC)] + \
               [(n, c_void_p) for n in ("robots", "spring_a", "spring_b", "spring_act",
                                        "spring_L0", "spring_k", "spring_coeff", "vert_ptr",
                                        "vert_idx", "act_ptr", "act_idx", "weights", "wgrad",
                                        "state", "adjoint", "ckpt", "scratch", "targets", "loss",
                                        "loss_grad", "traj", "traj_grad", "err")]

ROLLOUT_ERR_HORIZON, ROLLOUT_ERR_PACKING = 1, 2
OPT_MLP_MFMA = 1


_P = c_void_p
_I = c_int32
_SIGS = {
    "dl_abi_version": ([], c_int32),
    "dl_last_error": ([ctypes.c_char_p, c_size_t], c_int32),
    "dl_device_count": ([POINTER(c_int32)], c_int32),
    "dl_ms_center": ([_I, _I, _I, _P, _P, _P, _P, _P], c_int32),
    "dl_ms_center_grad": ([_I, _I, _I, _P, _P, _P, _P, _P], c_int32),
    "dl_ms_height": ([_I, _I, _I, _I, _I, _P, _P, _P, _P, _P], c_int32),
    "dl_ms_height_grad": ([_I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P], c_int32),
    "dl_ms_rotation": ([_I, _I, _I, _I, _P, _P, _P, _P, _P], c_int32),
    "dl_ms_rotation_grad": ([_I, _I, _I, _I, _P, _P, _P, _P, _P, _P], c_int32),
    "dl_nn_input": ([_I, POINTER(SensorLayoutC), _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P],
                    c_int32),
    "dl_nn_input_grad": ([_I, POINTER(SensorLayoutC), _I, _I, _P, _P, _P, _P, _P], c_int32),
    "dl_mlp_forward": ([_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P], c_int32),
    "dl_mlp_backward": ([_I, _I, _I, _I, _I, _I] + [_P] * 13, c_int32),
    "dl_set_option": ([_I, _I], c_int32),
    "dl_mlp_backward_static": ([_I, _I, _I, _I, _I, _I] + [_P] * 12, c_int32),
    "dl_ms_advance": ([_I, POINTER(MSTopology), POINTER(MSPhysics), _I, _P, _P, _P, _P, _P, _P],
                      c_int32),
    "dl_ms_advance_grad": ([_I, POINTER(MSTopology), POINTER(MSPhysics), _I] + [_P] * 9, c_int32),
    "dl_loss_velocity": ([_I] * 7 + [_P] * 5, c_int32),
    "dl_loss_velocity_grad": ([_I] * 7 + [_P, _P, c_double, _P, _P], c_int32),
    "dl_loss_rotation": ([_I] * 8 + [_P] * 7, c_int32),
    "dl_loss_rotation_grad": ([_I] * 8 + [_P, _P, _P, c_double, _P, _P, _P], c_int32),
    "dl_loss_height": ([_I] * 6 + [_P] * 7, c_int32),
    "dl_loss_height_grad": ([_I] * 6 + [_P, _P, c_double, _P, _P], c_int32),
    "dl_loss_pose": ([_I] * 7 + [_P] * 6, c_int32),
    "dl_loss_pose_grad": ([_I] * 7 + [_P, _P, c_double, _P, _P], c_int32),
    "dl_loss_actuation": ([_I] * 6 + [_P] * 7, c_int32),
    "dl_loss_actuation_grad": ([_I] * 6 + [_P, _P, _P, c_double, _P, _P], c_int32),
    "dl_loss_crawl": ([_I] * 5 + [_P] * 6, c_int32),
    "dl_loss_crawl_grad": ([_I] * 5 + [_P, _P, c_double, _P, _P], c_int32),
    "dl_loss_task": ([_I] * 4 + [_P] * 7, c_int32),
    "dl_loss_task_grad": ([_I] * 4 + [_P] * 4 + [c_double] + [_P] * 3, c_int32),
    "dl_weights_size": ([_I, _I, _I], c_int64),
    "dl_wgrad_size": ([_I, _I, _I], c_int64),
    "dl_scratch_size": ([_I] * 6, c_int64),
    "dl_ms_rollout_forward": ([POINTER(Rollout), _P], c_int32),
    "dl_ms_rollout_backward": ([POINTER(Rollout), _P], c_int32),
    "dl_wgrad_reduce": ([_I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P], c_int32),
    "dl_mpm_fill": ([_I, _P, c_int64, c_double, _P], c_int32),
    "dl_mpm_p2g": ([_I, _P, _I, _I] + [_P] * 6 + [_I] + [_P] * 5, c_int32),
    "dl_mpm_grid": ([_I, _P, _I] + [_P] * 4, c_int32),
    "dl_mpm_g2p": ([_I, _P, _I, _I] + [_P] * 6, c_int32),
    "dl_mpm_g2p_adj": ([_I, _P, _I, _I] + [_P] * 8, c_int32),
    "dl_mpm_grid_adj": ([_I, _P, _I] + [_P] * 6, c_int32),
    "dl_mpm_p2g_adj": ([_I, _P, _I, _I] + [_P] * 6 + [_I] + [_P] * 9, c_int32),
    "dl_mpm_sensor": ([_I] * 7 + [c_double] * 2 + [_P] * 9, c_int32),
    "dl_mpm_sensor_adj": ([_I] * 5 + [c_double] + [_P] * 6, c_int32),
    "dl_mpm_forward": ([_P, _P, _I, _I, _I, _P], c_int32),
    "dl_mpm_backward": ([_P, _P, _P, _I, _I, _P], c_int32),
}

_lib = None


def lib():
    """Load the library once (raises NativeError if it is not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        if L.dl_abi_version() != ABI_VERSION:
            raise NativeError(f"ABI version {L.dl_abi_version()} != {ABI_VERSION}")
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib().dl_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise NativeError(f"{name} failed ({rc}): {last_error()}")
    return rc


def dtype_code(dtype):
    if dtype == torch.float32:
        return F32
    if dtype == torch.float64:
        return F64
    raise NativeError(f"unsupported dtype {dtype}")


def ptr(t):
    """Device pointer of a contiguous CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise NativeError("tensor must live on the GPU")
    if not t.is_contiguous():
        raise NativeError("tensor must be contiguous")
    return t.data_ptr()


def stream():
    return torch.cuda.current_stream().cuda_stream
