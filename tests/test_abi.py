"""C ABI: the library loads without a GPU, exports every entry point the
header declares, and the ctypes structures match the C layouts."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from diffloco import _native as nat
from diffloco import mpm as dm

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "diffloco.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int32_t|int64_t)\s+(dl_\w+)\s*\(", text, re.M)))


def test_library_loads_and_version():
    L = nat.lib()
    assert L.dl_abi_version() == nat.ABI_VERSION


def test_every_declared_symbol_is_exported_and_bound():
    L = nat.lib()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(nat.declared_symbols())


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "diffloco.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(dl_ms_physics), sizeof(dl_sensor_layout),
         sizeof(dl_ms_topology), sizeof(dl_robot), sizeof(dl_rollout));
  printf("%zu %zu %zu %zu\n", offsetof(dl_rollout, phys), offsetof(dl_rollout, lay),
         offsetof(dl_rollout, robots), offsetof(dl_rollout, traj_grad));
  printf("%zu %zu\n", offsetof(dl_robot, weights), offsetof(dl_robot, target_stride));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dl_mpm_params), sizeof(dl_mpm_batch),
         sizeof(dl_mpm_slots), offsetof(dl_mpm_batch, params), offsetof(dl_mpm_batch, act_id),
         offsetof(dl_mpm_batch, wgrad));
  printf("%zu %zu %zu\n", offsetof(dl_mpm_batch, hid_grad), offsetof(dl_mpm_batch, windows),
         offsetof(dl_mpm_batch, n_windows));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = list(map(int, out))
    exp = [ctypes.sizeof(nat.MSPhysics), ctypes.sizeof(nat.SensorLayoutC),
           ctypes.sizeof(nat.MSTopology), ctypes.sizeof(nat.RobotC), ctypes.sizeof(nat.Rollout),
           nat.Rollout.phys.offset, nat.Rollout.lay.offset, nat.Rollout.robots.offset,
           nat.Rollout.traj_grad.offset, nat.RobotC.weights.offset,
           nat.RobotC.target_stride.offset,
           ctypes.sizeof(dm.MPMParamsC), ctypes.sizeof(dm.MPMBatchC), ctypes.sizeof(dm.MPMSlotsC),
           dm.MPMBatchC.params.offset, dm.MPMBatchC.act_id.offset, dm.MPMBatchC.wgrad.offset,
           dm.MPMBatchC.hid_grad.offset, dm.MPMBatchC.windows.offset,
           dm.MPMBatchC.n_windows.offset]
    assert got == exp
    assert nat.ROBOT_DTYPE.itemsize == got[3]


def test_size_helpers():
    L = nat.lib()
    n = 2 * 64 * 179 + 64 + 2 * 4 * 64 + 4
    assert L.dl_weights_size(179, 64, 4) == (n + 15) // 16 * 16
    assert all(L.dl_weights_size(I, 64, O) % 16 == 0 for I in (17, 179, 2384) for O in (1, 4, 7))
    assert L.dl_wgrad_size(179, 64, 4) == 64 * 179 + 64 + 4 * 64 + 4
    assert L.dl_scratch_size(54, 323, 179, 64, 4, 50) == (2 * 50 * (9 * 54 + 179 + 3 * 64 + 3 * 4 + 6)
                                                        + 10 * 323 + 4)


def test_invalid_arguments_are_rejected_without_a_gpu():
    """Argument validation runs before any device work: NULL descriptors and
    bad sizes return DL_E_ARG with a message."""
    L = nat.lib()
    assert L.dl_ms_rollout_forward(None, None) == -1
    assert "NULL" in nat.last_error()
    assert L.dl_ms_advance(0, None, None, 1, None, None, None, None, None, None) == -1
    assert L.dl_mlp_forward(0, 0, 0, 1, 1, 1, None, None, None, None, None, None, None, None) == -1
    D = nat.Rollout()
    D.dtype = 7
    assert L.dl_ms_rollout_forward(ctypes.byref(D), None) == -1
    assert "dtype" in nat.last_error()
    with pytest.raises(nat.NativeError):
        nat.call("dl_ms_center", 0, 0, 10, None, None, None, None, None)
    # batched MPM: descriptor checks come first
    assert L.dl_mpm_forward(None, None, 0, 1, 0, None) == -1
    mb = dm.MPMBatchC(dtype=0, B=1, P=1, n_act=1, horizon=10,
                      params=dm.quadruped3d_params(n_grid=30))
    assert L.dl_mpm_forward(ctypes.byref(mb), None, 0, 1, 0, None) == -1
    assert "multiple of 4" in nat.last_error()


def test_empty_batches_are_argument_errors():
    """A batch with no robots, or an MPM batch with no steps, is rejected
    with DL_E_ARG before any launch (not run as an empty grid)."""
    L = nat.lib()
    D = nat.Rollout()
    D.dtype, D.act, D.n_robots = nat.F32, nat.ACT_TANH, 0
    for fn in (L.dl_ms_rollout_forward, L.dl_ms_rollout_backward):
        assert fn(ctypes.byref(D), None) == -1
        assert "robot count" in nat.last_error()
    for B, horizon in ((0, 10), (1, 0)):
        mb = dm.MPMBatchC(dtype=0, B=B, P=64, n_act=1, horizon=horizon, tile_nodes=2048,
                          scatter=dm.SCATTER_ATOMIC, params=dm.quadruped3d_params(n_grid=32))
        assert L.dl_mpm_forward(ctypes.byref(mb), None, 0, 1, 0, None) == -1
        assert "bad sizes" in nat.last_error()
