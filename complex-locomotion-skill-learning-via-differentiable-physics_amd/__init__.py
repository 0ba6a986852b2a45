"""diffloco -- MI355X-native differentiable locomotion rollout.

Hand-written HIP kernels for gfx950 (``csrc/``, built into ``_lib/libdiffloco.so``)
behind a C ABI (``include/diffloco.h``) loaded with ctypes, and drop-in host
classes for the reference's solver / controller / environment interfaces:

* ``SolverMassSpring``  (multitask/solver_mass_spring.py)
* ``Model``             (multitask/nn.py)
* ``TaichiEnv``         (multitask/taichi_env.py)
* ``DiffPhyTrainer``    (multitask/diffphy_trainer.py, rollout part)
* ``ConfigSim``         (multitask/config_sim.py, derived keys)
* ``MassSpringBatch``   persistent batched rollout (many robots, one launch)
"""

from . import _native
from .config import ConfigSim, quadruped_config
from .fields import Field, Tape, lazy_grad
from .robot_design import (Design, VoxelRobot3D, build_quadruped, build_reference, build_voxels,
                           make_design, quadruped_cubes)

__all__ = ["ConfigSim", "quadruped_config", "Field", "Tape", "lazy_grad", "Design",
           "VoxelRobot3D", "build_quadruped", "build_reference", "build_voxels", "make_design",
           "quadruped_cubes", "SolverMassSpring", "Model", "TaichiEnv", "DiffPhyTrainer",
           "MassSpringBatch", "rollout_loss"]


def __getattr__(name):
    if name == "SolverMassSpring":
        from .solver_mass_spring import SolverMassSpring
        return SolverMassSpring
    if name == "Model":
        from .nn import Model
        return Model
    if name == "TaichiEnv":
        from .taichi_env import TaichiEnv
        return TaichiEnv
    if name == "DiffPhyTrainer":
        from .trainer import DiffPhyTrainer
        return DiffPhyTrainer
    if name in ("MassSpringBatch", "rollout_loss", "init_weights"):
        from . import rollout
        return getattr(rollout, name)
    raise AttributeError(name)
