"""FP64 CPU oracle for the differentiable locomotion rollout.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import, call
or execute anything under this directory; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the reference results are compared against (or as the timed
CPU baseline), never as the thing measured or shipped.

The oracle restates the reference's algorithm (Python + Taichi, FP64
everywhere: ``multitask/utils.py:12-15``) in FP64 on the CPU:

* ``builders``    -- robot builders (``multitask/robot_design.py``).
* ``mass_spring`` -- mass-spring step, sensors, MLP, losses, rollout
  (``multitask/solver_mass_spring.py``, ``multitask/taichi_env.py``,
  ``multitask/nn.py``, ``multitask/diffphy_trainer.py``).  Written with torch
  FP64 CPU tensors so that ``torch.autograd`` over the restated forward stands
  in for Taichi's generated adjoints (Taichi is not installed here).
* ``mpm``         -- MLS-MPM step, 2D and 3D (``multitask/solver_mpm.py``).
* ``cpu_rollout`` -- ctypes binding of ``cpu_rollout.c``: the quadruped
  training rollout restated in C with OpenMP (forward, hand-written adjoint,
  checkpoints with segment recompute, Adam), built in FP32 and FP64 by
  ``build_cpu.py``.  The FP32 build is the benchmark's CPU baseline; the
  FP64 build is pinned against ``mass_spring`` autograd
  (``tests/test_cpu_rollout.py``).

Pinning (what fixes each function independently of its own formula) is listed
in each module's header and exercised by ``tests/test_oracle_*.py``.
"""
