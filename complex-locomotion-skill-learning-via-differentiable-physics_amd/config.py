"""Simulation configuration (``multitask/config_sim.py`` equivalent, no I/O).

``ConfigSim.from_file(path)`` reads a reference experiment JSON (for example
``cfg3d/sim_quad.json``) and derives the same keys as
``ConfigSim._add_adaptive_configs`` (``multitask/config_sim.py:67-147``):
robot objects/springs, dt, sinusoid frequency, gravity and dashpot overrides,
the sensor-vector width.  Keys the reference hard-codes in its kernels are
exposed with the reference's values as defaults (friction 0.4,
``solver_mass_spring.py:136``; stiffness / 10, ``:64``; sin activations,
``nn.py:139,146``).

``quadruped_config`` builds the quadruped workload: a 4x1x2-cube body on
four 1x2x1 legs, 4 actuator groups, tanh MLP on a compact sensor vector,
windowed COM loss.
"""

import copy
import json
import math
import os

from .robot_design import build_quadruped, build_reference


class ConfigSim:
    def __init__(self, config, base_dir=None):
        self._config = copy.deepcopy(config)
        self._base_dir = base_dir
        self._add_adaptive_configs()

    @classmethod
    def from_file(cls, file_name):
        """Read an experiment file; robot 103's skeleton graph is read from
        ``skeleton.json`` beside it (the reference reads ``cfg3d/skeleton.json``,
        where its experiment files live) unless ``robot.skeleton_file`` names one."""
        with open(file_name) as fh:
            return cls(json.load(fh), base_dir=os.path.dirname(os.path.abspath(file_name)))

    def get_config(self):
        return self._config

    @property
    def config(self):
        return self._config

    def _add_adaptive_configs(self):
        c = self._config
        reference = c.get("preset", "reference") == "reference"
        rob, sim, proc, nn = c["robot"], c["simulator"], c["process"], c["nn"]
        if rob.get("simulator", "mass_spring") == "mpm":
            self._add_mpm_configs()
            return
        if rob.get("simulator", "mass_spring") != "mass_spring" or rob.get("dim", 3) != 3:
            raise NotImplementedError("mass-spring experiment files are 3D here")
        if rob.get("design") == "quadruped":
            design = build_quadruped(rob.get("spring_stiffness", 3e4), rob.get("spring_actuation", 0.1))
        else:
            skeleton = None
            if rob["robot_id"] == 103:
                skeleton = rob.get("skeleton_file") or os.path.join(self._base_dir or ".",
                                                                   "skeleton.json")
            design = build_reference(rob["robot_id"], rob["spring_stiffness"],
                                     rob["spring_actuation"], skeleton=skeleton)
        rob["design_obj"] = design
        rob["objects"] = [tuple(p) for p in design.objects.tolist()]
        rob["springs"] = [[int(a), int(b), float(L), float(k), float(s)] for a, b, L, k, s in zip(
            design.spring_a, design.spring_b, design.L0, design.stiffness, design.coeff)]
        rob["faces"] = design.faces
        rob["n_objects"] = design.n_vert
        rob["n_springs"] = design.n_spring
        rob["n_actuators"] = design.n_act
        if reference:
            # config_sim.py:112-121
            proc["dt"] = 0.004
            sim["gravity"] = -1.8
            sim["dashpot_damping"] = 0.1
        proc["spring_omega"] = 2 * math.pi / proc["dt"] / proc["run_period"]
        sim.setdefault("friction", 0.4)
        sim.setdefault("stiffness_scale", 0.1)
        sim.setdefault("drag_damping", 0.0)
        nn.setdefault("activation", "sin")
        nn.setdefault("init", "uniform")
        nn.setdefault("vcom", 0)
        nn.setdefault("vert_pad", 3)
        nn.setdefault("tv_comps", [0, 2])
        nn.setdefault("offset_scale", 0.2)
        n_tv = len(nn["tv_comps"])
        nn["n_input_states"] = (nn["n_sin_waves"] + 3 * nn["vcom"] +
                                design.n_vert * (3 + nn["vert_pad"]) + nn["duplicate_v"] * n_tv +
                                nn["duplicate_h"] + nn["duplicate_c"])
        nn["adam_a"] = nn["learning_rate"]
        c.setdefault("precision", "f64" if reference else "f32")


    def _add_mpm_configs(self):
        """config_sim.py:80-82, 104-129 for the MPM solver: the particle
        design (RobotDesignMPM, robot_design.py:38-110) from
        ``robot.design_file`` (relative to the experiment file) or an inline
        ``robot.design``; objects = particle positions, springs = their
        actuator ids, n_springs = their count (the reference overwrites it
        with the particle count), dt 1e-3, gravity -10, dashpot 0.2 (2D),
        dx = 1 / n_grid, and the sensor width of taichi_env.py:149-183."""
        from .mpm import particles_2d
        c = self._config
        rob, sim, proc, nn = c["robot"], c["simulator"], c["process"], c["nn"]
        design = rob.get("design")
        if design is None:
            path = rob["design_file"]
            if not os.path.isabs(path) and self._base_dir is not None:
                cand = os.path.join(self._base_dir, os.path.basename(path))
                path = cand if os.path.exists(cand) else path
            with open(path) as fh:
                design = json.load(fh)
        dim = rob.setdefault("dim", 2)
        if dim != 2:
            raise NotImplementedError("MPM experiment files describe 2D particle designs")
        pos, ids, n_act = particles_2d(design)
        rob["objects"] = [tuple(p) for p in pos.tolist()]
        rob["springs"] = [int(i) for i in ids]
        rob["faces"] = []
        rob["n_objects"] = len(pos)
        rob["n_springs"] = len(ids)
        rob["n_actuators"] = int(n_act)
        proc["dt"] = 0.001
        proc["spring_omega"] = 2 * math.pi / proc["dt"] / proc["run_period"]
        sim["gravity"] = -10.0
        sim["dashpot_damping"] = 0.2
        sim["n_particles"] = rob["n_objects"]
        sim["mpm"]["dx"] = 1 / sim["mpm"]["n_grid"]
        sim["mpm"]["inv_dx"] = 1 / sim["mpm"]["dx"]
        nn["n_input_states"] = (nn["n_sin_waves"] + dim * 2 * rob["n_objects"]
                                + nn["duplicate_v"] * (dim - 1) + nn["duplicate_h"]
                                + nn["duplicate_c"])
        nn["adam_a"] = nn["learning_rate"]
        c.setdefault("precision", "f64")


def quadruped_config(batch_size=32, steps=1000, window=500, precision="f32"):
    cfg = {
        "preset": "quadruped",
        "precision": precision,
        "train": {"task": ["task"], "window": window},
        "robot": {"dim": 3, "robot_id": 0, "simulator": "mass_spring", "design": "quadruped",
                  "control_length": 1, "spring_stiffness": 3e4, "spring_actuation": 0.1},
        "process": {"max_steps": steps, "dt": 0.004, "run_period": 100, "jump_period": 500,
                    "turn_period": 500, "max_speed": 0.08, "max_height": 0.25},
        "simulator": {"ground_height": 0.0, "gravity": -1.8, "drag_damping": 0.0,
                      "dashpot_damping": 30 * 0.004, "friction": 0.5, "stiffness_scale": 0.1},
        "nn": {"n_sin_waves": 10, "n_hidden": 64, "duplicate_v": 1, "duplicate_h": 1,
               "duplicate_c": 0, "batch_size": batch_size, "optimizer": "adam",
               "learning_rate": 3e-4, "adam_b1": 0.9, "adam_b2": 0.9, "activation": "tanh",
               "init": "normal", "vcom": 1, "vert_pad": 0, "tv_comps": [0, 1, 2],
               "offset_scale": 0.2},
    }
    return ConfigSim(cfg)
