"""Write the 3D mass-spring builder fixtures ``builder3d_fixtures.json``.

Runs the reference's own ``RobotDesignMassSpring3D`` (reference file
``multitask/robot_design.py:319-591``; it needs numpy, pandas and json, no
Taichi) in this container and records, for each case, the builder output:
vertex positions, springs ``[a, b, L0, stiffness, actuation]`` and faces,
exactly as the reference produces them (Python floats, written with repr
precision so they read back bit for bit).

Cases: the reference robots 100-104 (``build()``; robot 103 reads
``cfg3d/skeleton.json`` relative to the working directory, so the call runs
with the reference root as the working directory), the quadruped workload's
cube list, and three seeded random voxel robots of the 512-robot generator,
each replayed through ``add_mesh_square`` in cube order.

Usage (needs the reference checkout; nothing else reads it):

    python tests/golden/make_builder_fixtures.py /root/reference
"""

import json
import os
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]


def main(ref_root):
    sys.dont_write_bytecode = True            # nothing is written under the reference tree
    sys.path.insert(0, str(ref_root))
    from multitask.robot_design import RobotDesignMassSpring3D

    sys.path.insert(0, str(ROOT))
    from __graft_entry__ import load_package
    load_package()
    import numpy as np
    from diffloco.robot_design import quadruped_cubes, random_voxel_cubes

    def builder(rid, stiffness, actuation):
        return RobotDesignMassSpring3D({"robot": {"solver": "mass_spring", "id": rid,
                                                  "spring_stiffness": stiffness,
                                                  "spring_actuation": actuation}})

    cases = {}
    cwd = os.getcwd()
    for rid, (k, a) in {100: (1e4, 0.25), 101: (1e4, 0.25), 102: (1e4, 0.25),
                        103: (2e4, 0.2), 104: (1e4, 0.25)}.items():
        b = builder(rid, k, a)
        os.chdir(ref_root)
        try:
            objects, springs, faces = b.build()
        finally:
            os.chdir(cwd)
        cases[f"robot_{rid}"] = {"stiffness": k, "actuation": a, "cubes": None,
                                 "objects": [list(map(float, p)) for p in objects],
                                 "springs": [[int(s[0]), int(s[1]), float(s[2]), float(s[3]),
                                              float(s[4])] for s in springs],
                                 "faces": [list(map(int, f)) for f in faces]}

    def from_cubes(name, cubes, k=3e4, a=0.1):
        b = builder(0, k, a)
        for (i, j, kk), act, _ in cubes:
            b.add_mesh_square(i, j, kk, actuation=a if act else 0.0)
        cases[name] = {"stiffness": k, "actuation": a,
                       "cubes": [[list(c), bool(act), int(g)] for c, act, g in cubes],
                       "objects": [list(map(float, p)) for p in b.objects],
                       "springs": [[int(s[0]), int(s[1]), float(s[2]), float(s[3]), float(s[4])]
                                   for s in b.springs],
                       "faces": [list(map(int, f)) for f in b.faces]}

    from_cubes("quadruped", quadruped_cubes())
    rng = np.random.default_rng(2024)
    for n_cubes, n_groups in ((12, 4), (60, 16), (230, 40)):
        cubes, _ = random_voxel_cubes(rng, n_cubes, n_groups)
        from_cubes(f"random_{n_cubes}", cubes)

    out = HERE / "builder3d_fixtures.json"
    with open(out, "w") as fh:
        json.dump(cases, fh, separators=(",", ":"))
    print(f"wrote {out}: " + ", ".join(f"{n} N={len(c['objects'])} S={len(c['springs'])}"
                                       for n, c in cases.items()))


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference").resolve())
