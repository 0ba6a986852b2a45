"""Build the C restatement of the training rollout (oracle/cpu_rollout.c)
into oracle/_lib/libcpuroll_f32.so and libcpuroll_f64.so with gcc and
OpenMP.  Test infrastructure and CPU baseline only; run by
``__graft_entry__.build()`` or directly (``python oracle/build_cpu.py``)."""

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_lib"
SRC = HERE / "cpu_rollout.c"
# x86-64-v3 (AVX2 + FMA): the EPYC hosts of MI355X nodes and this build host
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-ffp-contract=fast", "-Wall",
         "-Wno-unused-function"]


def lib_path(real):
    return OUT / f"libcpuroll_{real}.so"


def build(verbose=False):
    OUT.mkdir(exist_ok=True)
    for real, ctype in (("f32", "float"), ("f64", "double")):
        out = lib_path(real)
        tmp = out.with_suffix(".so.tmp")
        cmd = [os.environ.get("CC", "gcc"), *FLAGS, f"-DREAL={ctype}", str(SRC), "-o", str(tmp),
               "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"gcc failed for {SRC.name} ({real}):\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, out)
        if verbose:
            print(f"built {out}")


if __name__ == "__main__":
    build(verbose=True)
