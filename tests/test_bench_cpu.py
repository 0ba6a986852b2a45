"""bench.py on the host: the one-GPU contract and the algorithmic byte models
its roofline figures are computed from (checked against the per-term sums
written out by hand for the quadruped workloads)."""

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""


def test_mass_spring_bytes_quadruped():
    # N = 54 vertices, S = 323 springs, I = 179, H = 64, O = 4; 11 780 MLP
    # parameters shared by 1024 robots
    N, S, I, H, O, R = 54, 323, 179, 64, 4, 1024
    P = I * H + H * O + H + O
    assert P == 11780
    f, b = bench.algorithmic_bytes(N, S, I, H, O, 50, R)
    # forward: x, v read + written (48 N), a checkpoint every 50 steps
    # (24 N / 50), one pass over springs and weights per design
    assert f == pytest.approx(2592 + 25.92 + (6460 + 47120) / 1024)
    # adjoint: recompute (48 N) + adjoint state (72 N), two passes
    assert b == pytest.approx(6480 + (12920 + 94240) / 1024)
    assert (round(f), round(b)) == (2670, 6585)
    # stored rows: each step's 9N + I + 2H + 2O + 6 reals written by the
    # forward and read by the adjoint instead of the recompute
    fs, bs = bench.algorithmic_bytes(N, S, I, H, O, 1000, R, stored=True)
    rows = 4 * (9 * 54 + 179 + 128 + 8 + 6)
    assert rows == 3228
    assert fs == pytest.approx(2592 + 24 * 54 / 1000 + (6460 + 47120) / 1024 + rows)
    assert bs == pytest.approx(72 * 54 + rows + (12920 + 94240) / 1024)
    assert (round(fs), round(bs)) == (5874, 7221)
    # FP64 doubles every term
    f8, b8 = bench.algorithmic_bytes(N, S, I, H, O, 50, R, dtype_bytes=8)
    assert (f8, b8) == pytest.approx((2 * f, 2 * b))


def test_mpm_bytes():
    gp = 0.249
    # segment recompute: forward twice, adjoint once, checkpoint every K
    assert bench.mpm_bytes(gp, 100) == pytest.approx(2 * 196 + 292 + (2 * 56 + 112) * gp + 0.96)
    assert round(bench.mpm_bytes(gp, 100)) == 741
    # stored states: no recompute, no checkpoint copies
    assert bench.mpm_bytes(gp, 100, 0.0) == pytest.approx(196 + 292 + (56 + 112) * gp)
    assert round(bench.mpm_bytes(gp, 100, 0.0)) == 530
    # half the segments recomputed: halfway between
    assert bench.mpm_bytes(gp, 100, 0.5) == pytest.approx(
        (bench.mpm_bytes(gp, 100) + bench.mpm_bytes(gp, 100, 0.0)) / 2)
