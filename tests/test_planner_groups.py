"""Host-only test of the table planner's pair groups: tests/plan_groups_check.cpp
calls sipgpu::rt::build_batch_image (the planner behind sipgpu_batch_plan_choice
with path 0, tail 1) on ladder- and ring-shaped batches and checks the task,
workspace, add-parts and tile structure it builds.  No device work."""
import os
import subprocess

from aces4_amd import sipgpu as sg

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_pair_group_plans(tmp_path):
    exe = str(tmp_path / "plan_groups_check")
    cmd = [HIPCC, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "aces4_amd", "csrc"), "-x", "hip",
           os.path.join(HERE, "plan_groups_check.cpp"), "-x", "none", sg.LIB_PATH, "-o", exe,
           "-Wl,-rpath," + os.path.dirname(sg.LIB_PATH)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "all ok" in r.stdout, r.stdout + r.stderr
    # the ladder splits into 16 groups of one pair per result block
    assert "blocks 16 pairs 256 subtasks 240 parts 16" in r.stdout
