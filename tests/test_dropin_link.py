"""Link test of the drop-in layer: tests/dropin_caller.cpp declares the
reference module's C prototypes in their by-reference form, links against
libsipgpu.so instead of the module and checks every symbol against plain
loops.  On the CPU the host-only symbols must work and the device-backed ones
must report an error; on the GPU every check must pass."""
import os
import subprocess

import pytest

from aces4_amd import sipgpu as sg

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.dirname(sg.LIB_PATH)


def build_caller(tmp_path):
    exe = str(tmp_path / "dropin_caller")
    # linked by path: the library the tests load, wherever it lives
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    os.path.join(HERE, "dropin_caller.cpp"), "-o", exe, sg.LIB_PATH,
                    "-Wl,-rpath," + LIBDIR], check=True)
    return exe


def test_dropin_caller_links_and_fails_loudly_without_gpu(tmp_path):
    if sg.device_count() > 0:
        pytest.skip("a GPU is present")
    exe = build_caller(tmp_path)
    r = subprocess.run([exe, "--no-gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_dropin_caller_on_gpu(tmp_path):
    exe = build_caller(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
