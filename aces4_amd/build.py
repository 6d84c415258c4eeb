"""Build libsipgpu.so (gfx950) in-tree with hipcc.

    python -m aces4_amd.build          # or: python aces4_amd/build.py

The library is written to aces4_amd/libsipgpu.so next to this file.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsipgpu.so")
SOURCES = ["plan.cpp", "permute.hip", "contract.hip", "calib.hip", "ozaki.hip", "ozaki_big.hip", "packed.hip", "synth.hip", "energy.hip", "blockops.hip", "runtime.cpp", "serve.cpp", "graph.cpp", "dispatch.cpp", "triples.hip", "triples.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "--offload-arch=gfx950",
    # Packed-FP32 instructions are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


# Extra flags per source.  MFMA kernels keep their accumulators in the
# unified VGPR file ("VGPR form"): the compiler then neither copies them
# between register banks around loops nor splits the budget between banks.
VGPR_FORM = ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]
PER_FILE = {"calib.hip": VGPR_FORM, "contract.hip": VGPR_FORM, "ozaki.hip": VGPR_FORM,
            "packed.hip": VGPR_FORM, "triples.hip": VGPR_FORM}
NOISE = "is not a recognized feature for this target"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if NOISE not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s\n%s" % (r.returncode, " ".join(cmd), err))
    return err


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, msgs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = ([HIPCC] + [f for f in FLAGS if f != "-shared"] + PER_FILE.get(src, []) +
               ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        msgs.append(_run(cmd))
        objs.append(obj)
    tmp = OUT + ".tmp"
    msgs.append(_run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950"] + objs + ["-o", tmp]))
    if verbose:
        text = "\n".join(m for m in msgs if m.strip())
        if text:
            print(text, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
