"""Build the HIP engine library in-tree (gfx950)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgeauxdock.so")
SOURCES = ["gd_engine.hip", "gd_io.cpp"]
HEADERS = ["gd_mc.hip", "gd_sur.hip", "gd_post.hip", "gd_device.h", "gd_kernels.h", "../../include/geauxdock.h",
           "../../include/geauxdock_io.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep packed FP32 (v_pk_*_f32) out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # branches on wave-uniform conditions (ballots in the pair loop) stay plain
    # scalar branches: without this the control-flow structurizer routes them
    # through flow blocks and copies every accumulator at the join
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
