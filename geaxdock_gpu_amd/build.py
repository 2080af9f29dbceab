"""Build the HIP engine library in-tree (gfx950).

Translation units, compiled in parallel and linked into libgeauxdock.so:
  gd_engine.hip        host side of the C ABI, surrogate, utility and post-MC
                       kernels;
  gd_mc_inst.hip x 4   the MC kernel instances for 1..4 ligand atoms per lane
                       (GD_MC_APL), the bulk of the device code;
  gd_io.cpp            loaders and writers (host only).
"""
import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgeauxdock.so")
HEADERS = ["gd_mc.h", "gd_sur.h", "gd_sur.hip", "gd_util.hip", "gd_post.hip", "gd_device.h",
           "gd_kernels.h", "../../include/geauxdock.h", "../../include/geauxdock_io.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep packed FP32 (v_pk_*_f32) out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # branches on wave-uniform conditions (ballots in the pair loop) stay plain
    # scalar branches: without this the control-flow structurizer routes them
    # through flow blocks and copies every accumulator at the join
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
]

# (object, source, extra flags)
UNITS = [("gd_engine.o", "gd_engine.hip", [])] + \
        [(f"gd_mc_apl{a}.o", "gd_mc_inst.hip", [f"-DGD_MC_APL={a}"]) for a in (1, 2, 3, 4)] + \
        [("gd_io.o", "gd_io.cpp", [])]


def _compile(hipcc, unit):
    obj, src, extra = unit
    out = os.path.join(OBJ, obj)
    cmd = [hipcc] + HIPCC_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src} {extra}:\n" + r.stdout + r.stderr)
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False, jobs=None):
    """Compile every unit (always; no staleness check) and link the library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or max(1, min(len(UNITS), os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda u: _compile(hipcc, u), UNITS))
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined"] + objs + \
        ["-o", LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
