"""Build the native libraries of the gfx950 putative matcher in-tree.

  libv3dmatch.so  HIP kernels + C ABI (hipcc, --offload-arch=gfx950)
  libv3dsynth.so  synthetic descriptor generator (gcc, host only)
  v3dm_match_sequence  batched matching of an image sequence (g++, over the C ABI)

Both land in v3dmatch/_lib/ next to the Python host package.  Run as a script or call
build_all().  Every call recompiles.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "v3dmatch", "_lib")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# per-source additions: the sampled check's probe keeps its MFMA accumulators in ordinary
# VGPRs (the default form moved every accumulator through AGPRs, 32 extra moves per tile)
HIP_SOURCE_FLAGS = {"v3dm_check.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
HIP_SOURCES = ["v3dm_fp32.hip", "v3dm_select.hip", "v3dm_short.hip", "v3dm_api.cpp", "v3dm_ext.cpp", "v3dm_batch.cpp",
               "v3dm_learned.cpp", "v3dm_features.cpp",
               "v3dm_voctree.cpp", "v3dm_ransac.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n  " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _hip_sources():
    extra = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip") and f not in HIP_SOURCES)
    return HIP_SOURCES + extra


def build_synth():
    os.makedirs(OUT, exist_ok=True)
    out = os.path.join(OUT, "libv3dsynth.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-std=c99", "-Wall",
          os.path.join(CSRC, "v3dm_synth.c"), "-o", out + ".tmp", "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_hip(jobs=8):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = _hip_sources()

    def compile_one(src):
        obj = os.path.join(OBJ, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        _run([HIPCC] + HIP_FLAGS + HIP_SOURCE_FLAGS.get(src, []) + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = os.path.join(OUT, "libv3dmatch.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_apps():
    """Host programs over the C ABI (g++, linked against the in-tree libv3dmatch.so)."""
    out = os.path.join(OUT, "v3dm_match_sequence")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(os.path.dirname(HERE), "include"),
          os.path.join(HERE, "apps", "match_sequence.cpp"), "-o", out + ".tmp", "-L", OUT, "-lv3dmatch",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    return [build_synth(), build_hip(), build_apps()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
