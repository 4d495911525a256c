"""Build the oracle libraries (test infrastructure only).

  liboracle.so     exact CPU restatement of the reference matcher (v3dm_oracle.c)
  libcpubase.so    FP32 OpenMP CPU port of the reference algorithm (cpu_baseline.c)

The reference's own matcher (V3D/CUDA/v3d_cudamatching.cu) needs nvcc, cuda.h and the
legacy cuBLAS v1 API, none of which exist in this image, so it has no oracle/_ref build.
The reference's two-view pose utilities do build with g++ from three of its source files:

  _ref/libref_pose.so   ref_pose.cpp (C entry points) + V3D/Geometry/v3d_poseutilities.cpp,
                        V3D/Math/v3d_mathutilities.cpp, V3D/Math/v3d_optimization.cpp,
                        compiled where they lie under the reference tree (build_ref; only
                        when that tree is present, output only into oracle/_ref/)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _gcc(src, out, extra):
    cmd = ["gcc", "-fPIC", "-shared", "-std=gnu99", "-Wall", "-fopenmp"] + extra + [
        os.path.join(HERE, src), "-o", os.path.join(HERE, out) + ".tmp", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + " ".join(cmd) + "\n" + r.stderr)
    os.replace(os.path.join(HERE, out) + ".tmp", os.path.join(HERE, out))
    return os.path.join(HERE, out)


def build_variant(define, out_dir):
    """Oracle with a planted mistake (-DV3DO_PLANT_*), for tests that show the checks fail."""
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, f"liboracle_{define.lower()}.so")
    cmd = ["gcc", "-fPIC", "-shared", "-std=gnu99", "-fopenmp", "-O2", "-ffp-contract=off", f"-D{define}",
           os.path.join(HERE, "v3dm_oracle.c"), "-o", out, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr)
    return out


REF_ROOT = "/root/reference/V3D"
REF_POSE_SOURCES = ["Geometry/v3d_poseutilities.cpp", "Math/v3d_mathutilities.cpp", "Math/v3d_optimization.cpp"]


def build_ref():
    """oracle/_ref/libref_pose.so from the reference's sources, or None when the reference
    tree is absent (the built library is used where it exists)."""
    if not all(os.path.isfile(os.path.join(REF_ROOT, f)) for f in REF_POSE_SOURCES):
        return None
    out_dir = os.path.join(HERE, "_ref")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "libref_pose.so")
    cmd = (["g++", "-std=c++11", "-O2", "-fPIC", "-shared", "-w", "-ffp-contract=off", "-I", REF_ROOT,
            os.path.join(HERE, "ref_pose.cpp")] + [os.path.join(REF_ROOT, f) for f in REF_POSE_SOURCES]
           + ["-o", out + ".tmp", "-Wl,--no-undefined"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("reference pose build failed:\n" + " ".join(cmd) + "\n" + r.stderr)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    # exact oracle: no contraction, no fast-math; AVX2 only vectorises the FP64 lanes
    a = _gcc("v3dm_oracle.c", "liboracle.so", ["-O3", "-mavx2", "-ffp-contract=off", "-fno-fast-math"])
    # timed CPU baseline: FP32 blocked GEMM, FMA allowed (an SGEMM-like summation)
    b = _gcc("cpu_baseline.c", "libcpubase.so", ["-O3", "-mavx2", "-mfma"])
    c = build_ref()
    return [a, b] + ([c] if c else [])


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
