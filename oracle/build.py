"""Build the oracle libraries (test infrastructure only).

  liboracle.so     exact CPU restatement of the reference matcher (v3dm_oracle.c)
  libcpubase.so    FP32 OpenMP CPU port of the reference algorithm (cpu_baseline.c)

The reference's own matcher (V3D/CUDA/v3d_cudamatching.cu) needs nvcc, cuda.h and the
legacy cuBLAS v1 API, none of which exist in this image, so no oracle/_ref build exists.
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


def build_all():
    # exact oracle: no contraction, no fast-math; AVX2 only vectorises the FP64 lanes
    a = _gcc("v3dm_oracle.c", "liboracle.so", ["-O3", "-mavx2", "-ffp-contract=off", "-fno-fast-math"])
    # timed CPU baseline: FP32 blocked GEMM, FMA allowed (an SGEMM-like summation)
    b = _gcc("cpu_baseline.c", "libcpubase.so", ["-O3", "-mavx2", "-mfma"])
    return [a, b]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
