"""Compile the drop-in caller programs of tests/cpp/ with g++ against the in-tree libraries."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "eth-v3d-lgpl_amd", "v3dmatch", "_lib")


def build(out_dir, name="dropin_match"):
    os.makedirs(out_dir, exist_ok=True)
    exe = os.path.join(out_dir, name)
    cmd = ["g++", "-std=c++11", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, name + ".cpp"), "-o", exe, "-L", LIB, "-lv3dmatch", "-lv3dsynth",
           f"-Wl,-rpath,{LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(" ".join(cmd) + "\n" + r.stderr)
    return exe
