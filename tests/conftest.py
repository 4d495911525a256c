import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "eth-v3d-lgpl_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def _ensure_built():
    """Build the in-tree libraries only when nothing is built yet."""
    need_pkg = not all(os.path.exists(os.path.join(PKG, "v3dmatch", "_lib", f))
                       for f in ("libv3dmatch.so", "libv3dsynth.so"))
    need_oracle = not all(os.path.exists(os.path.join(ROOT, "oracle", f))
                          for f in ("liboracle.so", "libcpubase.so"))
    if need_pkg:
        import importlib.util
        spec = importlib.util.spec_from_file_location("v3dm_build", os.path.join(PKG, "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build_all()
    if need_oracle:
        from oracle import build as obuild
        obuild.build_all()


_ensure_built()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
