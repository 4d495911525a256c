"""SIFT feature archives (".features" files) over the library's reader and writer.

The format is V3D::serializeDataToFile of a SerializableVector<SIFT_Feature>
(V3D/Base/v3d_serialization.h:953-1017, record V3D/Base/v3d_feature.h:41-49): "V3DBIN" +
uint32 count + 540 bytes per feature (int32 id, float x, y, direction x, y, scale,
magnitude, 128 descriptor floats, little-endian), or the "V3DTXT" text form.  Reading
and writing run in libv3dmatch.so (host code, no GPU needed).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._native import V3DMError, load_library


@dataclass
class Features:
    ids: np.ndarray        # (n,) int32
    xy: np.ndarray         # (n, 2) float32 positions
    direction: np.ndarray  # (n, 2) float32
    scale: np.ndarray      # (n,) float32
    magnitude: np.ndarray  # (n,) float32
    desc: np.ndarray       # (n, 128) float32

    def __len__(self):
        return int(self.ids.shape[0])


def _check(lib, rc):
    if rc != 0:
        raise V3DMError(rc, lib.v3dm_last_error(None).decode())


def count(path):
    """(number of features, binary?) of an archive."""
    lib = load_library()
    n, b = C.c_int(0), C.c_int(0)
    _check(lib, lib.v3dm_features_count(str(path).encode(), C.byref(n), C.byref(b)))
    return n.value, bool(b.value)


def read(path, cap=None):
    """The first min(count, cap) features of an archive."""
    lib = load_library()
    n_all, _ = count(path)
    n = n_all if cap is None else max(0, min(int(cap), n_all))
    f = Features(np.empty(n, np.int32), np.empty((n, 2), np.float32), np.empty((n, 2), np.float32),
                 np.empty(n, np.float32), np.empty(n, np.float32), np.empty((n, 128), np.float32))
    got = C.c_int(0)
    _check(lib, lib.v3dm_features_read(str(path).encode(), n, f.ids.ctypes.data, f.xy.ctypes.data,
                                       f.direction.ctypes.data, f.scale.ctypes.data, f.magnitude.ctypes.data,
                                       f.desc.ctypes.data, C.byref(got)))
    assert got.value == n
    return f


def write(path, desc, xy=None, direction=None, scale=None, magnitude=None, ids=None, binary=True):
    """Write n features (desc (n, 128)); missing fields are written as ids 0..n-1 / zeros."""
    lib = load_library()
    d = np.ascontiguousarray(desc, np.float32).reshape(-1, 128)
    n = d.shape[0]

    def arr(a, shape, dt):
        return None if a is None else np.ascontiguousarray(a, dt).reshape(shape)

    i = arr(ids, (n,), np.int32)
    p = arr(xy, (n, 2), np.float32)
    r = arr(direction, (n, 2), np.float32)
    s = arr(scale, (n,), np.float32)
    m = arr(magnitude, (n,), np.float32)
    ptr = (lambda a: None if a is None else a.ctypes.data)
    _check(lib, lib.v3dm_features_write(str(path).encode(), n, ptr(i), ptr(p), ptr(r), ptr(s), ptr(m), d.ctypes.data,
                                        1 if binary else 0))
