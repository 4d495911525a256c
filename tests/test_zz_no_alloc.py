"""GPU test of v3dm_match_device_noalloc (the plain form on device-resident descriptors,
allocating nothing; its work buffers live in a file-scope device array): through the C ABI,
its match list equals v3dm_find_putative_matches's on the same descriptors, on two calls in
a row (the second with smaller inputs, so that the shared array holds the first call's
data), and counts above the stated maximum are rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import synth

pytestmark = pytest.mark.gpu

MAX = 8192


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Hip:
    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        return p.value

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        return p

    def download(self, a, p):
        assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0

    def free(self, p):
        self.rt.hipFree(p)


def _lib():
    lib = v3dmatch.load_library()
    f = lib.v3dm_match_device_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    lib.v3dm_noalloc_last_error.restype = C.c_char_p
    return lib


def _noalloc(lib, hip, L, R, min_score=0.9, min_ratio=0.8):
    nL, nR = L.shape[0], R.shape[0]
    dL, dR = hip.upload(L), hip.upload(R)
    cap = max(min(nL, nR), 1)
    dP, dC = hip.alloc(cap * 2 * 4), hip.alloc(4)
    try:
        rc = lib.v3dm_match_device_noalloc(dL, nL, dR, nR, min_score, min_ratio, dP, dC, None)
        assert rc == 0, lib.v3dm_noalloc_last_error().decode()
        assert hip.rt.hipDeviceSynchronize() == 0
        cnt = np.zeros(1, np.int32)
        hip.download(cnt, dC)
        assert cnt[0] >= 0, "mutual check ordering wait ran out"
        pairs = np.zeros((max(int(cnt[0]), 1), 2), np.int32)
        if cnt[0] > 0:
            hip.download(pairs, dP)
        return pairs[:int(cnt[0])]
    finally:
        for p in (dL, dR, dP, dC):
            hip.free(p)


def _handle(L, R, min_score=0.9, min_ratio=0.8):
    m = v3dmatch.SIFT_Matching(0)
    m.enableRoundingWarning(False)
    m.allocate(max(L.shape[0], 32), max(R.shape[0], 32))
    m.setLeftFeatures(L.shape[0], L)
    m.setRightFeatures(R.shape[0], R)
    out = m.find_putative_matches_array(min_score, min_ratio)
    m.close()
    return out


def test_noalloc_equals_handle_path_on_two_calls():
    lib, hip = _lib(), _Hip()
    cases = [(MAX, 8000, 1, 0.9, 0.8), (1000, 777, 2, 0.9, 0.8), (4096, 4096, 3, 0.85, -1.0)]
    for nL, nR, seed, ms, mr in cases:
        say(f"case noalloc {nL}x{nR} seed {seed} minScore {ms} minRatio {mr}")
        qL, qR, _ = synth.make_pair(nL, nR, 0.4, config=1, pair=seed)
        L, R = synth.dequantise(qL), synth.dequantise(qR)
        got = _noalloc(lib, hip, L, R, ms, mr)
        want = _handle(L, R, ms, mr)
        say(f"  matches {len(got)} (handle path {len(want)})")
        assert np.array_equal(got, want)
        if nL <= 1000:
            eL, eR = oracle.round32(nL), oracle.round32(nR)
            assert np.array_equal(got, oracle.find_putative_matches(L[:eL], R[:eR], ms, mr))


def test_noalloc_rejects_sizes_above_the_maximum():
    lib, hip = _lib(), _Hip()
    dummy = hip.alloc(64)
    try:
        rc = lib.v3dm_match_device_noalloc(dummy, MAX + 1, dummy, 32, 0.9, 0.8, dummy, dummy, None)
        assert rc == 1 and b"MAX_FEATURES" in lib.v3dm_noalloc_last_error()
        rc = lib.v3dm_match_device_noalloc(dummy, 32, dummy, -1, 0.9, 0.8, dummy, dummy, None)
        assert rc == 1
        rc = lib.v3dm_match_device_noalloc(dummy, 32, dummy, 32, 0.9, 0.8, None, dummy, None)
        assert rc == 1
    finally:
        hip.free(dummy)
    # an empty side gives an empty result
    qL, qR, _ = synth.make_pair(64, 64, 0.4, config=1, pair=4)
    L = synth.dequantise(qL)
    assert _noalloc(lib, hip, L, np.zeros((31, 128), np.float32)).shape == (0, 2)
