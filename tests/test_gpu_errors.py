"""GPU tests of the C ABI's error paths, called through the library itself:

* calls out of order give V3DM_ERR_STATE (features, matching or results before allocate /
  before a match) and leave the handle usable;
* a result larger than the caller's buffer gives V3DM_ERR_CAPACITY with the full count in
  *nPairs, and the pairs that fit are the first ones of the full result;
* the mutual check's ordered compaction waits, boundedly, for the counts of the work-groups
  before it; the time-out path is forced from the host (v3dm_debug_mutual_wait: no
  predecessor count is ever seen, a limit of 8 sleeps), so every kernel still ends
  normally: the call fails with V3DM_ERR_HIP and the next call, with the wait restored,
  returns the oracle's match set.
"""
import ctypes as C

import numpy as np
import pytest

import v3dmatch
from oracle import oracle
from v3dmatch import synth
from v3dmatch._native import load_library

pytestmark = pytest.mark.gpu

V3DM_ERR_STATE = 2
V3DM_ERR_CAPACITY = 3
V3DM_ERR_HIP = 4


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def pair():
    qL, qR, _ = synth.make_pair(4096, 4096, 0.4, config=0, pair=5)
    L, R = synth.dequantise(qL), synth.dequantise(qR)
    want = oracle.find_putative_matches(L, R, 0.9, 0.8)
    return np.ascontiguousarray(L), np.ascontiguousarray(R), want


def _handle(lib):
    h = C.c_void_p()
    assert lib.v3dm_create(C.byref(h), 0) == 0
    return h


def _addr(a):
    return a.ctypes.data


def test_state_errors():
    lib = load_library()
    h = _handle(lib)
    try:
        x = np.zeros((32, 128), np.float32)
        pairs = np.zeros(64, np.int32)
        n = C.c_int(-1)
        say("case state: set features / match before allocate")
        assert lib.v3dm_set_left(h, 32, _addr(x)) == V3DM_ERR_STATE
        assert lib.v3dm_set_right(h, 32, _addr(x)) == V3DM_ERR_STATE
        rc = lib.v3dm_find_putative_matches(h, C.c_float(0.9), C.c_float(0.8),
                                            _addr(pairs), 32, C.byref(n))
        assert rc == V3DM_ERR_STATE
        assert b"allocate" in lib.v3dm_last_error(h)
        # the handle works after the errors
        assert lib.v3dm_allocate(h, 64, 64, 0) == 0
        assert lib.v3dm_set_left(h, 32, _addr(x)) == 0
        assert lib.v3dm_set_right(h, 32, _addr(x)) == 0
    finally:
        lib.v3dm_destroy(h)


def test_capacity_error(pair):
    L, R, want = pair
    lib = load_library()
    h = _handle(lib)
    try:
        assert lib.v3dm_allocate(h, 4096, 4096, 0) == 0
        assert lib.v3dm_set_left(h, 4096, _addr(L)) == 0
        assert lib.v3dm_set_right(h, 4096, _addr(R)) == 0
        full = np.zeros(2 * 4096, np.int32)
        n = C.c_int(-1)
        rc = lib.v3dm_find_putative_matches(h, C.c_float(0.9), C.c_float(0.8),
                                            _addr(full), 4096, C.byref(n))
        assert rc == 0
        total = n.value
        assert total == len(want) and total > 10
        got = full[:2 * total].reshape(-1, 2)
        assert np.array_equal(got, np.asarray(want, np.int32).reshape(-1, 2))
        small = np.full(2 * 10, -7, np.int32)
        n2 = C.c_int(-1)
        say(f"case capacity: cap 10 of {total}")
        rc = lib.v3dm_find_putative_matches(h, C.c_float(0.9), C.c_float(0.8),
                                            _addr(small), 10, C.byref(n2))
        assert rc == V3DM_ERR_CAPACITY
        assert n2.value == total
        assert np.array_equal(small.reshape(-1, 2), got[:10])
    finally:
        lib.v3dm_destroy(h)


def test_mutual_wait_runs_out(pair):
    L, R, want = pair
    m = v3dmatch.SIFT_Matching(0)
    try:
        m.enableRoundingWarning(False)
        m.allocate(4096, 4096)
        m.setLeftFeatures(4096, L)  # 4 work-groups in the mutual check
        m.setRightFeatures(4096, R)
        say("case mutual wait: forced time-out")
        m.debug_mutual_wait(8, True)
        with pytest.raises(v3dmatch.V3DMError) as e:
            m.find_putative_matches_array(0.9, 0.8)
        assert e.value.code == V3DM_ERR_HIP and "ran out" in e.value.message
        m.debug_mutual_wait()
        got = m.find_putative_matches_array(0.9, 0.8)
        assert np.array_equal(np.asarray(got, np.int32).reshape(-1, 2), np.asarray(want, np.int32).reshape(-1, 2))
    finally:
        m.close()
