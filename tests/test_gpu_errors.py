"""GPU tests of the C ABI's error paths, called through the library itself:

* calls out of order give V3DM_ERR_STATE (features, matching or results before allocate /
  before a match) and leave the handle usable;
* a result larger than the caller's buffer gives V3DM_ERR_CAPACITY with the full count in
  *nPairs, and the pairs that fit are the first ones of the full result;
* invalid sizes and arguments (negative or above-limit maxima, negative feature counts, a pair
  naming an image outside the batch, unknown modes and tile shapes, a batch result larger
  than the caller's buffer) give V3DM_ERR_INVALID / V3DM_ERR_CAPACITY and leave the handle
  usable;
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

V3DM_ERR_INVALID = 1
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


def test_invalid_arguments():
    lib = load_library()
    h = _handle(lib)
    try:
        say("case invalid: allocate sizes")
        assert lib.v3dm_allocate(h, -1, 64, 0) == V3DM_ERR_INVALID
        assert lib.v3dm_allocate(h, (1 << 20) + 32, 64, 0) == V3DM_ERR_INVALID
        assert b"2^20" in lib.v3dm_last_error(h)
        assert lib.v3dm_allocate(h, 64, 64, 0) == 0  # usable afterwards
        say("case invalid: batch images and pairs")
        desc = np.zeros((100, 128), np.uint8)
        bad = np.array([60, -1], np.int32)
        assert lib.v3dm_batch_set_images_u8(h, 2, _addr(bad), _addr(desc)) == V3DM_ERR_INVALID
        good = np.array([60, 40], np.int32)
        assert lib.v3dm_batch_set_images_u8(h, 2, _addr(good), _addr(desc)) == 0
        left, right = np.array([0, 5], np.int32), np.array([1, 0], np.int32)
        assert lib.v3dm_batch_set_pairs(h, 2, _addr(left), _addr(right)) == V3DM_ERR_INVALID
        assert b"outside" in lib.v3dm_last_error(h)
        assert lib.v3dm_batch_set_tuning(h, 1, 7, 0, 0) == V3DM_ERR_INVALID
        assert lib.v3dm_batch_set_tile_shape(h, 0, 64) == V3DM_ERR_INVALID
        assert lib.v3dm_batch_set_tile_shape(h, 9, 128) == V3DM_ERR_INVALID
    finally:
        lib.v3dm_destroy(h)


def test_batch_capacity_error():
    """A batch result larger than the caller's buffer: V3DM_ERR_CAPACITY with the full count;
    the same run read again with room for it returns every pair, equal to the result of a
    normal call."""
    from v3dmatch import synth
    lib = load_library()
    h = _handle(lib)
    try:
        counts, desc, left, right, _ = synth.make_batch(6, lo=300, hi=2000, config=61)
        desc = np.ascontiguousarray(desc)
        assert lib.v3dm_batch_set_images_u8(h, len(counts), _addr(counts), _addr(desc)) == 0
        assert lib.v3dm_batch_set_pairs(h, len(left), _addr(left), _addr(right)) == 0
        offs = np.zeros(len(left) + 1, np.int32)
        small = np.zeros(2 * 4, np.int32)
        n = C.c_int(-1)
        say("case batch capacity: cap 4")
        rc = lib.v3dm_batch_match(h, C.c_float(0.9), C.c_float(0.8), _addr(offs), _addr(small), 4, C.byref(n))
        assert rc == V3DM_ERR_CAPACITY and n.value > 4 and offs[-1] == n.value
        full = np.zeros(2 * n.value, np.int32)
        n2 = C.c_int(-1)
        assert lib.v3dm_batch_get_matches(h, _addr(offs), _addr(full), n.value, C.byref(n2)) == 0
        assert n2.value == n.value
        ref = np.zeros(2 * n.value, np.int32)
        n3 = C.c_int(-1)
        assert lib.v3dm_batch_match(h, C.c_float(0.9), C.c_float(0.8), _addr(offs), _addr(ref), n.value,
                                    C.byref(n3)) == 0
        assert n3.value == n.value and np.array_equal(full, ref)
        ms = C.c_float(0)
        assert lib.v3dm_batch_time_dominant_kernel(h, 0, C.byref(ms)) == V3DM_ERR_INVALID
    finally:
        lib.v3dm_destroy(h)
