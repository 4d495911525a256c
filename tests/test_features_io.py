""".features archives and the sequence caller's rules (host code; no GPU).

The binary layout is pinned by archives packed here byte by byte from the format's
definition (V3D/Base/v3d_serialization.h:351-432 BinaryOStreamArchive: integers as four
little-endian bytes, floats as their IEEE bytes; :851-859 serializeVector: uint32 count
then the elements; V3D/Base/v3d_feature.h:41-49 the record: id, position x, y, direction
x, y, scale, magnitude, 128 descriptor floats), independent of the library's writer.  The
text layout is pinned the same way from TextOStreamArchive (v3d_serialization.h:176-260:
tokens, "{" / "}" scopes, the "SIFT_Feature" tag, 6 significant digits).
"""
import struct
import subprocess

import numpy as np
import pytest

import v3dmatch
from v3dmatch import features, sequence


def _random(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.random((n, 128), dtype=np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (rng.integers(-1000, 1000, n).astype(np.int32), rng.uniform(0, 2000, (n, 2)).astype(np.float32),
            rng.standard_normal((n, 2)).astype(np.float32), rng.uniform(1, 30, n).astype(np.float32),
            rng.uniform(0, 1, n).astype(np.float32), d)


def _pack_binary(ids, xy, dr, sc, mg, d):
    out = [b"V3DBIN", struct.pack("<I", len(ids))]
    for k in range(len(ids)):
        out.append(struct.pack("<i", int(ids[k])))
        out.append(struct.pack("<6f", xy[k, 0], xy[k, 1], dr[k, 0], dr[k, 1], sc[k], mg[k]))
        out.append(struct.pack("<128f", *d[k].tolist()))
    return b"".join(out)


def test_binary_reader_on_packed_archive(tmp_path):
    ids, xy, dr, sc, mg, d = _random(37, 1)
    p = tmp_path / "a.features"
    p.write_bytes(_pack_binary(ids, xy, dr, sc, mg, d))
    assert p.stat().st_size == 6 + 4 + 37 * 540
    assert features.count(p) == (37, True)
    f = features.read(p)
    assert np.array_equal(f.ids, ids) and np.array_equal(f.xy, xy) and np.array_equal(f.direction, dr)
    assert np.array_equal(f.scale, sc) and np.array_equal(f.magnitude, mg) and np.array_equal(f.desc, d)
    g = features.read(p, cap=5)
    assert len(g) == 5 and np.array_equal(g.desc, d[:5])


def test_binary_writer_matches_packed_bytes(tmp_path):
    ids, xy, dr, sc, mg, d = _random(11, 2)
    p = tmp_path / "b.features"
    features.write(p, d, xy, dr, sc, mg, ids, binary=True)
    assert p.read_bytes() == _pack_binary(ids, xy, dr, sc, mg, d)


def test_text_archive_round_trip(tmp_path):
    ids, xy, dr, sc, mg, d = _random(9, 3)
    p = tmp_path / "t.features"
    features.write(p, d, xy, dr, sc, mg, ids, binary=False)
    text = p.read_text()
    tok = text.split()
    assert text.startswith("V3DTXT\n") and tok[1] == "9" and tok[2] == "{" and tok[-1] == "}"
    assert tok[3] == "SIFT_Feature" and tok[4] == "{" and tok[5] == str(ids[0]) and tok[5 + 135] == "}"
    assert float(tok[6]) == float(f"{xy[0, 0]:.6g}")
    assert features.count(p) == (9, False)
    f = features.read(p)
    six = lambda a: np.array([float(f"{v:.6g}") for v in np.ravel(a)], np.float32).reshape(np.shape(a))
    assert np.array_equal(f.ids, ids)
    assert np.array_equal(f.desc, six(d)) and np.array_equal(f.xy, six(xy)) and np.array_equal(f.scale, six(sc))


def test_text_reader_on_hand_written_archive(tmp_path):
    vals = [0.5] * 128
    body = "SIFT_Feature { 7 1.5 2.25 -1 0 3 0.125 " + " ".join("0.5" for _ in vals) + " } \n  "
    p = tmp_path / "h.features"
    p.write_text("V3DTXT\n2 { " + body + body.replace(" 7 ", " 8 ") + "} \n")
    f = features.read(p)
    assert list(f.ids) == [7, 8] and f.xy[1].tolist() == [1.5, 2.25] and f.magnitude[0] == 0.125
    assert np.all(f.desc == 0.5)


@pytest.mark.parametrize("content,what", [
    (b"NOTV3D....", "unknown archive magic tag"), (b"V3DBIN", "missing feature count"),
    (b"V3DBIN" + struct.pack("<I", 3) + b"\0" * 540, "truncated"), (b"V3DTXT\n2 [ ", "expected '{'"),
])
def test_malformed_archives(tmp_path, content, what):
    p = tmp_path / "bad.features"
    p.write_bytes(content)
    with pytest.raises(v3dmatch.V3DMError) as e:
        features.read(p)
    assert e.value.code == 7 and what in e.value.message


def test_missing_archive(tmp_path):
    with pytest.raises(v3dmatch.V3DMError) as e:
        features.count(tmp_path / "none.features")
    assert e.value.code == 7


def test_empty_archives(tmp_path):
    for binary in (True, False):
        p = tmp_path / f"e{int(binary)}.features"
        features.write(p, np.zeros((0, 128), np.float32), binary=binary)
        assert features.count(p) == (0, binary)
        assert len(features.read(p)) == 0


def test_caller_count_rule():
    # min(maxFeatures, size & 0xffe0) (match_sequence.cpp:183): the mask wraps at 65536
    cases = {(31, 7168): 0, (32, 7168): 32, (95, 7168): 64, (7200, 7168): 7168, (65535, 70000): 65504,
             (65536, 70000): 0, (70000, 70000): 4448, (70000, 100000): 4448}
    for (size, nmax), want in cases.items():
        assert sequence.caller_count(size, nmax) == want, (size, nmax)
    assert sequence.caller_count(70000, 100000, exact=True) == 69984


def test_sequence_pairs_follow_the_caller_loop():
    def brute(n, r):
        out = []
        cur = 0
        while cur < n:
            other = cur + 1
            while other < cur + 2 * r:
                if other < n:
                    out.append((cur, other))
                other += 1
            cur += 1
        return out
    for n in (0, 1, 2, 5, 9):
        for r in (0, 1, 2, 3, 7):
            assert sequence.sequence_pairs(n, r) == brute(n, r)
    assert sequence.sequence_pairs(4, 2) == [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def test_match_sequence_program_usage():
    exe = v3dmatch.lib_path("v3dm_match_sequence")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "Usage" in r.stderr
