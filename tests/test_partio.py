"""The per-frame .bgeo output (ziran2019_amd/partio.py): Houdini's classic
binary geometry format, version 5, as Partio v1.1.0's BGEO writer lays it
out for the reference's writePartio (Lib/Ziran/Math/Geometry/PartioIO.h:
141-181) and the fracture project's frame output (Projects/fracture/
FractureSimulation.h:111-144).  No .bgeo file ships with the reference: the
fixtures below are the format's bytes written out by hand for small cases
(big-endian header fields, the position's w = 1, the 0x00 0xff end)."""
import io

import numpy as np
import pytest

from ziran2019_amd import partio


def _bytes(X, attrs=()):
    b = io.BytesIO()
    partio.write_bgeo(b, X, attrs)
    return b.getvalue()


def test_one_particle_position_only():
    got = _bytes([[1.0, 2.0, 3.0]])
    want = bytes.fromhex(
        "4267656f" "56" "00000005" "00000001"          # 'Bgeo', 'V', version 5, 1 point
        + "00000000" * 7                               # primitives, groups, attributes: none
        + "3f800000" "40000000" "40400000" "3f800000"  # x y z and w = 1 (float32, big-endian)
        + "00ff")
    assert got == want


def test_empty_set():
    got = _bytes(np.zeros((0, 3)))
    assert got == bytes.fromhex("4267656f5600000005" + "00000000" * 8 + "00ff")


def test_attribute_headers_and_records():
    X = [[0.5, -1.0, 0.25], [2.0, 0.0, -0.5]]
    got = _bytes(X, [("pf", partio.VECTOR, [0.75, 1.0]), ("id", partio.INT, [7, -2])])
    want = bytes.fromhex(
        "4267656f" "56" "00000005" "00000002"
        + "00000000" * 3 + "00000002" + "00000000" * 3   # two point attributes
        + "0002" + "7066" + "0001" + "00000005" + "00000000"  # "pf": 1 component, vector, default 0
        + "0002" + "6964" + "0001" + "00000001" + "00000000"  # "id": 1 component, int, default 0
        + "3f000000" "bf800000" "3e800000" "3f800000" "3f400000" "00000007"
        + "40000000" "00000000" "bf000000" "3f800000" "3f800000" "fffffffe"
        + "00ff")
    assert got == want


def test_round_trip_rounds_to_float32():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(1000, 3))
    c = rng.uniform(size=1000)
    got_X, attrs = partio.read_bgeo(io.BytesIO(_bytes(X, [("pf", partio.VECTOR, c)])))
    assert np.array_equal(got_X, X.astype(np.float32))
    assert np.array_equal(attrs["pf"][:, 0], c.astype(np.float32))


def test_reader_rejects_other_files():
    with pytest.raises(ValueError):
        partio.read_bgeo(io.BytesIO(b"NotBgeo" + b"\0" * 40))
    good = _bytes([[1.0, 2.0, 3.0]])
    with pytest.raises(ValueError):
        partio.read_bgeo(io.BytesIO(good[:-1]))
