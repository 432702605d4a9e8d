"""Per-frame particle output in Houdini's classic binary geometry format
(.bgeo), as the reference writes it through Partio.

The reference calls ``Partio::write`` (Lib/Ziran/Math/Geometry/PartioIO.h:
141-181 writePartio, the "position" attribute of every particle, called by
MpmSimulationBase::writeState when write_partio is set, MpmSimulationBase.cpp:
731-745; Projects/fracture/FractureSimulation.h:111-144 adds "pf" and
"sigmaC" per frame).  Partio (wdas/partio, tag v1.1.0, the version
Deps/CMakeLists.txt pins; not vendored in the reference) writes a name
ending in ".bgeo" uncompressed with its BGEO writer, whose published layout
is Houdini's classic geometry format, version 5, big-endian:

    int32 magic 'Bgeo', int8 'V', int32 version 5
    int32 points, primitives 0, point groups 0, primitive groups 0,
          point attributes (all but "position"), vertex attributes 0,
          primitive attributes 0, detail attributes 0
    per point attribute (in the order added, "position" left out):
          uint16 name length, name bytes, uint16 component count,
          int32 type (0 float, 1 int, 5 vector), one int32 0 per component
          (the default value)
    per point: position as 4 float32 (x, y, z, w = 1), then each attribute's
          components (float32 or int32) in the header's order
    the end of the extra section: bytes 0x00 0xff

Positions and float attributes are rounded to float32, as Partio stores
them.  No golden .bgeo file ships with the reference, so the bytes are
pinned by this layout (header fields, sizes, the round trip), not by a
reference file.
"""
import os
import struct

import numpy as np

FLOAT, INT, VECTOR = 0, 1, 5
_MAGIC = (ord("B") << 24) | (ord("g") << 16) | (ord("e") << 8) | ord("o")


def write_bgeo(f, X, attributes=()):
    """Write particles to ``f`` (a path or a binary stream).  ``X``: (n, 3)
    positions; ``attributes``: (name, type, values) with type FLOAT, INT or
    VECTOR and values of shape (n,) or (n, count), in the order Partio's
    addAttribute calls would add them after "position"."""
    if isinstance(f, (str, os.PathLike)):
        with open(f, "wb") as fh:
            return write_bgeo(fh, X, attributes)
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    n = X.shape[0]
    # columns as the 32-bit patterns of their values (native order), written big-endian
    cols = [np.column_stack([X.astype(np.float32), np.ones((n, 1), dtype=np.float32)]).view(np.uint32)]
    header = [struct.pack(">icii", _MAGIC, b"V", 5, n), struct.pack(">iiiiiii", 0, 0, 0, len(attributes), 0, 0, 0)]
    for name, kind, values in attributes:
        if kind not in (FLOAT, INT, VECTOR):
            raise ValueError(f"unknown attribute type {kind}")
        v = np.asarray(values).reshape(n, -1)
        count = v.shape[1]
        raw = name.encode()
        header.append(struct.pack(">H", len(raw)) + raw + struct.pack(">Hi", count, kind) + b"\0" * (4 * count))
        cols.append(np.ascontiguousarray(v.astype(np.int32) if kind == INT else v.astype(np.float32)).view(np.uint32))
    f.write(b"".join(header))
    if n:
        # one record per point: the columns side by side, as raw 4-byte words
        rec = np.concatenate([c.reshape(n, -1) for c in cols], axis=1)
        f.write(rec.astype(">u4").tobytes())
    f.write(b"\x00\xff")


def read_bgeo(f):
    """Read a file written by ``write_bgeo`` (or by Partio with point
    attributes only): (X float32 (n, 3), {name: values})."""
    if isinstance(f, (str, os.PathLike)):
        with open(f, "rb") as fh:
            return read_bgeo(fh)
    data = f.read()
    magic, v, version, n = struct.unpack_from(">icii", data, 0)
    if magic != _MAGIC or v != b"V" or version != 5:
        raise ValueError("not a version 5 classic .bgeo file")
    prims, pgroups, primgroups, nattr, nvert, nprimattr, ndetail = struct.unpack_from(">iiiiiii", data, 13)
    if prims or pgroups or primgroups or nvert or nprimattr or ndetail:
        raise ValueError("only point attributes are read")
    off = 41
    attrs = []
    for _ in range(nattr):
        (ln,) = struct.unpack_from(">H", data, off)
        name = data[off + 2:off + 2 + ln].decode()
        off += 2 + ln
        count, kind = struct.unpack_from(">Hi", data, off)
        off += 6 + 4 * count
        attrs.append((name, kind, count))
    width = 4 + sum(c for _, _, c in attrs)
    rec = np.frombuffer(data, dtype=">u4", count=n * width, offset=off).reshape(n, width)
    off += 4 * n * width
    if data[off:off + 2] != b"\x00\xff" or off + 2 != len(data):
        raise ValueError("missing end of the extra section")
    words = rec.astype(np.uint32)  # the 32-bit patterns in native order
    X = words[:, :3].copy().view(np.float32)
    out, col = {}, 4
    for name, kind, count in attrs:
        block = words[:, col:col + count].copy()
        out[name] = block.view(np.int32) if kind == INT else block.view(np.float32)
        col += count
    return X, out
