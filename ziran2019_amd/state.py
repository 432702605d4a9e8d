"""Restart state of the MPM step in the reference's own framing.

SimulationBase::write / read store the state that MpmSimulationBase's
writeState produces: Scene::writeState -> DataManager::writeData
(Lib/Ziran/Sim/Scene.h:189-217, Lib/Ziran/CS/DataStructure/DataManager.h:
263-293, DataArray.h:100-117, Lib/Ziran/CS/Util/BinaryIO.h:82-101):

    int32  particle count
    uint64 number of attribute arrays
    per array:
        uint64 name length, name bytes
        int32  lg2_grain_size of its DisjointRanges (7, DisjointRanges.h:51)
        uint64 number of ranges, uint64 sizeof(Range) = 8, the ranges (int32 lower, upper)
        uint64 number of values, uint64 bytes per value, the values
    uint64 0, uint64 12 and uint64 0, uint64 8 (the empty triangle and segment
    index vectors of Scene::writeState)

DataManager::readData finds every array by its name, so the order of the
arrays is free; the reference writes them in the iteration order of a
std::unordered_map keyed by std::hash of the names (implementation-defined),
this module in the fixed order of ``ATTRIBUTES``.  Values are the raw bytes
of the reference's element types: Vector<double, 3> (24 B), Matrix<double,
3, 3> column-major (72 B), double, int32, and the PhaseField<double, 3>
record (8 doubles + the allow_damage byte, padded to 72 B,
Projects/fracture/PhaseField.h:13-21).  ``mu`` / ``lambda`` hold the
per-particle model parameters, which the reference keeps inside its model
structs, and the material arrays hold this engine's model and plasticity
choice; a reader that does not know an array skips it.
"""
import struct

import numpy as np

# name -> (numpy dtype, values per particle); the order of a written file
ATTRIBUTES = (
    ("m", np.float64, 1),
    ("X", np.float64, 3),
    ("V", np.float64, 3),
    ("C", np.float64, 9),
    ("F", np.float64, 9),
    ("element measure", np.float64, 1),
    ("mu", np.float64, 1),
    ("lambda", np.float64, 1),
    ("material", np.int32, 1),
    ("plastic parameters", np.float64, 4),
    ("plastic state", np.float64, 1),
    ("degradation", np.float64, 1),
    ("Fe_Nonequilibrated", np.float64, 9),
    ("phase field", np.uint8, 72),
)
_KIND = {name: (dt, k) for name, dt, k in ATTRIBUTES}
# the PhaseField<double,3> record: residual_phase, c, H, l0, one_over_sigma_c,
# pf_Fp, H_max, vol (doubles) then allow_damage (bool) and padding
PHASE_FIELD_FIELDS = ("residual_phase", "c", "H", "l0", "one_over_sigma_c", "pf_Fp", "H_max", "vol")


def pack_phase_field(pf):
    """(n, 72) bytes of PhaseField records from a dict of per-particle arrays."""
    n = len(pf["c"])
    rec = np.zeros((n, 9), dtype=np.float64)
    for k, name in enumerate(PHASE_FIELD_FIELDS):
        rec[:, k] = pf[name]
    raw = rec.view(np.uint8).reshape(n, 72).copy()
    raw[:, 64:] = 0
    raw[:, 64] = np.asarray(pf["allow_damage"], dtype=np.uint8)
    return raw


def unpack_phase_field(raw):
    raw = np.ascontiguousarray(raw, dtype=np.uint8).reshape(-1, 72)
    vals = raw[:, :64].copy().view(np.float64).reshape(-1, 8)
    out = {name: vals[:, k].copy() for k, name in enumerate(PHASE_FIELD_FIELDS)}
    out["allow_damage"] = raw[:, 64].astype(np.int32)
    return out


def write(f, count, arrays):
    """Writes ``arrays`` (name -> array with ``count`` rows) to the binary
    stream ``f`` in the order of ATTRIBUTES (names outside it after them)."""
    names = [a for a, _, _ in ATTRIBUTES if a in arrays] + [a for a in arrays if a not in _KIND]
    f.write(struct.pack("<iQ", int(count), len(names)))
    for name in names:
        a = np.ascontiguousarray(arrays[name])
        if a.shape[0] != count:
            raise ValueError(f"array {name} has {a.shape[0]} rows for {count} particles")
        per = a.nbytes // count if count else a.itemsize * int(np.prod(a.shape[1:], dtype=np.int64))
        nb = name.encode()
        f.write(struct.pack("<Q", len(nb)) + nb)
        f.write(struct.pack("<i", 7))
        if count:
            f.write(struct.pack("<QQii", 1, 8, 0, int(count)))
        else:
            f.write(struct.pack("<QQ", 0, 8))
        f.write(struct.pack("<QQ", int(count), int(per)))
        f.write(a.tobytes())
    f.write(struct.pack("<QQQQ", 0, 12, 0, 8))


def read(f):
    """Reads a state written by ``write`` (or by the C++ host): (count,
    name -> array); arrays of known names get their dtype and shape."""
    def take(fmt):
        size = struct.calcsize(fmt)
        b = f.read(size)
        if len(b) != size:
            raise ValueError("restart state: unexpected end of data")
        return struct.unpack(fmt, b)
    count, narr = take("<iQ")
    out = {}
    for _ in range(narr):
        (ln,) = take("<Q")
        name = f.read(ln).decode()
        take("<i")
        nranges, rbytes = take("<QQ")
        if rbytes != 8:
            raise ValueError(f"restart state: range size {rbytes} in {name}")
        length = sum(hi - lo for lo, hi in (take("<ii") for _ in range(nranges)))
        size, per = take("<QQ")
        if size != length:
            raise ValueError(f"restart state: {name} has {size} values for ranges of {length}")
        raw = np.frombuffer(f.read(size * per), dtype=np.uint8)
        if name in _KIND:
            dt, k = _KIND[name]
            if per != np.dtype(dt).itemsize * k:
                raise ValueError(f"restart state: {name} has {per} bytes per value")
            a = raw.view(dt).reshape(size, k) if k > 1 else raw.view(dt)
            out[name] = a.copy()
        else:
            out[name] = raw.reshape(size, per).copy()
    tail = take("<QQQQ")
    if tail != (0, 12, 0, 8):
        raise ValueError("restart state: bad trailer")
    return count, out
