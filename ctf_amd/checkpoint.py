"""Mara checkpoint files (HDF5) for the device path: the host-side mirror of the Lua
drivers' checkpoint code over ``libmara_h5.so`` (``include/mara_checkpoint.h``).

- ``UnigridDataField`` follows ``DataManagerHDF5`` (modules/unigrid.lua:9-212): a primitive
  array in Mara's AoS layout with its ``write(filename, group)`` and ``read(filename, group)``.
- ``checkpoint_write`` / ``checkpoint_read`` follow drvtrb's ``CheckpointWrite`` and
  ``CheckpointRead`` (ctf-core/drvtrb.lua:509-582). They take either a host primitive array
  or a device context, whose committed state is written and read on the device side.
- ``File`` gives dictionary-style access to scalar datasets, as ``hdf5.File`` indexing does
  in the drivers (lua-hdf5/LuaHDF5.lua:153-219): strings and numbers.

Everything here runs on the host; a device context only moves its state through
``maragpu_download_field`` / ``maragpu_upload_field``.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from typing import Any

import numpy as np

from . import native

HERE = os.path.dirname(os.path.abspath(__file__))
H5LIB = os.path.join(HERE, "libmara_h5.so")

PRIM_NAMES = ["rho", "pre", "vx", "vy", "vz", "Bx", "By", "Bz"]    # Rmhd::GetPrimNames, Mara/rmhd.cpp:28-34
TRUNCATE, UPDATE = 0, 1
MISSING, NUMBER, STRING, OTHER = 0, 1, 2, 3

_lib = None


class CheckpointError(RuntimeError):
    pass


def lib():
    """libmara_h5.so, loaded once (after libmaragpu.so, which it links)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(H5LIB):
        raise CheckpointError(f"{H5LIB} is missing: build it with ctf_amd.build (it needs HDF5)")
    native.lib()
    L = C.CDLL(H5LIB)
    cp, vp, sz = C.c_char_p, C.c_void_p, C.c_size_t
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    sigs = {
        "mara_h5_last_error": (cp, []),
        "mara_h5_write_prim": (C.c_int, [cp, cp, dp, ip, C.c_int, C.c_int, C.POINTER(cp), C.c_int]),
        "mara_h5_read_prim": (C.c_int, [cp, cp, dp, ip, C.c_int, C.c_int, C.POINTER(cp)]),
        "mara_h5_extent": (C.c_int, [cp, cp, ip, C.POINTER(C.c_longlong)]),
        "mara_h5_write_string": (C.c_int, [cp, cp, cp, cp, sz]),
        "mara_h5_write_number": (C.c_int, [cp, cp, cp, C.c_double]),
        "mara_h5_kind": (C.c_int, [cp, cp, cp]),
        "mara_h5_read_string": (C.c_int, [cp, cp, cp, vp, sz, C.POINTER(sz)]),
        "mara_h5_read_number": (C.c_int, [cp, cp, cp, dp]),
        "mara_h5_keys": (C.c_int, [cp, cp, vp, sz, C.POINTER(sz)]),
        "maragpu_checkpoint_write": (C.c_int, [vp, cp, cp, C.c_int]),
        "maragpu_checkpoint_read": (C.c_int, [vp, cp, cp]),
        "mara_json_quote": (C.c_int, [cp, sz, vp, sz, C.POINTER(sz)]),
        "mara_json_unquote": (C.c_int, [cp, sz, vp, sz, C.POINTER(sz)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _lib = L
    return L


EXPORTS = ["mara_h5_last_error", "mara_h5_write_prim", "mara_h5_read_prim", "mara_h5_extent",
           "mara_h5_write_string", "mara_h5_write_number", "mara_h5_kind", "mara_h5_read_string",
           "mara_h5_read_number", "mara_h5_keys", "maragpu_checkpoint_write", "maragpu_checkpoint_read",
           "mara_json_quote", "mara_json_unquote"]


def _b(s):
    return None if s is None else (s.encode() if isinstance(s, str) else s)


def _chk(rc):
    if rc != 0:
        raise CheckpointError(f"[HDF5] {lib().mara_h5_last_error().decode()} "
                              f"({native.lib().maragpu_strerror(rc).decode()})")


def _names(names):
    arr = (C.c_char_p * len(names))(*[_b(n) for n in names])
    return arr


def _two_pass(fn, *args):
    """Calls a sized-output entry point twice: for the size, then for the bytes."""
    n = C.c_size_t()
    _chk(fn(*args, None, 0, C.byref(n)))
    buf = C.create_string_buffer(max(n.value, 1))
    _chk(fn(*args, buf, n.value, C.byref(n)))
    return buf.raw[:n.value]


# -- JSON strings as the drivers' json module writes and reads them -------------------------
def json_quote(data: bytes) -> bytes:
    """dkjson's quotestring (modules/json.lua:289-304) of a byte string, quotes included."""
    return _two_pass(lib().mara_json_quote, bytes(data), len(data))


def json_unquote(text: bytes) -> bytes:
    """dkjson's scanstring (modules/json.lua:529-590) of a quoted JSON string."""
    return _two_pass(lib().mara_json_unquote, bytes(text), len(text))


# -- scalar datasets ----------------------------------------------------------------------
class File:
    """Scalar string and number datasets of a checkpoint, by name or "group/name"
    (hdf5.File / hdf5.Group indexing, lua-hdf5/LuaHDF5.lua:153-219). Assigning a str or bytes
    writes a fixed-length string, a number writes a double; reading returns bytes for a
    string, float for a number, None when missing."""

    def __init__(self, path: str):
        self.path = path

    @staticmethod
    def _split(key):
        g, _, n = key.rpartition("/")
        return g, n

    def __setitem__(self, key, value):
        g, n = self._split(key)
        L = lib()
        if isinstance(value, (str, bytes, bytearray)):
            data = _b(value) if isinstance(value, str) else bytes(value)
            _chk(L.mara_h5_write_string(_b(self.path), _b(g), _b(n), data, len(data)))
        elif isinstance(value, (int, float, np.floating, np.integer)) and not isinstance(value, bool):
            _chk(L.mara_h5_write_number(_b(self.path), _b(g), _b(n), float(value)))
        else:
            raise CheckpointError("DataSet:unrecognized type for writing")

    def kind(self, key):
        g, n = self._split(key)
        return lib().mara_h5_kind(_b(self.path), _b(g), _b(n))

    def __getitem__(self, key):
        g, n = self._split(key)
        k = self.kind(key)
        if k == MISSING:
            return None
        if k == STRING:
            return _two_pass(lib().mara_h5_read_string, _b(self.path), _b(g), _b(n))
        if k == NUMBER:
            v = C.c_double()
            _chk(lib().mara_h5_read_number(_b(self.path), _b(g), _b(n), C.byref(v)))
            return v.value
        raise CheckpointError(f"Indexable:object {key} has an unsupported type")

    def keys(self, group=""):
        raw = _two_pass(lib().mara_h5_keys, _b(self.path), _b(group))
        return raw.decode().split("\n") if raw else []

    def extent(self, dset):
        nd = C.c_int()
        dims = (C.c_longlong * 8)()
        _chk(lib().mara_h5_extent(_b(self.path), _b(dset), C.byref(nd), dims))
        return [int(dims[a]) for a in range(nd.value)]


# -- the primitive array --------------------------------------------------------------------
class UnigridDataField:
    """DataManagerHDF5 (modules/unigrid.lua:9-212) for one process: a primitive array of
    shape (N0 + 2Ng, N1 + 2Ng, N2 + 2Ng, Nq) whose interior is written to and read from
    <group>/<name> datasets of shape (N0, N1, N2)."""

    def __init__(self, N, Ng=3, dataset_names=PRIM_NAMES):
        self.N = tuple(int(n) for n in N)
        self.Ng = int(Ng)
        self.dataset_names = list(dataset_names)
        shape = tuple(n + 2 * self.Ng for n in self.N) + (len(self.dataset_names),)
        self._array = np.zeros(shape)

    def array(self):
        return self._array

    def _args(self, P):
        if P.dtype != np.float64 or not P.flags.c_contiguous or P.shape != self._array.shape:
            raise CheckpointError("primitive array must be contiguous float64 of the domain's shape")
        return native.dptr(P), (C.c_int * 3)(*self.N)

    def write(self, filename, group="prim", mode=TRUNCATE, P=None):
        P = self._array if P is None else P
        ptr, n = self._args(P)
        _chk(lib().mara_h5_write_prim(_b(filename), _b(group), ptr, n, self.Ng, len(self.dataset_names),
                                      _names(self.dataset_names), mode))

    def read(self, filename, group="prim", P=None):
        P = self._array if P is None else P
        ptr, n = self._args(P)
        _chk(lib().mara_h5_read_prim(_b(filename), _b(group), ptr, n, self.Ng, len(self.dataset_names),
                                     _names(self.dataset_names)))


def write_prim(target, path, group="prim", mode=TRUNCATE):
    """The primitive state to <group>/<rho .. Bz>: a device Context's committed state, or a
    host AoS array (N0 + 2Ng, N1 + 2Ng, N2 + 2Ng, 8) with Ng = 3."""
    if isinstance(target, native.Context):
        _chk(lib().maragpu_checkpoint_write(target.h, _b(path), _b(group), mode))
    else:
        P = np.ascontiguousarray(target, dtype=np.float64)
        UnigridDataField(tuple(n - 6 for n in P.shape[:3])).write(path, group, mode, P=P)


def read_prim(target, path, group="prim"):
    """<group>/<rho .. Bz> into a device Context's committed state (ghosts filled with their
    periodic images) or into the interior of a host AoS array (ghosts left as they are)."""
    if isinstance(target, native.Context):
        _chk(lib().maragpu_checkpoint_read(target.h, _b(path), _b(group)))
    else:
        if target.dtype != np.float64 or not target.flags.c_contiguous:
            raise CheckpointError("primitive array must be contiguous float64")
        UnigridDataField(tuple(n - 6 for n in target.shape[:3])).read(path, group, P=target)


# -- drvtrb's checkpoints -----------------------------------------------------------------
def checkpoint_write(target, path, status: dict, measure_log: Any, runargs: dict, program: str = " ",
                     version: str = "", driving: bytes | None = None):
    """CheckpointWrite (ctf-core/drvtrb.lua:509-549): prim/<field>, then the strings runargs,
    measure (JSON), program, version and, when driving, the serialized driving field as a
    JSON string, then one dataset per status entry under status/."""
    if os.path.exists(path):
        os.remove(path)
    write_prim(target, path)
    f = File(path)
    f["runargs"] = json.dumps(runargs)
    f["measure"] = json.dumps(measure_log)
    f["program"] = program
    f["version"] = version
    if driving is not None:
        f["driving"] = json_quote(driving)
    for k, v in status.items():
        f[f"status/{k}"] = v


def checkpoint_read(path, target):
    """CheckpointRead (ctf-core/drvtrb.lua:555-582): the state into `target` (a Context or a
    host array); returns (Status, MeasureLog). Status values are floats for numbers and str
    for strings, as val:value() gives them."""
    read_prim(target, path)
    f = File(path)
    measure = f["measure"]
    text = measure.decode() if measure is not None else "[]"
    # a table that was empty when written reads back as '[]' (drvtrb.lua:569-573)
    measure_log = {} if text == "[]" else json.loads(text)
    status = {}
    for k in f.keys("status"):
        v = f[f"status/{k}"]
        status[k] = v.decode() if isinstance(v, bytes) else v
    return status, measure_log


def read_driving(path) -> bytes | None:
    """The serialized driving field of a checkpoint (json.decode of its 'driving' string,
    ctf-core/drvtrb.lua:738-742), ready for Mara.set_driving / native.OuField.load."""
    blob = File(path)["driving"]
    return None if blob is None else json_unquote(blob)
