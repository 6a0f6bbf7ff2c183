"""ORACLE — test infrastructure only.

The reference's own Lua front end (oracle/_ref/libmaralua.so: the reference's vendored
Lua 5.2.1 with its `Mara` and `buffer` modules, compiled from the reference tree by
`make -C oracle ref`), driven from Python. A test runs Lua chunks in it exactly as a Mara
driver script does, hands it primitive arrays as buffer userdata and reads numbers back.
Beside `Mara`, the state can load the device library's Lua binding that INTEGRATION.md
shows (oracle/_ref/libmaragpu_lua.so, `MaraGPU`), so the reference and the device answer
the same calls in one interpreter.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LUA_SO = os.path.join(HERE, "_ref", "libmaralua.so")
STUB_SO = os.path.join(HERE, "_ref", "libmaragpu_lua.so")

_lib = None


def have_lua() -> bool:
    return os.path.exists(LUA_SO)


def have_stub() -> bool:
    return os.path.exists(STUB_SO)


def lib():
    global _lib
    if _lib is None:
        # global symbols: the device binding resolves Lua's API from this library
        L = C.CDLL(LUA_SO, mode=C.RTLD_GLOBAL)
        L.ref_lua_new.restype = C.c_void_p
        L.ref_lua_close.argtypes = [C.c_void_p]
        L.ref_lua_run.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int]
        L.ref_lua_require.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        L.ref_lua_set_buffer.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]
        L.ref_lua_get_buffer.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]
        L.ref_lua_get_buffer.restype = C.c_long
        L.ref_lua_get_numbers.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]
        L.ref_lua_get_numbers.restype = C.c_long
        _lib = L
    return _lib


class LuaError(RuntimeError):
    pass


class State:
    """One Lua state with Mara (and optionally MaraGPU) loaded."""

    def __init__(self, with_device=False):
        self.L = lib()
        self.h = self.L.ref_lua_new()
        if not self.h:
            raise LuaError("lua_newstate failed")
        self._stub = None
        if with_device:
            self._stub = C.CDLL(STUB_SO, mode=C.RTLD_GLOBAL)
            f = C.cast(self._stub.luaopen_MaraGPU, C.c_void_p)
            if self.L.ref_lua_require(self.h, b"MaraGPU", f):
                raise LuaError("cannot load MaraGPU")

    def close(self):
        if self.h:
            self.L.ref_lua_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, code: str):
        err = C.create_string_buffer(1024)
        if self.L.ref_lua_run(self.h, code.encode(), err, 1024):
            raise LuaError(err.value.decode())

    def set_buffer(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.L.ref_lua_set_buffer(self.h, name.encode(), a.ctypes.data_as(C.POINTER(C.c_double)), a.size)

    def get_buffer(self, name: str, shape) -> np.ndarray:
        out = np.empty(shape)
        n = self.L.ref_lua_get_buffer(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)), out.size)
        if n != out.size:
            raise LuaError(f"global {name} is not a buffer of {out.size} doubles ({n})")
        return out

    def numbers(self, name: str, n: int) -> np.ndarray:
        out = np.zeros(n)
        m = self.L.ref_lua_get_numbers(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)), n)
        if m < 0:
            raise LuaError(f"global {name} is not a table")
        return out[:min(m, n)]
