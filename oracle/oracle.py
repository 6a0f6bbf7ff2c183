"""ORACLE — test infrastructure only (the checker, never the product).

ctypes front end for the two CPU implementations of Mara's plain method:

* ``liboracle.so``        — this repository's CPU restatement (``mara_oracle.cpp``)
* ``_ref/libmararef.so``  — the reference's own Mara sources compiled by
                            ``oracle/Makefile`` (present only where the reference
                            tree was available at build time)

Both export the same functions (prefix ``oracle_`` / ``ref_``), so :class:`Lib`
drives either one.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` use this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmararef.so")

RECON = {"pcm": 0, "plm": 1, "weno5": 2}
IS_MODES = {"js96": 0, "b08": 1, "sz10": 2}
RIEMANN = {"hll": 0, "hllc": 1, "hlld": 2}
GODUNOV = {"plm-split": 0, "weno-split": 1, "plm-muscl": 2}
EIGEN = {"component": 0, "rmhd": 1}
ADVANCE = {"single": 1, "rk2": 2, "rk3": 3, "rk4": 4}
# weno.h ReconstructOperation
OPS = {"PLM_C2L": 0, "PLM_C2R": 1, "WENO5_FD_C2R": 2, "WENO5_FD_C2L": 3,
       "WENO5_FV_C2R": 4, "WENO5_FV_C2L": 5, "WENO5_FV_C2A": 6, "WENO5_FV_A2C": 7}


class Config(C.Structure):
    _fields_ = [("N", C.c_int * 3), ("Ng", C.c_int), ("x0", C.c_double * 3),
                ("x1", C.c_double * 3), ("gamma", C.c_double), ("recon", C.c_int),
                ("theta", C.c_double), ("IS", C.c_int), ("sz10A", C.c_double),
                ("riemann", C.c_int), ("rk", C.c_int), ("pfloor", C.c_double),
                ("ereset", C.c_int), ("preset", C.c_double), ("dreset", C.c_double),
                ("godunov", C.c_int), ("eigen", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("max_lambda", C.c_double), ("hlld_success", C.c_double),
                ("hlld_failure", C.c_double), ("c2p_iterations", C.c_double),
                ("c2p_solves", C.c_double), ("riemann_retries", C.c_double),
                ("riemann_failures", C.c_double), ("reserved", C.c_double)]


@dataclass
class Setup:
    """The Lua-facing configuration (Mara.set_* / config_solver) as one record."""
    N: tuple = (16, 16, 16)
    Ng: int = 3
    x0: tuple = (0.0, 0.0, 0.0)
    x1: tuple = (1.0, 1.0, 1.0)
    gamma: float = 4.0 / 3.0
    extrap: str = "plm"
    theta: float = 2.0
    IS: str = "js96"
    sz10A: float = 50.0
    riemann: str = "hlld"
    advance: str = "rk2"
    pfloor: float = -1.0
    ereset: bool = False
    preset: float = 1.0
    dreset: float = 1.0
    godunov: str = "plm-split"
    eigen: str = "component"

    def to_c(self) -> Config:
        c = Config()
        c.N[:] = list(self.N)
        c.Ng = self.Ng
        c.x0[:] = list(self.x0)
        c.x1[:] = list(self.x1)
        c.gamma = self.gamma
        c.recon = RECON[self.extrap]
        c.theta = self.theta
        c.IS = IS_MODES[self.IS]
        c.sz10A = self.sz10A
        c.riemann = RIEMANN[self.riemann]
        c.rk = ADVANCE[self.advance]
        c.pfloor = self.pfloor
        c.ereset = int(self.ereset)
        c.preset = self.preset
        c.dreset = self.dreset
        c.godunov = GODUNOV[self.godunov]
        c.eigen = EIGEN[self.eigen]
        return c


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def measures_dict(v):
    """Names the 33 values of the measure table (the layout of maragpu_measures)."""
    v = [float(x) for x in v]
    return {"mean_velocity": v[0:4], "mean_cons": v[4:12], "mean_prim": v[12:20],
            "mean_energies": {"magnetic": v[20], "kinetic": v[21], "internal": v[22], "total": v[23]},
            "mean_max_temperature": (v[24], v[25]), "mean_max_sonic_mach": (v[26], v[27]),
            "mean_min_alfvenic_mach": (v[28], v[29]), "mean_max_magnetic_field": (v[30], v[31]),
            "max_lorentz_factor": v[32]}


class Lib:
    """One of the two CPU implementations, loaded through ctypes."""

    def __init__(self, which: str = "oracle"):
        path = ORACLE_SO if which == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run make -C oracle"
                                    f"{' ref' if which != 'oracle' else ''})")
        self.which = which
        self.p = "oracle_" if which == "oracle" else "ref_"
        self.lib = C.CDLL(path)
        L = self.lib
        f = self._f
        f("create").restype = C.c_void_p
        f("create").argtypes = [C.POINTER(Config)]
        f("destroy").argtypes = [C.c_void_p]
        f("num_zones").restype = C.c_long
        f("num_zones").argtypes = [C.c_void_p]
        f("advance").restype = C.c_int
        f("advance").argtypes = [C.c_void_p, _dp, C.c_double, C.POINTER(Stats)]
        f("get_timestep").restype = C.c_double
        f("get_timestep").argtypes = [C.c_void_p, C.c_double]
        f("prim_to_cons_batch").argtypes = [C.c_long, _dp, _dp, C.c_double, _ip]
        f("cons_to_prim_batch").argtypes = [C.c_long, _dp, _dp, C.c_double, C.c_double, _ip, _ip]
        f("c2p_solver_batch").argtypes = [C.c_long, C.c_int, _dp, _dp, C.c_double, C.c_double,
                                          _ip, _ip]
        f("flux_eigen_batch").argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_int, C.c_double]
        f("quartic").restype = C.c_int
        f("quartic").argtypes = [C.c_double] * 5 + [_dp]
        f("reconstruct_batch").argtypes = [C.c_long, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                           C.c_double]
        f("riemann_batch").argtypes = [C.c_long, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double,
                                       C.c_double, _ip, _dp, _ip]
        if which == "oracle":
            L.oracle_set_threads.argtypes = [C.c_int]
            L.oracle_get_threads.restype = C.c_int
            L.oracle_mean_max_divB.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
            L.oracle_count_flops.argtypes = [C.c_void_p, _dp, C.c_long, _dp, C.c_int]
            L.oracle_c2p_from_start_batch.argtypes = [C.c_long, _dp, _dp, _dp, C.c_double,
                                                      C.c_double, _ip, _ip]
            L.oracle_eigensystem_batch.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _ip, C.c_int,
                                                   C.c_double]
            L.oracle_measures.argtypes = [_ip, C.c_int, _dp, C.c_double, _dp, _dp]

    def _f(self, name):
        return getattr(self.lib, self.p + name)

    # -- driving field and cooling ----------------------------------------
    def ou_new(self, P0, zeta, k1, seed):
        """A new StochasticVectorField3d (Stone spectrum) as the reference's serialized bytes."""
        f = self._f("ou_new")
        f.restype = C.c_size_t
        f.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_char_p, C.c_size_t]
        n = f(P0, zeta, k1, seed, None, 0)
        buf = C.create_string_buffer(n)
        f(P0, zeta, k1, seed, buf, n)
        return buf.raw

    def ou_advance(self, field: bytes, dt, times=1):
        f = self._f("ou_advance")
        f.argtypes = [C.c_char_p, C.c_size_t, C.c_double, C.c_int]
        buf = C.create_string_buffer(field, len(field))
        f(buf, len(field), dt, times)
        return buf.raw

    def ou_sample(self, field: bytes, xyz):
        f = self._f("ou_sample")
        f.argtypes = [C.c_char_p, C.c_size_t, C.c_long, _dp, _dp]
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        F = np.zeros_like(xyz)
        f(field, len(field), len(xyz), _d(xyz), _d(F))
        return F

    def drive(self, solver: "Solver", field: bytes, P, dt):
        """Drive (DrivingProcedure) P in place on the solver's domain; returns the mean power."""
        assert P.dtype == np.float64 and P.flags.c_contiguous
        power = C.c_double()
        if self.which == "oracle":
            s = solver.setup
            N = np.array(s.N, dtype=np.int32)
            x0 = np.array(s.x0, dtype=np.float64)
            x1 = np.array(s.x1, dtype=np.float64)
            self.lib.oracle_drive.argtypes = [_ip, C.c_int, _dp, _dp, C.c_char_p, C.c_size_t, _dp,
                                              C.c_double, C.POINTER(C.c_double)]
            self.lib.oracle_drive(_i(N), s.Ng, _d(x0), _d(x1), field, len(field), _d(P), dt, C.byref(power))
        else:
            self.lib.ref_drive.argtypes = [C.c_char_p, C.c_size_t, _dp, C.c_double, C.POINTER(C.c_double)]
            self.lib.ref_drive(field, len(field), _d(P), dt, C.byref(power))
        return power.value

    def cool(self, solver: "Solver", kind, ref, t0, P, dt, units=(1.0, 1.0, 1.0)):
        """CoolingModuleT4 ('T4') or E4 ('E4') on P in place; returns the energy removed."""
        assert P.dtype == np.float64 and P.flags.c_contiguous
        k = {"T4": 1, "E4": 2}[kind]
        removed = C.c_double()
        if self.which == "oracle":
            s = solver.setup
            N = np.array(s.N, dtype=np.int32)
            x0 = np.array(s.x0, dtype=np.float64)
            x1 = np.array(s.x1, dtype=np.float64)
            u = np.array(units, dtype=np.float64)
            self.lib.oracle_cool.argtypes = [_ip, C.c_int, _dp, _dp, C.c_double, _dp, C.c_int, C.c_double,
                                             C.c_double, _dp, C.c_double, C.POINTER(C.c_double)]
            self.lib.oracle_cool(_i(N), s.Ng, _d(x0), _d(x1), s.gamma, _d(u), k, ref, t0, _d(P), dt,
                                 C.byref(removed))
        else:
            assert tuple(units) == (1.0, 1.0, 1.0)
            self.lib.ref_cool.argtypes = [C.c_int, C.c_double, C.c_double, _dp, C.c_double,
                                          C.POINTER(C.c_double)]
            self.lib.ref_cool(k, ref, t0, _d(P), dt, C.byref(removed))
        return removed.value

    def measures(self, P, Ng=3, gamma=4.0 / 3.0, units=(1.0, 1.0, 1.0)):
        """The Mara.measure table (oracle only) of an AoS primitive array with Ng ghosts;
        units = PhysicalUnits(Length, Mass, Time)."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        N = np.array([s - 2 * Ng for s in P.shape[:3]], dtype=np.int32)
        u = np.array(units, dtype=np.float64)
        out = np.zeros(33)
        self.lib.oracle_measures(_i(N), Ng, _d(P), gamma, _d(u), _d(out))
        return measures_dict(out)

    # -- point functions --------------------------------------------------
    def prim_to_cons(self, P, gamma):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 8)
        U = np.zeros_like(P)
        err = np.zeros(len(P), dtype=np.int32)
        self._f("prim_to_cons_batch")(len(P), _d(P), _d(U), gamma, _i(err))
        return U, err

    def cons_to_prim(self, U, gamma, pfloor=-1.0, guess=None):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 8)
        P = np.zeros_like(U) if guess is None else np.array(guess, dtype=np.float64).reshape(-1, 8)
        err = np.zeros(len(U), dtype=np.int32)
        it = np.zeros(len(U), dtype=np.int32)
        self._f("cons_to_prim_batch")(len(U), _d(U), _d(P), gamma, pfloor, _i(err), _i(it))
        return P, err, it

    def c2p_solver(self, U, which, gamma, pfloor=-1.0):
        """which: 0 = anton2dzw, 1 = noble1dw, each from the cons-based estimate"""
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 8)
        P = np.zeros_like(U)
        err = np.zeros(len(U), dtype=np.int32)
        it = np.zeros(len(U), dtype=np.int32)
        self._f("c2p_solver_batch")(len(U), which, _d(U), _d(P), gamma, pfloor, _i(err), _i(it))
        return P, err, it

    def c2p_from_start(self, U, start, gamma, pfloor=-1.0):
        """anton2dzw from the Newton start (Z, W) of each state (oracle only)."""
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 8)
        start = np.ascontiguousarray(start, dtype=np.float64).reshape(-1, 2)
        P = np.zeros_like(U)
        err = np.zeros(len(U), dtype=np.int32)
        it = np.zeros(len(U), dtype=np.int32)
        self.lib.oracle_c2p_from_start_batch(len(U), _d(U), _d(start), _d(P), gamma, pfloor,
                                             _i(err), _i(it))
        return P, err, it

    def eigensystem(self, P, dim, gamma):
        """RMHD characteristic eigensystem (oracle only): L, R (n x 8 x 8), lam, ok."""
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 8)
        n = len(P)
        L = np.zeros((n, 8, 8))
        R = np.zeros((n, 8, 8))
        lam = np.zeros((n, 8))
        ok = np.zeros(n, dtype=np.int32)
        self.lib.oracle_eigensystem_batch(n, _d(P), _d(L), _d(R), _d(lam), _i(ok), dim, gamma)
        return L, R, lam, ok.astype(bool)

    def flux_eigen(self, P, dim, gamma):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, 8)
        F = np.zeros_like(P)
        ap = np.zeros(len(P))
        am = np.zeros(len(P))
        self._f("flux_eigen_batch")(len(P), _d(P), _d(F), _d(ap), _d(am), dim, gamma)
        return F, ap, am

    def quartic(self, d4, d3, d2, d1, d0):
        r = np.zeros(4)
        n = self._f("quartic")(d4, d3, d2, d1, d0, _d(r))
        return r[:n]

    def reconstruct(self, V, op, theta=2.0, IS="js96", sz10A=50.0):
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 6)
        out = np.zeros(len(V))
        self._f("reconstruct_batch")(len(V), _d(V), _d(out), OPS[op] if isinstance(op, str) else op,
                                     theta, IS_MODES[IS] if isinstance(IS, str) else IS, sz10A)
        return out

    def riemann(self, solver, PL, PR, dim, gamma, pfloor=-1.0):
        PL = np.ascontiguousarray(PL, dtype=np.float64).reshape(-1, 8)
        PR = np.ascontiguousarray(PR, dtype=np.float64).reshape(-1, 8)
        F = np.zeros_like(PL)
        err = np.zeros(len(PL), dtype=np.int32)
        lam = np.zeros(len(PL))
        ok = np.zeros(len(PL), dtype=np.int32)
        s = RIEMANN[solver] if isinstance(solver, str) else solver
        self._f("riemann_batch")(len(PL), s, _d(PL), _d(PR), _d(F), dim, gamma, pfloor,
                                 _i(err), _d(lam), _i(ok))
        return F, err, lam, ok

    # -- grid ----------------------------------------------------------------
    def solver(self, setup: Setup) -> "Solver":
        return Solver(self, setup)


class Solver:
    """Mara's Lua-facing solver object: advance(P, dt) / get_timestep(CFL)."""

    def __init__(self, lib: Lib, setup: Setup):
        self.lib = lib
        self.setup = setup
        self._cfg = setup.to_c()
        self.h = lib._f("create")(C.byref(self._cfg))
        self.nzones = lib._f("num_zones")(self.h)
        self.shape = tuple(n + 2 * setup.Ng for n in setup.N) + (8,)

    def __del__(self):
        try:
            self.lib._f("destroy")(self.h)
        except Exception:
            pass

    def advance(self, P, dt):
        """Advances P (AoS, ghosts included) in place when the step succeeds.
        Returns (errors, Stats)."""
        assert P.dtype == np.float64 and P.flags.c_contiguous and P.size == self.nzones * 8
        st = Stats()
        err = self.lib._f("advance")(self.h, _d(P), dt, C.byref(st))
        return err, st

    def get_timestep(self, cfl):
        return self.lib._f("get_timestep")(self.h, cfl)

    def mean_max_divB(self, P):
        m, x, a = C.c_double(), C.c_double(), C.c_double()
        P = np.ascontiguousarray(P, dtype=np.float64)
        self.lib.lib.oracle_mean_max_divB(self.h, _d(P), C.byref(m), C.byref(x), C.byref(a))
        return m.value, x.value, a.value

    def count_flops(self, P_synced, nsample=2000, device_path=False):
        """Mean flops per interface and per c2p zone (see oracle_count_flops); with
        device_path the HLLD nested c2p starts from the average state, as the device's."""
        out = np.zeros(28)
        P = np.ascontiguousarray(P_synced, dtype=np.float64)
        self.lib.lib.oracle_count_flops(self.h, _d(P), nsample, _d(out), int(bool(device_path)))
        return out


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def synchronize(P, Ng):
    """Periodic ghost fill of an AoS prim array (Mara/simple-cart.cpp:81-170)."""
    nx, ny, nz = (s - 2 * Ng for s in P.shape[:3])
    out = P.copy()
    for ax, n in enumerate((nx, ny, nz)):
        idx = [slice(None)] * 4
        src = [slice(None)] * 4
        idx[ax] = slice(0, Ng)
        src[ax] = slice(n, n + Ng)
        out[tuple(idx)] = out[tuple(src)]
        idx[ax] = slice(n + Ng, n + 2 * Ng)
        src[ax] = slice(Ng, 2 * Ng)
        out[tuple(idx)] = out[tuple(src)]
    return out
