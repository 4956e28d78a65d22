"""ctypes binding of liboracle.so (the C restatement in mdoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")


class OrcParams(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double * 3), ("dt", ctypes.c_double),
                ("cutoff", ctypes.c_double), ("buffer_length", ctypes.c_double),
                ("check_list_length", ctypes.c_int)]


class OrcHeatbath(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("aimed_temperature", ctypes.c_double),
                ("tau", ctypes.c_double), ("gamma", ctypes.c_double), ("zeta", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


HEATBATH = {"none": 0, "nose_hoover": 1, "langevin": 2}


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        dp, ip = P(ctypes.c_double), P(ctypes.c_int)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [P(OrcParams), ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp, ip]
        L.orc_get_state.argtypes = [ctypes.c_void_p, dp, dp]
        for fn in ("orc_make_pairlist", "orc_particle_number", "orc_total_particle_number"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.orc_is_pairlist_expired.argtypes = [ctypes.c_void_p, ip]
        for fn in ("orc_update_position_half", "orc_refresh_images", "orc_calculate_force"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
            getattr(L, fn).restype = None
        L.orc_calculate.argtypes = [ctypes.c_void_p, ctypes.c_int, ip]
        for fn in ("orc_kinetic_energy", "orc_potential_energy", "orc_virial"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
            getattr(L, fn).restype = ctypes.c_double
        L.orc_forces.argtypes = [ctypes.c_void_p, dp]
        L.orc_pair_number.argtypes = [ctypes.c_void_p]
        L.orc_pair_number.restype = ctypes.c_long
        L.orc_pairs_canonical.argtypes = [ctypes.c_void_p, P(ctypes.c_int32), ctypes.c_long]
        L.orc_pairs_canonical.restype = ctypes.c_long
        L.orc_bruteforce_pairs.argtypes = [P(OrcParams), ctypes.c_int, dp, ctypes.c_double,
                                           P(ctypes.c_int32), ctypes.c_long]
        L.orc_bruteforce_pairs.restype = ctypes.c_long
        L.orc_bruteforce_forces.argtypes = [P(OrcParams), ctypes.c_int, dp, dp, dp, dp]
        L.orc_bruteforce_forces.restype = None
        L.orc_constants.argtypes = [ctypes.c_double, dp, dp]
        L.orc_constants.restype = None
        L.orc_set_heatbath.argtypes = [ctypes.c_void_p, P(OrcHeatbath)]
        L.orc_set_heatbath.restype = None
        L.orc_get_heatbath.argtypes = [ctypes.c_void_p, P(OrcHeatbath)]
        L.orc_get_heatbath.restype = None
        L.orc_heatbath_zeta.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_heatbath_zeta.restype = None
        for fn in ("orc_heatbath_momenta", "orc_langevin"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
            getattr(L, fn).restype = None
        L.orc_temperature.argtypes = [ctypes.c_void_p]
        L.orc_temperature.restype = ctypes.c_double
        U4, U2 = ctypes.c_uint32 * 4, ctypes.c_uint32 * 2
        L.orc_philox4x32.argtypes = [U4, U2, U4]
        L.orc_philox4x32.restype = None
        L.orc_langevin_normals.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64, dp]
        L.orc_langevin_normals.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(L, dt=0.001, cutoff=3.0, buffer_length=0.3, check_list_length=500) -> OrcParams:
    p = OrcParams()
    for d in range(3):
        p.L[d] = float(L[d])
    p.dt, p.cutoff, p.buffer_length, p.check_list_length = dt, cutoff, buffer_length, check_list_length
    return p


def constants(cutoff=3.0):
    c0, c2 = ctypes.c_double(), ctypes.c_double()
    lib().orc_constants(cutoff, ctypes.byref(c0), ctypes.byref(c2))
    return c0.value, c2.value


class OracleError(RuntimeError):
    pass


class Oracle:
    """One periodic domain of the restated reference path."""

    def __init__(self, state, capacity_factor: float = 2.0, **kw):
        self.L = lib()
        self.prm = params(state.L, dt=state.dt, **kw)
        self.n = state.n
        cap = int(state.n * capacity_factor) + 4096
        self.h = self.L.orc_create(ctypes.byref(self.prm), cap)
        q = np.ascontiguousarray(state.q, dtype=np.float64)
        p = np.ascontiguousarray(state.p, dtype=np.float64)
        self._check(self.L.orc_set_state(self.h, self.n, _dp(q), _dp(p), None))

    def _check(self, rc):
        if rc != 0:
            raise OracleError(f"oracle status {rc}")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def make_pairlist(self):
        self._check(self.L.orc_make_pairlist(self.h))

    def calculate(self, nsteps: int) -> int:
        r = ctypes.c_int(0)
        self._check(self.L.orc_calculate(self.h, nsteps, ctypes.byref(r)))
        return r.value

    def is_expired(self) -> bool:
        e = ctypes.c_int(0)
        self._check(self.L.orc_is_pairlist_expired(self.h, ctypes.byref(e)))
        return bool(e.value)

    def state(self):
        q = np.zeros((self.n, 3))
        p = np.zeros((self.n, 3))
        self._check(self.L.orc_get_state(self.h, _dp(q), _dp(p)))
        return q, p

    def forces(self) -> np.ndarray:
        f = np.zeros((self.n, 3))
        self._check(self.L.orc_forces(self.h, _dp(f)))
        return f

    def observe(self):
        """(KE, PE, virial) sums as the reference's observers return them."""
        return (self.L.orc_kinetic_energy(self.h), self.L.orc_potential_energy(self.h),
                self.L.orc_virial(self.h))

    def refresh_images(self):
        self.L.orc_refresh_images(self.h)

    def pair_number(self) -> int:
        return int(self.L.orc_pair_number(self.h))

    def pairs(self) -> np.ndarray:
        m = self.L.orc_pairs_canonical(self.h, None, 0)
        if m < 0:
            raise OracleError(f"oracle status {m}")
        out = np.zeros((m, 3), dtype=np.int32)
        m2 = self.L.orc_pairs_canonical(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), m)
        if m2 != m:
            raise OracleError("pair count changed")
        return out

    def total_particle_number(self) -> int:
        return int(self.L.orc_total_particle_number(self.h))

    def set_heatbath(self, kind: str, aimed_temperature: float = 1.0, tau: float = 0.1,
                     gamma: float = 0.1, zeta: float = 0.0, seed: int = 0):
        hb = OrcHeatbath(HEATBATH[kind], aimed_temperature, tau, gamma, zeta, seed)
        self.L.orc_set_heatbath(self.h, ctypes.byref(hb))

    def heatbath(self) -> dict:
        hb = OrcHeatbath()
        self.L.orc_get_heatbath(self.h, ctypes.byref(hb))
        return {f: getattr(hb, f) for f, _ in OrcHeatbath._fields_}

    def heatbath_zeta(self, current_temperature: float):
        self.L.orc_heatbath_zeta(self.h, current_temperature)

    def heatbath_momenta(self):
        self.L.orc_heatbath_momenta(self.h)

    def langevin(self):
        self.L.orc_langevin(self.h)

    def temperature(self) -> float:
        return self.L.orc_temperature(self.h)


def philox4x32(ctr, key):
    U4, U2 = ctypes.c_uint32 * 4, ctypes.c_uint32 * 2
    out = U4()
    lib().orc_philox4x32(U4(*ctr), U2(*key), out)
    return list(out)


def langevin_normals(seed: int, tag: int, step: int) -> np.ndarray:
    out = np.zeros(4)
    lib().orc_langevin_normals(seed, tag, step, _dp(out))
    return out


def bruteforce_pairs(state, r2max: float) -> np.ndarray:
    prm = params(state.L)
    q = np.ascontiguousarray(state.q, dtype=np.float64)
    m = lib().orc_bruteforce_pairs(ctypes.byref(prm), state.n, _dp(q), r2max, None, 0)
    out = np.zeros((m, 3), dtype=np.int32)
    lib().orc_bruteforce_pairs(ctypes.byref(prm), state.n, _dp(q), r2max,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), m)
    return out


def bruteforce_forces(state):
    prm = params(state.L)
    q = np.ascontiguousarray(state.q, dtype=np.float64)
    f = np.zeros((state.n, 3))
    pe, vir = ctypes.c_double(), ctypes.c_double()
    lib().orc_bruteforce_forces(ctypes.byref(prm), state.n, _dp(q), _dp(f), ctypes.byref(pe),
                                ctypes.byref(vir))
    return f, pe.value, vir.value


def density_grid(q, L, grid_size: float):
    """BubbleHist of the Cavitation project for one unit at the origin
    (cavitation.cc:38-57 for the grid, 117-141 for the counts): g_d =
    int(L_d / grid_size + 0.5) boxes per axis, box width w = L_x / g_x on
    every axis, particle i in box int(x / w) + int(y / w) g_x + int(z / w)
    g_x g_y, counted in an unsigned char.  Returns (g, the g_x g_y g_z
    bytes as SaveDensity writes them, cavitation.cc:85-90).  A box index
    outside the grid raises, as the reference's assert stops."""
    g = [int(float(L[d]) / grid_size + 0.5) for d in range(3)]
    w = float(L[0]) / float(g[0])
    inv = 1.0 / w
    q = np.asarray(q, dtype=np.float64)
    ix = np.trunc(q * inv).astype(np.int64)
    k = ix[:, 0] + ix[:, 1] * g[0] + ix[:, 2] * g[0] * g[1]
    total = g[0] * g[1] * g[2]
    if (k < 0).any() or (k >= total).any():
        raise OracleError("a particle lies outside the density grid")
    counts = np.bincount(k, minlength=total)
    return g, (counts & 0xFF).astype(np.uint8)
