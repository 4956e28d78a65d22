"""ctypes binding of libmdx (include/mdx.h).

The library must exist: there is no CPU fallback.  Importing this module
does not touch the GPU; creating a :class:`System` does.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdx.so")

EXPORTS = (
    "mdx_default_params", "mdx_create", "mdx_destroy", "mdx_last_error", "mdx_set_state",
    "mdx_get_state", "mdx_particle_number", "mdx_get_images", "mdx_synchronize",
    "mdx_make_pairlist",
    "mdx_is_pairlist_expired",
    "mdx_update_position_half", "mdx_refresh_images", "mdx_calculate_force", "mdx_calculate",
    "mdx_observe", "mdx_forces", "mdx_pair_number", "mdx_pairs_canonical",
    "mdx_pairs_canonical_noalloc", "mdx_pairs_csr", "mdx_adjust_periodic_boundary",
    "mdx_scale_positions", "mdx_set_box", "mdx_get_cell_order", "mdx_density_grid", "mdx_timing",
    "mdx_timing_reset", "mdx_kick_profile", "mdx_lj_trajectory",
    "mdx_set_heatbath", "mdx_get_heatbath", "mdx_heatbath_zeta", "mdx_heatbath_momenta",
    "mdx_langevin",
    "mdx_ensemble_create", "mdx_ensemble_destroy", "mdx_ensemble_last_error", "mdx_ensemble_size",
    "mdx_ensemble_set_state", "mdx_ensemble_get_state", "mdx_ensemble_calculate",
    "mdx_ensemble_status", "mdx_ensemble_observe", "mdx_ensemble_get_config",
    "mdx_ensemble_set_config", "mdx_ensemble_autotune",
)

STATUS = {0: "MDX_OK", -1: "MDX_ERR_ARG", -2: "MDX_ERR_CAPACITY", -3: "MDX_ERR_INVALID_POS",
          -4: "MDX_ERR_TOO_FAST", -5: "MDX_ERR_BOX", -6: "MDX_ERR_HIP", -7: "MDX_ERR_STATE"}


class MdxParams(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double * 3), ("dt", ctypes.c_double), ("cutoff", ctypes.c_double),
                ("buffer_length", ctypes.c_double), ("check_list_length", ctypes.c_int),
                ("capacity", ctypes.c_int), ("device", ctypes.c_int), ("stream", ctypes.c_void_p)]


class MdxLjResult(ctypes.Structure):
    _fields_ = [("loss", ctypes.c_double), ("pressure_mean", ctypes.c_double),
                ("pe_mean", ctypes.c_double), ("dloss_deps", ctypes.c_double),
                ("dloss_dsigma", ctypes.c_double), ("forward_ms", ctypes.c_double),
                ("backward_ms", ctypes.c_double)]


class MdxEnsembleConfig(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int), ("grid_mult", ctypes.c_int), ("step_grid", ctypes.c_int),
                ("stage_cap", ctypes.c_int), ("ms_per_step", ctypes.c_double)]


class MdxHeatbath(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("aimed_temperature", ctypes.c_double),
                ("tau", ctypes.c_double), ("gamma", ctypes.c_double), ("zeta", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


HEATBATH = {"none": 0, "nose_hoover": 1, "langevin": 2}


class MdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    """Load libmdx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mdacp_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp, dp, ip, i64p = ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_int), P(ctypes.c_int64)
    sig = {
        "mdx_default_params": (None, [P(MdxParams)]),
        "mdx_create": (ctypes.c_int, [P(MdxParams), P(vp)]),
        "mdx_destroy": (ctypes.c_int, [vp]),
        "mdx_last_error": (ctypes.c_char_p, [vp]),
        "mdx_set_state": (ctypes.c_int, [vp, ctypes.c_int, dp, dp]),
        "mdx_get_state": (ctypes.c_int, [vp, dp, dp]),
        "mdx_particle_number": (ctypes.c_int, [vp, ip, ip]),
        "mdx_get_images": (ctypes.c_int, [vp, dp, dp, P(ctypes.c_int32), P(ctypes.c_int32)]),
        "mdx_synchronize": (ctypes.c_int, [vp]),
        "mdx_make_pairlist": (ctypes.c_int, [vp]),
        "mdx_is_pairlist_expired": (ctypes.c_int, [vp, ip]),
        "mdx_update_position_half": (ctypes.c_int, [vp]),
        "mdx_refresh_images": (ctypes.c_int, [vp]),
        "mdx_calculate_force": (ctypes.c_int, [vp]),
        "mdx_calculate": (ctypes.c_int, [vp, ctypes.c_int, ip]),
        "mdx_observe": (ctypes.c_int, [vp, dp, dp, dp]),
        "mdx_forces": (ctypes.c_int, [vp, dp]),
        "mdx_pair_number": (ctypes.c_int, [vp, i64p]),
        "mdx_pairs_canonical": (ctypes.c_int, [vp, P(ctypes.c_int32), ctypes.c_int64, i64p]),
        "mdx_pairs_canonical_noalloc": (ctypes.c_int, [vp, P(ctypes.c_int32), ctypes.c_int64,
                                                       i64p]),
        "mdx_pairs_csr": (ctypes.c_int, [vp, P(ctypes.c_int32), P(ctypes.c_int32),
                                         P(ctypes.c_int32), ctypes.c_int64, i64p]),
        "mdx_adjust_periodic_boundary": (ctypes.c_int, [vp]),
        "mdx_scale_positions": (ctypes.c_int, [vp, ctypes.c_double]),
        "mdx_set_box": (ctypes.c_int, [vp, dp]),
        "mdx_get_cell_order": (ctypes.c_int, [vp, P(ctypes.c_int32)]),
        "mdx_density_grid": (ctypes.c_int, [vp, ctypes.c_double, P(ctypes.c_int32),
                                            P(ctypes.c_uint8), ctypes.c_int64]),
        "mdx_timing": (ctypes.c_int, [vp, dp, dp, dp, i64p, i64p]),
        "mdx_timing_reset": (ctypes.c_int, [vp, ctypes.c_int]),
        "mdx_kick_profile": (ctypes.c_int, [vp, dp, dp]),
        "mdx_lj_trajectory": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, P(MdxLjResult), dp]),
        "mdx_set_heatbath": (ctypes.c_int, [vp, P(MdxHeatbath)]),
        "mdx_get_heatbath": (ctypes.c_int, [vp, P(MdxHeatbath)]),
        "mdx_heatbath_zeta": (ctypes.c_int, [vp, ctypes.c_double]),
        "mdx_heatbath_momenta": (ctypes.c_int, [vp]),
        "mdx_langevin": (ctypes.c_int, [vp]),
        "mdx_ensemble_create": (ctypes.c_int, [P(MdxParams), ctypes.c_int, ip, dp, P(vp)]),
        "mdx_ensemble_destroy": (ctypes.c_int, [vp]),
        "mdx_ensemble_last_error": (ctypes.c_char_p, [vp]),
        "mdx_ensemble_size": (ctypes.c_int, [vp, ip, ip]),
        "mdx_ensemble_set_state": (ctypes.c_int, [vp, dp, dp]),
        "mdx_ensemble_get_state": (ctypes.c_int, [vp, dp, dp]),
        "mdx_ensemble_calculate": (ctypes.c_int, [vp, ctypes.c_int, ip]),
        "mdx_ensemble_status": (ctypes.c_int, [vp, ip, ip]),
        "mdx_ensemble_observe": (ctypes.c_int, [vp, dp, dp, dp]),
        "mdx_ensemble_get_config": (ctypes.c_int, [vp, P(MdxEnsembleConfig)]),
        "mdx_ensemble_set_config": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "mdx_ensemble_autotune": (ctypes.c_int, [vp, ctypes.c_int, P(MdxEnsembleConfig)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class System:
    """One periodic LJ system resident on one GPU (an MDUnit's worth of
    Variables + MeshList + PairList, mdunit.cc:8-39)."""

    def __init__(self, L, dt=0.001, cutoff=3.0, buffer_length=0.3, check_list_length=500,
                 capacity=0, device=0, stream=None):
        self.lib = load()
        prm = MdxParams()
        self.lib.mdx_default_params(ctypes.byref(prm))
        for d in range(3):
            prm.L[d] = float(L[d])
        prm.dt, prm.cutoff, prm.buffer_length = dt, cutoff, buffer_length
        prm.check_list_length, prm.capacity, prm.device = check_list_length, capacity, device
        prm.stream = stream
        h = ctypes.c_void_p()
        rc = self.lib.mdx_create(ctypes.byref(prm), ctypes.byref(h))
        if rc != 0:
            raise MdxError(rc, self.lib.mdx_last_error(None).decode())
        self.h = h
        self.n = 0

    @classmethod
    def from_state(cls, state, **kw):
        s = cls(state.L, dt=state.dt, **kw)
        s.set_state(state.q, state.p)
        return s

    def _chk(self, rc):
        if rc != 0:
            raise MdxError(rc, self.lib.mdx_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mdx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, q, p):
        q = np.ascontiguousarray(q, dtype=np.float64)
        p = np.ascontiguousarray(p, dtype=np.float64)
        assert q.shape == p.shape and q.ndim == 2 and q.shape[1] == 3
        self.n = q.shape[0]
        self._chk(self.lib.mdx_set_state(self.h, self.n, _dp(q), _dp(p)))

    def state(self):
        q = np.empty((self.n, 3))
        p = np.empty((self.n, 3))
        self._chk(self.lib.mdx_get_state(self.h, _dp(q), _dp(p)))
        return q, p

    def particle_number(self):
        pn, tn = ctypes.c_int(), ctypes.c_int()
        self._chk(self.lib.mdx_particle_number(self.h, ctypes.byref(pn), ctypes.byref(tn)))
        return pn.value, tn.value

    def images(self):
        """(q, p, source id, shift code) of the periodic images (the
        reference's ghosts), in the library's image order."""
        pn, tn = self.particle_number()
        g = tn - pn
        q = np.empty((g, 3))
        p = np.empty((g, 3))
        ids = np.empty(g, dtype=np.int32)
        code = np.empty(g, dtype=np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        self._chk(self.lib.mdx_get_images(self.h, _dp(q), _dp(p), ids.ctypes.data_as(i32),
                                          code.ctypes.data_as(i32)))
        return q, p, ids, code

    def make_pairlist(self):
        self._chk(self.lib.mdx_make_pairlist(self.h))

    def is_expired(self) -> bool:
        e = ctypes.c_int()
        self._chk(self.lib.mdx_is_pairlist_expired(self.h, ctypes.byref(e)))
        return bool(e.value)

    def update_position_half(self):
        self._chk(self.lib.mdx_update_position_half(self.h))

    def refresh_images(self):
        self._chk(self.lib.mdx_refresh_images(self.h))

    def calculate_force(self):
        self._chk(self.lib.mdx_calculate_force(self.h))

    def calculate(self, nsteps: int) -> int:
        r = ctypes.c_int()
        self._chk(self.lib.mdx_calculate(self.h, nsteps, ctypes.byref(r)))
        return r.value

    def observe(self):
        ke, pe, vir = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.mdx_observe(self.h, ctypes.byref(ke), ctypes.byref(pe), ctypes.byref(vir)))
        return ke.value, pe.value, vir.value

    def forces(self):
        f = np.empty((self.n, 3))
        self._chk(self.lib.mdx_forces(self.h, _dp(f)))
        return f

    def pair_number(self) -> int:
        c = ctypes.c_int64()
        self._chk(self.lib.mdx_pair_number(self.h, ctypes.byref(c)))
        return c.value

    def pairs(self) -> np.ndarray:
        c = ctypes.c_int64()
        self._chk(self.lib.mdx_pairs_canonical(self.h, None, 0, ctypes.byref(c)))
        out = np.empty((c.value, 3), dtype=np.int32)
        c2 = ctypes.c_int64()
        self._chk(self.lib.mdx_pairs_canonical(
            self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), c.value, ctypes.byref(c2)))
        assert c2.value == c.value
        return out

    def pairs_csr(self):
        """MeshList's CSR half list: (key_pointer [tn], number_of_partners
        [tn], sorted_list) over reals by id then images (mdx_pairs_csr)."""
        _, tn = self.particle_number()
        i32 = ctypes.POINTER(ctypes.c_int32)
        kp = np.zeros(tn, dtype=np.int32)
        npart = np.zeros(tn, dtype=np.int32)
        c = ctypes.c_int64()
        self._chk(self.lib.mdx_pairs_csr(self.h, kp.ctypes.data_as(i32),
                                         npart.ctypes.data_as(i32), None, 0, ctypes.byref(c)))
        sl = np.zeros(max(c.value, 1), dtype=np.int32)
        c2 = ctypes.c_int64()
        self._chk(self.lib.mdx_pairs_csr(self.h, None, None, sl.ctypes.data_as(i32), c.value,
                                         ctypes.byref(c2)))
        return kp, npart, sl[:c.value]

    def adjust_periodic_boundary(self):
        self._chk(self.lib.mdx_adjust_periodic_boundary(self.h))

    def scale_positions(self, alpha: float):
        self._chk(self.lib.mdx_scale_positions(self.h, float(alpha)))

    def set_box(self, L):
        L3 = np.ascontiguousarray(L, dtype=np.float64)
        self._chk(self.lib.mdx_set_box(self.h, _dp(L3)))

    def cell_order(self):
        ids = np.zeros(self.n, dtype=np.int32)
        self._chk(self.lib.mdx_get_cell_order(self.h, ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return ids

    def density_grid(self, grid_size: float):
        """BubbleHist's counts per grid box (uint8, modulo 256) as an
        array [gz][gy][gx] (mdx_density_grid)."""
        dims = np.zeros(3, dtype=np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        self._chk(self.lib.mdx_density_grid(self.h, grid_size, dims.ctypes.data_as(i32), None, 0))
        out = np.zeros(int(np.prod(dims)), dtype=np.uint8)
        self._chk(self.lib.mdx_density_grid(self.h, grid_size, dims.ctypes.data_as(i32),
                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                            out.size))
        return out.reshape(dims[2], dims[1], dims[0])

    def timing_reset(self, enable=True):
        self._chk(self.lib.mdx_timing_reset(self.h, 1 if enable else 0))

    def timing(self):
        b, f, o = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        nb, ns = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.lib.mdx_timing(self.h, ctypes.byref(b), ctypes.byref(f), ctypes.byref(o),
                                      ctypes.byref(nb), ctypes.byref(ns)))
        return dict(build_ms=b.value, force_ms=f.value, other_ms=o.value, builds=nb.value,
                    steps=ns.value)

    def kick_profile(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.mdx_kick_profile(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_heatbath(self, kind: str, aimed_temperature: float = 1.0, tau: float = 0.1,
                     gamma: float = 0.1, zeta: float = 0.0, seed: int = 0):
        """SimulationInfo's heat bath (ControlTemperature / HeatbathType /
        AimedTemperature / HeatbathTau / HeatbathGamma) and Variables::Zeta."""
        hb = MdxHeatbath(HEATBATH[kind], aimed_temperature, tau, gamma, zeta, seed)
        self._chk(self.lib.mdx_set_heatbath(self.h, ctypes.byref(hb)))

    def heatbath(self) -> dict:
        hb = MdxHeatbath()
        self._chk(self.lib.mdx_get_heatbath(self.h, ctypes.byref(hb)))
        return {k: getattr(hb, k) for k, _ in MdxHeatbath._fields_}

    def heatbath_zeta(self, current_temperature: float):
        self._chk(self.lib.mdx_heatbath_zeta(self.h, current_temperature))

    def heatbath_momenta(self):
        self._chk(self.lib.mdx_heatbath_momenta(self.h))

    def langevin(self):
        self._chk(self.lib.mdx_langevin(self.h))

    def lj_trajectory(self, eps=1.0, sigma=1.0, nsteps=200, nobs=100, checkpoint_every=20,
                      p_target=1.5, pe_target=-4.0, grad=True, want_dq0=True):
        """Loss of the differentiable trajectory and, with grad, its gradient
        (dict with loss, pressure_mean, pe_mean, dloss_deps, dloss_dsigma,
        dloss_dq0 [n][3] by id, forward_ms, backward_ms)."""
        r = MdxLjResult()
        dq0 = np.zeros((self.n, 3)) if (grad and want_dq0) else None
        self._chk(self.lib.mdx_lj_trajectory(self.h, eps, sigma, nsteps, nobs, checkpoint_every,
                                             p_target, pe_target, 1 if grad else 0,
                                             ctypes.byref(r), _dp(dq0) if dq0 is not None else None))
        out = {k: getattr(r, k) for k, _ in MdxLjResult._fields_}
        out["dloss_dq0"] = dq0
        return out


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Ensemble:
    """Independent periodic systems ("replicas") advanced together on one
    GPU; each replica's trajectory is the one :class:`System` gives it alone.
    ``states`` is a sequence of objects with ``q``, ``p`` ([n][3]) and ``L``."""

    def __init__(self, states, dt=0.001, cutoff=3.0, buffer_length=0.3, check_list_length=500,
                 device=0, stream=None):
        self.lib = load()
        prm = MdxParams()
        self.lib.mdx_default_params(ctypes.byref(prm))
        prm.dt, prm.cutoff, prm.buffer_length = dt, cutoff, buffer_length
        prm.check_list_length, prm.device, prm.stream = check_list_length, device, stream
        self.sizes = np.array([s.q.shape[0] for s in states], dtype=np.int32)
        L3 = np.ascontiguousarray([[float(x) for x in s.L] for s in states], dtype=np.float64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.R = len(states)
        h = ctypes.c_void_p()
        rc = self.lib.mdx_ensemble_create(ctypes.byref(prm), self.R, _ip(self.sizes), _dp(L3),
                                          ctypes.byref(h))
        if rc != 0:
            raise MdxError(rc, self.lib.mdx_ensemble_last_error(None).decode())
        self.h = h
        self.set_state([s.q for s in states], [s.p for s in states])

    def _chk(self, rc):
        if rc != 0:
            raise MdxError(rc, self.lib.mdx_ensemble_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.mdx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, qs, ps):
        q = np.ascontiguousarray(np.concatenate(qs), dtype=np.float64)
        p = np.ascontiguousarray(np.concatenate(ps), dtype=np.float64)
        assert q.shape == (self.offsets[-1], 3) and p.shape == q.shape
        self._chk(self.lib.mdx_ensemble_set_state(self.h, _dp(q), _dp(p)))

    def state(self):
        """[(q, p)] per replica, by particle id."""
        n = int(self.offsets[-1])
        q = np.empty((n, 3))
        p = np.empty((n, 3))
        self._chk(self.lib.mdx_ensemble_get_state(self.h, _dp(q), _dp(p)))
        o = self.offsets
        return [(q[o[r]:o[r + 1]], p[o[r]:o[r + 1]]) for r in range(self.R)]

    def calculate(self, nsteps: int):
        """Runs nsteps; returns each replica's pairlist constructions so far."""
        reb = np.zeros(self.R, dtype=np.int32)
        self._chk(self.lib.mdx_ensemble_calculate(self.h, nsteps, _ip(reb)))
        return reb

    def status(self):
        st = np.zeros(self.R, dtype=np.int32)
        steps = np.zeros(self.R, dtype=np.int32)
        self._chk(self.lib.mdx_ensemble_status(self.h, _ip(st), _ip(steps)))
        return st, steps

    def observe(self):
        ke, pe, vir = (np.zeros(self.R) for _ in range(3))
        self._chk(self.lib.mdx_ensemble_observe(self.h, _dp(ke), _dp(pe), _dp(vir)))
        return ke, pe, vir

    def config(self):
        c = MdxEnsembleConfig()
        self._chk(self.lib.mdx_ensemble_get_config(self.h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in MdxEnsembleConfig._fields_}

    def set_config(self, tile: int, grid_mult: int = 2):
        self._chk(self.lib.mdx_ensemble_set_config(self.h, tile, grid_mult))

    def autotune(self, trial_steps: int = 50):
        c = MdxEnsembleConfig()
        self._chk(self.lib.mdx_ensemble_autotune(self.h, trial_steps, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in MdxEnsembleConfig._fields_}
