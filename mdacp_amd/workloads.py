"""Seeded synthetic initial states for the LJ step.

Every state is generated here, once, and handed identically to the GPU
library, the C oracle and the compiled reference harness, so that parity
never depends on the reference's own ``<random>`` usage
(``variables.cc:87-113`` uses ``std::mt19937``; ``confmaker.cc:193`` too).
The one exception is ``benchmark_state``: the reference's Benchmark-mode
initial conditions themselves, made by the C++ restatement in
``host/confmaker.cpp`` with the same ``<random>`` engines.

Positions follow the FCC site order of ``ConfigurationMaker::Execute``
(``confmaker.cc:9-44``): cells iz -> iy -> ix, basis (0,0,0), (0,h,h),
(h,0,h), (h,h,0), each coordinate shifted by +1e-7 so no particle sits on a
cell face.  Velocities are Maxwell-Boltzmann at temperature T from a
Box-Muller transform over numpy's PCG64 stream, with the mean momentum
removed and no rescale (mass = k_B = 1, so p = v).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os

import numpy as np

CUTOFF = 3.0          # CUTOFF_LENGTH, mdconfig.h:31
BUFFER_LENGTH = 0.3   # simulationinfo.cc:6
SEARCH_LENGTH = CUTOFF + BUFFER_LENGTH  # simulationinfo.cc:7
FCC_OFFSET = 1e-7     # confmaker.cc:71


@dataclasses.dataclass
class State:
    """Positions ``q`` (N,3), momenta ``p`` (N,3), box lengths ``L`` (3,)."""

    q: np.ndarray
    p: np.ndarray
    L: np.ndarray
    dt: float = 0.001
    name: str = ""

    @property
    def n(self) -> int:
        return int(self.q.shape[0])

    def copy(self) -> "State":
        return State(self.q.copy(), self.p.copy(), self.L.copy(), self.dt, self.name)


def fcc_lattice(ncell: int, a: float, offset: float = FCC_OFFSET) -> np.ndarray:
    """FCC sites of an ncell^3 block with lattice constant ``a``.

    Same site order and arithmetic as ``confmaker.cc:72-93``:
    x = ix*s + e, x = ix*s + hs + e.
    """
    hs = a * 0.5
    iz, iy, ix = np.meshgrid(np.arange(ncell), np.arange(ncell), np.arange(ncell), indexing="ij")
    bx = ix.reshape(-1).astype(np.float64) * a
    by = iy.reshape(-1).astype(np.float64) * a
    bz = iz.reshape(-1).astype(np.float64) * a
    basis = [(0, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0)]
    q = np.empty((bx.size, 4, 3), dtype=np.float64)
    for k, (hx, hy, hz) in enumerate(basis):
        q[:, k, 0] = (bx + hs) + offset if hx else bx + offset
        q[:, k, 1] = (by + hs) + offset if hy else by + offset
        q[:, k, 2] = (bz + hs) + offset if hz else bz + offset
    return q.reshape(-1, 3)


def maxwell_momenta(n: int, temperature: float, seed: int) -> np.ndarray:
    """Box-Muller normal momenta at ``temperature`` with zero net momentum."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u1 = rng.random(3 * n)
    u2 = rng.random(3 * n)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)
    p = (math.sqrt(temperature) * z).reshape(n, 3)
    p -= p.mean(axis=0)
    return p


def fcc_state(ncell: int, density: float, temperature: float, seed: int,
              dt: float = 0.001, name: str = "") -> State:
    """Periodic FCC block at ``density`` (4 sites per cubic cell)."""
    a = (4.0 / density) ** (1.0 / 3.0)
    # Use L/ncell as the lattice constant so the lattice tiles the box exactly.
    L = a * ncell
    if abs(round(L) - L) < 1e-9:
        L = float(round(L))
    a = L / ncell
    q = fcc_lattice(ncell, a)
    p = maxwell_momenta(q.shape[0], temperature, seed)
    return State(q, p, np.array([L, L, L], dtype=np.float64), dt, name)


def plain_small(seed: int = 1) -> State:
    """N=4000 FCC 10^3 cells, a=2.0, L=20, rho=0.5, T=1."""
    return fcc_state(10, 0.5, 1.0, seed, name="fcc4000_rho0.5")


def plain_full(seed: int = 2) -> State:
    """N=2,048,000 FCC 80^3 cells, a=2.0, L=160, rho=0.5, T=1."""
    return fcc_state(80, 0.5, 1.0, seed, name="fcc2048000_rho0.5")


def perturbed(state: State, amplitude: float, seed: int) -> State:
    """Uniform random displacement in [-amplitude, amplitude) per coordinate,
    wrapped into [0, L).  Gives a liquid-like pair set that is not fixed by
    lattice symmetry."""
    rng = np.random.Generator(np.random.PCG64(seed))
    s = state.copy()
    s.q = s.q + rng.uniform(-amplitude, amplitude, size=s.q.shape)
    s.q = np.mod(s.q, s.L)
    s.q[s.q >= s.L] = 0.0
    s.name = f"{state.name}_pert{amplitude}"
    return s


def slab_state(seed: int = 3, Lx: float = 200.0, Ly: float = 100.0, Lz: float = 100.0,
               liquid_density: float = 0.8, vapour_density: float = 0.05,
               temperature: float = 0.9, min_sep: float = 0.8,
               dt: float = 0.001) -> State:
    """Inhomogeneous box: liquid slab x in [Lx/4, 3Lx/4) at liquid_density,
    vapour elsewhere; positions uniformly random in each region, a
    candidate rejected when it comes closer than min_sep (periodic) to an
    accepted point; momenta Maxwell at ``temperature``.

    Random sequential addition in batches: each batch of candidates is
    tested against the accepted points (k-d tree), and of two candidates of
    one batch closer than min_sep the later one is dropped."""
    from scipy.spatial import cKDTree
    rng = np.random.Generator(np.random.PCG64(seed))
    L = np.array([Lx, Ly, Lz])
    x0, x1 = Lx * 0.25, Lx * 0.75
    n_liq = int(round(liquid_density * (x1 - x0) * Ly * Lz))
    n_vap = int(round(vapour_density * (Lx - (x1 - x0)) * Ly * Lz))
    pts = np.empty((0, 3))

    def fill(count, draw):
        nonlocal pts
        added = 0
        while added < count:
            m = int(min(max(4096, 1.5 * (count - added)), 250000))
            cand = draw(m)
            if len(pts):
                d, _ = cKDTree(pts, boxsize=L).query(cand, k=1, distance_upper_bound=min_sep)
                cand = cand[~(d < min_sep)]
            if len(cand) > 1:
                pairs = cKDTree(cand, boxsize=L).query_pairs(min_sep, output_type="ndarray")
                if len(pairs):
                    keep = np.ones(len(cand), dtype=bool)
                    keep[pairs.max(axis=1)] = False
                    cand = cand[keep]
            cand = cand[:count - added]
            pts = np.vstack([pts, cand])
            added += len(cand)

    def liquid(m):
        return np.column_stack([rng.uniform(x0, x1, m), rng.uniform(0, Ly, m), rng.uniform(0, Lz, m)])

    def vapour(m):
        # x uniformly in the complement [x1, Lx + x0), mapped back into the box
        return np.column_stack([rng.uniform(x1, Lx + x0, m) % Lx, rng.uniform(0, Ly, m),
                                rng.uniform(0, Lz, m)])

    fill(n_liq, liquid)
    fill(n_vap, vapour)
    q = pts
    p = maxwell_momenta(q.shape[0], temperature, seed + 1000)
    return State(q, p, L, dt, name="slab")


def replica_set(count: int = 256, seed: int = 4, nmin: int = 5, nmax: int = 20):
    """Independent FCC replicas: n ~ U{nmin..nmax}, rho ~ U[0.3,0.9],
    T ~ U[0.7,1.5]; replica r uses seed ``seed + r`` for its momenta."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = rng.integers(nmin, nmax + 1, size=count)
    rhos = rng.uniform(0.3, 0.9, size=count)
    temps = rng.uniform(0.7, 1.5, size=count)
    return [fcc_state(int(ns[r]), float(rhos[r]), float(temps[r]), seed + r,
                      name=f"replica{r}") for r in range(count)]


def write_state(path: str, s: State) -> None:
    """Binary layout read by the reference harness:
    int64 N; double L[3]; double dt; double q[N][3]; double p[N][3]."""
    with open(path, "wb") as f:
        np.array([s.n], dtype=np.int64).tofile(f)
        np.asarray(s.L, dtype=np.float64).tofile(f)
        np.array([s.dt], dtype=np.float64).tofile(f)
        np.ascontiguousarray(s.q, dtype=np.float64).tofile(f)
        np.ascontiguousarray(s.p, dtype=np.float64).tofile(f)


def benchmark_state(procs: int = 8, threads: int = 8, unit_length: float = 45.0,
                    density: float = 0.712, v0: float = 0.9, dt: float = 0.001) -> State:
    """The initial state of the reference's Benchmark mode (benchmark.cc:12-19)
    for a run of ``procs`` MPI processes x ``threads`` OpenMP threads: box
    ``unit_length`` x the unit grid (simulationinfo.cc:35-39), each unit
    filled by ExtractConfigurationMaker (confmaker.cc:174-255) and given
    velocities by SetInitialVelocity(v0) (variables.cc:87-113), units in id
    order.  Defaults: the published 64-unit run (N = 4,152,384, L = 180,
    density 0.712, InitialVelocity 0.9).  Made by libmdacp_gpu.so (host
    code, no GPU)."""
    lib = ctypes.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmdacp_gpu.so"))
    f = lib.mdacp_benchmark_state
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                  ctypes.POINTER(ctypes.c_long), ctypes.c_void_p]
    n = ctypes.c_long()
    L = np.zeros(3)
    if f(procs, threads, unit_length, density, v0, None, None, 0, ctypes.byref(n), L.ctypes.data):
        raise ValueError("benchmark_state: bad arguments (process and thread counts are powers of two)")
    q = np.empty((n.value, 3))
    p = np.empty((n.value, 3))
    if f(procs, threads, unit_length, density, v0, q.ctypes.data, p.ctypes.data, n.value,
         ctypes.byref(n), L.ctypes.data):
        raise ValueError("benchmark_state failed")
    return State(q, p, L, dt, name=f"benchmark_{procs}x{threads}_L{unit_length:g}_rho{density:g}")
