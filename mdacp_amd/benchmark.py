"""The reference's Benchmark mode (Benchmark::Run, benchmark.cc:12-52) on one
GPU through libmdx, for Python callers (bench.py, the tests).  The same loop
as the C++ driver ``mdx_benchmark`` (host/benchmark.cpp):

  MakePairList; ThermalizeLoop steps; then TotalLoop timed steps with an
  observation (time, temperature, pressure, energy per particle) after the
  steps i = 0, ObserveLoop, 2 ObserveLoop, ... of the timed loop, the
  observations inside the timed region; MUPS = TotalLoop N / seconds / 1e6.

Observables as MDManager computes them (mdmanager.cc:531-570), right after
the steps, with the images as Calculate leaves them.
"""
from __future__ import annotations

import time

import numpy as np

from .mdx import System


def run(state, thermalize: int = 150, loop: int = 1000, observe: int = 100, sync=None) -> dict:
    """Runs the Benchmark loop on ``state`` (e.g. workloads.benchmark_state()).
    ``sync``, if given, is called before the clock is read (a device sync)."""
    n = state.n
    V = float(state.L[0] * state.L[1] * state.L[2])
    rows = []
    with System.from_state(state) as g:
        g.make_pairlist()
        s_time = 0.0

        def advance(k):
            nonlocal s_time
            if k > 0:
                g.calculate(k)
                for _ in range(k):
                    s_time += state.dt   # MDManager::Calculate, mdmanager.cc:184

        advance(thermalize)
        if sync:
            sync()
        t0 = time.perf_counter()
        done = 0
        for i in range(0, loop, observe):
            advance(i + 1 - done)
            done = i + 1
            ke, pe, vir = g.observe()
            T = ke / n / 1.5
            P = (T - vir / n) * n / V
            rows.append((s_time, T, P, ke / n + pe / n))
        advance(loop - done)
        if sync:
            sync()
        sec = time.perf_counter() - t0
    return {"n": n, "rows": np.array(rows), "seconds": sec, "mups": loop * n / sec / 1e6}
