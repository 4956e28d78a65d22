"""Closed forms of the perfect FCC lattice, computed from the mathematics
only (no oracle, no reference code): lattice vectors of the FCC lattice
with cubic constant a are a*(i,j,k)/2 with i+j+k even.  Per particle, the
quadratically shifted LJ of variables.cc:21-27 gives

    PE/N     = 1/2 sum_{0<r<=rc} 4 (r^-12 - r^-6 + C2 r^2 + C0)
    virial/N = 1/2 sum_{0<r<=rc} r V'(r) / 3
    pairs/N  = 1/2 #{0 < r <= SL}

with C2 = 6 rc^-14 - 3 rc^-8 and C0 = -rc^-12 + rc^-6 - C2 rc^2, chosen so
that V(rc) = V'(rc) = 0.
"""
from __future__ import annotations

import math
from fractions import Fraction


def shift_constants(rc: float = 3.0):
    c2 = 6.0 * rc ** -14 - 3.0 * rc ** -8
    c0 = -rc ** -12 + rc ** -6 - c2 * rc * rc
    return c0, c2


def fcc_shells(a: float, rmax: float):
    """{r^2: multiplicity} of FCC lattice vectors with 0 < r <= rmax."""
    kmax = int(math.ceil(2 * rmax / a)) + 1
    shells = {}
    for i in range(-kmax, kmax + 1):
        for j in range(-kmax, kmax + 1):
            for k in range(-kmax, kmax + 1):
                if (i + j + k) % 2 or (i == 0 and j == 0 and k == 0):
                    continue
                m2 = Fraction(i * i + j * j + k * k, 4)   # r^2 / a^2
                r2 = float(m2) * a * a
                if r2 <= rmax * rmax * (1 + 1e-12):
                    shells[m2] = shells.get(m2, 0) + 1
    return {float(m2) * a * a: c for m2, c in shells.items()}


def fcc_per_particle(a: float, rc: float = 3.0, sl: float = 3.3):
    c0, c2 = shift_constants(rc)
    pe = vir = 0.0
    for r2, c in fcc_shells(a, rc).items():
        pe += c * 4.0 * (r2 ** -6 - r2 ** -3 + c2 * r2 + c0)
        vir += c * (-48.0 * r2 ** -6 + 24.0 * r2 ** -3 + 8.0 * c2 * r2)
    half_pairs = sum(fcc_shells(a, sl).values()) / 2
    within_rc = sum(fcc_shells(a, rc).values()) / 2
    return 0.5 * pe, 0.5 * vir / 3.0, half_pairs, within_rc
