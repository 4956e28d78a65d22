"""The reference's MeshList / Variables / MDUnit surface in the C++ mirror
and the C ABI under it: the CSR half list (mdx_pairs_csr, the getters of
meshlist.h:66-88), ChangeScale (mdunit.cc:217-222), Sort with SortParticle
(meshlist.cc:140-173), AdjustPeriodicBoundary and the unit's particle-number
calls.

Bars: the CSR's count equals the oracle's MeshList::GetPairNumber, and its
pairs, with images resolved, are the canonical pair set (bit-exact, each
pair across the box once from each side); an observer written against
the reference's interface reproduces the fixture's potential energy and
virial to 1e-12; a rescaled FCC lattice gives the closed forms of the
lattice sums to 1e-10; sorting leaves the trajectory (by particle id)
within 1e-9 of the unsorted one.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT, golden_state, load_golden
from lattice_sums import fcc_per_particle
from mdacp_amd import workloads as W
from mdacp_amd.mdx import MdxError, System
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

MIRROR = os.path.join(ROOT, "tests", "cpp", "_build", "mirror_driver")


def run_mirror(state, *args):
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.bin")
        W.write_state(path, state)
        argv = [MIRROR, args[0], path, *[str(a) for a in args[1:]]]
        argv = [a.replace("{tmp}", tmp) for a in argv]
        print("mirror_driver", args[0], f"N={state.n}", flush=True)
        r = subprocess.run(argv, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        out = {"sorted": 0}
        for ln in r.stdout.splitlines():
            if ln.startswith("#"):
                out["sorted"] += ln.strip() == "# Sorted!"   # MeshList::Sort's line
                continue
            k, *v = ln.split()
            if v:
                out[k] = [float(x) for x in v] if len(v) > 1 else float(v[0])
        files = {}
        for name in ("o_q.f64", "o_p.f64"):
            f = os.path.join(tmp, name)
            if os.path.exists(f):
                files[name] = np.fromfile(f).reshape(-1, 3)
        return out, files


def resolve(sys_, kp, npart, sl):
    """CSR entries -> canonical (id_lo, id_hi, code) triples, with images
    resolved through their source id and shift code."""
    pn, tn = sys_.particle_number()
    _, _, src, code = sys_.images()
    keys = np.repeat(np.arange(tn, dtype=np.int64), npart)
    part = sl.astype(np.int64)
    img = part >= pn
    j = np.where(img, src[np.minimum(part - pn, len(src) - 1)] if len(src) else part, part)
    c = np.where(img, code[np.minimum(part - pn, len(code) - 1)] if len(code) else 13, 13)
    # orient each pair so that the smaller id comes first; reversing the
    # pair reverses the image shift
    swap = keys > j
    lo = np.where(swap, j, keys)
    hi = np.where(swap, keys, j)
    c = np.where(swap, 26 - c, c)
    return np.stack([lo, hi, c], axis=1).astype(np.int32)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_csr_half_list_vs_oracle_and_canonical(name):
    s = golden_state(load_golden(name))
    o = O.Oracle(s)
    o.make_pairlist()
    with System.from_state(s) as g:
        g.make_pairlist()
        kp, npart, sl = g.pairs_csr()
        pn, tn = g.particle_number()
        assert len(sl) == o.pair_number()                     # MeshList::GetPairNumber
        assert kp[0] == 0 and (np.diff(kp) == npart[:-1]).all() and kp[-1] + npart[-1] == len(sl)
        assert (npart[pn:] == 0).all()
        keys = np.repeat(np.arange(tn), npart)
        assert (keys < sl).all()                              # key = the smaller index
        tri = resolve(g, kp, npart, sl)
        canon = g.pairs()
        u, counts = np.unique(tri, axis=0, return_counts=True)
        np.testing.assert_array_equal(u, canon)
        across = sl >= pn
        # a pair across the box is seen from both of its ends, a pair of
        # reals once
        assert (counts == np.where(u[:, 2] == 13, 1, 2)).all()
        assert across.sum() == 2 * (canon[:, 2] != 13).sum()


@pytest.mark.parametrize("name", ["fcc4000_pert", "fcc500_rho0.9_pert", "fcc864_rho0.3_pert"])
def test_reference_style_observer_through_mirror(name):
    """A DoubleObserver written against the reference's interface (key-partner
    pairs, ghosts weighted 1/2, observer.cc:18-79) gives the reference's PE
    and virial; the CSR getters agree with each other."""
    fx = load_golden(name)
    s = golden_state(fx)
    out, _ = run_mirror(s, "observe")
    assert out["csr_bad"] == 0
    assert out["pe"] == pytest.approx(fx["obs"][1], rel=1e-12)
    assert out["virial"] == pytest.approx(fx["obs"][2], rel=1e-12)
    assert out["canonical"] == int(fx["pair_count"])
    o = O.Oracle(s)
    o.make_pairlist()
    assert out["pairs"] == o.pair_number()
    assert out["total"] == o.total_particle_number()
    assert out["total_after_drop"] == s.n


@pytest.mark.parametrize("rho_to", [0.7, 0.3])
def test_change_scale_matches_lattice_closed_forms(rho_to):
    """FCC at rho 0.5 (N = 4000, L = 20) rescaled by MDUnit::ChangeScale to
    rho_to: the lattice sums of the new spacing (computed from the lattice,
    not from either code)."""
    s = W.fcc_state(10, 0.5, 1.0, 3)
    s.p[:] = 0.0
    alpha = (0.5 / rho_to) ** (1.0 / 3.0)
    out, _ = run_mirror(s, "scale", repr(alpha))
    pe, vir, half_pairs, _ = fcc_per_particle(2.0 * alpha)
    assert out["box"] == pytest.approx([20.0 * alpha] * 3, rel=1e-15)
    assert out["pe_per_particle"] == pytest.approx(pe, rel=1e-10)
    assert out["potential_per_particle"] == pytest.approx(pe, rel=1e-10)
    assert out["virial_per_particle"] == pytest.approx(vir, rel=1e-10)
    assert out["canonical"] == half_pairs * s.n


def test_sort_particle_keeps_the_trajectory():
    """SortParticle reorders the unit's particles every tenth construction;
    by particle id the trajectory is the unsorted one (to the rounding of a
    different summation order)."""
    s = W.perturbed(W.fcc_state(8, 0.6, 1.4, 71), 0.1, 72)
    a, fa = run_mirror(s, "run", 800, 1, "{tmp}/o")
    b, fb = run_mirror(s, "run", 800, 0, "{tmp}/o")
    assert a["constructions"] >= 11 and a["constructions"] == b["constructions"]
    # checked before each build, every tenth count, none at 0 (nothing binned yet)
    assert a["sorted"] == (a["constructions"] - 1) // 10 and b["sorted"] == 0
    d = fa["o_q.f64"] - fb["o_q.f64"]
    d -= s.L * np.round(d / s.L)
    assert np.abs(d).max() < 1e-9
    assert np.abs(fa["o_p.f64"] - fb["o_p.f64"]).max() < 1e-8


def test_set_box_and_scale_positions_through_c_abi():
    s = W.fcc_state(8, 0.5, 1.0, 5)
    with System.from_state(s) as g:
        g.make_pairlist()
        e0 = g.observe()[1] / s.n
        alpha = (0.5 / 0.9) ** (1.0 / 3.0)
        g.scale_positions(alpha)
        with pytest.raises(MdxError) as ei:
            g.forces()                    # the list is void after a rescale
        assert ei.value.code == -7
        g.set_box(s.L * alpha)
        g.make_pairlist()
        a = s.L[0] / 8
        assert g.observe()[1] / s.n == pytest.approx(fcc_per_particle(a * alpha)[0], rel=1e-10)
        assert e0 == pytest.approx(fcc_per_particle(a)[0], rel=1e-10)
        with pytest.raises(MdxError) as ei:
            g.set_box([6.0, 20.0, 20.0])  # below two search lengths
        assert ei.value.code == -5


def test_adjust_periodic_boundary_and_cell_order():
    s = W.perturbed(W.fcc_state(6, 0.5, 1.0, 7), 0.2, 8)
    q = s.q.copy()
    q[0] = [-0.25, s.L[1] + 0.5, 3.0]     # outside on two axes: wrapped
    with System.from_state(s) as g:
        g.set_state(q, s.p)
        g.adjust_periodic_boundary()
        qw, _ = g.state()
        assert qw[0] == pytest.approx([s.L[0] - 0.25, 0.5, 3.0], abs=1e-12)
        g.make_pairlist()
        order = g.cell_order()
        assert sorted(order.tolist()) == list(range(s.n))
        # cell order: the interior cell of each real is non-decreasing
        ms = np.floor(s.L / 3.3).astype(int)
        c = np.minimum(np.floor(qw[order] * (1.0 / (s.L / ms))).astype(int), ms - 1)
        lin = c[:, 0] + ms[0] * (c[:, 1] + ms[1] * c[:, 2])
        assert (np.diff(lin) >= 0).all()
        bad = s.q.copy()
        bad[1] = [3.0 * s.L[0], 1.0, 1.0]  # two boxes away: "Invalid position"
        g.set_state(bad, s.p)
        with pytest.raises(MdxError) as ei:
            g.adjust_periodic_boundary()
        assert ei.value.code == -3
