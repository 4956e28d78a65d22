"""Parity of the HIP path (libmdx through its C ABI) with the oracle and the
reference's golden vectors.

Bars: the canonical pair set bit-exact; forces within 1e-12 of the largest
force (max_i |dF_i| / max_i |F_i|, per-particle relative error being
ill-posed where forces cancel); PE / virial / KE within 1e-12 relative;
1000-step trajectories within the chaos-limited bounds stated per test.
"""
import hashlib

import numpy as np
import pytest

from conftest import GOLDEN_CASES, golden_state, load_golden
from lattice_sums import fcc_per_particle
from mdacp_amd import workloads as W
from mdacp_amd.mdx import MdxError, System
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12
OBS_TOL = 1e-12


def digest(pairs):
    return hashlib.sha256(np.ascontiguousarray(pairs, dtype="<i4").tobytes()).hexdigest()


def force_err(f, ref):
    return np.abs(f - ref).max() / max(np.abs(ref).max(), 1e-300)


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_pairs_forces_observables_vs_reference_fixture(name):
    fx = load_golden(name)
    s = golden_state(fx)
    with System.from_state(s) as g:
        g.make_pairlist()
        pairs = g.pairs()
        assert len(pairs) == int(fx["pair_count"])
        assert digest(pairs) == fx["pair_digest"].item().decode()
        if "pairs" in fx:
            np.testing.assert_array_equal(pairs, fx["pairs"])
        f = g.forces()
        ref = fx["forces"]
        if np.abs(ref).max() > 1e-6:
            assert force_err(f, ref) < FORCE_TOL
        else:  # perfect lattice: forces cancel to rounding
            assert np.abs(f).max() < 1e-10
        ke, pe, vir = g.observe()
        np.testing.assert_allclose([ke, pe, vir], fx["obs"], rtol=OBS_TOL, atol=1e-9)


@pytest.mark.parametrize("seed,amp,ncell,rho", [(31, 0.3, 12, 0.5), (32, 0.2, 9, 0.8),
                                                (33, 0.45, 8, 0.35), (34, 0.25, 7, 0.9)])
def test_pairs_and_forces_vs_oracle(seed, amp, ncell, rho):
    s = W.perturbed(W.fcc_state(ncell, rho, 1.0, seed), amp, seed + 100)
    o = O.Oracle(s)
    o.make_pairlist()
    with System.from_state(s) as g:
        g.make_pairlist()
        np.testing.assert_array_equal(g.pairs(), o.pairs())
        assert force_err(g.forces(), o.forces()) < FORCE_TOL
        np.testing.assert_allclose(g.observe(), o.observe(), rtol=OBS_TOL)
        assert g.particle_number()[1] == o.total_particle_number()


def test_rectangular_box_and_inhomogeneous_occupancy():
    # a dense slab in a vapour: skewed cell occupancy, rectangular box
    s = W.slab_state(seed=3, Lx=40.0, Ly=20.0, Lz=16.0)
    o = O.Oracle(s)
    o.make_pairlist()
    with System.from_state(s) as g:
        g.make_pairlist()
        np.testing.assert_array_equal(g.pairs(), o.pairs())
        assert force_err(g.forces(), o.forces()) < FORCE_TOL
        np.testing.assert_allclose(g.observe(), o.observe(), rtol=OBS_TOL)


def test_trajectory_vs_reference_fixture():
    """1000 NVE steps at N=4000 against the reference's own run: the same
    rebuild schedule, observables to 1e-9 relative (rounding differences
    in the force sums grow by the chaos of the fluid, ~1e-13 after 1000
    steps between the reference and the oracle), final positions to 1e-8."""
    fx = load_golden("traj_fcc4000")
    s = golden_state(fx)
    with System.from_state(s) as g:
        rows, nreb = [], 0
        for k in range(10):
            nreb += g.calculate(100)
            g.refresh_images()
            ke, pe, vir = g.observe()
            rows.append((100 * (k + 1), ke, pe, vir, nreb))
        rows = np.array(rows)
        ref = fx["series"]
        np.testing.assert_array_equal(rows[:, 4], ref[:, 4])
        np.testing.assert_allclose(rows[:, 1:4], ref[:, 1:4], rtol=1e-9)
        q, p = g.state()
    d = q - fx["q"]
    d -= s.L * np.round(d / s.L)
    assert np.abs(d).max() < 1e-8
    assert np.abs(p - fx["p"]).max() < 1e-7


def test_trajectory_vs_oracle_step_by_step():
    s = W.perturbed(W.fcc_state(8, 0.6, 1.3, 41), 0.2, 42)
    o = O.Oracle(s)
    with System.from_state(s) as g:
        for _ in range(5):
            ro = o.calculate(40)
            rg = g.calculate(40)
            assert rg == ro
            qo, po = o.state()
            qg, pg = g.state()
            d = qg - qo
            d -= s.L * np.round(d / s.L)
            assert np.abs(d).max() < 1e-10
            assert np.abs(pg - po).max() < 1e-9


def test_deterministic_bitwise():
    s = W.perturbed(W.plain_small(), 0.2, 51)
    out = []
    for _ in range(2):
        with System.from_state(s) as g:
            g.calculate(30)
            out.append(g.state())
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_full_size_fcc_properties():
    """N = 2,048,000 (80^3 FCC cells, rho 0.5): size-independent checks —
    39 partners per particle, the closed-form PE and virial, zero net
    force, and zero net momentum kept over 20 steps."""
    s = W.plain_full()
    pe_n, vir_n, half_pairs, _ = fcc_per_particle(2.0)
    with System.from_state(s) as g:
        g.make_pairlist()
        assert g.pair_number() == half_pairs * s.n
        ke, pe, vir = g.observe()
        assert pe / s.n == pytest.approx(pe_n, rel=1e-10)
        assert vir / s.n == pytest.approx(vir_n, rel=1e-10)
        f = g.forces()
        assert np.abs(f).max() < 1e-9
        p0 = s.p.sum(axis=0)
        g.calculate(20)
        q, p = g.state()
        np.testing.assert_allclose(p.sum(axis=0), p0, atol=1e-8)


def test_full_size_pairs_vs_oracle():
    """N = 2,048,000 after 0.25 sigma random displacements: the canonical
    pair set equals the oracle's (digest) and forces agree."""
    s = W.perturbed(W.plain_full(), 0.25, 61)
    o = O.Oracle(s)
    o.make_pairlist()
    op = o.pairs()
    of = o.forces()
    with System.from_state(s) as g:
        g.make_pairlist()
        gp = g.pairs()
        assert len(gp) == len(op)
        assert digest(gp) == digest(op)
        assert force_err(g.forces(), of) < FORCE_TOL


def test_full_size_run_through_rebuilds_vs_oracle():
    """N = 2,048,000 from the FCC start, stepped until the pairlist has been
    rebuilt at least twice (about 60 steps).  The total energy stays within
    3e-5 relative of its start: the bar the second-order test below sets at
    N = 4000 over 2000 steps from the same lattice, density and temperature
    (the integrator's relative error is intensive, so it carries over to the
    larger system and the shorter run).  Then, on the evolved state, the
    forces of the list built inside the run equal the oracle's from a fresh
    list, and a fresh build equals the oracle's pair set (digest) and forces."""
    s = W.plain_full()
    with System.from_state(s) as g:
        g.make_pairlist()
        ke, pe, _ = g.observe()
        e0, worst, rebuilds, steps = ke + pe, 0.0, 0, 0
        while rebuilds < 2 and steps < 400:
            rebuilds += g.calculate(10)
            steps += 10
            g.refresh_images()
            ke, pe, _ = g.observe()
            worst = max(worst, abs(ke + pe - e0) / abs(e0))
        print(f"full size: {steps} steps, {rebuilds} rebuilds, energy error {worst:.2e}", flush=True)
        assert rebuilds >= 2
        assert worst < 3e-5
        q, p = g.state()
        t = W.State(q, p, s.L.copy(), s.dt)
        o = O.Oracle(t)
        o.make_pairlist()
        of = o.forces()
        assert force_err(g.forces(), of) < FORCE_TOL       # the run's own list
        g.set_state(q, p)
        g.make_pairlist()
        gp, op = g.pairs(), o.pairs()
        assert len(gp) == len(op)
        assert digest(gp) == digest(op)
        assert force_err(g.forces(), of) < FORCE_TOL


def test_full_size_inhomogeneous_slab():
    """The inhomogeneous workload at full size: 200 x 100 x 100 box, liquid
    slab (rho 0.8, T 0.9) in vapour (rho 0.05), 850,000 particles, skewed
    cell occupancy (about 30 vs 2 per cell).  Pair set (digest), forces
    and observables against the oracle; then 50 steps keep the momentum."""
    s = W.slab_state()
    print(f"slab N={s.n}", flush=True)
    o = O.Oracle(s)
    o.make_pairlist()
    op, of, oo = o.pairs(), o.forces(), o.observe()
    with System.from_state(s) as g:
        g.make_pairlist()
        gp = g.pairs()
        assert len(gp) == len(op)
        assert digest(gp) == digest(op)
        assert force_err(g.forces(), of) < FORCE_TOL
        np.testing.assert_allclose(g.observe(), oo, rtol=OBS_TOL)
        p0 = s.p.sum(axis=0)
        g.calculate(50)
        q, p = g.state()
        assert np.isfinite(q).all() and np.isfinite(p).all()
        np.testing.assert_allclose(p.sum(axis=0), p0, atol=1e-8)


def test_error_paths():
    s = W.plain_small()
    s.q[9, 2] = np.nan
    with System.from_state(s) as g:
        with pytest.raises(MdxError) as e:
            g.make_pairlist()
        assert e.value.code == -3
    s = W.plain_small()
    s.p[3] = [0.0, 160.0, 0.0]
    with System.from_state(s) as g:
        g.make_pairlist()
        with pytest.raises(MdxError) as e:
            g.calculate(1)
        assert e.value.code == -4
    with System(W.plain_small().L) as g:
        with pytest.raises(MdxError) as e:
            g.calculate_force()
        assert e.value.code == -7


def test_pair_set_and_forces_permutation_invariant():
    """The same configuration loaded in another particle order: the same
    canonical pair set once the ids are mapped back, the same forces (to the
    force bar: a cell's particle order follows the ids, so sums may be
    ordered differently)."""
    s = W.perturbed(W.fcc_state(7, 0.6, 1.0, 81), 0.25, 82)
    perm = np.random.default_rng(83).permutation(s.n)
    t = W.State(s.q[perm].copy(), s.p[perm].copy(), s.L.copy(), s.dt)
    with System.from_state(s) as a, System.from_state(t) as b:
        a.make_pairlist()
        b.make_pairlist()
        pa, pb = a.pairs(), b.pairs()
        fa, fb = a.forces(), b.forces()
    # ids of b are positions in perm: id k of b is particle perm[k] of a
    lo, hi = perm[pb[:, 0]], perm[pb[:, 1]]
    code = pb[:, 2].copy()
    swap = lo > hi
    # a pair (i, j, code) read from the other side carries the opposite shift
    sx, sy, sz = code % 3 - 1, (code // 3) % 3 - 1, code // 9 - 1
    flipped = (1 - sx) + 3 * (1 - sy) + 9 * (1 - sz)
    mapped = np.stack([np.where(swap, hi, lo), np.where(swap, lo, hi),
                       np.where(swap, flipped, code)], axis=1).astype(np.int32)
    mapped = mapped[np.lexsort((mapped[:, 2], mapped[:, 1], mapped[:, 0]))]
    np.testing.assert_array_equal(mapped, pa)
    assert force_err(fb, fa[perm]) < FORCE_TOL


def test_energy_error_second_order_in_dt():
    """NVE drift-kick-drift at N=4000 over 2 time units: the largest total
    energy error shrinks 4x when dt halves (a second-order symplectic
    integrator; a first-order step order or a force/energy mismatch breaks
    this), stays below 3e-5 relative at dt = 0.001, and the energy follows
    the oracle's to 1e-9 relative at every sample."""
    errs = []
    for dt in (0.002, 0.001):
        s = golden_state(load_golden("fcc4000"))
        s.dt = dt
        o = O.Oracle(s)
        nsub = int(round(0.1 / dt))
        with System.from_state(s) as g:
            g.make_pairlist()
            ke, pe, _ = g.observe()
            e0, worst = ke + pe, 0.0
            for _ in range(20):
                g.calculate(nsub)
                o.calculate(nsub)
                g.refresh_images()
                o.refresh_images()
                ke, pe, _ = g.observe()
                ko, po, _ = o.observe()
                assert (ke + pe) == pytest.approx(ko + po, rel=1e-9)
                worst = max(worst, abs(ke + pe - e0) / abs(e0))
        errs.append(worst)
    assert errs[1] < 3e-5
    assert 3.5 < errs[0] / errs[1] < 4.5


def _edge_state(case):
    rng = np.random.default_rng(91)
    if case == "single":
        q, L = np.array([[1.0, 2.0, 3.0]]), 10.0
    elif case == "pair_across_boundary":
        q, L = np.array([[0.3, 5.0, 5.0], [9.2, 5.0, 5.0]]), 10.0       # 1.1 apart through x = 0
    elif case == "pair_at_cutoff":
        q, L = np.array([[2.0, 5.0, 5.0], [5.0, 5.0, 5.0]]), 10.0       # exactly rc apart
    elif case == "pair_at_search_length":
        q, L = np.array([[2.0, 5.0, 5.0], [5.3, 5.0, 5.0]]), 10.0       # 3.3 apart (rounded)
    elif case == "sub_tile":
        # 37 of the 64 sites of a spacing-2 cubic grid, jittered
        g = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(-1, 3) * 2.0
        q, L = g[rng.permutation(64)[:37]] + rng.uniform(-0.1, 0.1, size=(37, 3)) + 0.5, 8.0
    else:   # narrow_box: just wider than two search lengths, FCC 3^3 jittered
        s0 = W.perturbed(W.fcc_state(3, 4.0 / (6.7 / 3) ** 3, 1.0, 92), 0.1, 93)
        q, L = s0.q, 6.7
    p = rng.normal(0, 0.5, size=q.shape)
    p -= p.mean(axis=0)
    return W.State(q, p, np.array([L, L, L]), 0.001)


@pytest.mark.parametrize("case", ["single", "pair_across_boundary", "pair_at_cutoff",
                                  "pair_at_search_length", "sub_tile", "narrow_box"])
def test_edge_cases_vs_oracle(case):
    """Tiny and boundary inputs: one particle, a pair through the periodic
    boundary, pairs exactly at rc and at the search length, fewer particles
    than a tile, a box just wider than two search lengths.  Pair set
    bit-exact, forces and observables to the bars, and 200 steps along the
    oracle's trajectory."""
    s = _edge_state(case)
    o = O.Oracle(s)
    o.make_pairlist()
    with System.from_state(s) as g:
        g.make_pairlist()
        np.testing.assert_array_equal(g.pairs(), o.pairs())
        fo = o.forces()
        fg = g.forces()
        if np.abs(fo).max() > 1e-6:
            assert force_err(fg, fo) < FORCE_TOL
        else:
            # no force, or the rounding residue of the shifted force at rc
            # (zero in exact arithmetic): absolute bar
            assert np.abs(fg - fo).max() < 1e-12
        np.testing.assert_allclose(g.observe(), o.observe(), rtol=OBS_TOL, atol=1e-14)
        assert g.calculate(200) == o.calculate(200)
        qg, pg = g.state()
        qo, po = o.state()
        d = qg - qo
        d -= s.L * np.round(d / s.L)
        assert np.abs(d).max() < 1e-10 and np.abs(pg - po).max() < 1e-9


def _lowest_in_next_cell(c, inv):
    """The smallest double that bins into cell c + 1 (1-based cells, bin =
    floor(x inv) + 1): it may lie just below the face c / inv."""
    v = c / inv
    while np.floor(v * inv) < c:
        v = np.nextafter(v, np.inf)
    while np.floor(np.nextafter(v, -np.inf) * inv) >= c:
        v = np.nextafter(v, -np.inf)
    return v


def _face_pairs_state(seed=96, per_step=60):
    """Pairs at the search length whose partner sits on the y and z faces of
    its cell: partner j at (x0 +- c, yf, zf) with yf, zf the lowest positions
    binned into the next cells, real i at (x0, yf - a, zf - b), c =
    sqrt(SL^2 - dy^2 - dz^2) moved by -3 .. 3 ulp, so r^2 falls on both
    sides of SL^2 and j lies in a row above i's (or the same row when a or b
    is 0), where the build's x window around i is at its narrowest.  A box
    whose cell width L / ms is not a binary fraction."""
    rng = np.random.default_rng(seed)
    L, SL = 25.0, 3.0 + 0.3
    ms = int(np.floor(L / SL))
    inv = ms / L
    q = []
    for k in range(-3, 4):
        for _ in range(per_step):
            cy, cz = rng.integers(1, ms, size=2)
            yf, zf = _lowest_in_next_cell(cy, inv), _lowest_in_next_cell(cz, inv)
            a, b = rng.uniform(0.0, 3.2, size=2) * rng.integers(0, 2, size=2)
            if a * a + b * b > 0.95 * SL * SL:
                a, b = 0.5 * a, 0.5 * b
            yi, zi = yf - a, zf - b
            c = np.sqrt(SL * SL - (yf - yi) ** 2 - (zf - zi) ** 2)
            c = c + k * np.spacing(c)
            x0 = rng.uniform(SL + 0.1, L - SL - 0.1)
            q.append([x0, yi, zi])
            q.append([x0 + rng.choice([-1.0, 1.0]) * c, yf, zf])
    q = np.array(q)
    p = rng.normal(0, 0.5, size=q.shape)
    p -= p.mean(axis=0)
    return W.State(q, p, np.array([L, L, L]), 0.001)


def _windows_without_margins_miss(s, pairs):
    """How many listed pairs (i, j) of the same image (code 13) an x window
    with no rounding allowance would drop: |x_j - x_i| <= sqrt(SL^2 - gy^2 -
    gz^2), gy, gz the distances from i to the faces of j's row.  Run on the
    CPU, it shows that the case is sharp enough to catch a window that is
    too narrow."""
    SL = 3.0 + 0.3
    ms = np.floor(s.L / SL).astype(int)
    inv = ms / s.L

    def cell(x, d):
        return min(max(int(np.floor(x * inv[d])) + 1, 1), ms[d])

    miss = 0
    for lo, hi, code in pairs.tolist():
        if code != 13:
            continue
        for i, j in ((lo, hi), (hi, lo)):
            qi, qj = s.q[i], s.q[j]
            g2 = 0.0
            for d in (1, 2):
                ci, cj = cell(qi[d], d), cell(qj[d], d)
                if cj < ci:
                    g2 += (qi[d] - (ci - 1) / inv[d]) ** 2
                elif cj > ci:
                    g2 += (ci / inv[d] - qi[d]) ** 2
            h2 = SL * SL - g2
            miss += not (h2 >= 0 and abs(qj[0] - qi[0]) <= np.sqrt(h2))
    return miss


def test_pairs_at_search_length_on_cell_faces():
    """The build cuts each cell row of a real's 27 cells to the x window that
    its y-z distance to the row allows, widened for rounding; pairs at the
    search length against the cell faces are listed exactly as the oracle
    lists them, including those a window without that allowance drops."""
    s = _face_pairs_state()
    o = O.Oracle(s)
    o.make_pairlist()
    op = o.pairs()
    own = {(2 * m, 2 * m + 1) for m in range(s.n // 2)}
    listed = sum(1 for lo, hi, _ in op.tolist() if (lo, hi) in own)
    assert 0 < listed < s.n // 2        # built pairs on both sides of SL^2
    assert _windows_without_margins_miss(s, op) > 0
    with System.from_state(s) as g:
        g.make_pairlist()
        np.testing.assert_array_equal(g.pairs(), op)


def test_rejected_inputs():
    """No particles, and a box narrower than two search lengths, are refused
    with an error (the reference needs a domain of at least that width for
    its images; here the status names it)."""
    with pytest.raises(MdxError):
        with System(np.array([10.0, 10.0, 10.0])) as g:
            g.set_state(np.zeros((0, 3)), np.zeros((0, 3)))
    s = _edge_state("narrow_box")
    s.L = np.array([6.5, 6.5, 6.5])
    s.q = np.mod(s.q, 6.5)
    with pytest.raises(MdxError):
        with System.from_state(s) as g:
            g.make_pairlist()
