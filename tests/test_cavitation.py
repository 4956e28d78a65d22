"""The Cavitation project (cavitation.cc:238-287) and its BubbleHist density
grid (cavitation.cc:13-167).

Fixture cavitation_unit16 (tests/golden/make_golden_cavitation.py): the
reference's own program run as one unit on the Benchmark initial state at
UnitLength 16 (N = 2916): 100 Nose-Hoover steps, ChangeScale 1.1, 301 NVE
steps with seven BubbleHist analyses; the printed "# Thermalize" /
"# Observe" rows (6 significant digits) and the seven conf%04d.dat dumps.

Bars: the oracle's restatement of the run reproduces the fixture's dumps
byte for byte and its rows to the printed digits (CPU); its density grid
gives the closed-form counts of planted lattices (CPU); on the GPU,
mdx_density_grid equals the oracle's grid on the same positions byte for
byte, the run through the C ABI follows the oracle's to 1e-9 with equal
dumps, and mdx_benchmark (Mode=Cavitation, the C++ mirror's BubbleHist)
writes the fixture's dumps byte for byte and prints its rows.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from mdacp_amd import workloads as W
from oracle import pyoracle as O

EXE = os.path.join(ROOT, "mdacp_amd", "mdx_benchmark")
ROW_RTOL = 1e-5       # rows printed to 6 significant digits


def config(fx):
    kv = dict(l.split("=", 1) for l in str(fx["config"]).splitlines() if "=" in l)
    return {k: v.strip() for k, v in kv.items()}


def initial_state(cfg):
    return W.benchmark_state(1, 1, float(cfg["UnitLength"]), float(cfg["Density"]),
                             float(cfg["InitialVelocity"]))


def row(time, ke, pe, vir, n, L):
    """MDManager::Temperature / PotentialEnergy / Pressure / TotalEnergy
    (mdmanager.cc:531-570) from the sums."""
    T = ke / n / 1.5
    V = float(L[0] * L[1] * L[2])
    return [time, T, pe / n, (T - vir / n) * n / V, ke / n + pe / n]


def run_cavitation(cfg, make, advance, observe, rescale, density):
    """Cavitation::Run's schedule over an engine given as callables.  The
    rows are observed as MDManager does, with the images as Calculate
    leaves them (half a step behind their reals during the thermalization;
    fresh after the analysis's MakePairList)."""
    t_loop, loop = int(cfg["ThermalizeLoop"]), int(cfg["TotalLoop"])
    every, alpha = int(cfg["ObserveLoop"]), float(cfg["ChangeScale"])
    gsize = float(cfg["GridSize"])
    s = initial_state(cfg)
    dt, steps = s.dt, 0
    eng = make(s)
    th, ob, dumps = [], [], []
    for i in range(t_loop):
        advance(eng, 1)
        steps += 1
        if i % every == 0:
            th.append(row(steps * dt, *observe(eng), s.n, s.L))
    L = s.L * alpha                                   # MDManager::ChangeScale
    eng = rescale(eng, alpha, L)
    for i in range(loop + 1):
        advance(eng, 1)
        steps += 1
        if i % every == 0:
            dumps.append(density(eng, L, gsize))      # BubbleHist::Analyse
            ob.append(row(steps * dt, *observe(eng), s.n, L))
    return np.array(th), np.array(ob), np.stack(dumps)


def oracle_run(cfg):
    def make(s):
        o = O.Oracle(s)
        o.set_heatbath("nose_hoover", float(cfg["AimedTemperature"]), float(cfg["HeatbathTau"]))
        o.make_pairlist()
        return o

    def rescale(o, alpha, L):
        o.make_pairlist()
        q, p = o.state()
        dt = o.prm.dt
        o.close()
        o2 = O.Oracle(W.State(q * alpha, p, L.copy(), dt))   # heat bath off
        o2.make_pairlist()
        return o2

    def density(o, L, gsize):
        o.make_pairlist()
        q, _ = o.state()
        return O.density_grid(q, L, gsize)[1]

    return run_cavitation(cfg, make, lambda o, k: o.calculate(k), lambda o: o.observe(), rescale,
                          density)


@pytest.fixture(scope="module")
def oracle_result():
    fx = load_golden("cavitation_unit16")
    return fx, oracle_run(config(fx))


# ------------------------------------------------------------------ CPU


def test_fixture_is_consistent():
    fx = load_golden("cavitation_unit16")
    n, grid = int(fx["n"]), [int(g) for g in fx["grid"]]
    assert n == initial_state(config(fx)).n == 2916
    L1 = 16.0 * float(config(fx)["ChangeScale"])
    assert grid == [int(L1 / 3.0 + 0.5)] * 3
    assert float(fx["grid_size"]) == pytest.approx(L1 / grid[0], rel=1e-5)
    assert fx["dumps"].shape == (len(fx["observe"]), np.prod(grid))
    assert (fx["dumps"].astype(np.int64).sum(axis=1) == n).all()


def test_oracle_reproduces_reference_cavitation_run(oracle_result):
    fx, (th, ob, dumps) = oracle_result
    np.testing.assert_allclose(th, fx["thermalize"], rtol=ROW_RTOL)
    np.testing.assert_allclose(ob, fx["observe"], rtol=ROW_RTOL)
    np.testing.assert_array_equal(dumps, fx["dumps"])
    # a bubble forms in the stretched liquid: the spread of the counts grows
    assert fx["dumps"][-1].std() > fx["dumps"][0].std()


def test_oracle_density_grid_closed_forms():
    """Planted lattices with known counts per box."""
    L = np.array([12.0, 12.0, 12.0])
    g, data = O.density_grid(np.zeros((0, 3)), L, 3.0)
    assert g == [4, 4, 4] and (data == 0).all()
    rng = np.random.default_rng(5)
    want = rng.integers(0, 9, size=64)
    centres = (np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"),
                        -1).reshape(-1, 3)[:, ::-1] + 0.5) * 3.0     # box k = ix + 4 iy + 16 iz
    q = np.repeat(centres, want, axis=0) + rng.uniform(-1.4, 1.4, size=(want.sum(), 3))
    g, data = O.density_grid(q, L, 3.0)
    np.testing.assert_array_equal(data, want)
    # 2.9 -> int(12 / 2.9 + 0.5) = 4 boxes of 3.0; 5.0 -> 2 boxes of 6.0
    assert O.density_grid(q, L, 2.9)[0] == [4, 4, 4]
    g2, d2 = O.density_grid(q, L, 5.0)
    assert g2 == [2, 2, 2] and d2.astype(int).sum() == want.sum()
    # one box for all: the count modulo 256 (unsigned char)
    many = rng.uniform(0, 12, size=(700, 3))
    assert O.density_grid(many, L, 12.0)[1].tolist() == [700 % 256]
    # the box width comes from x for every axis: L_y = 10 gives 3 rows of
    # width 3, and y = 9.5 lands in row 3, i.e. the next layer's first row
    g3, d3 = O.density_grid(np.array([[0.5, 9.5, 0.5]]), np.array([12.0, 10.0, 10.0]), 3.0)
    assert g3 == [4, 3, 3] and np.flatnonzero(d3).tolist() == [12]
    with pytest.raises(O.OracleError):
        O.density_grid(np.array([[0.5, 0.5, 12.5]]), L, 3.0)


# ------------------------------------------------------------------ GPU


@pytest.mark.gpu
@pytest.mark.parametrize("gsize", [3.0, 2.2, 17.0])
def test_density_grid_matches_oracle(gsize):
    from mdacp_amd.mdx import System
    s = W.perturbed(W.fcc_state(8, 0.7, 1.0, 31), 0.3, 32)
    with System.from_state(s) as g:
        g.make_pairlist()
        q, _ = g.state()
        want_g, want = O.density_grid(q, s.L, gsize)
        got = g.density_grid(gsize)
        assert list(got.shape[::-1]) == want_g
        np.testing.assert_array_equal(got.ravel(), want)


@pytest.mark.gpu
def test_density_grid_rejects_particle_outside():
    from mdacp_amd.mdx import MdxError, System
    s = W.fcc_state(6, 0.5, 1.0, 3)
    with System.from_state(s) as g:
        g.make_pairlist()
        q, p = g.state()
        q[7] = [1.0, 1.0, s.L[2] + 2.0]   # beyond the last layer
        g.set_state(q, p)
        with pytest.raises(MdxError) as ei:
            g.density_grid(3.0)
        assert ei.value.code == -3
        with pytest.raises(MdxError) as ei:
            g.density_grid(0.0)
        assert ei.value.code == -1


@pytest.mark.gpu
def test_cavitation_through_c_abi_follows_oracle(oracle_result):
    from mdacp_amd.mdx import System
    fx, (th_o, ob_o, dumps_o) = oracle_result
    cfg = config(fx)

    def make(s):
        g = System.from_state(s)
        g.set_heatbath("nose_hoover", float(cfg["AimedTemperature"]), float(cfg["HeatbathTau"]))
        g.make_pairlist()
        return g

    def rescale(g, alpha, L):
        g.make_pairlist()
        g.scale_positions(alpha)          # MDUnit::ChangeScale
        g.set_box(L)
        g.set_heatbath("none")
        g.make_pairlist()
        return g

    def density(g, L, gsize):
        g.make_pairlist()
        return g.density_grid(gsize).ravel()

    print("cavitation N=2916 through the C ABI", flush=True)
    th, ob, dumps = run_cavitation(cfg, make, lambda g, k: g.calculate(k), lambda g: g.observe(),
                                   rescale, density)
    np.testing.assert_allclose(th, th_o, rtol=1e-9)
    np.testing.assert_allclose(ob, ob_o, rtol=1e-9)
    np.testing.assert_array_equal(dumps, dumps_o)
    np.testing.assert_array_equal(dumps, fx["dumps"])


@pytest.mark.gpu
def test_cavitation_program_matches_reference(tmp_path):
    fx = load_golden("cavitation_unit16")
    (tmp_path / "cav.cfg").write_text(str(fx["config"]))
    print("mdx_benchmark Mode=Cavitation N=2916", flush=True)
    r = subprocess.run([EXE, "cav.cfg", "--procs", "1", "--threads", "1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    print(out)
    rows = {tag: np.array([l.split()[:5] for l in out.splitlines() if l.endswith("# " + tag)],
                          dtype=np.float64) for tag in ("Thermalize", "Observe")}
    np.testing.assert_allclose(rows["Thermalize"], fx["thermalize"], rtol=ROW_RTOL)
    np.testing.assert_allclose(rows["Observe"], fx["observe"], rtol=ROW_RTOL)
    grid = [int(x) for x in fx["grid"]]
    assert f"# grid: {grid[0]},{grid[1]},{grid[2]}" in out
    assert f"# grid_count = {int(np.prod(grid))}" in out
    for k, want in enumerate(fx["dumps"]):
        got = np.fromfile(tmp_path / f"conf{k:04d}.dat", dtype=np.uint8)
        np.testing.assert_array_equal(got, want)
    assert not (tmp_path / f"conf{len(fx['dumps']):04d}.dat").exists()
