"""Heat baths of the HIP path (mdx_set_heatbath, mdx_calculate under
ControlTemperature) against the reference and the oracle.

Nose-Hoover is deterministic: its 1000-step trajectory is checked against
the reference's own thermostatted run (tests/golden/traj_nh_fcc4000.npz)
with the bounds of the NVE trajectory test, and step by step against the
oracle.  Langevin noise is a pure function of (seed, id, kick): the HIP
path and the oracle draw the same normals (Philox4x32-10, Box-Muller; the
two differ only by libm rounding of log / cos / sin), so their trajectories
agree closely; the dynamics are checked by the Ornstein-Uhlenbeck
stationary variance 2T / (2 - gamma dt) on a system without forces.
"""
import numpy as np
import pytest

from conftest import golden_state, load_golden
from mdacp_amd import workloads as W
from mdacp_amd.mdx import MdxError, System
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def wrapped(d, L):
    return d - L * np.round(d / L)


def test_nose_hoover_vs_reference_fixture():
    fx = load_golden("traj_nh_fcc4000")
    s = golden_state(fx)
    with System.from_state(s) as g:
        g.set_heatbath("nose_hoover", float(fx["aimed_temperature"]), float(fx["tau"]))
        rows, zeta, nreb = [], [], 0
        for k in range(10):
            nreb += g.calculate(100)
            g.refresh_images()
            ke, pe, vir = g.observe()
            rows.append((100 * (k + 1), ke, pe, vir, nreb))
            zeta.append(g.heatbath()["zeta"])
        q, p = g.state()
    rows = np.array(rows)
    ref = fx["series"]
    np.testing.assert_array_equal(rows[:, 4], ref[:, 4])
    np.testing.assert_allclose(rows[:, 1:4], ref[:, 1:4], rtol=1e-9)
    np.testing.assert_allclose(zeta, fx["zeta"], rtol=0, atol=1e-9)
    assert np.abs(wrapped(q - fx["q"], s.L)).max() < 1e-8
    assert np.abs(p - fx["p"]).max() < 1e-7


def test_nose_hoover_vs_oracle_step_by_step():
    s = W.perturbed(W.fcc_state(8, 0.6, 1.3, 71), 0.2, 72)
    o = O.Oracle(s)
    o.set_heatbath("nose_hoover", 0.8, 0.3, zeta=0.1)
    with System.from_state(s) as g:
        g.set_heatbath("nose_hoover", 0.8, 0.3, zeta=0.1)
        for _ in range(5):
            assert g.calculate(40) == o.calculate(40)
            qg, pg = g.state()
            qo, po = o.state()
            assert np.abs(wrapped(qg - qo, s.L)).max() < 1e-10
            assert np.abs(pg - po).max() < 1e-9
            assert g.heatbath()["zeta"] == pytest.approx(o.heatbath()["zeta"], abs=1e-11)


def test_heatbath_pieces_closed_form():
    """mdx_heatbath_zeta adds (t - T0) / tau^2 dt/2; mdx_heatbath_momenta
    scales every momentum by exp(-zeta dt/2)."""
    s = W.perturbed(W.fcc_state(6, 0.5, 1.0, 73), 0.1, 74)
    with System.from_state(s) as g:
        g.set_heatbath("nose_hoover", 1.2, 0.25, zeta=-0.3)
        g.heatbath_zeta(2.0)
        z = g.heatbath()["zeta"]
        assert z == pytest.approx(-0.3 + (2.0 - 1.2) / 0.25 ** 2 * (s.dt * 0.5), rel=1e-15)
        g.heatbath_momenta()
        _, p = g.state()
    np.testing.assert_allclose(p, s.p * np.exp(-(s.dt * 0.5) * z), rtol=1e-15, atol=0)


def test_langevin_vs_oracle():
    s = W.perturbed(W.fcc_state(8, 0.6, 1.3, 75), 0.2, 76)
    o = O.Oracle(s)
    o.set_heatbath("langevin", 1.1, gamma=0.7, seed=2024)
    with System.from_state(s) as g:
        g.set_heatbath("langevin", 1.1, gamma=0.7, seed=2024)
        for _ in range(5):
            assert g.calculate(40) == o.calculate(40)
            qg, pg = g.state()
            qo, po = o.state()
            assert np.abs(wrapped(qg - qo, s.L)).max() < 1e-9
            assert np.abs(pg - po).max() < 1e-8
        # the padding lane carries noise into KE, as in the reference
        assert g.observe()[0] == pytest.approx(o.observe()[0], rel=1e-10)


def test_langevin_stationary_variance_and_counter():
    """No forces (nearest neighbour 5.2 > 3.3): p relaxes to the
    Ornstein-Uhlenbeck variance 2T / (2 - gamma dt) per lane, all four
    lanes included; the noise stream restarts with mdx_set_state."""
    s = W.fcc_state(14, 0.01, 1.0, 77)
    gamma, T = 50.0, 1.2
    with System.from_state(s) as g:
        g.set_heatbath("langevin", T, gamma=gamma, seed=99)
        g.calculate(300)
        _, p = g.state()
        ke4 = g.observe()[0]
        want = 2 * T / (2 - gamma * s.dt)
        assert abs(p.var() - want) < 4 * want * np.sqrt(2.0 / p.size)
        assert ke4 / s.n == pytest.approx(2.0 * want, rel=0.03)   # four lanes x T/2
        g.set_state(s.q, s.p)
        g.calculate(3)
        a = g.state()[1]
        g.set_state(s.q, s.p)
        g.calculate(3)
        b = g.state()[1]
    np.testing.assert_array_equal(a, b)


def test_heatbath_arguments_rejected():
    s = W.fcc_state(5, 0.5, 1.0, 78)
    with System.from_state(s) as g:
        for kw in (dict(kind="nose_hoover", tau=0.0), dict(kind="langevin", gamma=-1.0),
                   dict(kind="nose_hoover", aimed_temperature=float("nan"))):
            with pytest.raises(MdxError) as ei:
                g.set_heatbath(**kw)
            assert ei.value.code == -1
        g.set_heatbath("none")
        g.calculate(5)
