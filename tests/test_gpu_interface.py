"""The C++ mirror of the reference interface (Variables / MeshList / PairList /
ForceCalculator / observers / MDUnit over libmdx) run the way the
reference's Benchmark project runs MDManager, against the reference's own
1000-step trajectory."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, golden_state, load_golden
from mdacp_amd import workloads as W

DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "shim_driver")


def test_host_interface_library_links():
    import ctypes
    from mdacp_amd import mdx
    mdx.load()
    ctypes.CDLL(os.path.join(ROOT, "mdacp_amd", "libmdacp_gpu.so"))
    assert os.access(DRIVER, os.X_OK)


@pytest.mark.gpu
def test_reference_interface_trajectory():
    fx = load_golden("traj_fcc4000")
    s = golden_state(fx)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.bin")
        W.write_state(path, s)
        print(f"shim_driver N={s.n} 1000 steps", flush=True)
        r = subprocess.run([DRIVER, path, "1000", "100", os.path.join(tmp, "o")],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        rows = np.array([[float(v) for v in ln.split()] for ln in r.stdout.splitlines()])
        q = np.fromfile(os.path.join(tmp, "o_q.f64")).reshape(-1, 3)
        p = np.fromfile(os.path.join(tmp, "o_p.f64")).reshape(-1, 3)
    ref = fx["series"]
    np.testing.assert_array_equal(rows[:, 4], ref[:, 4])     # same rebuild schedule
    np.testing.assert_allclose(rows[:, 1:4], ref[:, 1:4], rtol=1e-9)
    d = q - fx["q"]
    d -= s.L * np.round(d / s.L)
    assert np.abs(d).max() < 1e-8
    assert np.abs(p - fx["p"]).max() < 1e-7


@pytest.mark.gpu
def test_reference_interface_nose_hoover_trajectory():
    """MDUnit::Calculate of the mirror under SimulationInfo's Nose-Hoover heat
    bath (Zeta kept on the host as the reference keeps it) against the
    reference's thermostatted 1000-step run."""
    fx = load_golden("traj_nh_fcc4000")
    s = golden_state(fx)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.bin")
        W.write_state(path, s)
        print(f"shim_driver N={s.n} 1000 Nose-Hoover steps", flush=True)
        r = subprocess.run([DRIVER, path, "1000", "100", os.path.join(tmp, "o"),
                            repr(float(fx["aimed_temperature"])), repr(float(fx["tau"]))],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        rows = np.array([[float(v) for v in ln.split()] for ln in r.stdout.splitlines()])
        q = np.fromfile(os.path.join(tmp, "o_q.f64")).reshape(-1, 3)
        p = np.fromfile(os.path.join(tmp, "o_p.f64")).reshape(-1, 3)
    ref = fx["series"]
    np.testing.assert_array_equal(rows[:, 4], ref[:, 4])
    np.testing.assert_allclose(rows[:, 1:4], ref[:, 1:4], rtol=1e-9)
    np.testing.assert_allclose(rows[:, 5], fx["zeta"], rtol=0, atol=1e-9)
    d = q - fx["q"]
    d -= s.L * np.round(d / s.L)
    assert np.abs(d).max() < 1e-8
    assert np.abs(p - fx["p"]).max() < 1e-7
