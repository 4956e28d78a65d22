"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the
CPU (tests/cpp/sanitize.mk builds tests/cpp/_build/san/).

Covered: the C ABI's argument checks and every path that returns before a
launch (a NULL handle is MDX_ERR_ARG everywhere; a box below two search
lengths, bad parameters, empty replicas; without a device, a valid create
fails with MDX_ERR_HIP), the Benchmark-mode initial conditions and their C
entry, the oracle over steps, rebuilds and both heat baths, the mirror's
StopWatch, and mdx_benchmark's input parsing up to the device.  A
sanitizer report, or a result that differs from the regular build's,
fails the test: the initial state's SHA-256 equals the reference's
(fixture benchmark_unit8) and the oracle's results equal liboracle's.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from mdacp_amd import workloads as W
from oracle import pyoracle as O

MK = os.path.join(ROOT, "tests", "cpp", "sanitize.mk")
OUT = os.path.join(ROOT, "tests", "cpp", "_build", "san")
DRIVER = os.path.join(OUT, "sanitize_driver")
BENCH = os.path.join(OUT, "mdx_benchmark")
NO_DEVICE = not os.path.exists("/dev/kfd")
ENV = dict(os.environ,
           ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
if NO_DEVICE:
    ENV["SANITIZE_NO_DEVICE"] = "1"


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-f", MK, "-j8"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]


def run(argv, cwd, expect_rc=0):
    r = subprocess.run(argv, capture_output=True, text=True, timeout=300, env=ENV, cwd=cwd)
    report = r.stderr
    assert "AddressSanitizer" not in report and "LeakSanitizer" not in report, report[-4000:]
    assert "runtime error" not in report, report[-4000:]
    assert r.returncode == expect_rc, (r.returncode, r.stdout, report[-2000:])
    out = {}
    for ln in r.stdout.splitlines():
        k, *v = ln.split() or [""]
        try:
            out[k] = [float(x) for x in v] if len(v) > 1 else (float(v[0]) if v else None)
        except ValueError:
            pass        # the driver's other lines
    return out, r


def test_capi_argument_checks(built, tmp_path):
    out, _ = run([DRIVER, "capi"], tmp_path)
    assert out["failures"] == 0 and out["checks"] >= 50


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_benchmark_state_under_sanitizers(built, tmp_path):
    fx = load_golden("benchmark_unit8")
    out, _ = run([DRIVER, "conf", "1", "8", "12", "0.712", "0.9", str(tmp_path / "u8")], tmp_path)
    assert out["n"] == int(fx["n"])
    assert digest(np.fromfile(tmp_path / "u8_q.f64")) == str(fx["q_sha256"])
    assert digest(np.fromfile(tmp_path / "u8_p.f64")) == str(fx["p_sha256"])
    # another unit grid against the regular build
    out, _ = run([DRIVER, "conf", "2", "4", "10", "0.6", "1.1", str(tmp_path / "g8")], tmp_path)
    s = W.benchmark_state(2, 4, 10.0, 0.6, 1.1)
    assert out["n"] == s.n
    assert digest(np.fromfile(tmp_path / "g8_q.f64")) == digest(s.q)
    assert digest(np.fromfile(tmp_path / "g8_p.f64")) == digest(s.p)


@pytest.mark.parametrize("ncell,rho,seed", [(5, 0.7, 5), (6, 0.3, 9)])
def test_oracle_under_sanitizers(built, tmp_path, ncell, rho, seed):
    s = W.perturbed(W.fcc_state(ncell, rho, 1.0, seed), 0.25, seed + 7)
    path = tmp_path / "state.bin"
    W.write_state(str(path), s)
    out, _ = run([DRIVER, "oracle", str(path), "300"], tmp_path)
    o = O.Oracle(s)
    o.make_pairlist()
    assert out["pairs"] == o.pair_number() and out["canonical"] == len(o.pairs())
    np.testing.assert_array_equal(out["f0"], o.forces()[0])
    assert out["rebuilds"] == o.calculate(300)
    o.refresh_images()
    np.testing.assert_array_equal([out["ke"], out["pe"], out["virial"]], o.observe())
    o.set_heatbath("nose_hoover", 1.2, 0.1)
    o.calculate(20)
    assert out["zeta"] == o.heatbath()["zeta"]
    o.set_heatbath("langevin", 1.2, 0.1, 0.1, o.heatbath()["zeta"], 7)
    o.calculate(20)
    assert out["qsum"] == sum(o.state()[0].ravel().tolist())    # summed in id order


def test_stopwatch_under_sanitizers(built, tmp_path):
    out, _ = run([DRIVER, "stopwatch", str(tmp_path)], tmp_path)
    laps = np.loadtxt(tmp_path / "sw00000.dat")
    assert out["failures"] == 0 and laps.shape == (3,) and (laps >= 0).all()


def test_benchmark_driver_parsing_under_sanitizers(built, tmp_path):
    run([BENCH], tmp_path, expect_rc=2)                                  # usage
    run([BENCH, str(tmp_path / "missing.cfg")], tmp_path, expect_rc=2)  # cannot open
    cfg = tmp_path / "a.cfg"
    cfg.write_text("Mode=Collision\nUnitLength=10\n")
    _, r = run([BENCH, str(cfg)], tmp_path, expect_rc=2)
    assert "Mode=Benchmark and Mode=Cavitation" in r.stderr
    cfg.write_text("# comment only\nMode = Benchmark   # trailing comment\nDensity=0.7\n")
    _, r = run([BENCH, str(cfg)], tmp_path, expect_rc=2)
    assert "need UnitLength" in r.stderr
    cfg.write_text("Mode=Benchmark\nSystemSize=20\n")
    _, r = run([BENCH, str(cfg), "--procs", "2", "--threads", "1"], tmp_path, expect_rc=2)
    assert "cubic unit grid" in r.stderr
    if NO_DEVICE:
        # parsed, initial conditions made, then no device
        for mode in ("Benchmark", "Cavitation"):
            cfg.write_text(f"Mode={mode}\nUnitLength=10\nDensity=0.7\nTotalLoop=10\n")
            _, r = run([BENCH, str(cfg)], tmp_path, expect_rc=1)
            assert "not available" in r.stderr
