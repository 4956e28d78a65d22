"""The reference's Benchmark-mode initial conditions, restated in C++
(mdacp_amd/host/confmaker.cpp), against the reference's own
ExtractConfigurationMaker and SetInitialVelocity (fixtures written by
tests/golden/make_golden_benchmark.py from oracle/_ref/ref_benchmark), and
the oracle on that state against the reference run as one domain.
No GPU: the configuration maker is host code."""
import hashlib

import numpy as np

from conftest import load_golden
from mdacp_amd import workloads as W
from oracle import pyoracle as O


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_benchmark_state_bitwise_as_reference_unit8():
    fx = load_golden("benchmark_unit8")
    s = W.benchmark_state(int(fx["procs"]), int(fx["threads"]), float(fx["unit_length"]),
                          float(fx["density"]), float(fx["v0"]))
    assert s.n == int(fx["n"])
    np.testing.assert_array_equal(s.L, fx["L"])
    np.testing.assert_array_equal(s.q[:4], fx["q_head"])
    np.testing.assert_array_equal(s.p[-4:], fx["p_tail"])
    assert sha(s.q) == str(fx["q_sha256"])
    assert sha(s.p) == str(fx["p_sha256"])


def test_benchmark_state_published_decomposition():
    """8 processes x 8 threads: a 4 x 4 x 4 grid of 45-wide bricks, every
    brick filled once (N as published), each brick's momenta of zero mean,
    speeds V0 before the mean is removed."""
    fx = load_golden("benchmark_published")
    s = W.benchmark_state(8, 8, 45.0, 0.712, 0.9)
    assert s.n == int(fx["n"]) == 4152384
    np.testing.assert_array_equal(s.L, [180.0, 180.0, 180.0])
    per_unit = s.n // 64
    cells = set()
    for u in range(64):
        q = s.q[u * per_unit:(u + 1) * per_unit]
        lo = np.floor(q.min(axis=0) / 45.0).astype(int)
        assert np.all(q.max(axis=0) < (lo + 1) * 45.0)
        cells.add(tuple(lo))
        assert np.abs(s.p[u * per_unit:(u + 1) * per_unit].mean(axis=0)).max() < 1e-15
    assert len(cells) == 64


def test_oracle_single_domain_matches_reference_unit8():
    """The oracle on the 8-unit initial state as one periodic domain follows
    the reference run as one domain (KE, PE, virial sums, images refreshed)
    to 1e-9 over 1150 steps.  The state has particles exactly on the box
    faces; the reference images none of them at the first build (its edge
    test is strict), which costs the first pairlist their cross-boundary
    pairs, and the oracle follows that rule."""
    fx = load_golden("benchmark_unit8")
    ref = fx["single_domain"]
    s = W.benchmark_state(1, 8, 12.0, 0.712, 0.9)
    o = O.Oracle(s)
    done = 0
    for step in (1, 2, 25, 150, 151, 400, 1150):
        o.calculate(step - done)
        done = step
        o.refresh_images()
        ke, pe, vir = o.observe()
        np.testing.assert_allclose([ke, pe, vir], ref[step - 1, 1:], rtol=1e-9)
