"""Writes the Benchmark-mode fixtures of tests/golden/ from the compiled
reference.

Runs only in the build container, where the reference tree exists:
``make -C oracle ref`` builds oracle/_ref/ref_benchmark from the reference's
own sources (ExtractConfigurationMaker, MDManager::SetInitialVelocity and
the Benchmark::Run loop, reached through their public interface) plus
oracle/ref_benchmark.cc.

  benchmark_unit8.npz      8 units (1 process x 8 threads), UnitLength 12,
                           Density 0.712, InitialVelocity 0.9: N, L, SHA-256
                           digests of the reference's initial q and p
                           (float64 [n][3], units in id order), the first and
                           last particles; the Benchmark::Run observations
                           (time, temperature, pressure, energy per particle)
                           of the 8-unit run at full precision
                           (decomposed_series); and the same initial state
                           run by the reference as one periodic domain
                           (oracle/_ref/ref_harness run, one thread): step,
                           KE, PE and virial sums after every step, images
                           refreshed (single_domain).
  benchmark_published.npz  the published 64-unit run (8 processes x 8 threads,
                           UnitLength 45, Density 0.712, InitialVelocity 0.9,
                           N = 4,152,384): the rows the reference publishes
                           in profile/cpu_profile/mdunit64/mdacp-stdout
                           (printed to 6 significant digits) and its MUPS
                           (CPU and hybrid runs).

usage: python tests/golden/make_golden_benchmark.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import tempfile

import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_benchmark")
REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
sys.path.insert(0, ROOT)
from mdacp_amd import workloads as W  # noqa: E402

MPIRUN = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")
PUBLISHED_STDOUT = os.path.join(REF, "profile", "cpu_profile", "mdunit64", "mdacp-stdout")
PUBLISHED_STDOUT_GPU = os.path.join(REF, "profile", "gpu_profile", "mdunit64", "mdacp-stdout")


def cfg_text(unit_length, density=0.712, v0=0.9, thermalize=150, loop=1000, observe=100):
    return (f"Mode=Benchmark\nDensity={density}\nUnitLength={unit_length}\n"
            f"InitialVelocity={v0}\nThermalizeLoop={thermalize}\nTotalLoop={loop}\n"
            f"ObserveLoop={observe}\n")


def read_conf(path):
    raw = open(path, "rb").read()
    off, ids, qs, ps = 0, [], [], []
    while off < len(raw):
        uid, m = np.frombuffer(raw, dtype="<i4", count=2, offset=off)
        off += 8
        a = np.frombuffer(raw, dtype="<f8", count=6 * int(m), offset=off).reshape(int(m), 6)
        off += 48 * int(m)
        ids.append(int(uid))
        qs.append(a[:, :3])
        ps.append(a[:, 3:])
    assert ids == sorted(ids)
    return np.concatenate(qs), np.concatenate(ps)


def read_series(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    n = [int(l.split("=")[1]) for l in open(path) if l.startswith("# N =")][0]
    return n, np.array(rows, dtype=np.float64)


def mups(path):
    line = [l for l in open(path) if l.startswith("# N =") and "[MUPS]" in l][0].split()
    return float(line[line.index("[MUPS]") - 1])


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def unit8(tmp):
    cfg = os.path.join(tmp, "unit8.cfg")
    with open(cfg, "w") as f:
        f.write(cfg_text(12))
    env = dict(os.environ, OMP_NUM_THREADS="8")
    conf = os.path.join(tmp, "unit8.bin")
    subprocess.run([HARNESS, "conf", cfg, conf], check=True, env=env, capture_output=True, cwd=tmp)
    q, p = read_conf(conf)
    series = os.path.join(tmp, "unit8.txt")
    subprocess.run([HARNESS, "run", cfg, series], check=True, env=env, capture_output=True, cwd=tmp)
    n, rows = read_series(series)
    assert n == len(q)
    # the same state as one periodic domain
    st = W.benchmark_state(1, 8, 12.0, 0.712, 0.9)
    assert digest(st.q) == digest(q) and digest(st.p) == digest(p)
    W.write_state(os.path.join(tmp, "unit8_state.bin"), st)
    subprocess.run([REF_HARNESS, "run", os.path.join(tmp, "unit8_state.bin"),
                    os.path.join(tmp, "single"), "1150", "1"], check=True,
                   env=dict(os.environ, OMP_NUM_THREADS="1"), capture_output=True, cwd=tmp)
    single = np.fromfile(os.path.join(tmp, "single_series.f64")).reshape(-1, 5)[:, :4]
    np.savez(os.path.join(HERE, "benchmark_unit8.npz"), procs=1, threads=8, unit_length=12.0,
             density=0.712, v0=0.9, n=n, L=np.array([24.0, 24.0, 24.0]),
             q_sha256=np.array(digest(q)), p_sha256=np.array(digest(p)),
             q_head=q[:4], p_head=p[:4], q_tail=q[-4:], p_tail=p[-4:],
             decomposed_series=rows, single_domain=single)
    print("unit8:", n, "particles,", len(rows), "observations")


def published():
    pub = [l.split()[:4] for l in open(PUBLISHED_STDOUT) if l.rstrip().endswith("#observe")]
    pub_rows = np.array(pub, dtype=np.float64)
    n_pub = [int(l.split()[3]) for l in open(PUBLISHED_STDOUT) if l.startswith("# N =")][0]
    assert n_pub == 4152384
    np.savez(os.path.join(HERE, "benchmark_published.npz"), procs=8, threads=8, unit_length=45.0,
             density=0.712, v0=0.9, n=n_pub, published_rows=pub_rows,
             published_mups_cpu=mups(PUBLISHED_STDOUT),
             published_mups_hybrid=mups(PUBLISHED_STDOUT_GPU))
    print("published:", n_pub, "particles,", len(pub_rows), "rows")


def main():
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        unit8(tmp)
    published()


if __name__ == "__main__":
    main()
