"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Runs only in the build container, where the reference tree exists:
``make -C oracle ref`` builds oracle/_ref/ref_harness from the reference's
own sources plus oracle/ref_harness.cc, and this script feeds it seeded
states from mdacp_amd.workloads and stores its outputs as .npz fixtures:

  <case>.npz  state (q0, p0, L, dt), the reference's canonical pair set
              (full triples when small, else a SHA-256 digest plus per-particle
              partner counts), forces at dt=1, and (KE, PE, virial) sums;
  traj_<case>.npz  the reference's 1000-step trajectory: observables every
              100 steps (step, KE, PE, virial, pairlist constructions) and the
              final q, p;
  traj_nh_fcc4000.npz  the same under the reference's Nose-Hoover heat bath
              (AimedTemperature 1.5, HeatbathTau 0.1), with Variables::Zeta
              at every observation.

usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mdacp_amd import workloads as W  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
FULL_PAIRS_MAX = 40000


def cases():
    c1 = W.plain_small()
    yield "fcc4000", c1
    yield "fcc4000_pert", W.perturbed(c1, 0.3, 11)
    small07 = W.fcc_state(5, 0.7, 1.0, 5, name="fcc500_rho0.7")
    yield "fcc500_rho0.7_pert", W.perturbed(small07, 0.25, 12)
    small09 = W.fcc_state(5, 0.9, 1.2, 6, name="fcc500_rho0.9")
    yield "fcc500_rho0.9_pert", W.perturbed(small09, 0.2, 13)
    small03 = W.fcc_state(6, 0.3, 0.8, 7, name="fcc864_rho0.3")
    yield "fcc864_rho0.3_pert", W.perturbed(small03, 0.4, 14)


def pair_digest(triples: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(triples, dtype="<i4").tobytes()).hexdigest()


def run(mode, state, tmp, *args, heatbath=None):
    path = os.path.join(tmp, "state.bin")
    W.write_state(path, state)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    env.pop("MDX_REF_HEATBATH", None)
    if heatbath:
        env["MDX_REF_HEATBATH"] = heatbath
    out = os.path.join(tmp, "out")
    r = subprocess.run([HARNESS, mode, path, out, *map(str, args)], env=env,
                       capture_output=True, text=True, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(f"harness {mode} failed: {r.stdout}\n{r.stderr}")
    return out


def main():
    if not os.path.exists(HARNESS):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, s in cases():
            out = run("pairs", s, tmp)
            triples = np.fromfile(out + "_pairs.i32", dtype=np.int32).reshape(-1, 3)
            forces = np.fromfile(out + "_forces.f64").reshape(-1, 3)
            obs = np.fromfile(out + "_obs.f64")
            counts = np.bincount(triples[:, 0], minlength=s.n) + np.bincount(triples[:, 1], minlength=s.n)
            fx = dict(q0=s.q, p0=s.p, L=s.L, dt=np.float64(s.dt), forces=forces, obs=obs,
                      pair_count=np.int64(len(triples)), pair_digest=np.bytes_(pair_digest(triples)),
                      partner_counts=counts.astype(np.int32))
            if len(triples) <= FULL_PAIRS_MAX:
                fx["pairs"] = triples
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **fx)
            print(name, s.n, len(triples), obs)
        s = W.plain_small()
        out = run("run", s, tmp, 1000, 100)
        series = np.fromfile(out + "_series.f64").reshape(-1, 5)
        np.savez_compressed(os.path.join(HERE, "traj_fcc4000.npz"), q0=s.q, p0=s.p, L=s.L,
                            dt=np.float64(s.dt), series=series,
                            q=np.fromfile(out + "_q.f64").reshape(-1, 3),
                            p=np.fromfile(out + "_p.f64").reshape(-1, 3))
        print("traj_fcc4000", series[-1])
        T_aim, tau = 1.5, 0.1
        out = run("run", s, tmp, 1000, 100, heatbath=f"NoseHoover:{T_aim!r}:{tau!r}")
        series = np.fromfile(out + "_series.f64").reshape(-1, 5)
        np.savez_compressed(os.path.join(HERE, "traj_nh_fcc4000.npz"), q0=s.q, p0=s.p, L=s.L,
                            dt=np.float64(s.dt), aimed_temperature=np.float64(T_aim),
                            tau=np.float64(tau), series=series,
                            zeta=np.fromfile(out + "_zeta.f64"),
                            q=np.fromfile(out + "_q.f64").reshape(-1, 3),
                            p=np.fromfile(out + "_p.f64").reshape(-1, 3))
        print("traj_nh_fcc4000", series[-1])


if __name__ == "__main__":
    main()
