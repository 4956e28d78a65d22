"""Writes tests/golden/cavitation_unit16.npz: the reference's Cavitation
project (cavitation.cc:238-287) run by the reference's own program,
oracle/_ref/mdacp_cavitation (main.cc, projectmanager.cc, cavitation.cc and
the library, compiled from the reference tree by ``make -C oracle ref``), as
one process with one thread, so one periodic unit.

Runs only in the build container, where the reference tree exists.  The
configuration (CONFIG below): the Benchmark initial conditions at UnitLength
16 (N = 2916), ThermalizeLoop steps under the Nose-Hoover heat bath, the
box and positions scaled by ChangeScale, then TotalLoop + 1 NVE steps with
BubbleHist analysing the density grid every ObserveLoop steps.

Stored: the input text (config), the BubbleHist grid lines (grid, the boxes
per axis; grid_size), the "# Thermalize" and "# Observe" rows (time,
temperature, potential energy, pressure, total energy; printed by the
reference to 6 significant digits; the seconds column dropped), and every
conf%04d.dat file BubbleHist::SaveDensity wrote (dumps, one row per
analysis, the raw bytes).

usage: python tests/golden/make_golden_cavitation.py
"""
from __future__ import annotations

import glob
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PROGRAM = os.path.join(ROOT, "oracle", "_ref", "mdacp_cavitation")

CONFIG = """Mode=Cavitation
UnitLength=16
Density=0.712
InitialVelocity=0.9
AimedTemperature=0.9
HeatbathTau=0.1
ThermalizeLoop=100
TotalLoop=300
ObserveLoop=50
ChangeScale=1.1
GridSize=3.0
OutputFile=cav.out
"""


def parse(path):
    thermalize, observe, grid, grid_size, n = [], [], None, None, None
    for line in open(path):
        f = line.split()
        if line.startswith("# grid:"):
            grid = [int(x) for x in line.split(":")[1].split(",")]
        elif line.startswith("# grid_size ="):
            grid_size = float(f[-1])
        elif line.startswith("# Number of Particles =") and n is None:
            n = int(f[-1])
        elif line.rstrip().endswith("# Thermalize"):
            thermalize.append([float(x) for x in f[:5]])
        elif line.rstrip().endswith("# Observe"):
            observe.append([float(x) for x in f[:5]])
    return n, grid, grid_size, np.array(thermalize), np.array(observe)


def main():
    if not os.path.exists(PROGRAM):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        cfg = os.path.join(tmp, "cav.cfg")
        with open(cfg, "w") as f:
            f.write(CONFIG)
        subprocess.run([PROGRAM, "-i", cfg], check=True, cwd=tmp, capture_output=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        n, grid, grid_size, th, ob = parse(os.path.join(tmp, "cav.out"))
        files = sorted(glob.glob(os.path.join(tmp, "conf*.dat")))
        dumps = np.stack([np.fromfile(p, dtype=np.uint8) for p in files])
    assert len(files) == len(ob) and dumps.shape[1] == int(np.prod(grid))
    assert (dumps.astype(np.int64).sum(axis=1) == n).all()
    np.savez(os.path.join(HERE, "cavitation_unit16.npz"), config=np.array(CONFIG), n=n,
             grid=np.array(grid), grid_size=grid_size, thermalize=th, observe=ob, dumps=dumps)
    print(f"N = {n}, grid {grid}, {len(th)} thermalize rows, {len(ob)} dumps,"
          f" counts {dumps.min()}..{dumps.max()}")


if __name__ == "__main__":
    main()
