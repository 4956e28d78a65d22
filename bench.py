"""Throughput of the plain LJ step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: N = 2,048,000 particles on an 80^3-cell FCC lattice (a = 2.0,
L = 160, rho = 0.5), Maxwell momenta at T = 1 (seed 2), quadratically
shifted LJ at rc = 3.0, search length 3.3, dt = 0.001, FP64 throughout.
A step is MDManager::Calculate (mdmanager.cc:154-186): pairlist expiry
check, rebuild when expired, half drift, image refresh, force kick, half
drift.  State is resident in HBM before the timed region; the timed region
is exactly K steps between two device synchronisations, rebuilds included.
Before it, 240 calibration steps (several pairlist lifetimes) measure the
rebuild interval k and the rebuild time t_r in the same process; `value`
is N over the per-step cost (window time less the window's own rebuild
time) / K + t_r / k, so a short window that happens to hold no rebuild
still carries its amortized share (the reference's MUPS is taken over
whole runs, benchmark.cc:45-51).  `window` keeps the raw timed figure.

Prints one JSON line.  `roofline` covers the force kick (the dominant
kernel): algorithmic bytes per launch over its mean launch time from HIP
events recorded on the library's stream during the timed region.
`cpu_baseline` is the reference's own CPU path (oracle/_ref, compiled from
the reference sources) on a bounded sample of the same workload, with its
rebuilds amortized over the same interval k, on every core this process is
given (its affinity mask and OpenMP allotment); `host` states the CPU model,
sockets, physical cores and that allotment.

Two more objects ride on the same line:
  `batched`  256 independent replicas (FCC n ~ U{5..20} cells, rho ~ U[0.3,
             0.9], T ~ U[0.7, 1.5], seeds 4+r) advanced together, one graph
             launch per step, tile shape autotuned; aggregate atom-steps/s
             over 500 steps, against looping the single-system path over
             the same replicas (each timed alone, summed).
  `adjoint`  the differentiable trajectory at N=32,000 (FCC 20^3, rho 0.7,
             T 1, seed 5): 200 steps, loss over the last 100, checkpoints
             every 20; forward and reverse-sweep device time, fwd+adjoint
             steps per second and the adjoint/forward cost ratio.
  `heatbath` the N=2.05M workload stepped under the Nose-Hoover (target
             1.0, tau 0.1) and Langevin (target 1.0, gamma 0.1) heat baths,
             200 steps each after 20 warm-up steps: atom-steps/s.
  `benchmark_mode`  the reference's own Benchmark mode (benchmark.cc:12-52)
             at the configuration of its published 64-unit runs: the initial
             state of 8 MPI processes x 8 OpenMP threads (N = 4,152,384,
             L = 180, density 0.712, InitialVelocity 0.9) on this one GPU,
             150 thermalizing steps, then 1000 timed steps observed every
             100: MUPS = TotalLoop N / seconds / 1e6 beside the MUPS the
             reference publishes for that run (CPU: 4 nodes, 64 cores;
             hybrid: 8 Tesla K40t), and the first and last observations.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector FMA at 2 flops)
FLOPS_PER_HALF_PAIR = 28     # FP64 flops of one pair term, the divide counted as one
METRIC = "atom-steps/s at N=2.05M LJ (GPU vs host CPU); fwd+adjoint step/s; % HBM peak"
WORKLOAD = ("plain LJ, FCC 80^3 cells N=2048000, L=160, rho=0.5, T=1 (seed 2), "
            "rc=3.0, search 3.3, dt=0.001, fp64, drift-kick-drift with pairlist expiry")


def host_cpus() -> dict:
    """What the host offers: its CPU model, sockets, physical cores and
    hardware threads (from /proc/cpuinfo), the CPUs this process may run on
    (its affinity mask) and the OpenMP thread allotment it was given."""
    model, cores, sockets, logical = "unknown", set(), set(), 0
    family = mnum = None
    hypervisor = False
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model == "unknown":
                    model = v
                elif k == "cpu family" and family is None:
                    family = v
                elif k == "model" and mnum is None:
                    mnum = v
                elif k == "flags" and " hypervisor" in f" {v} ":
                    hypervisor = True
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    sockets.add(phys)
                    phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
            sockets.add(phys)
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except Exception:
        allowed = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if family is not None:
        model = f"{model} (family {family}, model {mnum})"
    return {"model": model, "virtual_machine": hypervisor, "sockets": len(sockets) or None,
            "physical_cores": len(cores) or None, "hardware_threads": logical or None,
            "affinity_cpus": allowed, "omp_num_threads_env": int(omp) if omp.isdigit() else None}


def cpu_threads(info: dict) -> int:
    """OpenMP domains for the reference: every core this process may use
    (its affinity mask, no more than the physical cores, and no more than an
    OMP_NUM_THREADS allotment set for it), rounded down to a power of two
    so that ParaInfo (parainfo.cc:27-94) splits the box into equal bricks,
    and at least 4 (one domain holds at most 1,000,000 particles,
    mdconfig.h:11).  At N = 2,048,000, L = 160 a brick of 16 or 64 domains
    is 40 x 40 x 80 or 40 x 40 x 40, far above twice the search length."""
    cap = info["affinity_cpus"]
    if info["physical_cores"]:
        cap = min(cap, info["physical_cores"])
    if info["omp_num_threads_env"]:
        cap = min(cap, info["omp_num_threads_env"])
    threads = 1
    while threads * 2 <= cap:
        threads *= 2
    return max(threads, 4)


def cpu_baseline(state, steps: int, warmup: int, interval: float):
    """Reference CPU path on the host cores: `steps` Calculate() calls, each
    timed alone, after a pairlist build and `warmup` steps (setup excluded).
    Rebuilds are amortized as on the GPU: the mean step without a rebuild
    plus the mean extra time of a rebuilding step over `interval` steps
    (the rebuild interval measured on the GPU for the same trajectory; the
    two paths rebuild on the same steps)."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    from mdacp_amd import workloads as W
    info = host_cpus()
    threads = cpu_threads(info)
    if os.path.exists(harness):
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "state.bin")
            W.write_state(path, state)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                       OMP_PLACES="cores")
            r = subprocess.run([harness, "bench", path, os.path.join(tmp, "o"), str(steps),
                                str(warmup)], env=env, capture_output=True, text=True, timeout=900,
                               cwd=tmp)
            if r.returncode == 0:
                line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1]
                res = json.loads(line)
                rb, sec, sec_rb = res["rebuild_steps"], res["seconds"], res["seconds_rebuild_steps"]
                plain = (sec - sec_rb) / max(steps - rb, 1)
                extra = (sec_rb / rb - plain) if rb else 0.0
                amort = plain + extra / interval
                return {"value": state.n / amort, "unit": "atom-steps/s", "cores": threads,
                        "kind": "reference",
                        "sample": f"{steps} steps of the same N=2048000 workload after 1 build + "
                                  f"{warmup} warm-up steps, each step timed alone, {rb} of them "
                                  f"rebuilding; rebuilds amortized over the GPU-measured interval "
                                  f"of {interval:.1f} steps; {threads} OpenMP domains of the "
                                  f"MDACP AVX2 build on {threads} cores of {info['model']}",
                        "window_value": state.n * steps / sec,
                        "ms_per_step_without_rebuild": plain * 1e3,
                        "ms_per_rebuild": extra * 1e3, "rebuild_steps": rb,
                        "host": info}
            sys.stderr.write(f"reference harness failed ({r.returncode}): {r.stderr[-400:]}\n")
    # restated path (single thread)
    from oracle import pyoracle as O
    o = O.Oracle(state)
    o.make_pairlist()
    o.calculate(warmup)
    t0 = time.perf_counter()
    o.calculate(steps)
    sec = time.perf_counter() - t0
    return {"value": state.n * steps / sec, "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} timed steps of the same workload, single thread, {info['model']}",
            "host": info}


def traffic_from_profile():
    path = os.path.join(ROOT, "profiles", "kick_traffic_fcc2048000.json")
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def copy_probe(nbytes: int):
    """A plain device-to-device copy of known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters into bytes."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 8, "launches": 3}), flush=True)


def batched(steps: int, plain_steps: int):
    import torch
    from mdacp_amd import workloads as W
    from mdacp_amd.mdx import Ensemble, System

    states = W.replica_set()
    ntot = sum(s.n for s in states)
    e = Ensemble(states, device=0)
    # trials long enough to hold the first rebuilds (after about 35 steps
    # from the lattice), whose branch the grid multiple trades against
    cfg = e.autotune(50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.calculate(steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    st, done = e.status()
    e.close()
    # the same replicas through the single-system path, one after another
    looped = 0.0
    for s in states:
        with System.from_state(s, device=0) as p:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            p.calculate(plain_steps)
            torch.cuda.synchronize()
            looped += time.perf_counter() - t1
    value = ntot * steps / sec
    plain_value = ntot * plain_steps / looped
    return {"replicas": len(states), "particles": ntot, "steps": steps,
            "value": value, "unit": "atom-steps/s", "ms_per_step": sec * 1e3 / steps,
            "all_replicas_completed": bool((st == 0).all() and (done == steps).all()),
            "autotuned": {"tile": cfg["tile"], "grid_mult": cfg["grid_mult"],
                          "step_grid": cfg["step_grid"], "stage_cap": cfg["stage_cap"]},
            "looped_single_system": {"value": plain_value, "unit": "atom-steps/s",
                                     "steps": plain_steps, "seconds": looped},
            "speedup_vs_looped": value / plain_value}


def adjoint():
    from mdacp_amd import workloads as W
    from mdacp_amd.mdx import System

    s = W.fcc_state(20, 0.7, 1.0, 5)
    opts = dict(nsteps=200, nobs=100, checkpoint_every=20, p_target=1.5, pe_target=-4.0)
    with System.from_state(s, device=0) as sy:
        sy.lj_trajectory(grad=True, **opts)            # warm-up
    with System.from_state(s, device=0) as sy:
        g = sy.lj_trajectory(grad=True, **opts)
    fwd, bwd = g["forward_ms"], g["backward_ms"]
    return {"particles": s.n, "steps": 200, "checkpoint_every": 20,
            "forward_ms": fwd, "backward_ms": bwd,
            "fwd_adjoint_steps_per_s": 200 / ((fwd + bwd) * 1e-3),
            "adjoint_over_forward": bwd / fwd, "loss": g["loss"],
            "dloss_deps": g["dloss_deps"], "dloss_dsigma": g["dloss_dsigma"]}


def heatbath(state, steps: int = 200, warmup: int = 20):
    import torch
    from mdacp_amd.mdx import System

    out = {}
    for kind, kw in (("nose_hoover", dict(tau=0.1)), ("langevin", dict(gamma=0.1, seed=1))):
        with System.from_state(state, device=0) as sy:
            sy.set_heatbath(kind, 1.0, **kw)
            sy.make_pairlist()
            sy.calculate(warmup)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sy.calculate(steps)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
        out[kind] = {"value": state.n * steps / sec, "unit": "atom-steps/s",
                     "ms_per_step": sec * 1e3 / steps, "steps": steps}
    return out


# MUPS the reference publishes for its 64-unit Benchmark run (N = 4,152,384):
# profile/cpu_profile/mdunit64/mdacp-stdout:36 and
# profile/gpu_profile/mdunit64/mdacp-stdout:74
PUBLISHED_MUPS_CPU = 78.0973
PUBLISHED_MUPS_HYBRID = 138.482


def benchmark_mode():
    import torch
    from mdacp_amd import benchmark as B
    from mdacp_amd import workloads as W

    state = W.benchmark_state(8, 8, 45.0, 0.712, 0.9)
    r = B.run(state, sync=torch.cuda.synchronize)
    return {"config": "Benchmark mode, 8 MPI x 8 OpenMP units of the reference on one GPU, "
                      "UnitLength 45, Density 0.712, InitialVelocity 0.9, 150 + 1000 steps",
            "particles": r["n"], "seconds": r["seconds"], "mups": r["mups"],
            "published_mups_cpu_64_cores": PUBLISHED_MUPS_CPU,
            "published_mups_hybrid_8_k40t": PUBLISHED_MUPS_HYBRID,
            "over_published_cpu": r["mups"] / PUBLISHED_MUPS_CPU,
            "over_published_hybrid": r["mups"] / PUBLISHED_MUPS_HYBRID,
            "first_observation": list(r["rows"][0]), "last_observation": list(r["rows"][-1])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--calibration-steps", type=int, default=240,
                    help="steps before the timed window that measure the rebuild interval and cost")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-heatbath", action="store_true")
    ap.add_argument("--no-benchmark-mode", action="store_true")
    ap.add_argument("--only", choices=("batched", "adjoint"),
                    help="run one form alone and print its object (for profiling)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES between two device buffers (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU sharding is a design, not code)", file=sys.stderr)
        sys.exit(2)
    if args.only == "batched":
        print(json.dumps({"batched": batched(args.batched_steps, args.batched_steps)}), flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint": adjoint()}), flush=True)
        return

    import torch
    from mdacp_amd import workloads as W
    from mdacp_amd.mdx import System

    state = W.plain_full()
    n = state.n
    sys_ = System.from_state(state, device=0)
    sys_.make_pairlist()
    sys_.calculate(args.warmup)
    # Rebuild interval and cost, measured in this process over a span of
    # several pairlist lifetimes before the timed window: a short window
    # (20 steps) may hold no rebuild at all, yet the step's cost includes one
    # rebuild per interval.
    sys_.timing_reset(True)
    cal_rebuilds = sys_.calculate(args.calibration_steps)
    cal = sys_.timing()
    sys_.timing_reset(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rebuilds = sys_.calculate(args.steps)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    tm = sys_.timing()
    kick_ms, kick_bytes = sys_.kick_profile()
    pairs = sys_.pair_number()
    seen_steps = args.calibration_steps + args.steps
    seen_rebuilds = cal_rebuilds + rebuilds
    interval = seen_steps / max(seen_rebuilds, 1)
    ms_rebuild = (cal["build_ms"] + tm["build_ms"]) / max(cal["builds"] + tm["builds"], 1)
    # the window's wall time less its own rebuilds, plus one rebuild per
    # interval: the step's cost over a whole pairlist lifetime
    window_ms = sec * 1e3 / args.steps
    step_ms = (sec * 1e3 - tm["build_ms"]) / args.steps + ms_rebuild / interval
    value = n / (step_ms * 1e-3)
    achieved = kick_bytes / (kick_ms * 1e-3) / 1e9
    traffic = traffic_from_profile()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "particles": n, "pairs_within_search": pairs,
                   "rebuilds_in_timed_region": rebuilds},
        # value = N / ms_per_step: the timed window's steps less its rebuilds,
        # plus the rebuild cost amortized over the measured interval
        "window": {"seconds": sec, "ms_per_step": window_ms, "raw_value": n * args.steps / sec,
                   "rebuilds": rebuilds},
        "rebuild": {"interval_steps": interval, "ms_per_rebuild": ms_rebuild,
                    "steps_observed": seen_steps, "rebuilds_observed": seen_rebuilds},
        "split_ms": {"force_kick_per_step": tm["force_ms"] / args.steps,
                     "pairlist_build_per_rebuild": ms_rebuild,
                     "pairlist_build_per_step_amortized": ms_rebuild / interval,
                     "other_step_kernels_per_step": tm["other_ms"] / args.steps,
                     "window_totals": {"pairlist_build": tm["build_ms"], "force_kick": tm["force_ms"],
                                       "other_step_kernels": tm["other_ms"],
                                       "builds": tm["builds"]}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_kick_tiles<3>", "avg_launch_ms": kick_ms,
                     "algorithmic_bytes_per_launch": kick_bytes},
        # the same launch against the FP64 vector peak: algorithmic flops of
        # the half-list pair terms (the full list evaluates each pair twice)
        "fp64": {"achieved": FLOPS_PER_HALF_PAIR * pairs / (kick_ms * 1e-3) / 1e12,
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": FLOPS_PER_HALF_PAIR * pairs / (kick_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                 "algorithmic_flops_per_launch": FLOPS_PER_HALF_PAIR * pairs},
    }
    sys_.close()
    if not args.no_batched:
        out["batched"] = batched(args.batched_steps, args.batched_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint()
    if not args.no_heatbath:
        out["heatbath"] = heatbath(state)
    if not args.no_benchmark_mode:
        out["benchmark_mode"] = benchmark_mode()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(state, args.cpu_steps, 2, interval)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
