#!/usr/bin/env python3
"""GCMC trial moves/s on one MI355X (batched mixed FP32/FP64 vs plain FP64).

Workload (batched isotherm): 64 independent state points in one persistent
launch -- T in {273, 293, 323} K (22/21/21), P log-spaced 1e-3..10 bar --
each on its own seed-jittered copy of the 424-atom cubic framework, CO2 guest,
rcut 12.5 A, Ewald precision 1e-6 (871 k-vectors), mixed precision,
hipGraph-captured sweeps.  One bench step = one sweep = `--moves` trial moves
per state point.  Synthetic inputs, deterministic generator.

The state points are first filled to their equilibrium loading (sweeps in
chunks of 25 until the mean guest count moves by less than 0.3 %), then
`--warmup` untimed and `--steps` timed sweeps.

Printed as one JSON line (rank 0, N=1):
  value          -- attempted trial moves per second, whole job (all walkers)
  roofline       -- dominant kernel (k_run) against HBM: algorithmic bytes
                    B = 17 (N_fw + N_guest_sites) + 32 n_k per trial
  cpu_baseline   -- the reference's serial build on one state point of the
                    same workload, this box's host, one core (and, apart,
                    16 concurrent processes: the MPI build's model)
  fp64_single    -- plain FP64, one state point, same device kernel
  fp64_batched   -- plain FP64 on the same 64 state points, started from the
                    equilibrated mixed walkers' molecules
  widom_grid     -- Widom insertions on a grid (config-0 framework, CO2):
                    attempts/s on the device and of the reference's serial
                    build on this box's host
  config1_supercell, config3_gradient, config4_wang_landau
                 -- the 3392-atom supercell (one walker, FP64 and mixed),
                    dU/dtheta of 256 configurations, and the 128-walker
                    Wang-Landau launch (skipped with --no-configs)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
METRIC = "GCMC trial moves/s, 1 MI355X: batched FP32/FP64 vs plain FP64, ΔU≤1e-6 rel"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed sweeps")
    ap.add_argument("--warmup", type=int, default=5, help="untimed sweeps after the equilibration")
    ap.add_argument("--fill", type=int, default=400,
                    help="most equilibration sweeps before the warm-up (stops once the mean loading settles)")
    ap.add_argument("--no-configs", action="store_true", help="skip the config 1, 3 and 4 measurements")
    ap.add_argument("--moves", type=int, default=1000, help="trial moves per state point per sweep")
    ap.add_argument("--walkers", type=int, default=64)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="threads per workgroup (0: autotune)")
    ap.add_argument("--team", type=int, default=0, help="workgroups per walker (0: autotune)")
    ap.add_argument("--no-widom", action="store_true")
    ap.add_argument("--eager", action="store_true",
                    help="plain launches instead of graph replays (for counter collection)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (calibrates the HBM byte counters) and exit")
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed",
                    help="precision of the timed batch (the metric is quoted in mixed precision)")
    ap.add_argument("--phases", action="store_true",
                    help="with libfastmc_hip_phases.so (python -m fastmc_amd._build --phases): print the per-phase "
                         "step clock of walker 0 over the timed sweeps instead of the benchmark line")
    ap.add_argument("--only", choices=("config1", "config3", "config4"),
                    help="run only this measurement and print its JSON (for profiling it alone)")
    return ap.parse_args()


def copy_calibration():
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE per
    dispatch against 256 MiB is the scale applied to the kernel's counters."""
    import torch
    n = 64 * 1024 * 1024
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 4 * n, "dispatches": 5}), flush=True)


def device_sync():
    try:
        hip = C.CDLL("libamdhip64.so")
        hip.hipDeviceSynchronize()
    except OSError:
        pass


def make_batch(nwalk: int, root: str):
    from fastmc_amd import inputs as I
    from fastmc_amd.system import read_system
    dirs = I.config2(root, n=nwalk)
    systems = [read_system(d) for d in dirs]
    return dirs, systems


def algorithmic_bytes(nfw: int, n_guest_sites: float, nk: int) -> float:
    return 17.0 * (nfw + n_guest_sites) + 32.0 * nk


def reference_cpu(state_dir: str, seconds: float):
    """Time the reference serial build (oracle/_ref/gcmc_ser.x) on one state point."""
    exe = os.path.join(ROOT, "oracle", "_ref", "gcmc_ser.x")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="fmc_cpu_")
    try:
        for f in ("FIELD", "CONFIG", "CONTROL"):
            shutil.copy(os.path.join(state_dir, f), os.path.join(work, f))
        steps = 4000
        res = None
        for _ in range(4):  # grow the sample until it takes about `seconds`
            ctl = open(os.path.join(work, "CONTROL")).read()
            ctl = re.sub(r"(?m)^steps .*$", f"steps    {steps}", ctl)
            ctl = re.sub(r"(?m)^equilibration .*$", "equilibration    0", ctl)
            open(os.path.join(work, "CONTROL"), "w").write(ctl)
            for p in os.listdir(work):
                if p.startswith("branch"):
                    shutil.rmtree(os.path.join(work, p), ignore_errors=True)
            t0 = time.time()
            r = subprocess.run([exe], cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               timeout=max(60, 6 * seconds))
            wall = time.time() - t0
            out = open(os.path.join(work, "OUTPUT")).read()
            m = re.search(r"Time elapsed for\s+(\d+)\s+gcmc steps\s*:\s*([0-9.Ee+-]+)", out)
            if r.returncode != 0 or not m:
                return None
            n, t = int(m.group(1)), float(m.group(2))
            res = (n, t, wall)
            if t >= 0.6 * seconds:
                break
            steps = int(steps * min(8.0, max(1.5, seconds / max(t, 1e-3))))
        n, t, wall = res
        return {"value": n / t, "unit": "trial moves/s", "cores": 1, "kind": "reference",
                "sample": f"reference serial build (amdflang -O3) on one state point of the batch "
                          f"(T=293 K, P=0.1 bar), {n} GCMC steps from an empty "
                          f"framework, {t:.2f} s loop time"}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def reference_cpu_concurrent(state_dir: str, steps: int, nproc: int):
    """The reference's parallel model (ranks are independent chains,
    gcmc.f:388): nproc serial processes at once on the host's cores;
    aggregate moves/s = sum of steps / slowest wall time."""
    exe = os.path.join(ROOT, "oracle", "_ref", "gcmc_ser.x")
    if not os.path.exists(exe) or nproc < 1:
        return None
    works, procs = [], []
    try:
        for i in range(nproc):
            work = tempfile.mkdtemp(prefix=f"fmc_cpu{i}_")
            works.append(work)
            for f in ("FIELD", "CONFIG", "CONTROL"):
                shutil.copy(os.path.join(state_dir, f), os.path.join(work, f))
            ctl = open(os.path.join(work, "CONTROL")).read()
            ctl = re.sub(r"(?m)^steps .*$", f"steps    {steps}", ctl)
            ctl = re.sub(r"(?m)^equilibration .*$", "equilibration    0", ctl)
            open(os.path.join(work, "CONTROL"), "w").write(ctl)
        t0 = time.time()
        for work in works:
            procs.append(subprocess.Popen([exe], cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        for pr in procs:
            pr.wait(timeout=600)
        wall = time.time() - t0
        total = 0
        for work in works:
            out = open(os.path.join(work, "OUTPUT")).read()
            m = re.search(r"Time elapsed for\s+(\d+)\s+gcmc steps", out)
            if not m:
                return None
            total += int(m.group(1))
        return {"value": total / wall, "unit": "trial moves/s", "cores": nproc, "kind": "reference",
                "sample": f"{nproc} concurrent reference serial processes (the MPI build's independent-chain "
                          f"model), {steps} GCMC steps each on the same state point, {wall:.1f} s wall"}
    finally:
        for pr in procs:
            if pr.poll() is None:
                pr.kill()
        for work in works:
            shutil.rmtree(work, ignore_errors=True)


def widom_device(sysm, griddim: float = 0.5, nstep: int = 20):
    """Widom grid (fmc_widom) on one walker of the config-0 system, FP64 and mixed."""
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    out = {}
    for name, prec in (("fp64", FMC_FP64), ("mixed", FMC_MIXED)):
        h = Handle(sysm, precision=prec, seeds=[(12, 34, 56, 78)])
        try:
            h.widom(1, 2, 2.0)  # warm-up (kernel load, buffers)
            t0 = time.perf_counter()
            r = h.widom(1, nstep, griddim)
            dt = time.perf_counter() - t0
        finally:
            h.close()
        out[name] = {"attempts_per_s": r["attempts"] / dt, "seconds": dt, "henry": r["henry"]}
        out["attempts"] = r["attempts"]
        out["grid"] = list(r["ngrid"])
    return out


def widom_reference_cpu(state_dir: str, griddim: float = 2.0, nstep: int = 5):
    """The reference's serial build in Widom mode ("henry") on the same system, this host, one core."""
    exe = os.path.join(ROOT, "oracle", "_ref", "gcmc_ser.x")
    if not os.path.exists(exe):
        return None
    work = tempfile.mkdtemp(prefix="fmc_widcpu_")
    try:
        for f in ("FIELD", "CONFIG", "CONTROL"):
            shutil.copy(os.path.join(state_dir, f), os.path.join(work, f))
        ctl = open(os.path.join(work, "CONTROL")).read()
        ctl = ctl.replace("finish", f"henry\nwidom insertions {nstep}\ngrid spacing {griddim}\nfinish")
        open(os.path.join(work, "CONTROL"), "w").write(ctl)
        r = subprocess.run([exe], cwd=work, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
        out = open(os.path.join(work, "OUTPUT")).read()
        m = re.search(r"Time elapsed for\s+(\d+)\s+gcmc steps\s*:\s*([0-9.Ee+-]+)", out)
        if r.returncode != 0 or not m:
            return None
        n, t = int(m.group(1)), float(m.group(2))
        return {"attempts_per_s": n / t, "cores": 1, "kind": "reference",
                "sample": f"reference serial build in Widom mode, grid spacing {griddim} A x {nstep} "
                          f"orientations ({n} attempts, {t:.2f} s)"}
    finally:
        shutil.rmtree(work, ignore_errors=True)


def oracle_cpu(sysm, seconds: float):
    from oracle.oracle import Oracle
    o = Oracle(sysm, seed=(31, 41, 59, 26))
    n, t = 0, 0.0
    step = 500
    while t < seconds and n < 2_000_000:
        t0 = time.time()
        o.run(step)
        t += time.time() - t0
        n += step
    return {"value": n / t, "unit": "trial moves/s", "cores": 1, "kind": "port",
            "sample": f"C oracle, one state point, {n} steps"}


def note(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_state(src, dst, walkers):
    """Give dst's walkers the guest molecules of src's (same chemistry):
    fmc_load_guests rebuilds counts, S(k) and energies from the coordinates."""
    for w in range(walkers):
        for g in range(1, len(src.sysm.guests) + 1):
            dst.load_guests(g, src.guests(g, w=w), w=w)


def mean_guests(h, walkers):
    return sum(h.stats(w).nmols[0] for w in range(walkers)) / walkers


def equilibrate(h, walkers, moves, max_sweeps, chunk=25, tol=3e-3):
    """Equilibration of the state points before anything is timed: sweeps in
    chunks until the mean loading moves by less than tol between chunks."""
    done, prev = 0, None
    while done < max_sweeps:
        h.run_graph(chunk, moves)
        done += chunk
        n = mean_guests(h, walkers)
        if prev is not None and abs(n - prev) <= tol * max(prev, 1.0):
            break
        prev = n
    return done, mean_guests(h, walkers)


def timed_sweeps(h, sweeps, moves, warmup):
    """warmup untimed sweeps, then `sweeps` timed graph-replayed sweeps (the
    graph captured and replayed once untimed first)."""
    if warmup > 0:
        h.run_graph(warmup, moves)
    h.run_graph(sweeps, moves)
    h.reset_stats()
    device_sync()
    t0 = time.perf_counter()
    h.run_graph(sweeps, moves, count_stats=True)
    device_sync()
    return time.perf_counter() - t0


def config1_supercell(steps: int = 10000, fill: int = 60000):
    """Config 1: the 2x2x2 supercell (3392 atoms, 6650 k-vectors), CO2 at
    10 bar, one walker, FP64 and mixed, timed at the same loading."""
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    from fastmc_amd.system import read_system
    d = tempfile.mkdtemp(prefix="fmc_cfg1_")
    try:
        I.config1(d, pressure_bar=10.0)
        s = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    out = {"workload": "config 1: 2x2x2 supercell (3392 framework atoms, a = 51.6 A, 6650 k-vectors), CO2 at "
                       f"10 bar, one walker, timed for {steps} steps after a {fill}-step fill from the empty "
                       "framework (the mixed walker takes the FP64 walker's molecules)"}
    hf = Handle(s, precision=FMC_FP64, seeds=[(5, 6, 7, 8)])
    hm = Handle(s, precision=FMC_MIXED, seeds=[(5, 6, 7, 8)])
    try:
        for _ in range(fill // 20000):
            hf.run(20000)
        copy_state(hf, hm, 1)
        for name, h in (("fp64", hf), ("mixed", hm)):
            h.set_block(0)  # threads and workgroups per walker, tuned on a snapshot
            h.run(2000)
            device_sync()
            t0 = time.perf_counter()
            h.run(steps)
            device_sync()
            dt = time.perf_counter() - t0
            out[name] = {"trial_moves_per_s": steps / dt, "us_per_step": dt / steps * 1e6,
                         "guests": int(h.stats(0).nmols[0]), "resums": int(h.stats(0).resums),
                         "threads_per_walker": h.get_block(), "workgroups_per_walker": h.get_team()}
        out["mixed_vs_fp64"] = out["mixed"]["trial_moves_per_s"] / out["fp64"]["trial_moves_per_s"]
    finally:
        hf.close()
        hm.close()
    return out


def config3_gradient(nconf: int = 256, reps: int = 9):
    """Config 3: dU/dtheta (424 framework charges, LJ eps / sigma of the
    types) of 256 configurations sampled from config 0 at 1 bar."""
    from fastmc_amd import adjoint as A
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    from fastmc_amd.system import read_system
    d = tempfile.mkdtemp(prefix="fmc_cfg3_")
    try:
        I.config0(d, pressure_bar=1.0)
        s = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nmols, xyz = A.sample_configurations(s, nconf, every=200, warmup=20000)
    q, eps, sig = A.theta0(s)
    out = {"workload": f"config 3: U and dU/dtheta (424 framework charges, LJ eps and sigma of the 4 framework "
                       f"types) of {nconf} configurations sampled every 200 steps from config 0 at 1 bar "
                       f"({int(nmols[:, 0].min())}-{int(nmols[:, 0].max())} CO2 each), one fmc_energy_grad call"}
    h = Handle(s, precision=FMC_FP64)
    try:
        # each call timed on its own (it ends with the results on the host);
        # the median of the calls, the two precisions alternating
        times = {"mixed": [], "fp64": []}
        for prec in (FMC_MIXED, FMC_FP64):
            h.energy_grad(q, eps, sig, nmols, xyz, prec)
        device_sync()
        for _ in range(reps):
            for prec, name in ((FMC_MIXED, "mixed"), (FMC_FP64, "fp64")):
                t0 = time.perf_counter()
                h.energy_grad(q, eps, sig, nmols, xyz, prec)
                times[name].append(time.perf_counter() - t0)
        for name, v in times.items():
            dt = sorted(v)[len(v) // 2]
            out[name] = {"configurations_per_s": nconf / dt, "ms_per_call": dt * 1e3}
    finally:
        h.close()
    return out


def config4_wang_landau(walkers: int = 128, steps: int = 2000):
    """Config 4: Wang-Landau in N(CO2) 0-400 for CO2/N2 on the supercell,
    128 walkers over 8 windows in one launch, mixed precision."""
    from fastmc_amd import inputs as I
    from fastmc_amd.lib import FMC_MIXED, Handle
    from fastmc_amd.system import read_system
    d = tempfile.mkdtemp(prefix="fmc_cfg4_")
    try:
        I.config4(d)
        s = read_system(d)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    seeds = [(1 + (7 * i) % 177, 1 + (13 * i) % 177, 1 + (29 * i) % 177, (37 * i) % 168) for i in range(walkers)]
    h = Handle(s, precision=FMC_MIXED, seeds=seeds, n=walkers)
    try:
        h.wl_setup(maxn=400, nwindows=8, visittol=10000, flatcoeff=0.8, prectol=1e-6)
        h.set_block(0)
        h.run(2 * steps)  # preload to each window's lower edge, first flat-histogram steps
        device_sync()
        t0 = time.perf_counter()
        h.run(steps)
        device_sync()
        dt = time.perf_counter() - t0
        n = [h.stats(w).nmols[0] for w in range(walkers)]
        shape = {"threads_per_walker": h.get_block(), "workgroups_per_walker": h.get_team()}
    finally:
        h.close()
    return {"workload": "config 4: Wang-Landau in N(CO2) 0-400 for CO2/N2 (N2 grand-canonical at 0.8 bar) on "
                        "the 2x2x2 supercell, T = 293 K, 128 walkers over 8 N-windows in one launch, ln f from 1 at "
                        f"80 % flatness, mixed precision; timed {steps} steps after {2 * steps} (preload)",
            "walker_steps_per_s": walkers * steps / dt, "us_per_step": dt / steps * 1e6,
            "co2_min_max": [int(min(n)), int(max(n))], **shape}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.only:
        fn = {"config1": config1_supercell, "config3": config3_gradient, "config4": config4_wang_landau}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported here "
              "(state points shard across GPUs as whole walkers, with no per-trial communication)",
              file=sys.stderr)
        sys.exit(2)
    from fastmc_amd import lib as fmc_lib
    if args.phases:
        from fastmc_amd._build import LIB_PHASES
        fmc_lib.load_library(LIB_PHASES)
    from fastmc_amd.lib import FMC_FP64, FMC_MIXED, Handle
    from fastmc_amd.system import make_walker

    tmp = tempfile.mkdtemp(prefix="fmc_bench_")
    try:
        dirs, systems = make_batch(args.walkers, tmp)
        base = systems[0]
        keep = []
        walkers = []
        for i, s in enumerate(systems):
            walkers.append(make_walker(base, seed=(1 + (7 * i) % 170, 1 + (13 * i) % 170, 1 + (29 * i) % 170,
                                                   (37 * i) % 168),
                                       temperature=s.temperature, pressures=s.pressures,
                                       framework_xyz=s.framework_xyz(), keep=keep))
        h = Handle(base, walkers=walkers, precision=FMC_MIXED if args.precision == "mixed" else FMC_FP64)
        # the state points are filled to their equilibrium loading first
        note(f"{args.walkers} walkers; equilibrating")
        fill_sweeps, fill_n = equilibrate(h, args.walkers, args.moves, args.fill) if args.fill > 0 else (0, 0.0)
        # then the launch shape: threads per workgroup and workgroups per
        # walker, tuned on a snapshot of the equilibrated walkers (0 = tune)
        if args.team > 0:
            h.set_team(args.team)
        h.set_block(args.block)
        block, team = h.get_block(), h.get_team()
        note(f"equilibrated after {fill_sweeps} sweeps, mean N {fill_n:.1f}; {team} x {block} threads per walker; "
             f"timing {args.steps} sweeps")
        if args.phases:
            for _ in range(args.warmup):
                h.run(args.moves)
            fmc_lib.phase_clock(clear=True)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                h.run(args.moves)
            device_sync()
            wall = time.perf_counter() - t0
            pc = fmc_lib.phase_clock()
            pc["wall_us_per_step"] = wall / (args.steps * args.moves) * 1e6
            pc["config"] = {"walkers": args.walkers, "threads_per_walker": block, "workgroups_per_walker": team,
                            "mean_guests": round(mean_guests(h, args.walkers), 2), "launches": args.steps,
                            "steps_per_launch": args.moves}
            print(json.dumps({"phases": pc}), flush=True)
            h.close()
            return
        if args.eager:
            for _ in range(args.warmup):
                h.run(args.moves)
            h.reset_stats()
            device_sync()
            t0 = time.perf_counter()
            ms_tot = 0.0
            for _ in range(args.steps):
                h.run(args.moves, count_stats=True)
                ms_tot += h.last_timing()[0]
            device_sync()
            wall = time.perf_counter() - t0
        else:
            wall = timed_sweeps(h, args.steps, args.moves, args.warmup)
        gpu_ms, launches = h.last_timing()
        if args.eager:
            gpu_ms, launches = ms_tot, args.steps
        moves = args.walkers * args.moves * args.steps
        value = moves / wall
        # guest sites present on average during the timed sweeps
        nsites = base.guests[0].natoms
        tot_n, nmax, nmin = 0.0, 0, 10 ** 9
        for w in range(args.walkers):
            st = h.stats(w)
            tot_n += st.sum_n[0] / max(st.counted, 1)
            nmax = max(nmax, st.nmols[0])
            nmin = min(nmin, st.nmols[0])
        n_avg = tot_n / args.walkers
        B = algorithmic_bytes(h.nframework, n_avg * nsites, h.newld)
        launch_s = gpu_ms / 1e3 / max(launches, 1)
        achieved = B * args.walkers * args.moves / launch_s
        traffic = valu = None
        pmc = os.path.join(ROOT, "profiles", "k_run_batched_isotherm_counters.json")
        if os.path.exists(pmc):
            try:
                cj = json.load(open(pmc))
                per_trial = cj.get("hbm_bytes_per_trial")
                traffic = per_trial * args.walkers * args.moves if per_trial else None
                valu = cj.get("valu")
            except Exception:
                traffic = valu = None
        import math
        mid_i = min(range(len(systems)), key=lambda i: (abs(systems[i].temperature - 293.0),
                                                        abs(math.log10(systems[i].pressures[0] / 1e5) + 1.0)))
        fp64 = fp64_batched = None
        if not args.no_fp64:
            note(f"mixed {value:.4g} trial moves/s; FP64 batch")
            # plain FP64 on the same 64 state points at the same loading
            hd = Handle(base, walkers=walkers, precision=FMC_FP64)
            hd.set_team(team)
            hd.set_block(block)
            copy_state(h, hd, args.walkers)
            wd = timed_sweeps(hd, min(args.steps, 10), args.moves, 1)
            fp64_batched = {"value": args.walkers * args.moves * min(args.steps, 10) / wd, "unit": "trial moves/s",
                            "config": "plain FP64 on the same 64 state points, started from the mixed walkers' "
                                      "molecules, same launch shape and graph sweeps",
                            "mean_guests": round(mean_guests(hd, args.walkers), 2)}
            hd.close()
            mid = systems[mid_i]  # T = 293 K, P = 0.1 bar: the config-0 state point
            hs = Handle(base, walkers=[make_walker(base, seed=(11, 23, 37, 41), temperature=mid.temperature,
                                                   pressures=mid.pressures, framework_xyz=mid.framework_xyz(),
                                                   keep=keep)], precision=FMC_FP64)
            hs.set_block(0)
            hs.run_graph(5, args.moves)
            hs.run_graph(5, args.moves)
            t1 = time.perf_counter()
            hs.run_graph(5, args.moves)
            device_sync()
            w1 = time.perf_counter() - t1
            fp64 = {"value": 5 * args.moves / w1, "unit": "trial moves/s",
                    "config": f"plain FP64, one state point (T={mid.temperature:g} K, "
                              f"P={mid.pressures[0] / 1e5:.4g} bar), same device kernel, 1 walker",
                    "threads_per_walker": hs.get_block(), "workgroups_per_walker": hs.get_team()}
            hs.close()
        cpu = cpu_mpi = None
        if not args.no_cpu:
            note("CPU baseline")
            cpu = reference_cpu(dirs[mid_i], args.cpu_seconds)
            if cpu is None:
                cpu = oracle_cpu(systems[mid_i], args.cpu_seconds)
            elif cpu["kind"] == "reference":
                # the host's CPU share for one GPU, at most 16 cores
                nproc = max(1, min(16, os.cpu_count() or 1))
                steps = int(cpu["value"] * args.cpu_seconds * 0.5)
                cpu_mpi = reference_cpu_concurrent(dirs[mid_i], steps, nproc)
        h.close()
        widom = None
        if not args.no_widom:
            from fastmc_amd import inputs as I
            from fastmc_amd.system import read_system
            wdir = os.path.join(tmp, "widom")
            os.makedirs(wdir, exist_ok=True)
            I.config0(wdir, pressure_bar=1.0)
            note("Widom grid")
            widom = {"workload": "config-0 framework (424 atoms, 871 k-vectors), CO2, grid spacing 0.5 A, "
                                 "20 orientations per point, empty framework, one walker"}
            widom.update(widom_device(read_system(wdir)))
            if not args.no_cpu:
                widom["cpu_reference"] = widom_reference_cpu(wdir)
                if widom["cpu_reference"]:
                    widom["fp64_vs_cpu_reference"] = (widom["fp64"]["attempts_per_s"] /
                                                      widom["cpu_reference"]["attempts_per_s"])
        extra = {}
        if not args.no_configs:
            note("config 1")
            extra["config1_supercell"] = config1_supercell()
            note("config 3")
            extra["config3_gradient"] = config3_gradient()
            note("config 4")
            extra["config4_wang_landau"] = config4_wang_landau()
        line = {
            "metric": METRIC, "value": value, "unit": "trial moves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "fp32+fp64", "data": "synthetic",
            "config": {"workload": "batched isotherm: 64 state points (T 273/293/323 K x P 1e-3..10 bar), "
                                   "424-atom jittered cubic frameworks, CO2, rcut 12.5 A, 871 k-vectors, "
                                   "mixed precision, hipGraph sweeps, equilibrated before timing",
                       "walkers": args.walkers, "moves_per_walker_per_step": args.moves,
                       "threads_per_walker": block, "workgroups_per_walker": team,
                       "mean_guests": round(n_avg, 2),
                       "guests_min_max": [nmin, nmax],
                       "equilibration_sweeps": fill_sweeps, "mean_guests_after_equilibration": round(fill_n, 2)},
            "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK, "traffic": traffic,
                         "kernel": "k_run", "bytes_per_trial": B,
                         "avg_launch_ms": launch_s * 1e3, "valu": valu},
            "cpu_baseline": cpu,
            "cpu_baseline_concurrent": cpu_mpi,
            "fp64_single": fp64,
            "batched_vs_fp64_single": (value / fp64["value"]) if fp64 else None,
            "fp64_batched": fp64_batched,
            "mixed_vs_fp64_batched": (value / fp64_batched["value"]) if fp64_batched else None,
            "widom_grid": widom,
        }
        line.update(extra)
        print(json.dumps(line), flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
