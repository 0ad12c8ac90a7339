"""Reference isotherm loadings for the config-0 state points, from the
compiled reference (run in the build container only; the reference never
travels):

    make -C oracle ref                   # builds oracle/_ref/gcmc_ser.x from the reference sources
    python tests/golden/make_isotherm_ref.py [--runs 8] [--eq 100000] [--prod 100000]
    python tests/golden/make_isotherm_ref.py --runs 16 --temperature 273 --pressures 1.0 --append

For each state point (T = 293 K at 0.1, 1 and 10 bar by default; --append adds
points to the existing fixture) the config-0 inputs (fastmc_amd.inputs.config0)
are written to a scratch directory per run and the reference's serial GCMC
program runs there, `runs` times concurrently on the host's cores (the
reference seeds duni from the clock, so the runs are independent chains; the
starts are spaced by one second so the seeds differ, and the seeds printed in
OUTPUT are recorded).  From each OUTPUT the final statistics block is read:
"<N>: mean +/- standard error" (the reference's block-average error,
gcmc.f:1458-1557).  The fixture holds, per pressure, every run's mean,
standard error and seeds, plus the mean over runs.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from fastmc_amd import inputs as I  # noqa: E402

EXE = os.path.join(REPO, "oracle", "_ref", "gcmc_ser.x")
PRESSURES_BAR = (0.1, 1.0, 10.0)


def parse(out: str):
    m = re.search(r"<N>:\s+([-0-9.Ee+]+)\s+\+/-\s+([-0-9.Ee+]+)", out)
    s = re.search(r"i,j,k,l values for duni\(\)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)", out)
    if not m or not s:
        raise RuntimeError("OUTPUT without final statistics or seeds")
    return float(m.group(1)), float(m.group(2)), [int(x) for x in s.groups()]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=8)
    ap.add_argument("--eq", type=int, default=100000)
    ap.add_argument("--prod", type=int, default=100000)
    ap.add_argument("--temperature", type=float, default=293.0)
    ap.add_argument("--pressures", type=float, nargs="*", default=list(PRESSURES_BAR))
    ap.add_argument("--append", action="store_true", help="add the points to the existing fixture")
    args = ap.parse_args()
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run `make -C oracle ref` first")
    path = os.path.join(HERE, "isotherm_ref_cfg0.json")
    if args.append:
        result = json.load(open(path))
        assert (result["equilibration"], result["production"]) == (args.eq, args.prod)
    else:
        result = {"system": "config 0 (424-atom framework, CO2)", "equilibration": args.eq,
                  "production": args.prod, "points": []}
    for p in args.pressures:
        dirs, procs = [], []
        for r in range(args.runs):
            d = tempfile.mkdtemp(prefix=f"fmc_iso_{p}_{r}_")
            I.config0(d, pressure_bar=p, temperature=args.temperature)
            ctl = open(os.path.join(d, "CONTROL")).read()
            ctl = re.sub(r"(?m)^steps .*$", f"steps    {args.prod}", ctl)
            ctl = re.sub(r"(?m)^equilibration .*$", f"equilibration    {args.eq}", ctl)
            open(os.path.join(d, "CONTROL"), "w").write(ctl)
            dirs.append(d)
            procs.append(subprocess.Popen([EXE], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
            time.sleep(1.1)
        runs = []
        for d, pr in zip(dirs, procs):
            pr.wait()
            mean, err, seeds = parse(open(os.path.join(d, "OUTPUT")).read())
            runs.append({"mean_n": mean, "stderr_n": err, "seeds": seeds})
            shutil.rmtree(d, ignore_errors=True)
        n = [x["mean_n"] for x in runs]
        mu = sum(n) / len(n)
        sd = (sum((x - mu) ** 2 for x in n) / max(1, len(n) - 1)) ** 0.5
        result["points"].append({"temperature": args.temperature, "pressure_bar": p, "runs": runs, "mean_n": mu,
                                 "sd_between_runs": sd,
                                 "mean_stderr_n": sum(x["stderr_n"] for x in runs) / len(runs)})
        print(args.temperature, "K", p, "bar:", mu, "+/-", sd, "between runs; per-run stderr",
              result["points"][-1]["mean_stderr_n"], flush=True)
    json.dump(result, open(path, "w"), indent=1)


if __name__ == "__main__":
    main()
