"""Known answers for the jump / swap / switch moves from the compiled
reference (run in the build container only; the reference never travels):

    make -C oracle ref                   # builds oracle/_ref/gcmc_ser.x from the reference sources
    python tests/golden/make_moves_ref.py

The reference's own CALF-20 CO2/N2 test system (tests/golden/calf20, the
reference test's FIELD and CONFIG) with the move set of its test CONTROL
(insert, delete, displace, swap, switch and jump at 1/6 each,
test/T293.0P*/CONTROL) runs from an empty framework for a few thousand
steps.  From OUTPUT: the duni seeds the reference printed, the step count,
"Total accepted steps" and each move kind's acceptance ratio (f15.9,
gcmc.f:1461-1491).  Replaying the same seeds, a faithful restatement makes
every decision the reference made, so it reproduces these numbers exactly.
"""
from __future__ import annotations

import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)

EXE = os.path.join(REPO, "oracle", "_ref", "gcmc_ser.x")
MOVES = """  move insert 0.167
  move delete 0.167
  move displace 0.167
  move jump 0.167
  move swap 0.167
  move translate 0.0
  move rotate 0.0
  move switch 0.167"""
CASES = [("calf20_p1_4_moves6", 3000), ("calf20_p1_4_moves6_long", 12000)]


def control(steps):
    import make_golden
    ctl = make_golden.CALF_CONTROL
    ctl = re.sub(r"(?m)^steps .*$", f"steps    {steps}", ctl)
    ctl = re.sub(r"(?m)^equilibration .*$", "equilibration    0", ctl)
    blocks = ctl.split("&guest")
    out = [blocks[0]]
    for b in blocks[1:]:
        b = re.sub(r"(?ms)^  move insert.*?^  move switch [0-9.]+$", MOVES, b)
        out.append(b)
    return "&guest".join(out)


def run_case(name, steps):
    d = tempfile.mkdtemp(prefix=f"fmc_moves_{name}_")
    try:
        for f in ("FIELD", "CONFIG"):
            shutil.copy(os.path.join(HERE, "calf20", f), os.path.join(d, f))
        ctl = control(steps)
        open(os.path.join(d, "CONTROL"), "w").write(ctl)
        subprocess.run([EXE], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=3600, check=False)
        out = open(os.path.join(d, "OUTPUT")).read()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    seeds = [int(x) for x in re.search(r"i,j,k,l values for duni\(\)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)", out).groups()]
    nsteps = int(re.search(r"Time elapsed for\s+(\d+)\s+gcmc steps", out).group(1))
    total = int(re.search(r"Total accepted steps :\s+(\d+)", out).group(1))
    ratios = {k: float(v) for k, v in re.findall(r"\s(insertion|deletion|displacement|jump|swap|switch|translation|rotation) ratio:\s+([0-9.]+)", out)}
    return {"name": name, "control": ctl, "seeds": seeds, "steps": nsteps, "total_accepted": total, "ratios": ratios}


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run `make -C oracle ref` first")
    res = [run_case(*c) for c in CASES]
    for r in res:
        print({k: v for k, v in r.items() if k != "control"}, flush=True)
    json.dump(res, open(os.path.join(HERE, "moves_ref.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
