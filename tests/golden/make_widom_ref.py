"""Widom-grid known answers from the compiled reference (run in the build
container only; the reference never travels):

    make -C oracle ref                   # builds oracle/_ref/gcmc_ser.x from the reference sources
    python tests/golden/make_widom_ref.py

Each case writes its inputs to a scratch directory, adds the reference's
Widom directives to CONTROL ("henry", "widom insertions N", "grid spacing
D"; utility_pack.f:390,436-442, readinputs.f:598-599) and runs the
reference's serial program there.  From OUTPUT it records the duni seeds the
reference printed, the probability-grid size, and per guest the Henry's
constant of the node report (E20.9, gcmc.f:1319-1321) and the number of
non-overlapping attempts ("Total accepted steps" is guest 1's; the final
"Total steps counted" is printed per guest).
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
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

from fastmc_amd import inputs as I  # noqa: E402

EXE = os.path.join(REPO, "oracle", "_ref", "gcmc_ser.x")
WIDOM = "henry\nwidom insertions {n}\ngrid spacing {d}\nfinish"


def write_calf20(d):
    import make_golden
    make_golden.write_calf20(d)


CASES = [
    # name, writer, insertions per point, grid spacing (A)
    ("cfg0_p1", lambda d: I.config0(d, pressure_bar=1.0), 5, 2.0),
    ("calf20_co2n2", write_calf20, 3, 1.5),
]


def run_case(name, writer, nwid, griddim):
    d = tempfile.mkdtemp(prefix=f"fmc_widom_{name}_")
    try:
        writer(d)
        ctl = open(os.path.join(d, "CONTROL")).read()
        assert "finish" in ctl
        open(os.path.join(d, "CONTROL"), "w").write(ctl.replace("finish", WIDOM.format(n=nwid, d=griddim)))
        subprocess.run([EXE], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=3600, check=False)
        out = open(os.path.join(d, "OUTPUT")).read()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    seeds = [int(x) for x in re.search(r"i,j,k,l values for duni\(\)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)", out).groups()]
    grid = [int(x) for x in re.search(r"Probability grid size\s+:\s+(\d+)\s+(\d+)\s+(\d+)", out).groups()]
    node = out[out.index("Data reported from node"):out.index("final stats")]
    henry = [float(x) for x in re.findall(r"Henry's Constant \(mol /kg /bar\):\s+([-0-9.E+]+)", node)]
    final = out[out.index("final stats"):]
    counted = [int(x) for x in re.findall(r"Total steps counted:\s+(\d+)", final)]
    return {"name": name, "insertions_per_point": nwid, "grid_spacing": griddim, "seeds": seeds, "grid": grid,
            "henry": henry, "counted": counted}


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run `make -C oracle ref` first")
    res = [run_case(*c) for c in CASES]
    for r in res:
        print(r, flush=True)
    json.dump(res, open(os.path.join(HERE, "widom_ref.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
