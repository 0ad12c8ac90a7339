"""numguests.out known answers from the compiled reference (run in the build
container only; the reference never travels):

    make -C oracle ref                   # builds oracle/_ref/gcmc_ser.x from the reference sources
    python tests/golden/make_numguests_ref.py

The reference's CALF-20 CO2/N2 system with its test move set (insert,
delete, displace, swap, switch, jump at 1/6 each) runs with 'numguests 10';
the duni seeds come from OUTPUT, and branch01/numguests01.out and
numguests02.out (one line per guest every 10 steps, format
(i9,i7,2f20.6,7i4), gcmc.f:1138-1163) are kept as text.  A config-0 run
(one guest: branch01/numguests.out) covers the single-file name.
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

EXE = os.path.join(REPO, "oracle", "_ref", "gcmc_ser.x")


def set_numg(d, n):
    p = os.path.join(d, "CONTROL")
    ctl = re.sub(r"(?m)^numguests .*$", f"numguests {n}", open(p).read())
    open(p, "w").write(ctl)


def write_calf20(d):
    import make_stats_ref
    make_stats_ref.write_calf20_moves6(d, 0, 2000)
    set_numg(d, 10)


def write_cfg0(d):
    import make_stats_ref
    make_stats_ref.write_cfg0(d, 0, 1500)
    set_numg(d, 7)


CASES = [("calf20_numg", write_calf20), ("cfg0_numg", write_cfg0)]


def run_case(name, writer):
    d = tempfile.mkdtemp(prefix=f"fmc_numg_{name}_")
    try:
        writer(d)
        control = open(os.path.join(d, "CONTROL")).read()
        subprocess.run([EXE], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=3600, check=False)
        out = open(os.path.join(d, "OUTPUT")).read()
        files = {f: open(os.path.join(d, "branch01", f)).read()
                 for f in sorted(os.listdir(os.path.join(d, "branch01"))) if f.startswith("numguests")}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    seeds = [int(x) for x in re.search(r"i,j,k,l values for duni\(\)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)", out).groups()]
    total = int(re.search(r"Time elapsed for\s+(\d+)\s+gcmc steps", out).group(1))
    return {"name": name, "control": control, "seeds": seeds, "total": total, "files": files}


def main():
    if not os.path.exists(EXE):
        sys.exit(f"{EXE} missing: run `make -C oracle ref` first")
    res = [run_case(*c) for c in CASES]
    for r in res:
        print({k: (v if k != "files" else {f: len(t.splitlines()) for f, t in v.items()}) for k, v in r.items()
               if k != "control"}, flush=True)
    json.dump(res, open(os.path.join(HERE, "numguests_ref.json"), "w"), indent=1)


if __name__ == "__main__":
    main()
