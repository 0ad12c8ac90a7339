"""The reference's own isotherm test as a fixture (read in the build
container only; the reference never travels): for each of its ten state
points (test/T293.0P*/CONTROL: T = 293 K, CO2 and N2 partial pressures,
750000 equilibration and 500000 production steps, insert / delete /
displace / swap / switch / jump at 1/6 each) the partial pressures, and the
expected average numbers of CO2 and N2 per cell that run_test.py holds
(reference_dict, run_test.py:27-38, keyed by the total pressure in bar).
The test system's FIELD and CONFIG are the fixture files in calf20/.

    python tests/golden/make_calf20_isotherm.py
"""
from __future__ import annotations

import glob
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def main():
    src = open(os.path.join(REF, "run_test.py")).read()
    block = src[src.index("reference_dict = {"):src.index("}", src.index("reference_dict = {"))]
    expected = {float(k): [float(a), float(b)]
                for k, a, b in re.findall(r"([0-9.]+):\s*\[([0-9.]+),\s*([0-9.]+)\]", block)}
    points = []
    for d in sorted(glob.glob(os.path.join(REF, "test", "T293*"))):
        ctl = open(os.path.join(d, "CONTROL")).read()
        p = [float(x) for x in re.findall(r"pressure \(bar\)\s+([0-9.]+)", ctl)]
        total = round(sum(p), 6)
        key = min(expected, key=lambda k: abs(k - total))
        assert abs(key - total) < 2e-6, (d, total, key)
        points.append({"dir": os.path.basename(d), "pressures_bar": p, "total_bar": key,
                       "expected_co2_n2": expected[key],
                       "steps": int(re.search(r"(?m)^steps\s+(\d+)", ctl).group(1)),
                       "equilibration": int(re.search(r"(?m)^equilibration\s+(\d+)", ctl).group(1)),
                       "control": ctl})
    json.dump({"source": "run_test.py:27-38 and test/T293.0P*/CONTROL", "points": points},
              open(os.path.join(HERE, "calf20_isotherm_ref.json"), "w"), indent=1)
    print(len(points), "points")


if __name__ == "__main__":
    main()
