"""Shared test helpers: golden fixtures and the workloads they were recorded on."""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import shutil
import tempfile

from fastmc_amd import inputs as I
from fastmc_amd.system import read_system

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
KB_KCAL = 1.9872041e-3

WRITERS = {
    "cfg0_p0.1": lambda d: I.config0(d, pressure_bar=0.1),
    "cfg0_p10": lambda d: I.config0(d, pressure_bar=10.0),
    "cfg0_moves5": lambda d: I.config0(d, pressure_bar=3.0, moves={
        "insert": 0.2, "delete": 0.2, "displace": 0.2, "jump": 0.0, "swap": 0.0,
        "translate": 0.2, "rotate": 0.2, "switch": 0.0}),
    "cfg1_p10": lambda d: I.config1(d, pressure_bar=10.0),
}

CASES = ["cfg0_p0.1", "cfg0_p10", "cfg0_moves5", "calf20_co2n2", "cfg1_p10"]

_cache = {}


def fixture(name: str) -> dict:
    if name not in _cache:
        with gzip.open(os.path.join(GOLDEN, f"{name}.json.gz"), "rt") as fh:
            _cache[name] = json.load(fh)
    return _cache[name]


def case_dir(name: str) -> str:
    """Write the case's inputs to a fresh temp dir and check them against the fixture hashes."""
    fx = fixture(name)
    d = tempfile.mkdtemp(prefix=f"fmc_{name}_")
    if name in WRITERS:
        WRITERS[name](d)
    else:
        for f in ("FIELD", "CONFIG"):
            shutil.copy(os.path.join(GOLDEN, "calf20", f), os.path.join(d, f))
        with open(os.path.join(d, "CONTROL"), "w") as fh:
            fh.write(fx["control"])
    for f, h in fx["input_sha256"].items():
        with open(os.path.join(d, f), "rb") as fh:
            assert hashlib.sha256(fh.read()).hexdigest() == h, f"{name}: regenerated {f} differs"
    return d


def case_system(name: str):
    d = case_dir(name)
    s = read_system(d)
    shutil.rmtree(d, ignore_errors=True)
    return s


def du_tol(ref: float, temp: float, rel: float = 1e-10) -> float:
    """FP64 parity bound on a trial energy change: rel * max(|dU|, kB T) (kcal/mol)."""
    return rel * max(abs(ref), KB_KCAL * temp)


def compare_trace(tr, fx, nsteps=None, rel=1e-10, temp=None):
    """Compare a trace (numpy TRACE_DTYPE) with a golden fixture; returns (n_bad, max_scaled_err)."""
    t = fx["trace"]
    n = len(t["step"]) if nsteps is None else nsteps
    temp = fx["params"]["temp"] if temp is None else temp
    bad = []
    worst = 0.0
    for i in range(n):
        r = tr[i]
        same = (r["iguest"] == t["iguest"][i] and r["move"] == t["move"][i] and
                r["choice"] == t["choice"][i] and r["nmols"] == t["nmols"][i] and
                r["overlap"] == t["overlap"][i] and r["accepted"] == t["accepted"][i] and
                r["drew_rande"] == t["drew_rande"][i])
        xs = t["xyz"][i]
        dx = max(abs(a - b) for a, b in zip(xs, r["xyz"][: len(xs)]))
        err = 0.0
        if not t["overlap"][i]:
            err = abs(r["estep"] - t["estep"][i]) / du_tol(t["estep"][i], temp, 1.0)
            worst = max(worst, err)
        if not same or dx > 1e-8 or err > rel:
            bad.append(i)
    return bad, worst
