"""The reference's three notebook jobs (tests/golden/jobs, copied from the
reference's tests/*.json) set up as the reference's workflow does: defaults
then the job file, Geodata on the reference's own DTMs (tests/golden/dtm),
and the printed results the notebooks hold (tests/golden/notebook_stats.json).

IndianCreek_Test2.json has a trailing comma that json.load rejects (the
reference's read_json_file, parameters.py:46-92, would fail on it as the file
stands); the notebook ran before it was added, so it is dropped here.
"""
from __future__ import annotations

import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
DTM = os.path.join(GOLDEN, "dtm")
NOTEBOOK = json.load(open(os.path.join(GOLDEN, "notebook_stats.json")))
JOBS = [k for k in NOTEBOOK if not k.startswith("_")]


def job_overrides(name):
    txt = open(os.path.join(GOLDEN, "jobs", NOTEBOOK[name]["job"])).read()
    return json.loads(re.sub(r",\s*}", "}", txt))


def geodata(name):
    """(params, state, geodata) after Geodata.do() for the named job."""
    from slm_amd.geodata import Geodata, State
    from slm_amd.parameters import load_parameters
    params = load_parameters(overrides=job_overrides(name))
    state = State(params["state"])
    geo = Geodata(state, params["geodata"], dtm_path=DTM)
    geo.do()
    return params, state, geo


def stats_mismatches(name, df, skip=()):
    """Entries of the printed trajectory table that the float32 table df
    (useful.compute_stats) does not reproduce digit for digit."""
    bad = []
    for direction, cols in NOTEBOOK[name]["stats"].items():
        for col, printed in cols.items():
            for stat, want in zip(("min", "mean", "max"), printed):
                if (direction, col, stat) in skip:
                    continue
                got = "%.6f" % float(np.float32(df[col][direction, stat]))
                if got != want:
                    bad.append((direction, col, stat, got, want))
    return bad
