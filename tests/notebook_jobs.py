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


def pass1_counts(name, device):
    """Mapping pass 1 of the named job (mapping.py:240-295) on the GPU after
    Preprocess and Trace: mid-slope and ridge maps, then the coarse channel
    mapping and segmentation at the JSON's coarse thresholds.  Returns (number
    of segments, number of subsegments) as segment.py prints them."""
    import torch
    from slm_amd import hip
    from slm_amd.mapping import ChannelMapper
    from slm_amd.parameters import section
    from slm_amd.preprocess import Preprocess
    from slm_amd.trace import Trace
    from slm_amd.useful import Data, Info, get_bbox
    params, state, geo = geodata(name)
    pre = Preprocess(state, section(params, "preprocess"), geo)
    pre.do()
    tr = Trace(state, section(params, "trace"), geo, pre)
    tr.do()
    mp = section(params, "mapping")
    mask_array = state.merge_active_masks()
    bbox, nxb, nyb = get_bbox(~mask_array)
    info = Info(state, section(params, "trace"), geo.roi_pixel_size, mapping=mp)
    info.set_xy(nxb, nyb, geo.pad_width)
    info.set_thresholds(channel_threshold=np.float32(mp.coarse_channel_threshold),
                        segmentation_threshold=np.uint32(mp.coarse_segmentation_threshold))
    data = Data(info=info, bbox=bbox, pad=geo.pad_width, mapping_array=np.zeros(mask_array.shape, np.uint32),
                mask_array=mask_array, uv_array=pre.uv_array, slc_array=tr.slc_array, slt_array=tr.slt_array,
                sla_array=tr.sla_array)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    mask = t(data.mask_array)
    mapping = torch.zeros(data.mask_array.shape, dtype=torch.int32, device=device)
    hip.map_midslopes_ridges(t(data.sla_array), mask, mapping, int(info.is_midslope), int(info.is_ridge),
                             mp.midslope_filter_sigma, mp.midslope_threshold, mp.ridge_filter_sigma,
                             mp.ridge_threshold, sync=True)
    cm = ChannelMapper(info, mask, t(data.uv_array), mapping, slt=t(data.slt_array[:, :, 0]),
                       slc=t(data.slc_array[:, :, 0].astype(np.float64)))
    assert cm.map_channels_segments()
    lab = cm.label.cpu().numpy()
    return int(cm.n_segments), int(np.unique(lab[lab != 0]).shape[0])
