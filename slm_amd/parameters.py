"""JSON parameter files, merged as the reference does.

read_json_file restates python/streamlines/parameters.py:46-92: files are read
in order and each later file overrides single keys of a section, never whole
sections.  Sections become attribute namespaces, as Core.__init__ injects them
(python/streamlines/core.py:82-96).
"""
from __future__ import annotations

import json
import os
from types import SimpleNamespace

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS = os.path.join(HERE, "defaults.json")


def read_json_file(paths):
    """Merge JSON files (paths with or without the .json suffix) section by section."""
    merged = {}
    for path in paths:
        if not path.endswith(".json"):
            path = path + ".json"
        with open(path) as fp:
            params = json.load(fp)
        for key, item in params.items():
            if isinstance(item, dict):
                merged.setdefault(key, {}).update(item)
            else:
                merged[key] = item
    return merged


def load_parameters(job_json=None, overrides=None):
    """defaults.json, then the job file, then an override dict {section: {key: value}}."""
    paths = [DEFAULTS] + ([job_json] if job_json else [])
    merged = read_json_file(paths)
    for sec, kv in (overrides or {}).items():
        merged.setdefault(sec, {}).update(kv)
    return merged


def section(params: dict, name: str) -> SimpleNamespace:
    return SimpleNamespace(**params.get(name, {}))
