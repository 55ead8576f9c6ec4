"""Autotuning cache: the fastest tile schedule (unit order, workgroup count) of a plan shape on a
device, measured once by msosm_plan_autotune and kept in a JSON file.

The key is the device name and the plan shape (FFT length, DMs, blocks per push, input format,
outputs, fold bins). ``tuned_schedule`` returns the cached choice or measures it on the given device
input and stores it.
"""
from __future__ import annotations

import json
import os

from . import Plan

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "msosm_amd", "schedules.json")


def plan_key(device_name: str, plan: Plan, nbins: int = 0) -> str:
    return f"{device_name}|L={2 * plan.M}|D={plan.D}|B={plan.B}|in={plan.in_format}|out={plan.out_flags}|nbins={nbins}"


def _load(path: str) -> dict:
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def tuned_schedule(plan: Plan, input_dev: int, device_name: str, nbins: int = 0, pushes: int = 3,
                   cache: str = DEFAULT_CACHE) -> dict:
    """Apply the cached schedule for this plan shape, or measure it (the measurement resets the plan
    and its fold) and cache it. Returns {"workgroups", "order", "ms", "cached"}."""
    key = plan_key(device_name, plan, nbins)
    table = _load(cache)
    if key in table:
        ent = table[key]
        plan.set_schedule(ent["workgroups"], ent["order"])
        return dict(ent, cached=True)
    w, o, ms = plan.autotune(input_dev, pushes)
    ent = {"workgroups": w, "order": o, "ms": ms}
    table[key] = ent
    try:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tmp = cache + ".tmp"
        with open(tmp, "w") as fh:
            json.dump(table, fh, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    except OSError:
        pass  # a read-only home keeps the measured choice for this plan only
    return dict(ent, cached=False)
