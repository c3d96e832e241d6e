"""Persisted launch shapes found by the TV-L1 autotuner (of_tvl1_autotune).

The table maps a workload key "WxH/s<nscales>/w<warps>/<arch>" to the launch
variant (iterations per launch, rows per wave, persistent) of every pyramid
scale, finest first, plus the per-variant timings the choice was made from.
"""
from __future__ import annotations

import json
import os

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuning",
                     "tvl1_launch_shapes.json")
ARCH = "gfx950"


def key(width: int, height: int, nscales: int, warps: int, arch: str = ARCH) -> str:
    return f"{width}x{height}/s{nscales}/w{warps}/{arch}"


def load() -> dict:
    if not os.path.exists(TABLE):
        return {}
    with open(TABLE) as f:
        return json.load(f)


def lookup(width, height, nscales, warps, arch=ARCH):
    return load().get(key(width, height, nscales, warps, arch))


def store(width, height, nscales, warps, shapes, ms_table=None, arch=ARCH) -> None:
    data = load()
    entry = {"shapes": [list(map(int, s)) for s in shapes]}
    if ms_table is not None:
        entry["ms_per_scale_by_shape"] = [[round(float(v), 4) for v in row] for row in ms_table]
    data[key(width, height, nscales, warps, arch)] = entry
    os.makedirs(os.path.dirname(TABLE), exist_ok=True)
    with open(TABLE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def apply(tv, shapes) -> None:
    """Installs one launch rule per scale (keyed by the scale's pixel count);
    each entry is (K, R) or (K, R, persistent)."""
    tv.clear_launch_rules()
    for s, shape in enumerate(shapes):
        k, r = shape[0], shape[1]
        p = shape[2] if len(shape) > 2 else 0
        nx, ny = tv.scale_size(s)
        tv.set_launch_rule(nx * ny, int(k), int(r), int(p))
