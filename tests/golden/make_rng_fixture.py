"""Writes rng_streams.json: the first 16 raw values and uniforms of every
workload stream (seed 0), so a change to the stream definition is caught."""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from morganapublic_amd import workloads as W  # noqa: E402

out = {}
for sid in range(1, 9):
    out[str(sid)] = {"raw": [int(v) for v in W.stream_raw(0, sid, 16)],
                     "uniform": [float(v) for v in W.stream_uniform(0, sid, 16)],
                     "normal": [float(v) for v in W.stream_normal(0, sid, 8)]}
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_streams.json"),
               "w"), indent=1)
