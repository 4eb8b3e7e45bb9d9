"""Writes synth_streams.json: the first 64 values of every synthetic input
stream of the benchmark and parity workloads (catches generator drift).
Run: python tests/golden/make_synth_streams.py"""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from nt2_amd import synth  # noqa: E402


def bits(a):
    return [int(v) for v in np.ascontiguousarray(a[:64]).view(np.uint32 if a.dtype == np.float32 else np.uint64)]


def streams():
    out = {}
    x, y = synth.saxpy_inputs(64)
    out["saxpy.x"], out["saxpy.y"] = bits(x), bits(y)
    d = synth.bs_inputs(64, with_w=True)
    for k, v in d.items():
        out["bs." + k] = bits(v)
    for fn in sorted(synth.SWEEP):
        out["sweep." + fn] = bits(synth.sweep_args(fn, 64))
    a, b = synth.sumdot_inputs(64)
    out["sumdot.a"], out["sumdot.b"] = bits(a), bits(b)
    X, s, k, p = synth.colfold_inputs(rows=64, cols=4)
    out["colfold.X"], out["colfold.s"], out["colfold.k"], out["colfold.p"] = bits(X.ravel()), bits(s), bits(k), bits(p)
    out["raw.seed0.id1"] = [int(v) for v in synth.raw_bits(0, 1, 0, 16)]
    return out


if __name__ == "__main__":
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "synth_streams.json")
    json.dump(streams(), open(p, "w"))
    print("wrote", p)
