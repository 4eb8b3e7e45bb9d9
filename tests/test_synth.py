"""Synthetic input streams: pinned by fixtures, slice-independent, in range."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from nt2_amd import synth  # noqa: E402
import make_synth_streams  # noqa: E402


def test_streams_match_fixture():
    want = json.load(open(os.path.join(HERE, "golden", "synth_streams.json")))
    got = make_synth_streams.streams()
    assert got.keys() == want.keys()
    for k in want:
        assert got[k] == want[k], k


def test_splitmix_known_values():
    # SplitMix64 finaliser on a known state (Vigna's reference: state 0 after one increment)
    z = np.array([0x9E3779B97F4A7C15], dtype=np.uint64)
    assert int(synth._mix(z)[0]) == 0xE220A8397B1DCDAF


def test_slices_are_independent_of_chunking():
    a = synth.uniform(0, 3, 5000, 5.0, 30.0)
    b = np.concatenate([synth.unit(0, 3, 1234), synth.unit(0, 3, 5000 - 1234, start=1234)])
    assert np.array_equal(a, (5.0 + 25.0 * b).astype(np.float32))


def test_distributions():
    d = synth.bs_inputs(100000, with_w=True)
    assert d["S"].min() >= 5 and d["S"].max() <= 30
    assert d["K"].min() >= 1 and d["K"].max() <= 100
    assert d["T"].min() >= 0.25 and d["T"].max() <= 10
    assert np.all(d["r"] == np.float32(0.02)) and np.all(d["sig"] == np.float32(0.30))
    a, _ = synth.sumdot_inputs(200000)
    assert abs(a.mean() - 1.0) < 1e-5 and abs(a.std() - 1e-3) < 2e-5
    lg = synth.sweep_args("log", 100000)
    assert lg.min() >= 1e-30 and lg.max() <= 1e30
