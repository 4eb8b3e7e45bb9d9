"""Every coefficient table and constant of NT2's functors on the path, as the
reference's own headers hold them (tests/golden/nt2_coefficients.json, read
from the hexadecimal words by tests/golden/make_nt2_coefficients.py), against
the literals of the oracle (oracle/nt2ref.hpp) and of the GPU functors
(include/nt2hip/math.hpp): each table must appear in each source as a
contiguous run of exactly those values, in the reference's order."""
import json
import math
import os
import re
import struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = json.load(open(os.path.join(ROOT, "tests", "golden", "nt2_coefficients.json")))
SOURCES = {"oracle": os.path.join(ROOT, "oracle", "nt2ref.hpp"), "gpu": os.path.join(ROOT, "include", "nt2hip", "math.hpp")}

# tables of the cited headers that no evaluated functor uses
NOT_ON_PATH = {
    ("d_expo_reduction.hpp", 152): "double exp10 (exp2/exp10 are evaluated in float only)",
    ("d_expo_reduction.hpp", 158): "double exp10",
    ("f_trig_evaluation.hpp", 53): "float tan",
    ("d_trig_evaluation.hpp", 55): "double tan",
    ("d_trig_evaluation.hpp", 58): "double tan",
}
# constants on the path and the types whose value must appear in both sources
CONSTANTS = {
    "fd": ["Fiveo_8", "Half", "Invlog_2", "Invpio_2", "Invsqrt_2pi", "Log_2hi", "Log_2lo", "Maxlog", "Medium_pi",
           "Mhalf", "Minlog", "Mone", "Mtwo", "One", "Pio2_1", "Pio2_1t", "Pio2_2", "Pio2_2t", "Pio2_3", "Pio2_3t",
           "Pio_2", "Pio_4", "Sqrt_2o_2", "Two", "Twoopi"],
    "f": ["Invlog10_2", "Log10_2hi", "Log10_2lo", "Log_2", "Maxlog10", "Maxlog2", "Minlog10", "Minlog2",
          "Mlogtwo2nmb", "Smallestposval", "Twotonmb", "Twothird"],
}
# not on the path: log10/log2 (Invlog_10, Log10_e*, Log2_em1, Log_2olog_10,
# Mlog10two2nmb, Mlog2two2nmb), degree and pi-scaled trigonometry (Oneo_180,
# Oneo_90, Pio_180, Quarter, float Pi), Three (an integer mask), Nan.

HEXF = re.compile(r"(-?)\s*0x([0-9a-fA-F]+)(?:\.([0-9a-fA-F]*))?p([+-]?\d+)f?"
                  r"|(-?)\s*0x([0-9a-fA-F]+)(ull|ULL|ll|LL|u|U)?\b"
                  r"|(?<![\w.])(-?)(\d+\.\d*(?:[eE][+-]?\d+)?)f?\b")


def literals(path):
    """Numeric literals of a source in order: hexadecimal floats, hexadecimal
    words (8 digits: float bits; 16 or an ll suffix: double bits) and decimal
    floats, with a unary minus applied."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    out = []
    for m in HEXF.finditer(text):
        if m.group(2) is not None:
            out.append(float.fromhex(f"{m.group(1)}0x{m.group(2)}.{m.group(3) or ''}p{m.group(4)}"))
        elif m.group(6) is not None:
            dg, suf = m.group(6), (m.group(7) or "").lower()
            n = int(dg, 16)
            if len(dg) > 8 or suf in ("ll", "ull"):
                v = struct.unpack("<d", struct.pack("<Q", n & (2 ** 64 - 1)))[0]
            else:
                v = struct.unpack("<f", struct.pack("<I", n & 0xFFFFFFFF))[0]
            out.append(-v if m.group(5) else v)
        else:
            v = float(m.group(9))
            out.append(-v if m.group(8) else v)
    return out


def contains(seq, sub):
    n = len(sub)
    return any(seq[i:i + n] == sub for i in range(len(seq) - n + 1))


def table_key(t):
    return (t["file"].split("/")[-1], t["line"])


def test_json_words_decode_to_their_values():
    assert len(DATA["tables"]) == 28
    for t in DATA["tables"]:
        for p in t["parts"]:
            for w, v in zip(p["words"], p["values"]):
                n = int(w, 16)
                got = (struct.unpack("<f", struct.pack("<I", n))[0] if p["type"] == "float"
                       else struct.unpack("<d", struct.pack("<Q", n))[0])
                assert got == float.fromhex(v)


def test_every_table_on_the_path_is_in_both_sources():
    lits = {k: literals(p) for k, p in SOURCES.items()}
    missing = []
    for t in DATA["tables"]:
        if table_key(t) in NOT_ON_PATH:
            continue
        for p in t["parts"]:
            want = [float.fromhex(v) for v in p["values"]]
            for name, seq in lits.items():
                if not contains(seq, want):
                    missing.append((name,) + table_key(t))
    assert not missing, missing


def test_a_one_ulp_slip_in_a_table_is_caught():
    seq = literals(SOURCES["gpu"])
    for t in DATA["tables"]:
        if table_key(t) in NOT_ON_PATH:
            continue
        want = [float.fromhex(v) for v in t["parts"][0]["values"]]
        slipped = list(want)
        k = len(slipped) // 2
        slipped[k] = math.nextafter(slipped[k], math.inf)
        assert contains(seq, want) and not contains(seq, slipped), table_key(t)


def test_constants_on_the_path_are_in_both_sources():
    vals = {k: set(abs(v) for v in literals(p)) for k, p in SOURCES.items()}
    missing = []
    for types, names in CONSTANTS.items():
        for nm in names:
            c = DATA["constants"][nm]
            for ty in types:
                v = abs(float.fromhex(c["float" if ty == "f" else "double"]))
                for src, vs in vals.items():
                    if v not in vs:
                        missing.append((src, nm, ty))
    assert not missing, missing
    # double Pi enters only erf/erfc's tiny-argument branch, 2 x / sqrt(Pi):
    # the oracle holds Pi, the GPU functor sqrt(Pi) correctly rounded
    pi = float.fromhex(DATA["constants"]["Pi"]["double"])
    assert pi in vals["oracle"] and math.sqrt(pi) in vals["gpu"]


def test_erf_interval_thresholds_are_in_both_sources():
    """The scalar erf/erfc intervals (A0(v) in erf.hpp, erfc.hpp: 1e-10, 0.65,
    2.2, 6, and float erf's 0.4 shift) as doubles and, for the float functor,
    rounded to float."""
    assert len(DATA["decimals"]) == 9
    vals = {k: set(abs(v) for v in literals(p)) for k, p in SOURCES.items()}
    for d in DATA["decimals"]:
        v = abs(float.fromhex(d["double"]))
        vf = struct.unpack("<f", struct.pack("<f", v))[0]
        for src, vs in vals.items():
            assert v in vs or vf in vs, (src, d)
