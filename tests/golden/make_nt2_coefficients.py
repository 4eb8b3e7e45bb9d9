"""Write tests/golden/nt2_coefficients.json: the coefficient tables and
constants of NT2's functors on the evaluated path, read as the hexadecimal
words the reference's own headers hold.

Run in a checkout that has the reference at /root/reference (the JSON is
committed; the tests read only the JSON and this repository's sources):

    python3 tests/golden/make_nt2_coefficients.py

Decimal thresholds A0(v) of the scalar erf/erfc functors are listed too.

Tables: every NT2_HORNER_COEFF_T(type, n, (w0, ..., wn-1)) and
NT2_HORNER_RAT(type, n, m, x, (num), (den)) in the headers below, in file
order, each with its file, line and the words converted to the float or
double they encode (8 hex digits: float; 16 digits or an ll/ull suffix:
double).  Constants: every BOOST_SIMD_CONSTANT_REGISTER(Name, type, int,
float_word, double_word) under modules/.
"""
from __future__ import annotations

import json
import os
import re
import struct
import sys

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nt2_coefficients.json")

HEADERS = [
    "modules/core/exponential/include/nt2/exponential/functions/scalar/impl/expo/f_expo_reduction.hpp",
    "modules/core/exponential/include/nt2/exponential/functions/scalar/impl/expo/d_expo_reduction.hpp",
    "modules/core/exponential/include/nt2/exponential/functions/scalar/impl/logs/f_kernel.hpp",
    "modules/core/exponential/include/nt2/exponential/functions/scalar/impl/logs/d_log.hpp",
    "modules/core/euler/include/nt2/euler/functions/details/erf_kernel.hpp",
    "modules/core/hyperbolic/include/nt2/hyperbolic/functions/details/tanh_kernel.hpp",
    "modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/impl/trigo/f_trig_evaluation.hpp",
    "modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/impl/trigo/d_trig_evaluation.hpp",
    "modules/core/statistics/include/nt2/statistics/functions/generic/fastnormcdf.hpp",
]

WORD = re.compile(r"0x([0-9a-fA-F]+)(ull|ULL|ll|LL|ul|UL|u|U|l|L)?\b")


def strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", lambda m: re.sub(r"[^\n]", " ", m.group(0)), text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def word_value(digits: str, suffix: str | None) -> tuple[str, str, float]:
    dbl = len(digits) > 8 or (suffix or "").lower() in ("ll", "ull")
    v = int(digits, 16)
    if dbl:
        f = struct.unpack("<d", struct.pack("<Q", v))[0]
        return "double", f"0x{v:016x}", f
    f = struct.unpack("<f", struct.pack("<I", v))[0]
    return "float", f"0x{v:08x}", f


def balanced(text: str, open_at: int) -> int:
    """Index just past the parenthesis that closes the one at open_at."""
    depth = 0
    for i in range(open_at, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return i + 1
    raise ValueError("unbalanced")


def tables(rel: str) -> list[dict]:
    raw = open(os.path.join(REF, rel)).read()
    text = strip_comments(raw)
    out = []
    for m in re.finditer(r"NT2_HORNER_(COEFF_T|RAT)\s*\(", text):
        end = balanced(text, m.end() - 1)
        call = text[m.end():end - 1]
        groups = [g for g in re.findall(r"\(([^()]*)\)", call) if "0x" in g]
        line = text.count("\n", 0, m.start()) + 1
        parts = []
        for g in groups:
            ws = [word_value(d, s) for d, s in WORD.findall(g)]
            parts.append({"type": ws[0][0], "words": [w[1] for w in ws], "values": [w[2].hex() for w in ws]})
        kind = "horner" if m.group(1) == "COEFF_T" else "rational"
        out.append({"file": rel, "line": line, "kind": kind, "parts": parts})
    return out


def constants() -> dict:
    res = {}
    root = os.path.join(REF, "modules")
    pat = re.compile(r"BOOST_SIMD_CONSTANT_REGISTER\s*\(\s*(\w+)\s*,\s*[\w:]+\s*,\s*([^,]+?)\s*,\s*"
                     r"0x([0-9a-fA-F]+)[uUlL]*\s*,\s*0x([0-9a-fA-F]+)[uUlL]*\s*\)")
    for dp, _, fs in os.walk(root):
        for f in fs:
            if not f.endswith(".hpp"):
                continue
            p = os.path.join(dp, f)
            text = strip_comments(open(p, errors="replace").read())
            for m in pat.finditer(text):
                name = m.group(1)
                fw, dw = int(m.group(3), 16), int(m.group(4), 16)
                res[name] = {
                    "file": os.path.relpath(p, REF),
                    "float_word": f"0x{fw:08x}",
                    "float": struct.unpack("<f", struct.pack("<I", fw & 0xffffffff))[0].hex(),
                    "double_word": f"0x{dw:016x}",
                    "double": struct.unpack("<d", struct.pack("<Q", dw))[0].hex(),
                }
    return dict(sorted(res.items()))


# the scalar functors' decimal constants A0(v) (interval thresholds)
DECIMAL_HEADERS = [
    "modules/core/euler/include/nt2/euler/functions/scalar/erf.hpp",
    "modules/core/euler/include/nt2/euler/functions/scalar/erfc.hpp",
]


def decimals(rel: str) -> list[dict]:
    text = strip_comments(open(os.path.join(REF, rel)).read())
    out = []
    for m in re.finditer(r"\bA0\(\s*(-?[0-9]*\.?[0-9]+(?:[eE][+-]?[0-9]+)?)\s*\)", text):
        out.append({"file": rel, "line": text.count("\n", 0, m.start()) + 1, "text": m.group(1),
                    "double": float(m.group(1)).hex()})
    return out


def main() -> int:
    if not os.path.isdir(REF):
        print(f"{REF} is not here: the committed {OUT} stays as it is", file=sys.stderr)
        return 1
    data = {"tables": [t for h in HEADERS for t in tables(h)], "constants": constants(),
            "decimals": [d for h in DECIMAL_HEADERS for d in decimals(h)]}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
        f.write("\n")
    print(f"{len(data['tables'])} tables, {len(data['constants'])} constants -> {OUT}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
