"""The product never reaches the oracle: no import, load or link of oracle/
from nt2_amd/ or include/, and bench.py touches it only in its CPU-baseline leg."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _files(d, exts):
    for base, _, fs in os.walk(os.path.join(ROOT, d)):
        for f in fs:
            if f.endswith(exts):
                yield os.path.join(base, f)


def test_product_sources_do_not_reference_the_oracle():
    pat = re.compile(r"nt2ref|oracle", re.I)
    for f in list(_files("nt2_amd", (".py", ".hip", ".hpp", ".cpp", ".h", "Makefile"))) + list(
            _files("include", (".h", ".hpp"))):
        text = open(f).read()
        assert not pat.search(text), f


def test_bench_uses_oracle_only_for_cpu_baseline():
    text = open(os.path.join(ROOT, "bench.py")).read()
    body = text.split("def _cpu_baseline", 1)
    assert "nt2ref" not in body[0]
    after = body[1].split("\ndef ", 1)[1]
    assert "nt2ref" not in after
