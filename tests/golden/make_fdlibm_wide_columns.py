"""Writes tests/golden/fdlibm_wide_columns_f32.npy: floats above 2^18 pi for
which fdlibm's float reduction takes 12 or 13 columns (every float with 13,
every 8th with 12), as uint32 bits.  The GPU parity test runs them through
the EXACT sin/cos kernels, whose column masks for these cases the random and
strided tests almost never reach.  Compiles fdlibm_wide_columns.cpp (g++)
against the oracle's 2/pi table and walks every float above 2^18 pi
(about a minute)."""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "wide")
        subprocess.run(["g++", "-O2", "-std=gnu++17", "-ffp-contract=off", "-o", exe,
                        os.path.join(HERE, "fdlibm_wide_columns.cpp"), os.path.join(ROOT, "oracle", "rem_pio2.cpp")],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = [line.split() for line in out.splitlines() if line.strip()]
    assert all(int(j) in (12, 13) for j, _ in rows), "a float outside the 13-column bound"
    bits = np.array([int(b, 16) for _, b in rows], dtype=np.uint32)
    js = np.array([int(j) for j, _ in rows])
    np.save(os.path.join(HERE, "fdlibm_wide_columns_f32.npy"), bits)
    print(f"{bits.size} floats ({(js == 13).sum()} with 13 columns, {(js == 12).sum()} with 12)")


if __name__ == "__main__":
    main()
