"""The EXACT float sin/cos band (2^6 pi, 2^28]: the GPU reduces these
arguments with two double FMAs.  Up to 2^18 pi NT2 runs its double medium
reduction (generic/rem_pio2_medium.hpp:68-83); above, the kernels give the
correctly rounded reduced argument.  tests/cpp/trig_band_check compares the
two-FMA reduction on every float of the band with the oracle's restatement
of NT2's ladder and, above 2^18 pi, with a binary128 reduction; the kernels
themselves are compared with the oracle on the GPU (test_gpu_parity.py)."""
import os
import subprocess

CPP = os.path.join(os.path.dirname(__file__), "cpp")


def test_band_reduction_on_every_float():
    exe = os.path.join(CPP, "trig_band_check")
    subprocess.run(["make", "-s", "-C", CPP, "trig_band_check"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "NT2 part: 100663296 floats" in out.stdout and ": 0 differ\n" in out.stdout, out.stdout
    assert "correctly rounded part: 70709285 floats" in out.stdout, out.stdout
    assert "0 differ, 0 undecided" in out.stdout, out.stdout
