"""The EXACT float sin/cos reduction above 2^6 pi.

Up to 2^18 pi NT2 runs its double medium reduction
(generic/rem_pio2_medium.hpp:68-83) and the GPU two double FMAs:
tests/cpp/trig_band_check compares them on every float of (2^6 pi, 2^18 pi]
(and, above, up to 2^28, the two-FMA reduction with a binary128 one).  Above
2^18 pi NT2 runs fdlibm's float Payne-Hanek
(modules/core/trigonometric/include/nt2/trigonometric/functions/scalar/rem_pio2.hpp:83-233);
tests/cpp/fdlibm_check compares the GPU's register-only restatement
(include/nt2hip/math.hpp reduce_fdlibm_f, compiled for the host) with the
oracle's statement-by-statement one on every finite float above 2^18 pi,
and reduce_fdlibm_d (fdlibm's double Payne-Hanek, which NT2 runs above
2^20 pi/2) on 76 million sampled doubles of every binade and near multiples
of pi/2.
The kernels themselves are compared with the oracle on the GPU
(test_gpu_parity.py)."""
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


def test_fdlibm_reduction_on_every_float():
    exe = os.path.join(CPP, "fdlibm_check")
    subprocess.run(["make", "-s", "-C", CPP, "fdlibm_check"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "909570084 floats" in out.stdout and "(stride 1): 0 differ" in out.stdout, out.stdout
    assert "doubles above 2^20 pi/2 (sampled): 0 differ" in out.stdout, out.stdout
