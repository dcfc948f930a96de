"""The device arithmetic on the CPU: tests/device_arith_check.cpp compiles the kernels' own header
(mass-prime-siever_amd/csrc/mps_device.hpp, host and device functions alike) for the host and
checks it against 128-bit integer arithmetic and Euler's criterion:
  * MontField::mul, r2, from, fpow / fpow2 / fpow_2e, dbl and fermat2 for odd p up to 2^62 - 1
    (Math.hpp:29-79 is the reference's Montgomery arithmetic), operands 0 and p - 1 included, and
    dependent chains of 2000 products;
  * F64Field::mul, mul_small, key and hash_word for p below 2^50, over the signed range (-p, p)
    that its products return, -0.0 and +-(p - 1) included, and fpow2_small;
  * mod16_f / mod_small / mod_small_f for every divisor below 2^16, mod16, divmod52, mulmod_q52 at
    the order-mode moduli and up to 2^52, exact_div_odd, isqrt_ceil64;
  * jacobi32 / jacobi64 / jacobi_small / jacobi_signed against Euler's criterion (prime moduli up to
    2^62) and multiplicativity in the modulus (composite moduli);
  * the BSGS bucket tests (bucket_needs_look, bucket_may_hold, match_mask, fp_pair) against
    slot-by-slot compares on buckets of 0 to 8 fingerprints, raw 0xFFFF lookups included, and the
    no-miss properties.
The GPU runs the same functions: tests/test_gpu_planted.py checks their results end to end."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc is not installed")
    exe = str(tmp_path_factory.mktemp("arith") / "device_arith_check")
    cmd = [HIPCC if os.path.exists(HIPCC) else "hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
           "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "mass-prime-siever_amd", "csrc"), os.path.join(HERE, "device_arith_check.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def test_device_arithmetic_exact(checker):
    r = subprocess.run([checker, "20"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok:" in r.stdout
    assert "every divisor below 2^16" in r.stdout
