"""Host-side check of the kernels' special-form arithmetic (xpir_amd/csrc/arith.hpp):
the division-free Shoup quotient and the special-form lazy Shoup product against the
textbook 128-bit forms, for every modulus of the reference (crypto/NFLParams.cpp), on
boundary words and random words (tests/cpp/arith_check.cpp, built with hipcc as host
code; nothing runs on a GPU)."""
import os
import subprocess

from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_special_form_arithmetic(tmp_path):
    L = Oracle.lib()
    moduli = [str(L.xo_modulus(i)) for i in range(17)]
    exe = tmp_path / "arith_check"
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-o", str(exe), os.path.join(ROOT, "tests", "cpp", "arith_check.cpp")], check=True,
                   capture_output=True)
    r = subprocess.run([str(exe), "200000", *moduli], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stderr + r.stdout
